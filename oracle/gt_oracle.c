/*
 * gt_oracle.c -- CPU restatement of gpu_tracking's per-frame detection
 * pipeline (the oracle).  TEST INFRASTRUCTURE ONLY; see gt_oracle.h.
 *
 * Build: gcc -O2 -ffp-contract=off (no implicit contraction; the explicit
 * fmaf calls are the only fused operations, matching the HIP kernels).
 */
#include "gt_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* WGSL `1./255.` as an f32 constant (preprocess_cols.wgsl:54). */
#define OR_TRUNC_THRESHOLD 0.003921568859368562698364257812f /* 0x3B808081 */

static inline int32_t or_f2i(float x) {
    /* f32 -> i32 with NaN -> 0 and saturation. */
    if (x != x) return 0;
    if (x >= 2147483648.0f) return INT32_MAX;
    if (x <= -2147483648.0f) return INT32_MIN;
    return (int32_t)x;
}

static inline float or_max(float a, float b) { return b > a ? b : a; }

void or_derive_trackpy(uint32_t diameter, uint32_t separation, int32_t smoothing_size,
                       float noise_size, float minmass, uint32_t max_iterations,
                       float shift_threshold, float snr, float minmass_snr,
                       int adaptive_background, int truncate, int32_t nrows,
                       int32_t ncols, or_params *out) {
    /* gpu_setup.rs:536-547: smoothing_size defaults to diameter; dilation side
     * is (2*sep/sqrt(2)) as u32; margin = max(d/2, sep/2 - 1, s/2) in u32
     * arithmetic (sep/2 - 1 wraps for sep < 2) cast to i32. */
    uint32_t s = smoothing_size > 0 ? (uint32_t)smoothing_size : diameter;
    float dil = 2.0f * (float)separation / sqrtf(2.0f);
    uint32_t k = dil <= 0.0f ? 0u : (dil >= 4294967296.0f ? 0xFFFFFFFFu : (uint32_t)dil);
    uint32_t m0 = diameter / 2u, m1 = separation / 2u - 1u, m2 = s / 2u;
    uint32_t m = m0;
    if (m1 > m) m = m1;
    if (m2 > m) m = m2;
    memset(out, 0, sizeof(*out));
    out->nrows = nrows;
    out->ncols = ncols;
    out->smoothing = (int32_t)s;
    out->sigma2 = noise_size;
    out->circle = (int32_t)diameter;
    out->dilation = (int32_t)k;
    out->max_iterations = max_iterations;
    out->threshold = shift_threshold;
    out->minmass = minmass;
    out->margin = (int32_t)m;
    out->snr = snr;
    out->minmass_snr = minmass_snr;
    out->rough_snr_factor = adaptive_background ? 0.5f : 1.0f;
    out->image_std = 0.0f;
    out->truncate = truncate;
    out->threshold_rel = 0.0f;
}

void or_gauss_weights(int32_t s, float sigma2, float *w, float *norm) {
    /* preprocess_rows.wgsl:25,30-47: x = f32(i); exp(-x*x / (2*sigma2)). */
    int32_t r = s / 2;
    float n = 0.0f;
    for (int32_t i = -r; i <= r; i++) {
        float x = (float)i;
        float g = expf(-x * x / (2.0f * sigma2));
        w[i + r] = g;
        n = n + g;
    }
    *norm = n;
}

void or_bandpass(const float *frame, const or_params *p, float *processed, float *scratch) {
    const int32_t H = p->nrows, W = p->ncols, s = p->smoothing, r = s / 2;
    float w[1024];
    float norm;
    or_gauss_weights(s, p->sigma2, w, &norm);
    float *tg = scratch, *tc = scratch + (size_t)H * W;
    /* Row pass (preprocess_rows.wgsl:21-54): Gaussian taps skip out-of-bounds
     * columns but the norm keeps every tap; boxcar clamps to the edge. */
    for (int32_t u = 0; u < H; u++) {
        const float *row = frame + (size_t)u * W;
        for (int32_t v = 0; v < W; v++) {
            float gs = 0.0f, cs = 0.0f;
            for (int32_t i = -r; i <= r; i++) {
                int32_t pv = v + i;
                int inb = 1;
                if (pv < 0) { pv = 0; inb = 0; }
                else if (pv >= W) { pv = W - 1; inb = 0; }
                if (inb) gs = fmaf(w[i + r], row[pv], gs);
                cs = cs + row[pv];
            }
            tg[(size_t)u * W + v] = gs / norm;
            tc[(size_t)u * W + v] = cs;
        }
    }
    /* Column pass (preprocess_cols.wgsl:18-59). */
    const float box_area = (float)(s * s);
    for (int32_t u = 0; u < H; u++) {
        for (int32_t v = 0; v < W; v++) {
            float gs = 0.0f, cs = 0.0f;
            for (int32_t i = -r; i <= r; i++) {
                int32_t pu = u + i;
                int inb = 1;
                if (pu < 0) { pu = 0; inb = 0; }
                else if (pu >= H) { pu = H - 1; inb = 0; }
                if (inb) gs = fmaf(w[i + r], tg[(size_t)pu * W + v], gs);
                cs = cs + tc[(size_t)pu * W + v];
            }
            float res = gs / norm - cs / box_area;
            if (p->truncate && res < OR_TRUNC_THRESHOLD) res = 0.0f;
            processed[(size_t)u * W + v] = res;
        }
    }
}

size_t or_local_max(const float *processed, const or_params *p, uint32_t *cand, size_t cap,
                    float *M) {
    const int32_t H = p->nrows, W = p->ncols, k = p->dilation;
    /* window start: i32(f32(k)/2 - 0.5) (max_rows.wgsl:21, extract_max.wgsl:40) */
    const int32_t a = or_f2i((float)k / 2.0f - 0.5f);
    for (int32_t u = 0; u < H; u++) {
        for (int32_t v = 0; v < W; v++) {
            float m = processed[(size_t)u * W + v];
            int32_t pv = v - a;
            for (int32_t j = 0; j < k; j++, pv++)
                if (pv >= 0 && pv < W) m = or_max(m, processed[(size_t)u * W + pv]);
            M[(size_t)u * W + v] = m;
        }
    }
    const float thr = p->image_std * p->snr * p->rough_snr_factor;
    size_t n = 0;
    for (int32_t u = 0; u < H; u++) {
        for (int32_t v = 0; v < W; v++) {
            if (u < p->margin || u >= H - p->margin || v < p->margin || v >= W - p->margin)
                continue;
            const float c = processed[(size_t)u * W + v];
            if (c < thr) continue;
            int ismax = 1;
            int32_t pu = u - a;
            for (int32_t i = 0; i < k; i++, pu++) {
                if (pu >= 0 && pu < H && M[(size_t)pu * W + v] > c) { ismax = 0; break; }
            }
            if (!ismax) continue;
            if (n < cap) cand[n] = (uint32_t)((size_t)u * W + v);
            n++;
        }
    }
    return n;
}

typedef struct { float x, y, mass, max, count; } or_walkres;

/* get_center (walk.wgsl:67-102) */
static or_walkres or_get_center(const float *P, const or_params *p, int32_t u, int32_t v,
                                int32_t rint, float r2) {
    or_walkres res = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    for (int32_t i = -rint; i <= rint; i++) {
        const float x = (float)i, x2 = x * x;
        for (int32_t j = -rint; j <= rint; j++) {
            const float y = (float)j, y2 = y * y;
            if (x2 + y2 > r2) continue;
            const int32_t pu = u + i, pv = v + j;
            if (pu < 0 || pu >= p->nrows || pv < 0 || pv >= p->ncols) continue;
            const float d = P[(size_t)pu * p->ncols + pv];
            res.x = fmaf(d, x, res.x);
            res.y = fmaf(d, y, res.y);
            res.mass = res.mass + d;
            res.count = res.count + 1.0f;
            res.max = or_max(res.max, d);
        }
    }
    res.x = res.x / res.mass;
    res.y = res.y / res.mass;
    return res;
}

size_t or_walk(const float *P, const or_params *p, const uint32_t *cand, size_t ncand,
               or_row *rows) {
    return or_walk_r(P, p, cand, NULL, ncand, rows);
}

size_t or_walk_r(const float *P, const or_params *p, const uint32_t *cand, const float *cand_r,
                 size_t ncand, or_row *rows) {
    /* walk.wgsl:104-169 (radius: r <= 0 -> rint = (circle-1)/2, r2 = rint^2;
     * r > 0 -> rint = ceil(r), r2 = r*r; the move band is rint), then main
     * :171-229 (minmass; row.r = r when r > 0) */
    size_t n = 0;
    for (size_t c = 0; c < ncand; c++) {
        const float r = cand_r ? cand_r[c] : -1.0f;
        int32_t rint;
        float r2;
        if (r <= 0.0f) {
            rint = (p->circle - 1) / 2;
            const float rf = (float)rint;
            r2 = rf * rf;
        } else {
            rint = or_f2i(ceilf(r));
            r2 = r * r;
        }
        const int32_t margin = rint;
        int32_t pu = (int32_t)(cand[c] / (uint32_t)p->ncols);
        int32_t pv = (int32_t)(cand[c] % (uint32_t)p->ncols);
        or_walkres res = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        for (uint32_t it = 0; it < p->max_iterations; it++) {
            res = or_get_center(P, p, pu, pv, rint, r2);
            int changed = 0;
            if (res.x > p->threshold && pu < p->nrows - 1 - margin) { pu += 1; changed = 1; }
            else if (res.x < -p->threshold && pu > margin) { pu -= 1; changed = 1; }
            if (res.y > p->threshold && pv < p->ncols - 1 - margin) { pv += 1; changed = 1; }
            else if (res.y < -p->threshold && pv > margin) { pv -= 1; changed = 1; }
            if (!changed) break;
        }
        res.x = res.x + (float)pu;
        res.y = res.y + (float)pv;
        float mm;
        if (p->minmass > 0.0f) mm = p->minmass;
        else mm = p->minmass_snr * res.count * p->image_std * p->rough_snr_factor;
        if (res.mass >= mm) {
            or_row *o = &rows[n++];
            memset(o, 0, sizeof(*o));
            o->x = res.x;
            o->y = res.y;
            o->mass = res.mass;
            o->count = res.count;
            o->max_intensity = res.max;
            if (r > 0.0f) o->r = r;
            o->seed = cand[c];
        }
    }
    return n;
}

void or_characterize(const float *P, const float *raw, const or_params *p, or_row *rows,
                     size_t n, int use_f64, int recompute_mass) {
    const int32_t H = p->nrows, W = p->ncols;
    for (size_t k = 0; k < n; k++) {
        or_row *row = &rows[k];
        /* characterize.wgsl:62-77 */
        float r = row->r;
        int32_t rint;
        if (r <= 0.0f) { rint = (p->circle - 1) / 2; r = (float)rint; }
        else rint = or_f2i(ceilf(r));
        const float r2 = r * r;
        const int32_t u = or_f2i(nearbyintf(row->x)), v = or_f2i(nearbyintf(row->y));
        if (recompute_mass) {
            /* get_mass (characterize.wgsl:30-58) */
            float m = 0.0f;
            for (int32_t i = -rint; i <= rint; i++) {
                const float x = (float)i, x2 = x * x;
                for (int32_t j = -rint; j <= rint; j++) {
                    const float y = (float)j, y2 = y * y;
                    if (x2 + y2 > r2) continue;
                    const int32_t pu = u + i, pv = v + j;
                    if (pu < 0 || pu >= H || pv < 0 || pv >= W) continue;
                    m = m + P[(size_t)pu * W + pv];
                }
            }
            row->mass = m;
        }
        const float mass = row->mass;
        float rg32 = 0.0f, raw_mass = 0.0f, signal = 0.0f, es32 = 0.0f, ec32 = 0.0f;
        double rg64 = 0.0, es64 = 0.0, ec64 = 0.0;
        for (int32_t i = -rint; i <= rint; i++) {
            const float x = (float)i, x2 = x * x;
            for (int32_t j = -rint; j <= rint; j++) {
                const float y = (float)j, y2 = y * y;
                const float mask = x2 + y2;
                if (mask > r2) continue;
                const int32_t pu = u + i, pv = v + j;
                if (pu < 0 || pu >= H || pv < 0 || pv >= W) continue;
                const size_t idx = (size_t)pu * W + pv;
                const float d = P[idx];
                raw_mass = raw_mass + raw[idx];
                signal = or_max(signal, d);
                if (i != 0 || j != 0) {
                    /* sin(2 atan2(y,x)) = 2xy/(x^2+y^2), cos(2 atan2(y,x)) = (x^2-y^2)/(x^2+y^2) */
                    const double q = (double)(i * i + j * j);
                    const double s2 = 2.0 * (double)i * (double)j / q;
                    const double c2 = ((double)(i * i) - (double)(j * j)) / q;
                    if (use_f64) {
                        es64 = fma((double)d, s2, es64);
                        ec64 = fma((double)d, c2, ec64);
                    } else {
                        es32 = fmaf(d, (float)s2, es32);
                        ec32 = fmaf(d, (float)c2, ec32);
                    }
                } else {
                    if (use_f64) ec64 += (double)d;
                    else ec32 = ec32 + d;
                }
                if (use_f64) rg64 = fma((double)d, (double)mask, rg64);
                else rg32 = fmaf(d, mask, rg32);
            }
        }
        /* the centre sample is read without a bounds check in the reference
         * (characterize.wgsl:79,131); clamp so the read stays inside the frame */
        int32_t cu = u < 0 ? 0 : (u >= H ? H - 1 : u);
        int32_t cv = v < 0 ? 0 : (v >= W ? W - 1 : v);
        const float den = mass - P[(size_t)cu * W + cv];
        if (use_f64) {
            row->ecc = (float)(sqrt(es64 * es64 + ec64 * ec64) / (double)den);
            row->Rg = (float)sqrt(rg64 / (double)mass);
        } else {
            es32 = es32 * es32;
            ec32 = ec32 * ec32;
            row->ecc = sqrtf(es32 + ec32) / den;
            row->Rg = sqrtf(rg32 / mass);
        }
        row->raw_mass = raw_mass;
        row->signal = signal;
    }
}

void or_filter_close(const or_row *rows, size_t n, float separation, uint8_t *keep) {
    const float r2 = separation * separation;
    for (size_t i = 0; i < n; i++) {
        keep[i] = 1;
        const float mi = rows[i].mass;
        for (size_t j = 0; j < n; j++) {
            if (j == i) continue;
            const float dx = rows[j].x - rows[i].x, dy = rows[j].y - rows[i].y;
            float d2 = 0.0f;
            d2 = d2 + dx * dx;
            d2 = d2 + dy * dy;
            if (!(d2 < r2)) continue;
            const float mj = rows[j].mass;
            if (mi < mj || (mi == mj && i > j)) { keep[i] = 0; break; }
        }
    }
}

static int or_cmp_float(const void *a, const void *b) {
    float x = *(const float *)a, y = *(const float *)b;
    return (x > y) - (x < y);
}

static inline int32_t or_round_away(float x) { return or_f2i(roundf(x)); }

void or_doughnut(const float *raw, int32_t H, int32_t W, const or_row *rows, size_t n,
                 float radius, float bg_radius, float gap_radius, float *out3) {
    /* circle_inds(radius) and annulus_inds(bg_radius, gap_radius + radius)
     * (kernels.rs:138-171); the row's centre is rounded half away from zero
     * (execute_gpu.rs:752); out-of-frame samples are skipped (:754-774). */
    const int32_t cb = or_f2i(ceilf(radius));
    const float rr = radius * radius;
    const float inner = gap_radius + radius;
    const int32_t ab = or_f2i(ceilf(bg_radius));
    const float or2 = bg_radius * bg_radius, ir2 = inner * inner;
    size_t bgcap = (size_t)(2 * ab + 1) * (size_t)(2 * ab + 1);
    float *bg = (float *)malloc(sizeof(float) * (bgcap ? bgcap : 1));
    for (size_t k = 0; k < n; k++) {
        const int32_t mu = or_round_away(rows[k].x), mv = or_round_away(rows[k].y);
        float sig = 0.0f;
        size_t nsig = 0, nbg = 0;
        for (int32_t i = -cb; i <= cb; i++)
            for (int32_t j = -cb; j <= cb; j++) {
                if (!((float)(i * i + j * j) <= rr)) continue;
                int64_t pu = (int64_t)mu + i, pv = (int64_t)mv + j;
                if (pu < 0 || pu >= H || pv < 0 || pv >= W) continue;
                sig = sig + raw[pu * W + pv];
                nsig++;
            }
        for (int32_t i = -ab; i <= ab; i++)
            for (int32_t j = -ab; j <= ab; j++) {
                const float q = (float)(i * i + j * j);
                if (!(q <= or2 && q >= ir2)) continue;
                int64_t pu = (int64_t)mu + i, pv = (int64_t)mv + j;
                if (pu < 0 || pu >= H || pv < 0 || pv >= W) continue;
                bg[nbg++] = raw[pu * W + pv];
            }
        qsort(bg, nbg, sizeof(float), or_cmp_float);
        const float med = nbg ? bg[nbg / 2] : 0.0f;
        out3[3 * k + 0] = sig;
        out3[3 * k + 1] = med;
        out3[3 * k + 2] = sig - med * (float)nsig;
    }
    free(bg);
}

static int or_cmp_row(const void *a, const void *b) {
    uint32_t x = ((const or_row *)a)->seed, y = ((const or_row *)b)->seed;
    return (x > y) - (x < y);
}

void or_sort_rows(or_row *rows, size_t n) { qsort(rows, n, sizeof(or_row), or_cmp_row); }

size_t or_detect_frame(const float *frame, const or_params *p, int characterize, or_row *rows,
                       size_t cap, size_t *ncand_out) {
    const size_t hw = (size_t)p->nrows * p->ncols;
    float *proc = (float *)malloc(sizeof(float) * hw * 3);
    uint32_t *cand = (uint32_t *)malloc(sizeof(uint32_t) * (hw ? hw : 1));
    or_bandpass(frame, p, proc, proc + hw);
    size_t nc = or_local_max(proc, p, cand, hw, proc + hw);
    if (ncand_out) *ncand_out = nc;
    or_row *tmp = (or_row *)malloc(sizeof(or_row) * (nc ? nc : 1));
    size_t nr = or_walk(proc, p, cand, nc, tmp);
    if (characterize) or_characterize(proc, frame, p, tmp, nr, 1, 0);
    if (nr <= cap) memcpy(rows, tmp, sizeof(or_row) * nr);
    else nr = (size_t)-1;
    free(tmp);
    free(cand);
    free(proc);
    return nr;
}

typedef struct {
    const void *frames;
    int dtype;
    size_t nframes;
    const or_params *p;
    int characterize;
    int nthreads, tid;
    or_row *rows;
    size_t cap;
    uint64_t *counts, *ncands;
    int status;
} or_job;

static void *or_worker(void *arg) {
    or_job *j = (or_job *)arg;
    const size_t hw = (size_t)j->p->nrows * j->p->ncols;
    float *f32 = (float *)malloc(sizeof(float) * (hw ? hw : 1));
    j->status = 0;
    for (size_t f = (size_t)j->tid; f < j->nframes; f += (size_t)j->nthreads) {
        const float *fr;
        if (j->dtype == 0) {
            const uint16_t *src = (const uint16_t *)j->frames + f * hw;
            for (size_t i = 0; i < hw; i++) f32[i] = (float)src[i];
            fr = f32;
        } else {
            fr = (const float *)j->frames + f * hw;
        }
        size_t nc = 0;
        size_t nr = or_detect_frame(fr, j->p, j->characterize, j->rows + f * j->cap, j->cap, &nc);
        if (nr == (size_t)-1) { j->status = -1; nr = 0; }
        j->counts[f] = nr;
        if (j->ncands) j->ncands[f] = nc;
    }
    free(f32);
    return NULL;
}

int or_detect_batch(const void *frames, int dtype, size_t nframes, const or_params *p,
                    int characterize, int nthreads, or_row *rows, size_t cap_per_frame,
                    uint64_t *counts, uint64_t *ncands) {
    if (nthreads < 1) nthreads = 1;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * nthreads);
    or_job *jobs = (or_job *)malloc(sizeof(or_job) * nthreads);
    for (int t = 0; t < nthreads; t++) {
        or_job j = {frames, dtype, nframes, p, characterize, nthreads, t, rows, cap_per_frame,
                    counts, ncands, 0};
        jobs[t] = j;
        pthread_create(&th[t], NULL, or_worker, &jobs[t]);
    }
    int st = 0;
    for (int t = 0; t < nthreads; t++) {
        pthread_join(th[t], NULL);
        if (jobs[t].status) st = -1;
    }
    free(jobs);
    free(th);
    return st;
}

void or_bandpass_f64(const float *frame, const or_params *p, double *processed, double *scratch) {
    const int32_t H = p->nrows, W = p->ncols, s = p->smoothing, r = s / 2;
    double w[1024], norm = 0.0;
    for (int32_t i = -r; i <= r; i++) {
        w[i + r] = exp(-(double)i * (double)i / (2.0 * (double)p->sigma2));
        norm += w[i + r];
    }
    double *tg = scratch, *tc = scratch + (size_t)H * W;
    for (int32_t u = 0; u < H; u++)
        for (int32_t v = 0; v < W; v++) {
            double gs = 0.0, cs = 0.0;
            for (int32_t i = -r; i <= r; i++) {
                int32_t pv = v + i, inb = 1;
                if (pv < 0) { pv = 0; inb = 0; }
                else if (pv >= W) { pv = W - 1; inb = 0; }
                if (inb) gs += w[i + r] * (double)frame[(size_t)u * W + pv];
                cs += (double)frame[(size_t)u * W + pv];
            }
            tg[(size_t)u * W + v] = gs / norm;
            tc[(size_t)u * W + v] = cs;
        }
    for (int32_t u = 0; u < H; u++)
        for (int32_t v = 0; v < W; v++) {
            double gs = 0.0, cs = 0.0;
            for (int32_t i = -r; i <= r; i++) {
                int32_t pu = u + i, inb = 1;
                if (pu < 0) { pu = 0; inb = 0; }
                else if (pu >= H) { pu = H - 1; inb = 0; }
                if (inb) gs += w[i + r] * tg[(size_t)pu * W + v];
                cs += tc[(size_t)pu * W + v];
            }
            double res = gs / norm - cs / (double)(s * s);
            if (p->truncate && res < 1.0 / 255.0) res = 0.0;
            processed[(size_t)u * W + v] = res;
        }
}

void or_centroid_f64(const double *P, const or_params *p, int32_t pu0, int32_t pv0, double *dx,
                     double *dy, double *mass) {
    const int32_t rint = (p->circle - 1) / 2;
    double sx = 0.0, sy = 0.0, m = 0.0;
    for (int32_t i = -rint; i <= rint; i++)
        for (int32_t j = -rint; j <= rint; j++) {
            if (i * i + j * j > rint * rint) continue;
            int32_t pu = pu0 + i, pv = pv0 + j;
            if (pu < 0 || pu >= p->nrows || pv < 0 || pv >= p->ncols) continue;
            double d = P[(size_t)pu * p->ncols + pv];
            sx += d * i;
            sy += d * j;
            m += d;
        }
    *dx = sx / m;
    *dy = sy / m;
    *mass = m;
}

/* ------------------------------------------------------------------------- */
/* LoG style                                                                 */

void or_log_radii(float min_r, float max_r, uint32_t n, int log_spacing, float *out) {
    /* gpu_setup.rs:958-984: the start value advances by repeated f32
     * addition; with log spacing each radius is 10^start. */
    if (n == 0) return;
    if (log_spacing) {
        float start = log10f(min_r);
        const float end = log10f(max_r);
        const float diff = (end - start) / (float)(n - 1);
        for (uint32_t k = 0; k + 1 < n; k++) {
            const float t = start;
            start = start + diff;
            out[k] = powf(10.0f, t);
        }
        out[n - 1] = powf(10.0f, start);
    } else {
        float start = min_r;
        const float diff = (max_r - min_r) / (float)(n - 1);
        for (uint32_t k = 0; k + 1 < n; k++) {
            const float t = start;
            start = start + diff;
            out[k] = t;
        }
        out[n - 1] = start;
    }
}

int32_t or_log_taps(float sigma, float *g, float *d, float *norm1, float *norm3) {
    /* separable_log.wgsl:22-89: rint = i32(4 sigma + 0.5); tap weights
     * exp(-x^2 / (2 sigma^2)) and exp(-x^2 / (2 sigma^2)) * (x^2 - sigma^2);
     * norms 1/(sqrt(2 pi) sigma) and 1/(sqrt(2 pi) sigma sigma^2). */
    const float sigma2 = sigma * sigma;
    const int32_t R = or_f2i(4.0f * sigma + 0.5f);
    for (int32_t i = -R; i <= R; i++) {
        const float x = (float)i, x2 = x * x;
        const float e = expf(-x2 / (2.0f * sigma2));
        if (g) g[i + R] = e;
        if (d) d[i + R] = e * (x2 - sigma2);
    }
    const float sqrt2pi = 2.5066282746310002f;
    if (norm1) *norm1 = 1.0f / (sqrt2pi * sigma);
    if (norm3) *norm3 = 1.0f / (sqrt2pi * sigma * sigma2);
    return R;
}

/* One 1-D pass of separable_log.wgsl: along rows (dim 0) or columns (dim 1),
 * taps w[0..2R], out-of-frame taps skipped, result = sum * norm. */
static void or_log_pass(const float *in, int32_t H, int32_t W, int dim, const float *w, int32_t R,
                        float norm, float *out) {
    for (int32_t u = 0; u < H; u++)
        for (int32_t v = 0; v < W; v++) {
            float s = 0.0f;
            for (int32_t i = -R; i <= R; i++) {
                const int32_t pu = dim == 0 ? u + i : u, pv = dim == 0 ? v : v + i;
                if (pu < 0 || pu >= H || pv < 0 || pv >= W) continue;
                s = fmaf(w[i + R], in[(size_t)pu * W + pv], s);
            }
            out[(size_t)u * W + v] = s * norm;
        }
}

void or_log_plane(const float *raw, int32_t H, int32_t W, float sigma, float *L, float *tmp,
                  float *tmp2) {
    /* Laplace<2> over SeparableConvolution<2> (utils.rs:114-296) with
     * separable_log.wgsl: for diff_dim 0 then 1, a dim-0 pass of the raw
     * frame into tmp and a dim-1 pass of tmp subtracted from the output
     * (zeroed by the first pass).  The second-derivative taps go on the
     * pass whose dim equals diff_dim. */
    int32_t R = or_log_taps(sigma, NULL, NULL, NULL, NULL);
    float *g = (float *)malloc(sizeof(float) * (2 * R + 1));
    float *d = (float *)malloc(sizeof(float) * (2 * R + 1));
    float n1, n3;
    or_log_taps(sigma, g, d, &n1, &n3);
    const size_t hw = (size_t)H * W;
    /* diff_dim 0 */
    or_log_pass(raw, H, W, 0, d, R, n3, tmp);
    or_log_pass(tmp, H, W, 1, g, R, n1, tmp2);
    for (size_t k = 0; k < hw; k++) L[k] = 0.0f - tmp2[k];
    /* diff_dim 1 */
    or_log_pass(raw, H, W, 0, g, R, n1, tmp);
    or_log_pass(tmp, H, W, 1, d, R, n3, tmp2);
    for (size_t k = 0; k < hw; k++) L[k] = L[k] - tmp2[k];
    free(g);
    free(d);
}

/* is_max_in_plane (logspace_max.wgsl:62-92) */
static int or_is_max_in_plane(const float *P, int32_t H, int32_t W, int32_t i, int32_t j,
                              float center) {
    for (int32_t a = -1; a < 2; a++)
        for (int32_t b = -1; b < 2; b++) {
            if (a == 0 && b == 0) continue;
            const int32_t x = i + a, y = j + b;
            if (x < 0 || x >= H || y < 0 || y >= W) return 0;
            if (P[(size_t)x * W + y] >= center) return 0;
        }
    return 1;
}

size_t or_log_max(const float *bottom, const float *middle, const float *top,
                  const float *processed, int32_t H, int32_t W, float snr_threshold,
                  uint32_t *cand, size_t cap) {
    /* logspace_max.wgsl:94-154: processed >= std*snr*rough, centre > 0 and
     * strictly above its 8 neighbours in its own plane and in the planes
     * below / above (NULL = no such plane); every neighbour in the frame. */
    size_t n = 0;
    for (int32_t i = 0; i < H; i++)
        for (int32_t j = 0; j < W; j++) {
            const size_t idx = (size_t)i * W + j;
            if (processed[idx] < snr_threshold) continue;
            const float c = middle[idx];
            if (c <= 0.0f) continue;
            if (!or_is_max_in_plane(middle, H, W, i, j, c)) continue;
            if (bottom && !or_is_max_in_plane(bottom, H, W, i, j, c)) continue;
            if (top && !or_is_max_in_plane(top, H, W, i, j, c)) continue;
            if (n < cap) cand[n] = (uint32_t)idx;
            n++;
        }
    return n;
}

/* disk_overlap / blob_overlap (execute_gpu.rs:597-630), f32 as in Rust */
static float or_make_acos(float d, float r1, float r1s, float r2s) {
    float ratio = (d * d + r1s - r2s) / (2.0f * d * r1);
    ratio = ratio < -1.0f ? -1.0f : (ratio > 1.0f ? 1.0f : ratio);
    return acosf(ratio);
}

float or_blob_overlap(float r1, float r2, float distance) {
    if (distance > r1 + r2) return 0.0f;
    if (distance <= fabsf(r1 - r2)) return 1.0f;
    const float d = distance;
    const float r1s = r1 * r1, r2s = r2 * r2;
    const float acos1 = or_make_acos(d, r1, r1s, r2s);
    const float acos2 = or_make_acos(d, r2, r2s, r1s);
    const float a = -d + r2 + r1;
    const float b = d - r2 + r1;
    const float c = d + r2 - r1;
    const float e = d + r2 + r1;
    const float area = r1s * acos1 + r2s * acos2 - 0.5f * sqrtf(fabsf(a * b * c * e));
    const float pi = 3.14159265358979323846f;
    return area / (pi * (r1s < r2s ? r1s : r2s));
}

size_t or_prune_blobs(or_row *rows, size_t n, float overlap_threshold) {
    /* post_process (execute_gpu.rs:699-712): when overlap_threshold < 1,
     * a stable sort by r descending, mass descending, then prune_blobs_log
     * (:592-669): each surviving blob in that order drops every surviving
     * other blob within distance^2 < (2 r)^2 whose overlap exceeds the
     * threshold.  Survivors are compacted in sorted order; returns their
     * count.  O(n^2). */
    if (!(overlap_threshold < 1.0f)) return n;
    /* stable insertion-merge sort on (r desc, mass desc) */
    or_row *tmp = (or_row *)malloc(sizeof(or_row) * (n ? n : 1));
    for (size_t w = 1; w < n; w *= 2) {
        for (size_t lo = 0; lo < n; lo += 2 * w) {
            size_t mid = lo + w < n ? lo + w : n, hi = lo + 2 * w < n ? lo + 2 * w : n;
            size_t a = lo, b = mid, k = lo;
            while (a < mid && b < hi) {
                const or_row *x = &rows[a], *y = &rows[b];
                /* take y first only if it sorts strictly before x */
                int ybefore = (y->r > x->r) || (y->r == x->r && y->mass > x->mass);
                tmp[k++] = ybefore ? rows[b++] : rows[a++];
            }
            while (a < mid) tmp[k++] = rows[a++];
            while (b < hi) tmp[k++] = rows[b++];
        }
        memcpy(rows, tmp, sizeof(or_row) * n);
    }
    free(tmp);
    uint8_t *keep = (uint8_t *)malloc(n ? n : 1);
    memset(keep, 1, n);
    for (size_t q = 0; q < n; q++) {
        if (!keep[q]) continue;
        const float rad = rows[q].r * 2.0f, r2 = rad * rad;
        for (size_t k = 0; k < n; k++) {
            if (k == q || !keep[k]) continue;
            const float dx = rows[k].x - rows[q].x, dy = rows[k].y - rows[q].y;
            const float d2 = 0.0f + dx * dx + dy * dy;
            if (!(d2 < r2)) continue;
            if (or_blob_overlap(rows[q].r, rows[k].r, sqrtf(d2)) > overlap_threshold) keep[k] = 0;
        }
    }
    size_t m = 0;
    for (size_t k = 0; k < n; k++)
        if (keep[k]) rows[m++] = rows[k];
    free(keep);
    return m;
}

static size_t or_adaptive_passes(const float *frame, const or_params *pp, float sd,
                                 uint32_t n_iter, or_row *pend, size_t np, or_row *acc);

/* The walked LoG rows of every scale (uncharacterized unless asked), in
 * (scale, raster) order with seed = scale * H * W + raster; proc receives the
 * bandpassed frame (3 * H * W floats of room).  Returns the row array
 * (caller frees) and its length in *total. */
static or_row *or_log_rows(const float *frame, const or_params *p, const float *radii,
                           uint32_t n_radii, int characterize, float *proc, size_t *total_out,
                           size_t *ncand_out) {
    const int32_t H = p->nrows, W = p->ncols;
    const size_t hw = (size_t)H * W;
    or_bandpass(frame, p, proc, proc + hw);
    float *planes = (float *)malloc(sizeof(float) * hw * 3);
    float *t1 = (float *)malloc(sizeof(float) * hw), *t2 = (float *)malloc(sizeof(float) * hw);
    uint32_t *cand = (uint32_t *)malloc(sizeof(uint32_t) * (hw ? hw : 1));
    float *cr = (float *)malloc(sizeof(float) * (hw ? hw : 1));
    size_t total = 0, ncand = 0;
    or_row *all = NULL;
    const float sq2 = sqrtf(2.0f);
    /* threshold_rel (no reference counterpart; the reference allocates a
     * global_max buffer it never writes, gpu_setup.rs:481-488): a first pass
     * takes the largest LoG value of the frame over every scale, then a
     * candidate stays when its value exceeds threshold_rel times it. */
    float rel_thr = 0.0f;
    const int use_rel = p->threshold_rel > 0.0f;
    if (use_rel) {
        float gmax = 0.0f;
        for (uint32_t m = 0; m < n_radii; m++) {
            or_log_plane(frame, H, W, radii[m] / sq2, planes, t1, t2);
            for (size_t i = 0; i < hw; i++)
                if (planes[i] > gmax) gmax = planes[i];
        }
        rel_thr = p->threshold_rel * gmax;
    }
    for (uint32_t m = 0; m < n_radii; m++) {
        /* planes needed: m-1, m, m+1 (ring of 3 by scale index) */
        if (m == 0) or_log_plane(frame, H, W, radii[0] / sq2, planes, t1, t2);
        if (m + 1 < n_radii)
            or_log_plane(frame, H, W, radii[m + 1] / sq2, planes + ((m + 1) % 3) * hw, t1, t2);
        const float *mid = planes + (m % 3) * hw;
        const float *bot = m > 0 ? planes + ((m + 2) % 3) * hw : NULL;
        const float *top = m + 1 < n_radii ? planes + ((m + 1) % 3) * hw : NULL;
        const float thr = p->image_std * p->snr * p->rough_snr_factor;
        size_t nc = or_log_max(bot, mid, top, proc, H, W, thr, cand, hw);
        if (use_rel) {
            size_t kept = 0;
            for (size_t c = 0; c < nc; c++)
                if (mid[cand[c]] > rel_thr) cand[kept++] = cand[c];
            nc = kept;
        }
        ncand += nc;
        for (size_t c = 0; c < nc; c++) cr[c] = radii[m];
        all = (or_row *)realloc(all, sizeof(or_row) * (total + nc + 1));
        size_t nr = or_walk_r(proc, p, cand, cr, nc, all + total);
        if (characterize) or_characterize(proc, frame, p, all + total, nr, 1, 0);
        for (size_t k = 0; k < nr; k++) all[total + k].seed += (uint32_t)(m * hw);
        total += nr;
    }
    if (ncand_out) *ncand_out = ncand;
    free(cr);
    free(cand);
    free(t1);
    free(t2);
    free(planes);
    *total_out = total;
    return all;
}

size_t or_log_detect_frame(const float *frame, const or_params *p, const float *radii,
                           uint32_t n_radii, int characterize, float overlap_threshold,
                           or_row *rows, size_t cap, size_t *ncand_out) {
    /* execute_gpu.rs:237-286 (bandpass, then per scale the LoG plane, the
     * scale-space maximum of the previous scale and its walk) and the LoG
     * branch of post_process. */
    const size_t hw = (size_t)p->nrows * p->ncols;
    float *proc = (float *)malloc(sizeof(float) * hw * 3);
    size_t total = 0;
    or_row *all = or_log_rows(frame, p, radii, n_radii, characterize, proc, &total, ncand_out);
    size_t n = or_prune_blobs(all, total, overlap_threshold);
    if (n <= cap) memcpy(rows, all, sizeof(or_row) * n);
    else n = (size_t)-1;
    free(all);
    free(proc);
    return n;
}

size_t or_log_detect_frame_adaptive(const float *frame, const or_params *p_in, const float *radii,
                                    uint32_t n_radii, uint32_t n_iter, int characterize,
                                    float overlap_threshold, or_row *rows, size_t cap) {
    or_params p = *p_in;
    const size_t hw = (size_t)p.nrows * p.ncols;
    float mean, sd;
    or_frame_std(frame, hw, &mean, &sd);
    p.image_std = sd;
    float *proc = (float *)malloc(sizeof(float) * hw * 3);
    size_t total = 0;
    or_row *all = or_log_rows(frame, &p, radii, n_radii, 0, proc, &total, NULL);
    or_row *acc = (or_row *)malloc(sizeof(or_row) * (total ? total : 1));
    const size_t na = or_adaptive_passes(frame, &p, sd, n_iter, all, total, acc);
    if (characterize) or_characterize(proc, frame, &p, acc, na, 1, 0);
    size_t n = or_prune_blobs(acc, na, overlap_threshold);
    if (n <= cap) memcpy(rows, acc, sizeof(or_row) * n);
    else n = (size_t)-1;
    free(acc);
    free(all);
    free(proc);
    return n;
}

static int or_is_marker(float v) {
    uint32_t u;
    memcpy(&u, &v, sizeof(u));
    return u == 0xFFFFFFFFu;
}

/* one level of the eight-way tree, in place over t[0..len) */
static size_t or_tree_level(float *t, size_t len) {
    const size_t s = (len + 7) / 8;
    for (size_t i = 0; i < s; i++) {
        float sum = 0.0f;
        for (size_t j = 0; j < 8; j++) {
            const size_t k = i + s * j;
            if (k < len) sum = sum + t[k];
        }
        t[i] = sum;
    }
    return s;
}

void or_frame_std(const float *x, size_t n, float *mean_out, float *std_out) {
    or_frame_std_div(x, n, (uint32_t)n, (uint32_t)(n - 1), mean_out, std_out);
}

void or_frame_std_div(const float *x, size_t n, uint32_t div_mean, uint32_t div_var,
                      float *mean_out, float *std_out) {
    const size_t s0 = (n + 7) / 8;
    float *t = (float *)malloc(sizeof(float) * (s0 > 0 ? s0 : 1));
    /* mean: sum_image_outplace (markers skipped), then the in-place levels */
    for (size_t i = 0; i < s0; i++) {
        float sum = 0.0f;
        for (size_t j = 0; j < 8; j++) {
            const size_t k = i + s0 * j;
            if (k < n && !or_is_marker(x[k])) sum = sum + x[k];
        }
        t[i] = sum;
    }
    size_t len = s0;
    while (len > 24) len = or_tree_level(t, len);
    float m = 0.0f;
    for (size_t i = 0; i < len; i++) m = m + t[i];
    m = m / (float)div_mean;
    /* variance: std_image (squared deviations, 0 at markers) fused into the
     * first in-place level over n values */
    for (size_t i = 0; i < s0; i++) {
        float sum = 0.0f;
        for (size_t j = 0; j < 8; j++) {
            const size_t k = i + s0 * j;
            if (k < n) {
                float v = 0.0f;
                if (!or_is_marker(x[k])) {
                    const float d = x[k] - m;
                    v = d * d;
                }
                sum = sum + v;
            }
        }
        t[i] = sum;
    }
    len = s0;
    while (len > 24) len = or_tree_level(t, len);
    float v = 0.0f;
    for (size_t i = 0; i < len; i++) v = v + t[i];
    v = v / (float)div_var;
    *mean_out = m;
    *std_out = sqrtf(v);
    free(t);
}

/* adaptive_filter.wgsl:56-87 set_nan: the walk disk (radius (circle-1)/2,
 * r <= 0 rows) around the row's centre (WGSL round: half to even) marked
 * with the NaN marker */
static void or_mask_row(float *fr, const or_params *p, const or_row *row) {
    int32_t rint;
    if (row->r > 0.0f) rint = or_f2i(ceilf(row->r));
    else rint = (p->circle - 1) / 2;
    const int32_t u = or_f2i(rintf(row->x)), v = or_f2i(rintf(row->y));
    const float r = (float)rint, r2 = r * r;
    uint32_t marker = 0xFFFFFFFFu;
    float mk;
    memcpy(&mk, &marker, sizeof(mk));
    for (int32_t i = -rint; i <= rint; i++) {
        const float x = (float)i, x2 = x * x;
        for (int32_t j = -rint; j <= rint; j++) {
            const float y = (float)j, y2 = y * y;
            if (x2 + y2 > r2) continue;
            const int32_t pu = u + i, pv = v + j;
            if (pu < 0 || pu >= p->nrows || pv < 0 || pv >= p->ncols) continue;
            fr[(size_t)pu * p->ncols + pv] = mk;
        }
    }
}

/* The adaptive passes (execute_gpu.rs:288-330, adaptive_filter.wgsl:89-125)
 * over walked rows pend[0..np) (consumed): accepted rows go to acc in
 * acceptance order; returns their number. */
static size_t or_adaptive_passes(const float *frame, const or_params *pp, float sd,
                                 uint32_t n_iter, or_row *pend, size_t np, or_row *acc) {
    const or_params p = *pp;
    const size_t hw = (size_t)p.nrows * p.ncols;
    or_row *next = (or_row *)malloc(sizeof(or_row) * (np ? np : 1));
    size_t na = 0;
    float mean;
    float *masked = (float *)malloc(sizeof(float) * (hw ? hw : 1));
    memcpy(masked, frame, sizeof(float) * hw);
    uint32_t counter = (uint32_t)hw;
    float cur = sd;
    for (uint32_t it = 0; it < n_iter; it++) {
        if (it > 0) or_frame_std_div(masked, hw, counter, counter - 1u, &mean, &cur);
        size_t nn = 0;
        for (size_t k = 0; k < np; k++) {
            const or_row *row = &pend[k];
            float mm;
            if (p.minmass > 0.0f) mm = p.minmass;
            else mm = p.minmass_snr * row->count * cur;
            if (row->mass < mm || row->max_intensity < p.snr * cur) {
                next[nn++] = *row;
            } else {
                acc[na++] = *row;
                or_mask_row(masked, &p, row);
                counter -= (uint32_t)row->count;
            }
        }
        memcpy(pend, next, sizeof(or_row) * nn);
        np = nn;
    }
    free(masked);
    free(next);
    return na;
}

size_t or_detect_frame_adaptive(const float *frame, const or_params *p_in, uint32_t n_iter,
                                int characterize, or_row *rows, size_t cap) {
    or_params p = *p_in;
    const size_t hw = (size_t)p.nrows * p.ncols;
    float mean, sd;
    or_frame_std(frame, hw, &mean, &sd);
    p.image_std = sd;
    float *proc = (float *)malloc(sizeof(float) * hw * 3);
    uint32_t *cand = (uint32_t *)malloc(sizeof(uint32_t) * (hw ? hw : 1));
    or_bandpass(frame, &p, proc, proc + hw);
    const size_t nc = or_local_max(proc, &p, cand, hw, proc + hw);
    or_row *pend = (or_row *)malloc(sizeof(or_row) * (nc ? nc : 1));
    or_row *acc = (or_row *)malloc(sizeof(or_row) * (nc ? nc : 1));
    const size_t np = or_walk(proc, &p, cand, nc, pend);
    const size_t na = or_adaptive_passes(frame, &p, sd, n_iter, pend, np, acc);
    if (characterize) or_characterize(proc, frame, &p, acc, na, 1, 0);
    size_t out = na;
    if (na <= cap) memcpy(rows, acc, sizeof(or_row) * na);
    else out = (size_t)-1;
    free(acc);
    free(pend);
    free(cand);
    free(proc);
    return out;
}

/* ---- illumination correction -------------------------------------------- */

int32_t or_illum_taps(float sigma, float *w) {
    const int32_t rint = (int32_t)(4.0f * sigma + 0.5f);
    const float sigma2 = sigma * sigma;
    for (int32_t i = -rint; i <= rint; i++) {
        const float x = (float)i;
        const float x2 = x * x;
        w[i + rint] = expf(-x2 / (2.0f * sigma2));
    }
    return rint;
}

/* one pass of smooth_mean_frame.wgsl:38-72; dim 0 runs along a column (u is
 * the row), dim 1 along a row (u is the column) */
static void or_illum_pass(const float *in, float *out, int32_t H, int32_t W, const float *w,
                          int32_t rint, int dim) {
    for (int32_t r = 0; r < H; r++) {
        for (int32_t c = 0; c < W; c++) {
            const int32_t u = dim == 0 ? r : c;
            const int32_t bound = dim == 0 ? H : W;
            float gs = 0.0f, ns = 0.0f;
            for (int32_t i = -rint; i <= rint; i++) {
                const int32_t pu = u + i;
                if (pu < 0 || pu >= bound) continue;
                const float x = dim == 0 ? in[(size_t)pu * W + c] : in[(size_t)r * W + pu];
                gs = fmaf(w[i + rint], x, gs);
                ns = ns + w[i + rint];
            }
            float norm = ns;
            if (dim == 0) norm = norm * 1.0f;     /* n_frames = 1 */
            out[(size_t)r * W + c] = gs / norm;
        }
    }
}

void or_illum_smooth(const float *in, int32_t H, int32_t W, float sigma, float *out, float *tmp) {
    const int32_t rmax = (int32_t)(4.0f * sigma + 0.5f);
    float *w = (float *)malloc(sizeof(float) * (size_t)(2 * rmax + 1));
    const int32_t rint = or_illum_taps(sigma, w);
    or_illum_pass(in, tmp, H, W, w, rint, 0);
    or_illum_pass(tmp, out, H, W, w, rint, 1);
    free(w);
}

void or_frames_mean(const float *frames, size_t T, size_t n, float *out) {
    for (size_t i = 0; i < n; i++) out[i] = 0.0f;
    for (size_t t = 0; t < T; t++)
        for (size_t i = 0; i < n; i++) out[i] = out[i] + frames[t * n + i];
    const float nf = (float)T;
    for (size_t i = 0; i < n; i++) out[i] = out[i] / nf;
}

void or_illum_divide(const float *frame, const float *profile, size_t n, float *out) {
    for (size_t i = 0; i < n; i++) out[i] = frame[i] / profile[i];
}
