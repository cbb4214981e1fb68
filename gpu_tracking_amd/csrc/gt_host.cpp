// gt_host.cpp -- host logic of the runtime (no device calls).
#include "gt_host.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <unordered_map>

namespace {

int32_t f2i_sat(float x) {
    // f32 -> i32 as Rust's `as i32`: NaN -> 0, saturating.
    if (x != x) return 0;
    if (x >= 2147483648.0f) return 2147483647;
    if (x <= -2147483648.0f) return -2147483647 - 1;
    return static_cast<int32_t>(x);
}

uint32_t f2u_sat(float x) {
    // f32 -> u32 as Rust's `as u32`.
    if (!(x > 0.0f)) return 0u;
    if (x >= 4294967296.0f) return 0xFFFFFFFFu;
    return static_cast<uint32_t>(x);
}

}  // namespace

void gt_host_params_default(gt_params *p) {
    // TrackingParams::default() (src/gpu_setup.rs:61-97)
    std::memset(p, 0, sizeof(*p));
    p->style = GT_STYLE_TRACKPY;
    p->diameter = 9;
    p->separation = 10;
    p->threshold = 0.0f;
    p->invert = 0;
    p->percentile = 0.0f;
    p->topn = 0;
    p->maxsize = 0.0f;
    p->preprocess = 1;
    p->filter_close = 1;
    p->n_radii = 10;
    p->characterize = 0;
    p->doughnut_correction = 0;
    p->max_iterations = 10;
    p->minmass = 0.0f;
    p->truncate_preprocessed = 1;
    p->include_r_in_output = 0;
    p->shift_threshold = 0.6f;
    p->noise_size = 1.0f;
    p->illumination_correction_per_frame = 0;
}

bool gt_host_derive(const gt_params *p, uint32_t H, uint32_t W, gt_derived *d, std::string *why) {
    (void)H;
    (void)W;
    *d = gt_derived();
    if (p->style != GT_STYLE_TRACKPY && p->style != GT_STYLE_LOG) {
        *why = "unknown style";
        return false;
    }
    if (p->has_adaptive_background) {
        if (p->adaptive_background == 0) {
            // execute_gpu.rs:299-315 unwraps the first of zero passes
            *why = "adaptive_background needs at least one pass";
            return false;
        }
    }
    if ((p->has_snr || p->has_minmass_snr || p->has_adaptive_background) &&
        static_cast<uint64_t>(H) * W <= 24) {
        *why = "snr, minmass_snr and adaptive_background need frames of more than 24 pixels "
               "(noise estimate)";
        return false;
    }
    if (p->has_illumination_sigma) {
        // the smoothing's taps reach i32(4 sigma + 0.5) pixels each side
        if (!(p->illumination_sigma > 0.0f) || !(p->illumination_sigma <= 65536.0f)) {
            *why = "illumination_sigma must be positive and at most 65536";
            return false;
        }
        d->illum_sigma = p->illumination_sigma;
        d->illum_per_frame = p->illumination_correction_per_frame ? 1 : 0;
    }
    if (!(p->noise_size > 0.0f)) {
        *why = "noise_size must be positive";
        return false;
    }
    d->style = p->style;
    d->sigma2 = p->noise_size;
    if (p->style == GT_STYLE_LOG) {
        // gpu_setup.rs:549-557: smoothing side 2 * u32(max_radius + 0.5) + 1
        if (!(p->min_radius > 0.0f) || !(p->max_radius > 0.0f) || !std::isfinite(p->min_radius) ||
            !std::isfinite(p->max_radius)) {
            *why = "min_radius and max_radius must be positive and finite";
            return false;
        }
        if (p->n_radii < 1 || p->n_radii > GT_LOG_MAX_SCALES) {
            *why = "n_radii must be between 1 and 64";
            return false;
        }
        const uint32_t s = p->has_smoothing_size ? p->smoothing_size
                                                 : 2u * f2u_sat(p->max_radius + 0.5f) + 1u;
        if (s == 0) {
            *why = "smoothing_size must be at least 1";
            return false;
        }
        if (s / 2 > GT_MAX_R) {
            *why = "smoothing_size above 33 is not supported";
            return false;
        }
        d->smoothing = s;
        d->circle = 0;
        d->separation = 0;
        d->dilation = 0;
        d->margin = 0;
        // gpu_setup.rs:958-984; the start value advances by f32 addition
        const uint32_t n = p->n_radii;
        d->radii.assign(n, 0.0f);
        if (p->log_spacing) {
            float start = ::log10f(p->min_radius);
            const float end = ::log10f(p->max_radius);
            const float diff = (end - start) / static_cast<float>(n - 1);
            for (uint32_t k = 0; k + 1 < n; ++k) {
                const float t = start;
                start = start + diff;
                d->radii[k] = ::powf(10.0f, t);
            }
            d->radii[n - 1] = ::powf(10.0f, start);
        } else {
            float start = p->min_radius;
            const float diff = (p->max_radius - p->min_radius) / static_cast<float>(n - 1);
            for (uint32_t k = 0; k + 1 < n; ++k) {
                const float t = start;
                start = start + diff;
                d->radii[k] = t;
            }
            d->radii[n - 1] = start;
        }
        for (float r : d->radii) {
            const float sigma = r / ::sqrtf(2.0f);
            if (!(r > 0.0f) || f2i_sat(4.0f * sigma + 0.5f) > GT_LOG_MAX_R) {
                *why = "every radius of the scan must be positive and at most 16.9";
                return false;
            }
        }
        d->overlap_threshold = p->overlap_threshold;
        if (!(p->threshold_rel >= 0.0f) || !std::isfinite(p->threshold_rel)) {
            *why = "threshold_rel must be a finite number >= 0";
            return false;
        }
        d->threshold_rel = p->threshold_rel;
    } else {
        if (p->threshold_rel != 0.0f) {
            *why = "threshold_rel applies to the LoG style only";
            return false;
        }
        if (p->diameter == 0) {
            *why = "diameter must be at least 1";
            return false;
        }
        // gpu_setup.rs:535-547
        const uint32_t s = p->has_smoothing_size ? p->smoothing_size : p->diameter;
        if (s == 0) {
            *why = "smoothing_size must be at least 1";
            return false;
        }
        if (s / 2 > GT_MAX_R) {
            *why = "smoothing_size above 33 is not supported";
            return false;
        }
        d->smoothing = s;
        d->circle = p->diameter;
        d->separation = p->separation;
        d->dilation = f2u_sat(2.0f * static_cast<float>(p->separation) / std::sqrt(2.0f));
        uint32_t m = p->diameter / 2u;
        const uint32_t m1 = p->separation / 2u - 1u;  // wraps for separation < 2, as the u32 code does
        if (m1 > m) m = m1;
        if (s / 2u > m) m = s / 2u;
        d->margin = static_cast<int32_t>(m);
    }
    d->max_iterations = p->max_iterations;
    d->shift_threshold = p->shift_threshold;
    d->minmass = p->minmass;
    // gpu_setup.rs:572-578: rough_snr_factor 0.5 with adaptive_background
    d->snr = p->has_snr ? p->snr : 0.0f;
    d->minmass_snr = p->has_minmass_snr ? p->minmass_snr : 0.0f;
    d->rough = p->has_adaptive_background ? 0.5f : 1.0f;
    d->adaptive = p->has_adaptive_background ? p->adaptive_background : 0u;
    d->need_std = (p->has_snr || p->has_minmass_snr || p->has_adaptive_background) ? 1 : 0;
    d->truncate = p->truncate_preprocessed ? 1 : 0;
    d->characterize = p->characterize ? 1 : 0;
    d->filter_close = p->filter_close ? 1 : 0;
    d->doughnut = p->doughnut_correction ? 1 : 0;
    d->include_r = p->include_r_in_output ? 1 : 0;
    // execute_gpu.rs:925-931 and :743-750: radius = diameter / 2; the ring
    // spans [gap_radius + radius, bg_radius (default 2 * radius)].
    const float radius = static_cast<float>(p->diameter) / 2.0f;
    d->bg_radius = p->has_bg_radius ? p->bg_radius : 2.0f * radius;
    d->bg_radius_given = p->has_bg_radius ? 1 : 0;
    d->gap_radius = p->has_gap_radius ? p->gap_radius : 0.0f;
    return true;
}

void gt_host_fill_logtable(const gt_derived *d, GtLogTable *t) {
    std::memset(t, 0, sizeof(*t));
    t->n = static_cast<int32_t>(d->radii.size());
    t->snr_thr = 0.0f * d->snr * d->rough;
    int32_t off = 0;
    for (int32_t k = 0; k < t->n; ++k) {
        // separable_log.wgsl:22-89 with sigma = radius / sqrt(2)
        // (execute_gpu.rs:253): taps exp(-x^2 / (2 sigma^2)) and that times
        // (x^2 - sigma^2), norms 1/(sqrt(2 pi) sigma), 1/(sqrt(2 pi) sigma^3)
        const float sigma = d->radii[k] / ::sqrtf(2.0f);
        const float sigma2 = sigma * sigma;
        const int32_t R = f2i_sat(4.0f * sigma + 0.5f);
        t->radius[k] = d->radii[k];
        t->R[k] = R;
        t->off[k] = off;
        for (int32_t i = -R; i <= R; ++i) {
            const float x = static_cast<float>(i), x2 = x * x;
            const float e = ::expf(-x2 / (2.0f * sigma2));
            t->taps[off + i + R] = e;
            t->taps[off + 2 * R + 1 + i + R] = e * (x2 - sigma2);
        }
        off += 2 * (2 * R + 1);
        const float sqrt2pi = 2.5066282746310002f;
        t->n1[k] = 1.0f / (sqrt2pi * sigma);
        t->n3[k] = 1.0f / (sqrt2pi * sigma * sigma2);
        if (R > t->Rm) t->Rm = R;
        // walk disk rows: the same float test as the kernels' per-pixel one
        const float r = d->radii[k], r2 = r * r;
        const int32_t rint = f2i_sat(std::ceil(r));
        for (int32_t i = -GT_LOG_WR; i <= GT_LOG_WR; ++i) {
            int32_t jm = -1;
            if (rint <= GT_LOG_WR && i >= -rint && i <= rint) {
                const float x = static_cast<float>(i), x2 = x * x;
                for (int32_t j = 0; j <= rint; ++j) {
                    const float y = static_cast<float>(j), y2 = y * y;
                    if (!(x2 + y2 > r2)) jm = j;
                }
            }
            t->wjm[k][i + GT_LOG_WR] = static_cast<int8_t>(jm);
        }
    }
}

namespace {

template <class Fma>
bool one_correction_exact_with(float b, Fma fma) {
    const float rb = 1.0f / b;
    for (uint32_t m = 0; m < (1u << 23); ++m) {
        const uint32_t u = 0x3F800000u | m;
        float a;
        std::memcpy(&a, &u, 4);
        const float q = a * rb;
        const float r = fma(-q, b, a);
        const float q1 = fma(r, rb, q);
        if (q1 != a / b) return false;
    }
    return true;
}

__attribute__((target("fma"))) bool one_correction_exact_hw(float b) {
    return one_correction_exact_with(b, [](float x, float y, float z) { return __builtin_fmaf(x, y, z); });
}

bool one_correction_exact_lib(float b) {
    return one_correction_exact_with(b, [](float x, float y, float z) { return std::fma(x, y, z); });
}

template <class Fma>
bool split_exact_with(float b, float rh, float rl, Fma fma) {
    for (uint32_t m = 0; m < (1u << 23); ++m) {
        const uint32_t u = 0x3F800000u | m;
        float a;
        std::memcpy(&a, &u, 4);
        if (fma(a, rh, a * rl) != a / b) return false;
    }
    return true;
}

__attribute__((target("fma"))) bool split_exact_hw(float b, float rh, float rl) {
    return split_exact_with(b, rh, rl, [](float x, float y, float z) { return __builtin_fmaf(x, y, z); });
}

}  // namespace

bool gt_host_split_exact(float b, float *rh, float *rl) {
    const double r = 1.0 / static_cast<double>(b);
    *rh = static_cast<float>(r);
    *rl = static_cast<float>(r - static_cast<double>(*rh));
    if (!(b >= 1.0f && b < 16777216.0f)) return false;
    static std::mutex mu;
    static std::map<uint32_t, bool> cache;
    uint32_t key;
    std::memcpy(&key, &b, 4);
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = cache.find(key);
        if (it != cache.end()) return it->second;
    }
    const float h = *rh, l = *rl;
    const bool ok = __builtin_cpu_supports("fma")
                        ? split_exact_hw(b, h, l)
                        : split_exact_with(b, h, l, [](float x, float y, float z) { return std::fma(x, y, z); });
    std::lock_guard<std::mutex> lk(mu);
    cache[key] = ok;
    return ok;
}

bool gt_host_one_correction_exact(float b) {
    if (!(b >= 1.0f && b < 16777216.0f)) return false;
    static std::mutex mu;
    static std::map<uint32_t, bool> cache;
    uint32_t key;
    std::memcpy(&key, &b, 4);
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = cache.find(key);
        if (it != cache.end()) return it->second;
    }
    const bool ok = __builtin_cpu_supports("fma") ? one_correction_exact_hw(b)
                                                  : one_correction_exact_lib(b);
    std::lock_guard<std::mutex> lk(mu);
    cache[key] = ok;
    return ok;
}

void gt_host_fill_kparams(const gt_derived *d, uint32_t H, uint32_t W, GtKParams *kp) {
    std::memset(kp, 0, sizeof(*kp));
    kp->H = static_cast<int32_t>(H);
    kp->W = static_cast<int32_t>(W);
    const int32_t R = static_cast<int32_t>(d->smoothing / 2);
    kp->R = R;
    // preprocess_rows.wgsl:30-47 evaluates exp per tap in f32; the table is
    // computed once here with the same f32 argument.
    float norm = 0.0f;
    for (int32_t i = -R; i <= R; ++i) {
        const float x = static_cast<float>(i);
        const float g = std::exp(-x * x / (2.0f * d->sigma2));
        kp->w[i + R] = g;
        norm = norm + g;
    }
    kp->gnorm = norm;
    kp->box_area = static_cast<float>(static_cast<int32_t>(d->smoothing) * static_cast<int32_t>(d->smoothing));
    kp->rgnorm = 1.0f / kp->gnorm;
    kp->rbox = 1.0f / kp->box_area;
    // Integer frames let the bandpass divide by the constant norms with a
    // reciprocal and one fma correction: that needs dividends that are not
    // tiny (every nonzero weight >= 2^-46) and the sequence checked exact for
    // both divisors; otherwise the general kernel divides.
    kp->fast_div = 1;
    for (int32_t i = 0; i <= 2 * R; ++i)
        if (kp->w[i] != 0.0f && kp->w[i] < 1.4210854715202004e-14f) kp->fast_div = 0;
    if (kp->fast_div && !(gt_host_one_correction_exact(kp->gnorm) &&
                          gt_host_one_correction_exact(kp->box_area)))
        kp->fast_div = 0;
    // two operations per division where both norms pass the whole-binade
    // check of the split reciprocal (the dividends' a * r_lo stay normal:
    // nonzero weights >= 2^-46 as above)
    float gh, gl, bh, bl;
    if (kp->fast_div && gt_host_split_exact(kp->gnorm, &gh, &gl) &&
        gt_host_split_exact(kp->box_area, &bh, &bl)) {
        kp->fast_div = 2;
        kp->rgnorm = gh;
        kp->rgnorm_lo = gl;
        kp->rbox = bh;
        kp->rbox_lo = bl;
    }
    kp->truncate = d->truncate;
    kp->k = static_cast<int32_t>(d->dilation);
    kp->a = f2i_sat(static_cast<float>(kp->k) / 2.0f - 0.5f);
    kp->margin = d->margin;
    kp->max_thr = 0.0f * d->snr * d->rough;
    kp->rint = (static_cast<int32_t>(d->circle) - 1) / 2;
    kp->r2 = static_cast<float>(kp->rint) * static_cast<float>(kp->rint);
    kp->max_iter = d->max_iterations;
    kp->shift_thr = d->shift_threshold;
    kp->minmass = d->minmass;
    kp->minmass_snr = d->minmass_snr;
    kp->image_std = 0.0f;
    kp->rough = d->rough;
    kp->snr = d->snr;
    kp->thr_rel = d->threshold_rel;
    kp->fstd = nullptr;   // set by the context when the noise estimate runs
    kp->characterize = d->characterize;
    kp->circle = static_cast<int32_t>(d->circle);
}

uint32_t gt_host_column_names(const gt_params *p, const char **names, const char **types) {
    // column_names (src/execute_gpu.rs:482-519)
    uint32_t n = 0;
    auto push = [&](const char *a, const char *b) {
        if (names) names[n] = a;
        if (types) types[n] = b;
        ++n;
    };
    push("frame", "int");
    push("y", "float");
    push("x", "float");
    push("mass", "float");
    if (p->include_r_in_output) push("r", "float");
    if (p->characterize) {
        push("Rg", "float");
        push("raw", "float");
        push("signal", "float");
        push("ecc", "float");
    }
    if (p->doughnut_correction) {
        push("raw_mass", "float");
        push("raw_bg_median", "float");
        push("raw_mass_corrected", "float");
    }
    if (p->has_search_range) {
        push("particle", "int");
        push("duration", "int");
    }
    return n;
}

void gt_host_widen(const void *src, int32_t dtype, size_t n, float *dst) {
    switch (dtype) {
    case GT_U16: {
        const uint16_t *s = static_cast<const uint16_t *>(src);
        for (size_t i = 0; i < n; ++i) dst[i] = static_cast<float>(s[i]);
        break;
    }
    case GT_U8: {
        const uint8_t *s = static_cast<const uint8_t *>(src);
        for (size_t i = 0; i < n; ++i) dst[i] = static_cast<float>(s[i]);
        break;
    }
    default: std::memcpy(dst, src, n * sizeof(float)); break;
    }
}

gt_status gt_host_plan_frames(int32_t characterize_points, uint64_t nframes, const uint64_t *keys,
                              uint64_t nkeys, const float *points, uint64_t npoints,
                              uint32_t point_cols, int32_t has_frames, int32_t has_r,
                              gt_host_plan *plan, std::string *why) {
    plan->frames.clear();
    plan->pt_begin.clear();
    plan->pt_end.clear();
    plan->pts.clear();
    if (characterize_points && points) {
        const uint32_t need = 2u + (has_frames ? 1u : 0u) + (has_r ? 1u : 0u);
        if (point_cols != need) {
            *why = "unsupported characterize array dimensions.\naccepted dimensions: (Nx2) or (Nx3).";
            return GT_ERR_ARRAY_DIMENSIONS;
        }
        const uint32_t cy = has_frames ? 1 : 0, cx = cy + 1, cr = cx + 1;
        plan->pts.resize(npoints);
        for (uint64_t i = 0; i < npoints; ++i) {
            const float *row = points + i * point_cols;
            gt_result_row r;
            std::memset(&r, 0, sizeof(r));
            r.x = row[cy];
            r.y = row[cx];
            if (has_r) r.r = row[cr];
            plan->pts[i] = r;
        }
        if (!has_frames) {
            // FrameSubsetter without a frame column cycles the same positions
            // over every frame (execute_gpu.rs:1010-1017).
            if (nframes == 0) return GT_OK;
            for (uint64_t f = 0; f < nframes; ++f) {
                plan->frames.push_back(f);
                plan->pt_begin.push_back(0);
                plan->pt_end.push_back(npoints);
            }
            return GT_OK;
        }
        // With a frame column: groups in ascending frame order; a frame value
        // that decreases is an error; the first group starts at frame 0 even
        // when it is empty (linking.rs:104-138).
        float cur = 0.0f;
        size_t begin = 0;
        for (uint64_t i = 0; i < npoints; ++i) {
            const float fr = points[i * point_cols];
            if (fr > cur) {
                plan->frames.push_back(static_cast<uint64_t>(cur));
                plan->pt_begin.push_back(begin);
                plan->pt_end.push_back(i);
                begin = i;
                cur = fr;
            }
            if (fr < cur) {
                *why = "the points provided for characterization aren't sorted in ascending frame number.";
                return GT_ERR_NON_SORTED_CHARACTERIZATION;
            }
        }
        if (npoints > begin) {
            plan->frames.push_back(static_cast<uint64_t>(cur));
            plan->pt_begin.push_back(begin);
            plan->pt_end.push_back(npoints);
        }
        return GT_OK;
    }
    if (keys) {
        plan->frames.assign(keys, keys + nkeys);
    } else {
        plan->frames.resize(nframes);
        for (uint64_t f = 0; f < nframes; ++f) plan->frames[f] = f;
    }
    return GT_OK;
}

gt_host_post::gt_host_post(const gt_params *p, const gt_derived *d, uint32_t H, uint32_t W,
                           const uint64_t *reset_points, uint64_t nreset)
    : p_(*p), d_(*d), H_(H), W_(W) {
    ncols_ = gt_host_column_names(p, nullptr, nullptr);
    if (p->has_search_range) {
        if (!(p->search_range > 0.0f) || !std::isfinite(p->search_range)) {
            error_ = "search_range must be positive and finite";
            return;
        }
        linker_.reset(new gt_linker(p->search_range, p->has_memory ? p->memory : 0u));
        if (reset_points && nreset) {
            reset_.assign(reset_points, reset_points + nreset);
            std::sort(reset_.begin(), reset_.end());
        }
    }
}

gt_status gt_host_post::process(std::vector<gt_detection> &rows, uint64_t frame_index,
                                const float *raw, std::atomic<int32_t> *interrupt,
                                std::string *err, const float *post4) {
    (void)interrupt;
    (void)err;
    if (post4 && d_.style == GT_STYLE_TRACKPY) {
        // filter_close and the doughnut ran on the GPU: order the rows by
        // seed carrying their results, keep the survivors
        // (seed, position) keys: a frame's seeds are distinct pixels
        keys_.resize(rows.size());
        for (size_t i = 0; i < rows.size(); ++i)
            keys_[i] = (static_cast<uint64_t>(rows[i].seed) << 32) | static_cast<uint32_t>(i);
        std::sort(keys_.begin(), keys_.end());
        tmp_rows_.clear();
        dough_.clear();
        for (uint64_t key : keys_) {
            const uint32_t i = static_cast<uint32_t>(key);
            const float *q = post4 + 4 * static_cast<size_t>(i);
            if (q[3] == 0.0f) continue;
            tmp_rows_.push_back(rows[i].row);
            if (d_.doughnut) {
                dough_.push_back(q[0]);
                dough_.push_back(q[1]);
                dough_.push_back(q[2]);
            }
        }
        return assemble(frame_index, interrupt, err);
    }
    std::sort(rows.begin(), rows.end(),
              [](const gt_detection &a, const gt_detection &b) { return a.seed < b.seed; });
    tmp_rows_.resize(rows.size());
    for (size_t i = 0; i < rows.size(); ++i) tmp_rows_[i] = rows[i].row;
    size_t n = tmp_rows_.size();
    if (d_.style == GT_STYLE_TRACKPY && d_.filter_close && n > 0) {
        keep_.assign(n, 1);
        gt_filter_close(tmp_rows_.data(), n, static_cast<float>(d_.separation), keep_.data());
        size_t m = 0;
        for (size_t i = 0; i < n; ++i)
            if (keep_[i]) tmp_rows_[m++] = tmp_rows_[i];
        n = m;
        tmp_rows_.resize(n);
    }
    if (d_.style == GT_STYLE_LOG && n > 0) {
        n = gt_host_prune_blobs(tmp_rows_.data(), n, d_.overlap_threshold);
        tmp_rows_.resize(n);
    }
    if (d_.doughnut) {
        dough_.assign(3 * n, 0.0f);
        if (n && d_.style == GT_STYLE_LOG) {
            // per detection: radius r, ring [gap + r, bg_radius or 2 r]
            // (execute_gpu.rs:740-747 with :925-931)
            for (size_t i = 0; i < n; ++i) {
                const float r = tmp_rows_[i].r;
                gt_doughnut(raw, H_, W_, &tmp_rows_[i], 1, r,
                            d_.bg_radius_given ? d_.bg_radius : 2.0f * r, d_.gap_radius,
                            &dough_[3 * i]);
            }
        } else if (n) {
            gt_doughnut(raw, H_, W_, tmp_rows_.data(), n, static_cast<float>(d_.circle) / 2.0f,
                        d_.bg_radius, d_.gap_radius, dough_.data());
        }
    }
    return assemble(frame_index, interrupt, err);
}

// Linking and the output rows of one frame (tmp_rows_ and dough_ hold the
// frame's surviving rows in seed order and their doughnut columns).
gt_status gt_host_post::assemble(uint64_t frame_index, std::atomic<int32_t> *interrupt,
                                 std::string *err) {
    const size_t n = tmp_rows_.size();
    if (linker_) {
        // the frame's positions go to the linking thread (started with the
        // first frame); an error it met ends the run here
        LinkJob job;
        job.frame = frame_index;
        job.row0 = ncols_ ? out_.size() / ncols_ : 0;
        job.xy.resize(2 * n);
        for (size_t i = 0; i < n; ++i) {
            job.xy[2 * i] = tmp_rows_[i].x;
            job.xy[2 * i + 1] = tmp_rows_[i].y;
        }
        {
            std::lock_guard<std::mutex> lk(link_mu_);
            if (link_status_ != GT_OK) {
                *err = link_err_;
                return link_status_;
            }
            link_q_.push_back(std::move(job));
        }
        if (!link_thread_.joinable()) {
            link_interrupt_ = interrupt;
            link_thread_ = std::thread([this] { link_loop(); });
        }
        link_cv_.notify_one();
    }
    const float fidx = static_cast<float>(frame_index);
    const size_t base = out_.size();
    if (!out_.resize(base + n * ncols_)) {
        *err = "out of host memory for the output rows";
        return GT_ERR_GPU_DEVICE;
    }
    float *o = out_.data() + base;
    for (size_t i = 0; i < n; ++i) {
        const gt_result_row &r = tmp_rows_[i];
        *o++ = fidx;
        *o++ = r.x;
        *o++ = r.y;
        *o++ = r.mass;
        if (d_.include_r) *o++ = r.r;
        if (d_.characterize) {
            *o++ = r.Rg;
            *o++ = r.raw_mass;
            *o++ = r.signal;
            *o++ = r.ecc;
        }
        if (d_.doughnut) {
            *o++ = dough_[3 * i + 0];
            *o++ = dough_[3 * i + 1];
            *o++ = dough_[3 * i + 2];
        }
        if (linker_) {
            *o++ = 0.0f;    // particle and duration, filled in by finish()
            *o++ = 0.0f;
        }
    }
    if (o != out_.data() + out_.size()) {
        *err = "internal error: the row layout does not match the column list";
        return GT_ERR_INVALID_ARGUMENT;
    }
    return GT_OK;
}

gt_host_post::~gt_host_post() {
    std::string ignored;
    (void)stop_linking(&ignored);
}

void gt_host_post::link_loop() {
    std::vector<uint64_t> ids;
    for (;;) {
        LinkJob job;
        {
            std::unique_lock<std::mutex> lk(link_mu_);
            link_cv_.wait(lk, [&] { return !link_q_.empty() || link_stop_; });
            if (link_q_.empty()) return;
            job = std::move(link_q_.front());
            link_q_.pop_front();
            if (link_status_ != GT_OK) continue;   // drain after an error
        }
        // linker_reset_points (execute_gpu.rs:792-797): a new linker, ids
        // start over
        if (std::binary_search(reset_.begin(), reset_.end(), job.frame)) {
            seg_books_.push_back(linker_->finish());
            linker_.reset(new gt_linker(p_.search_range, p_.has_memory ? p_.memory : 0u));
        }
        const size_t n = job.xy.size() / 2;
        ids.resize(n);
        std::string err;
        const gt_status ls = linker_->advance(job.xy.data(), n, ids.data(), link_interrupt_, &err);
        if (ls != GT_OK) {
            std::lock_guard<std::mutex> lk(link_mu_);
            link_status_ = ls;
            link_err_ = err;
            continue;
        }
        if (row_ids_.size() < job.row0 + n) {
            row_ids_.resize(job.row0 + n, 0);
            row_seg_.resize(job.row0 + n, 0);
        }
        for (size_t i = 0; i < n; ++i) {
            row_ids_[job.row0 + i] = ids[i];
            row_seg_[job.row0 + i] = static_cast<uint32_t>(seg_books_.size());
        }
    }
}

gt_status gt_host_post::stop_linking(std::string *err) {
    if (!link_thread_.joinable()) return GT_OK;
    {
        std::lock_guard<std::mutex> lk(link_mu_);
        link_stop_ = true;
    }
    link_cv_.notify_all();
    link_thread_.join();
    if (link_status_ != GT_OK) *err = link_err_;
    return link_status_;
}

gt_status gt_host_post::finish(gt_output *out, std::string *err) {
    if (linker_) {
        const gt_status ls = stop_linking(err);
        if (ls != GT_OK) return ls;
        // particle column, then the duration column (execute_gpu.rs:
        // 1165-1181), the last two.  Each segment between reset points keeps
        // its own durations (the reference reads every row's id in the last
        // segment's table).
        seg_books_.push_back(linker_->finish());
        const uint32_t nc = ncols_;
        const size_t nrows = nc ? out_.size() / nc : 0;
        for (size_t r = 0; r < nrows && r < row_ids_.size(); ++r) {
            float *row = &out_[r * nc];
            const auto &books = seg_books_[row_seg_[r]];
            const uint64_t id = row_ids_[r];
            row[nc - 2] = static_cast<float>(id);
            row[nc - 1] = id < books.size() ? static_cast<float>(books[id].duration) : 0.0f;
        }
    }
    // One static name/type list per combination of the four optional column
    // groups, so the pointers handed out stay valid and are never rewritten.
    struct Lists {
        const char *names[16][16];
        const char *types[16][16];
    };
    static const Lists lists = [] {
        Lists l;
        std::memset(&l, 0, sizeof(l));
        for (int m = 0; m < 16; ++m) {
            gt_params q;
            gt_host_params_default(&q);
            q.include_r_in_output = m & 1;
            q.characterize = (m >> 1) & 1;
            q.doughnut_correction = (m >> 2) & 1;
            q.has_search_range = (m >> 3) & 1;
            gt_host_column_names(&q, l.names[m], l.types[m]);
        }
        return l;
    }();
    const int combo = (p_.include_r_in_output ? 1 : 0) | (p_.characterize ? 2 : 0) |
                      (p_.doughnut_correction ? 4 : 0) | (p_.has_search_range ? 8 : 0);
    const char *const *names = lists.names[combo];
    const char *const *types = lists.types[combo];
    const uint32_t nc = ncols_;
    out->ncols = nc;
    out->nrows = nc ? out_.size() / nc : 0;
    if (!out_.size() && !out_.resize(1)) return GT_ERR_GPU_DEVICE;   // a block to free, even empty
    out->data = out_.release();
    out->col_names = names;
    out->col_types = types;
    return GT_OK;
}

extern "C" gt_status gt_filter_close(const gt_result_row *rows, uint64_t n, float separation,
                                     uint8_t *keep) {
    if (n == 0) return GT_OK;
    if (!rows || !keep) return gt_runtime_fail(GT_ERR_INVALID_ARGUMENT, "null argument");
    const float r2 = separation * separation;
    // Uniform grid with cells of side `separation`; +-2 cells are searched so
    // rounding of x / separation near a cell edge cannot hide a neighbour.
    const float cell = separation > 0.0f ? separation : 1.0f;
    std::unordered_map<int64_t, std::vector<uint32_t>> grid;
    grid.reserve(static_cast<size_t>(n));
    auto key = [](int64_t cx, int64_t cy) { return (cx << 32) ^ (cy & 0xFFFFFFFFll); };
    std::vector<int64_t> cxs(n), cys(n);
    std::vector<uint8_t> finite(n);
    for (uint64_t i = 0; i < n; ++i) {
        finite[i] = std::isfinite(rows[i].x) && std::isfinite(rows[i].y);
        if (!finite[i]) continue;
        cxs[i] = static_cast<int64_t>(std::floor(static_cast<double>(rows[i].x) / cell));
        cys[i] = static_cast<int64_t>(std::floor(static_cast<double>(rows[i].y) / cell));
        grid[key(cxs[i], cys[i])].push_back(static_cast<uint32_t>(i));
    }
    for (uint64_t i = 0; i < n; ++i) {
        keep[i] = 1;
        if (!finite[i]) continue;  // NaN/inf distances never fall below r^2
        const float mi = rows[i].mass;
        bool drop = false;
        for (int64_t dx = -2; dx <= 2 && !drop; ++dx)
            for (int64_t dy = -2; dy <= 2 && !drop; ++dy) {
                auto it = grid.find(key(cxs[i] + dx, cys[i] + dy));
                if (it == grid.end()) continue;
                for (uint32_t j : it->second) {
                    if (j == i) continue;
                    const float ddx = rows[j].x - rows[i].x, ddy = rows[j].y - rows[i].y;
                    float d2 = 0.0f;
                    d2 = d2 + ddx * ddx;
                    d2 = d2 + ddy * ddy;
                    if (!(d2 < r2)) continue;
                    const float mj = rows[j].mass;
                    if (mi < mj || (mi == mj && i > j)) {
                        drop = true;
                        break;
                    }
                }
            }
        if (drop) keep[i] = 0;
    }
    return GT_OK;
}

// circle_inds(radius) / annulus_inds(bg_radius, gap_radius + radius)
// (kernels.rs:138-171), in their i-major, j-minor order.
void gt_host_doughnut_offsets(float radius, float bg_radius, float gap_radius,
                              std::vector<std::pair<int32_t, int32_t>> *sig,
                              std::vector<std::pair<int32_t, int32_t>> *ring) {
    sig->clear();
    ring->clear();
    const int32_t cb = f2i_sat(std::ceil(radius));
    const float rr = radius * radius;
    for (int32_t i = -cb; i <= cb; ++i)
        for (int32_t j = -cb; j <= cb; ++j)
            if (static_cast<float>(i * i + j * j) <= rr) sig->emplace_back(i, j);
    const float inner = gap_radius + radius;
    const int32_t ab = f2i_sat(std::ceil(bg_radius));
    const float or2 = bg_radius * bg_radius, ir2 = inner * inner;
    for (int32_t i = -ab; i <= ab; ++i)
        for (int32_t j = -ab; j <= ab; ++j) {
            const float q = static_cast<float>(i * i + j * j);
            if (q <= or2 && q >= ir2) ring->emplace_back(i, j);
        }
}

extern "C" gt_status gt_doughnut(const float *raw, uint32_t H, uint32_t W, const gt_result_row *rows,
                                 uint64_t n, float radius, float bg_radius, float gap_radius,
                                 float *out3) {
    if (n == 0) return GT_OK;
    if (!raw || !rows || !out3) return gt_runtime_fail(GT_ERR_INVALID_ARGUMENT, "null argument");
    std::vector<std::pair<int32_t, int32_t>> sig, ring;
    gt_host_doughnut_offsets(radius, bg_radius, gap_radius, &sig, &ring);
    std::vector<float> bg;
    bg.reserve(ring.size());
    for (uint64_t k = 0; k < n; ++k) {
        // centre rounded half away from zero (f32::round, execute_gpu.rs:752)
        const int64_t mu = f2i_sat(std::round(rows[k].x)), mv = f2i_sat(std::round(rows[k].y));
        float s = 0.0f;
        size_t ns = 0;
        for (const auto &o : sig) {
            const int64_t pu = mu + o.first, pv = mv + o.second;
            if (pu < 0 || pu >= H || pv < 0 || pv >= W) continue;
            s = s + raw[pu * W + pv];
            ++ns;
        }
        bg.clear();
        for (const auto &o : ring) {
            const int64_t pu = mu + o.first, pv = mv + o.second;
            if (pu < 0 || pu >= H || pv < 0 || pv >= W) continue;
            bg.push_back(raw[pu * W + pv]);
        }
        float med = 0.0f;
        if (!bg.empty()) {
            std::nth_element(bg.begin(), bg.begin() + bg.size() / 2, bg.end());
            med = bg[bg.size() / 2];
        }
        out3[3 * k + 0] = s;
        out3[3 * k + 1] = med;
        out3[3 * k + 2] = s - med * static_cast<float>(ns);
    }
    return GT_OK;
}

namespace {

// blob_overlap / disk_overlap (execute_gpu.rs:597-630), f32 throughout as
// the Rust code (no contraction).
float make_acos(float d, float r1, float r1s, float r2s) {
    float ratio = (d * d + r1s - r2s) / (2.0f * d * r1);
    ratio = std::min(std::max(ratio, -1.0f), 1.0f);
    return ::acosf(ratio);
}

float blob_overlap(float r1, float r2, float distance) {
    if (distance > r1 + r2) return 0.0f;
    if (distance <= std::fabs(r1 - r2)) return 1.0f;
    const float d = distance;
    const float r1s = r1 * r1, r2s = r2 * r2;
    const float acos1 = make_acos(d, r1, r1s, r2s);
    const float acos2 = make_acos(d, r2, r2s, r1s);
    const float a = -d + r2 + r1;
    const float b = d - r2 + r1;
    const float c = d + r2 - r1;
    const float e = d + r2 + r1;
    const float area = r1s * acos1 + r2s * acos2 - 0.5f * ::sqrtf(std::fabs(a * b * c * e));
    const float pi = 3.14159265358979323846f;
    return area / (pi * std::min(r1s, r2s));
}

}  // namespace

size_t gt_host_prune_blobs(gt_result_row *rows, size_t n, float overlap_threshold) {
    if (!(overlap_threshold < 1.0f) || n == 0) return n;
    // Stable sort, r descending then mass descending.  The reference's
    // partial_cmp().unwrap() would panic on NaN; here NaN sorts last.
    auto kf = [](float v) { return v != v ? -INFINITY : v; };
    std::stable_sort(rows, rows + n, [&](const gt_result_row &a, const gt_result_row &b) {
        const float ar = kf(a.r), br = kf(b.r);
        if (ar != br) return ar > br;
        return kf(a.mass) > kf(b.mass);
    });
    // prune_blobs_log: every surviving blob in order drops every surviving
    // other blob with distance^2 < (2 r)^2 whose overlap exceeds the
    // threshold.  Neighbours come from a uniform grid whose cell is the
    // largest query radius, so the 3 x 3 cells around a query cover it.
    float rmax = 0.0f;
    for (size_t i = 0; i < n; ++i)
        if (rows[i].r > rmax) rmax = rows[i].r;
    // (slightly enlarged so rounding of x / cell cannot push a neighbour two
    // cells away)
    const double cell = rmax > 0.0f ? 2.0001 * static_cast<double>(rmax) : 1.0;
    std::unordered_map<int64_t, std::vector<uint32_t>> grid;
    grid.reserve(n);
    auto key = [](int64_t cx, int64_t cy) { return (cx << 32) ^ (cy & 0xFFFFFFFFll); };
    std::vector<int64_t> cxs(n), cys(n);
    std::vector<uint8_t> finite(n), keep(n, 1);
    for (size_t i = 0; i < n; ++i) {
        finite[i] = std::isfinite(rows[i].x) && std::isfinite(rows[i].y);
        if (!finite[i]) continue;
        cxs[i] = static_cast<int64_t>(std::floor(static_cast<double>(rows[i].x) / cell));
        cys[i] = static_cast<int64_t>(std::floor(static_cast<double>(rows[i].y) / cell));
        grid[key(cxs[i], cys[i])].push_back(static_cast<uint32_t>(i));
    }
    for (size_t q = 0; q < n; ++q) {
        if (!keep[q] || !finite[q]) continue;
        const float rad = rows[q].r * 2.0f, r2 = rad * rad;
        for (int64_t dx = -1; dx <= 1; ++dx)
            for (int64_t dy = -1; dy <= 1; ++dy) {
                auto it = grid.find(key(cxs[q] + dx, cys[q] + dy));
                if (it == grid.end()) continue;
                for (uint32_t k : it->second) {
                    if (k == q || !keep[k]) continue;
                    const float ddx = rows[k].x - rows[q].x, ddy = rows[k].y - rows[q].y;
                    float d2 = 0.0f;
                    d2 = d2 + ddx * ddx;
                    d2 = d2 + ddy * ddy;
                    if (!(d2 < r2)) continue;
                    if (blob_overlap(rows[q].r, rows[k].r, ::sqrtf(d2)) > overlap_threshold)
                        keep[k] = 0;
                }
            }
    }
    size_t m = 0;
    for (size_t i = 0; i < n; ++i)
        if (keep[i]) rows[m++] = rows[i];
    return m;
}

extern "C" gt_status gt_prune_blobs(gt_result_row *rows, uint64_t n, float overlap_threshold,
                                    uint64_t *n_out) {
    if (!n_out || (n && !rows)) return gt_runtime_fail(GT_ERR_INVALID_ARGUMENT, "null argument");
    *n_out = gt_host_prune_blobs(rows, static_cast<size_t>(n), overlap_threshold);
    return GT_OK;
}

namespace {

// FrameSubsetter's linking groups (linking.rs:94-155) of rows [frame, y, x]:
// consecutive rows of one frame value form a group; a group's key is its
// frame value cast to an unsigned integer (Rust's saturating `as usize`);
// when the first row's frame is above 0 an empty group keyed 0 comes first;
// no rows, no groups.  Frames must not decrease.
struct FrameGroup {
    uint64_t key, begin, end;
};

uint64_t frame_key(float f) {
    if (!(f > 0.0f)) return 0;
    if (f >= 18446744073709551616.0f) return UINT64_MAX;
    return static_cast<uint64_t>(f);
}

gt_status frame_groups(const float *points, uint64_t n, std::vector<FrameGroup> *out) {
    out->clear();
    float cur = 0.0f;
    uint64_t begin = 0;
    for (uint64_t i = 0; i < n; ++i) {
        const float fr = points[3 * i];
        if (fr > cur) {
            out->push_back({frame_key(cur), begin, i});
            begin = i;
            cur = fr;
        }
        if (fr < cur)
            return gt_runtime_fail(GT_ERR_NON_SORTED_CHARACTERIZATION,
                                   "the points provided for characterization aren't sorted in "
                                   "ascending frame number.");
    }
    if (n > begin) out->push_back({frame_key(cur), begin, n});
    return GT_OK;
}

void group_xy(const float *points, const FrameGroup &g, std::vector<float> *xy) {
    xy->resize(2 * (g.end - g.begin));
    for (uint64_t i = g.begin; i < g.end; ++i) {
        (*xy)[2 * (i - g.begin)] = points[3 * i + 1];
        (*xy)[2 * (i - g.begin) + 1] = points[3 * i + 2];
    }
}

}  // namespace

extern "C" gt_status gt_link(const float *points, uint64_t n, float search_range, uint32_t memory,
                             uint64_t *ids) {
    // linker_all over FrameSubsetter's linking groups (linking.rs:818-831)
    if (n && (!points || !ids)) return gt_runtime_fail(GT_ERR_INVALID_ARGUMENT, "null argument");
    if (!(search_range > 0.0f) || !std::isfinite(search_range))
        return gt_runtime_fail(GT_ERR_INVALID_ARGUMENT, "search_range must be positive and finite");
    std::vector<FrameGroup> groups;
    gt_status st = frame_groups(points, n, &groups);
    if (st != GT_OK) return st;
    gt_linker lk(search_range, memory);
    std::vector<float> xy;
    std::string err;
    for (const FrameGroup &g : groups) {
        group_xy(points, g, &xy);
        st = lk.advance(xy.data(), g.end - g.begin, ids + g.begin, nullptr, &err);
        if (st != GT_OK) return gt_runtime_fail(st, "%s", err.c_str());
    }
    return GT_OK;
}

extern "C" gt_status gt_connect(const float *points1, uint64_t n1, const float *points2,
                                uint64_t n2, float search_range, uint64_t *ids1, uint64_t *ids2) {
    // connect_all (linking.rs:960-1012): the two sets' frame groups in
    // ascending frame order; a frame in both sets is linked set 1 -> set 2
    // (Linker::connect, :405-432), a frame in one set takes fresh ids
    // (fill_base).  Where the reference's sync_iterators (:945-955) fills
    // the frame that overtook the one being caught up to twice, that frame
    // is filled once, when it is reached.
    if ((n1 && (!points1 || !ids1)) || (n2 && (!points2 || !ids2)))
        return gt_runtime_fail(GT_ERR_INVALID_ARGUMENT, "null argument");
    if (!(search_range > 0.0f) || !std::isfinite(search_range))
        return gt_runtime_fail(GT_ERR_INVALID_ARGUMENT, "search_range must be positive and finite");
    std::vector<FrameGroup> g1, g2;
    gt_status st = frame_groups(points1, n1, &g1);
    if (st != GT_OK) return st;
    st = frame_groups(points2, n2, &g2);
    if (st != GT_OK) return st;
    gt_linker lk(search_range, 0);
    std::vector<float> xy1, xy2;
    std::string err;
    size_t i = 0, j = 0;
    while (i < g1.size() || j < g2.size()) {
        if (j == g2.size() || (i < g1.size() && g1[i].key < g2[j].key)) {
            lk.fresh(g1[i].end - g1[i].begin, ids1 + g1[i].begin);
            ++i;
        } else if (i == g1.size() || g2[j].key < g1[i].key) {
            lk.fresh(g2[j].end - g2[j].begin, ids2 + g2[j].begin);
            ++j;
        } else {
            group_xy(points1, g1[i], &xy1);
            group_xy(points2, g2[j], &xy2);
            st = lk.connect(xy1.data(), g1[i].end - g1[i].begin, xy2.data(),
                            g2[j].end - g2[j].begin, ids1 + g1[i].begin, ids2 + g2[j].begin, &err);
            if (st != GT_OK) return gt_runtime_fail(st, "%s", err.c_str());
            ++i;
            ++j;
        }
    }
    return GT_OK;
}
