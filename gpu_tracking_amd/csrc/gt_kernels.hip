// gt_kernels.hip -- gfx950 kernels of the per-frame detection pipeline.
//
//   k_bandpass   preprocess_rows + preprocess_cols (src/shaders/preprocess_rows.wgsl:14-55,
//                preprocess_cols.wgsl:11-60) fused in one LDS-tiled kernel: the u16/f32
//                tile plus its halo is read once with lanes along columns (coalesced),
//                the row pass writes (gauss, box) pairs to LDS and the column pass reads
//                them back.
//   k_localmax   max_rows + extract_max (max_rows.wgsl:10-32, extract_max.wgsl:35-78)
//                fused: separable k x k grey dilation from an LDS tile, then a
//                wavefront ballot compacts the maxima with one atomic per wave.
//   k_walk       walk + characterize (walk.wgsl:67-229, characterize.wgsl:30-155), one
//                lane per candidate; detections are appended with one atomic per wave.
//
// Arithmetic follows the oracle statement for statement: taps are accumulated
// in order i = -R..R with explicit fmaf, divisions and square roots are IEEE
// correctly rounded (hipcc's default), and the file is compiled with
// -ffp-contract=off so no other operation is fused.
#include <hip/hip_runtime.h>
#include <math.h>

#include <algorithm>

#include "gt_internal.h"

namespace {

// WGSL `1./255.` (preprocess_cols.wgsl:54): f32 bit pattern 0x3B808081.
constexpr float kTrunc = 0.003921568859368562698364257812f;

constexpr int BP_TW = 64, BP_TH = 32, BP_THREADS = 256;
constexpr int LM_TW = 64, LM_THREADS = 256;
constexpr int WK_THREADS = 256;

template <typename T>
__device__ __forceinline__ float to_f32(T v) { return static_cast<float>(v); }

__device__ __forceinline__ int32_t f2i(float x) {
    // NaN -> 0, saturating (what v_cvt_i32_f32 does; C++ leaves it undefined).
    if (x != x) return 0;
    if (x >= 2147483648.0f) return 2147483647;
    if (x <= -2147483648.0f) return (-2147483647 - 1);
    return static_cast<int32_t>(x);
}

__device__ __forceinline__ float maxf_ref(float a, float b) { return b > a ? b : a; }

// Appends the lanes with `flag` set to a list guarded by *counter: one
// atomicAdd per wavefront.  Must be reached by every lane of the wave.
__device__ __forceinline__ uint32_t wave_append(bool flag, uint32_t *counter) {
    const unsigned long long mask = __ballot(flag);
    if (mask == 0ull) return 0u;
    const uint32_t lane = __lane_id();
    const uint32_t prefix = __popcll(mask & ((1ull << lane) - 1ull));
    const int leader = __ffsll(static_cast<long long>(mask)) - 1;
    uint32_t base = 0u;
    if (static_cast<int>(lane) == leader) base = atomicAdd(counter, static_cast<uint32_t>(__popcll(mask)));
    base = __shfl(base, leader);
    return base + prefix;
}

// ---------------------------------------------------------------------------
// Bandpass: separable Gaussian (zero padded, each pass divided by the full
// tap sum) minus the s x s boxcar mean (edge clamped), truncated below 1/255.
template <typename T, int RT>
__global__ __launch_bounds__(BP_THREADS) void k_bandpass(const T *__restrict__ frames,
                                                         float *__restrict__ proc, GtKParams p) {
    extern __shared__ float lds[];
    const int R = RT > 0 ? RT : p.R;
    const int NT = 2 * R + 1;
    const int RW = BP_TW + 2 * R, RH = BP_TH + 2 * R;
    float *Sraw = lds;
    float *Sg = Sraw + RH * RW;
    float *Sc = Sg + RH * BP_TW;
    __shared__ float Sw[2 * GT_MAX_R + 1];
    const int H = p.H, W = p.W;
    const size_t HW = static_cast<size_t>(H) * W;
    const int x0 = blockIdx.x * BP_TW, y0 = blockIdx.y * BP_TH;
    const T *f = frames + static_cast<size_t>(blockIdx.z) * HW;
    float *out = proc + static_cast<size_t>(blockIdx.z) * HW;
    const int tid = threadIdx.x;
    if (tid < NT) Sw[tid] = p.w[tid];
    // Stage the tile and its halo; out-of-frame coordinates clamp to the edge,
    // which is exactly the boxcar's "nearest" rule, and the Gaussian skips
    // those taps below by testing the unclamped coordinate.
    for (int idx = tid; idx < RH * RW; idx += BP_THREADS) {
        const int ry = idx / RW, rx = idx - ry * RW;
        const int gy = min(max(y0 - R + ry, 0), H - 1);
        const int gx = min(max(x0 - R + rx, 0), W - 1);
        Sraw[idx] = to_f32(f[static_cast<size_t>(gy) * W + gx]);
    }
    __syncthreads();
    const float gnorm = p.gnorm;
    // Row pass over every staged row (halo rows included).
    for (int idx = tid; idx < RH * BP_TW; idx += BP_THREADS) {
        const int ry = idx / BP_TW, cx = idx - ry * BP_TW;
        const int gx = x0 + cx;
        const float *src = Sraw + ry * RW + cx;
        float gs = 0.0f, cs = 0.0f;
        if (RT > 0) {
#pragma unroll
            for (int t = 0; t < 2 * RT + 1; ++t) {
                const int px = gx - RT + t;
                const float v = src[t];
                if (px >= 0 && px < W) gs = fmaf(p.w[t], v, gs);
                cs = cs + v;
            }
        } else {
            for (int t = 0; t < NT; ++t) {
                const int px = gx - R + t;
                const float v = src[t];
                if (px >= 0 && px < W) gs = fmaf(Sw[t], v, gs);
                cs = cs + v;
            }
        }
        Sg[idx] = gs / gnorm;
        Sc[idx] = cs;
    }
    __syncthreads();
    // Column pass.
    const float box = p.box_area;
    for (int idx = tid; idx < BP_TH * BP_TW; idx += BP_THREADS) {
        const int ty = idx / BP_TW, cx = idx - ty * BP_TW;
        const int gy = y0 + ty, gx = x0 + cx;
        if (gy >= H || gx >= W) continue;
        float gs = 0.0f, cs = 0.0f;
        if (RT > 0) {
#pragma unroll
            for (int t = 0; t < 2 * RT + 1; ++t) {
                const int py = gy - RT + t;
                const float g = Sg[(ty + t) * BP_TW + cx];
                const float c = Sc[(ty + t) * BP_TW + cx];
                if (py >= 0 && py < H) gs = fmaf(p.w[t], g, gs);
                cs = cs + c;
            }
        } else {
            for (int t = 0; t < NT; ++t) {
                const int py = gy - R + t;
                const float g = Sg[(ty + t) * BP_TW + cx];
                const float c = Sc[(ty + t) * BP_TW + cx];
                if (py >= 0 && py < H) gs = fmaf(Sw[t], g, gs);
                cs = cs + c;
            }
        }
        float res = gs / gnorm - cs / box;
        if (p.truncate && res < kTrunc) res = 0.0f;
        out[static_cast<size_t>(gy) * W + gx] = res;
    }
}

// ---------------------------------------------------------------------------
// Local maximum: a pixel inside the margin is a candidate when no in-frame
// value of its k x k window [-a, k-1-a]^2 is greater than it, and it is not
// below the snr threshold.
// Untiled local maximum for dilation windows too large for the LDS tile.
__global__ __launch_bounds__(LM_THREADS) void k_localmax_direct(const float *__restrict__ proc,
                                                                GtKParams p,
                                                                uint32_t *__restrict__ cand,
                                                                uint32_t cap,
                                                                uint32_t *__restrict__ count) {
    const int H = p.H, W = p.W, k = p.k, a = p.a;
    const size_t HW = static_cast<size_t>(H) * W;
    const int fr = blockIdx.z;
    const float *P = proc + static_cast<size_t>(fr) * HW;
    const int gx = blockIdx.x * LM_TW + (threadIdx.x % LM_TW);
    const int gy = blockIdx.y * 4 + (threadIdx.x / LM_TW);
    bool flag = false;
    if (gy < H && gx < W && gy >= p.margin && gy < H - p.margin && gx >= p.margin &&
        gx < W - p.margin) {
        const float c = P[static_cast<size_t>(gy) * W + gx];
        if (!(c < p.max_thr)) {
            flag = true;
            for (int i = 0; i < k && flag; ++i) {
                const int yy = gy - a + i;
                if (yy < 0 || yy >= H) continue;
                for (int j = 0; j < k; ++j) {
                    const int xx = gx - a + j;
                    if (xx >= 0 && xx < W && P[static_cast<size_t>(yy) * W + xx] > c) {
                        flag = false;
                        break;
                    }
                }
            }
        }
    }
    const uint32_t pos = wave_append(flag, &count[fr * GT_COUNTER_STRIDE]);
    if (flag && pos < cap)
        cand[static_cast<size_t>(fr) * cap + pos] = static_cast<uint32_t>(gy) * static_cast<uint32_t>(W) + gx;
}

// ---------------------------------------------------------------------------
// Walk + characterize.
struct WalkRes {
    float x, y, mass, max, count;
};

// Where the processed samples of a disk come from: the frame in global
// memory, or the block's LDS copy of a region around its tile.  Both give
// the same values; only the load differs.
struct GlobalSrc {
    const float *p;
    int W;
    __device__ __forceinline__ float at(int u, int v) const {
        return p[static_cast<size_t>(u) * W + v];
    }
};
struct LdsSrc {
    const float *p;  // LDS
    int oy, ox, pitch;
    __device__ __forceinline__ float at(int u, int v) const {
        return p[(u - oy) * pitch + (v - ox)];
    }
};

// get_center (walk.wgsl:67-102): centroid offset, mass, count and maximum of
// the processed frame over the disk x^2 + y^2 <= r2 around (u, v).
template <class S>
__device__ __forceinline__ void disk_accumulate(WalkRes &res, const S &src, int H, int W, int u,
                                                int v, int i, int j, float r2) {
    const float x = static_cast<float>(i), x2 = x * x;
    const float y = static_cast<float>(j), y2 = y * y;
    if (x2 + y2 > r2) return;
    const int pu = u + i, pv = v + j;
    if (pu < 0 || pu >= H || pv < 0 || pv >= W) return;
    const float d = src.at(pu, pv);
    res.x = fmaf(d, x, res.x);
    res.y = fmaf(d, y, res.y);
    res.mass = res.mass + d;
    res.count = res.count + 1.0f;
    res.max = maxf_ref(res.max, d);
}

template <int RI, class S>
__device__ __forceinline__ WalkRes get_center(const S &src, int H, int W, int u, int v,
                                              int rint_rt, float r2_rt) {
    WalkRes res = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    if constexpr (RI > 0) {
        constexpr float r2 = static_cast<float>(RI) * static_cast<float>(RI);
#pragma unroll
        for (int i = -RI; i <= RI; ++i)
#pragma unroll
            for (int j = -RI; j <= RI; ++j) disk_accumulate(res, src, H, W, u, v, i, j, r2);
    } else {
        for (int i = -rint_rt; i <= rint_rt; ++i)
            for (int j = -rint_rt; j <= rint_rt; ++j)
                disk_accumulate(res, src, H, W, u, v, i, j, r2_rt);
    }
    res.x = res.x / res.mass;
    res.y = res.y / res.mass;
    return res;
}

// One walk step's move rule (walk.wgsl:141-158); returns whether it moved.
__device__ __forceinline__ bool walk_move(const WalkRes &res, int &pu, int &pv, int H, int W,
                                          int margin, float thr) {
    bool changed = false;
    if (res.x > thr && pu < H - 1 - margin) { pu += 1; changed = true; }
    else if (res.x < -thr && pu > margin) { pu -= 1; changed = true; }
    if (res.y > thr && pv < W - 1 - margin) { pv += 1; changed = true; }
    else if (res.y < -thr && pv > margin) { pv -= 1; changed = true; }
    return changed;
}

// characterize (characterize.wgsl:61-143) at round-half-even of the walked
// centre, with the Rg and eccentricity moment sums accumulated in f64.
template <typename T, class S>
__device__ void characterize_row(const S &src, const T *__restrict__ Rw, int H, int W, int circle,
                                 gt_result_row &row) {
    float r = row.r;
    int rint;
    if (r <= 0.0f) {
        rint = (circle - 1) / 2;
        r = static_cast<float>(rint);
    } else {
        rint = f2i(ceilf(r));
    }
    const float r2 = r * r;
    const int u = f2i(rintf(row.x)), v = f2i(rintf(row.y));
    float raw_mass = 0.0f, signal = 0.0f;
    double rg = 0.0, es = 0.0, ec = 0.0;
    for (int i = -rint; i <= rint; ++i) {
        const float x = static_cast<float>(i), x2 = x * x;
        for (int j = -rint; j <= rint; ++j) {
            const float y = static_cast<float>(j), y2 = y * y;
            const float mask = x2 + y2;
            if (mask > r2) continue;
            const int pu = u + i, pv = v + j;
            if (pu < 0 || pu >= H || pv < 0 || pv >= W) continue;
            const float d = src.at(pu, pv);
            raw_mass = raw_mass + to_f32(Rw[static_cast<size_t>(pu) * W + pv]);
            signal = maxf_ref(signal, d);
            if (i != 0 || j != 0) {
                // sin(2 atan2(j, i)) and cos(2 atan2(j, i)) in closed form
                const double q = static_cast<double>(i * i + j * j);
                const double s2 = 2.0 * static_cast<double>(i) * static_cast<double>(j) / q;
                const double c2 = (static_cast<double>(i * i) - static_cast<double>(j * j)) / q;
                es = fma(static_cast<double>(d), s2, es);
                ec = fma(static_cast<double>(d), c2, ec);
            } else {
                ec = ec + static_cast<double>(d);
            }
            rg = fma(static_cast<double>(d), static_cast<double>(mask), rg);
        }
    }
    // the reference reads the centre sample without a bounds check
    const int cu = min(max(u, 0), H - 1), cv = min(max(v, 0), W - 1);
    const float den = row.mass - src.at(cu, cv);
    row.ecc = static_cast<float>(sqrt(es * es + ec * ec) / static_cast<double>(den));
    row.Rg = static_cast<float>(sqrt(rg / static_cast<double>(row.mass)));
    row.raw_mass = raw_mass;
    row.signal = signal;
}

// minmass rule (walk.wgsl:200-213)
__device__ __forceinline__ bool passes_minmass(const GtKParams &p, const WalkRes &res) {
    float mm;
    if (p.minmass > 0.0f) mm = p.minmass;
    else mm = p.minmass_snr * res.count * p.image_std * p.rough;
    return res.mass >= mm;
}

__device__ __forceinline__ void fill_row(gt_detection &d, const WalkRes &res, uint32_t seed,
                                         uint32_t frame) {
    d.row.x = res.x;
    d.row.y = res.y;
    d.row.mass = res.mass;
    d.row.r = 0.0f;
    d.row.max_intensity = res.max;
    d.row.Rg = 0.0f;
    d.row.raw_mass = 0.0f;
    d.row.signal = 0.0f;
    d.row.ecc = 0.0f;
    d.row.count = res.count;
    d.seed = seed;
    d.frame = frame;
}

// Walk from global memory, one lane per candidate (generic path, used when
// the dilation window is too large for the fused tile kernel).
template <typename T, int RI>
__global__ __launch_bounds__(WK_THREADS) void k_walk(const float *__restrict__ proc,
                                                     const T *__restrict__ raw, GtKParams p,
                                                     const uint32_t *__restrict__ cand,
                                                     uint32_t cap,
                                                     const uint32_t *__restrict__ ccount,
                                                     gt_detection *__restrict__ out,
                                                     uint32_t out_cap,
                                                     uint32_t *__restrict__ ocount,
                                                     uint32_t frame0) {
    const int fr = blockIdx.y;
    const int H = p.H, W = p.W;
    const size_t HW = static_cast<size_t>(H) * W;
    const GlobalSrc src{proc + static_cast<size_t>(fr) * HW, W};
    const T *Rw = raw + static_cast<size_t>(fr) * HW;
    const uint32_t n = min(ccount[fr], cap);
    const uint32_t *fc = cand + static_cast<size_t>(fr) * cap;
    gt_detection *fout = out + static_cast<size_t>(fr) * out_cap;
    const int margin = RI > 0 ? RI : p.rint;
    const uint32_t lane = __lane_id();
    const uint32_t stride = gridDim.x * WK_THREADS;
    // i0 is the first candidate of this wavefront: the loop condition is
    // uniform across the wave, so wave_append sees every lane.
    for (uint32_t i0 = blockIdx.x * WK_THREADS + (threadIdx.x & ~63u); i0 < n; i0 += stride) {
        const uint32_t i = i0 + lane;
        bool keep = false;
        gt_detection d;
        if (i < n) {
            const uint32_t seed = fc[i];
            int pu = static_cast<int>(seed / static_cast<uint32_t>(W));
            int pv = static_cast<int>(seed - static_cast<uint32_t>(pu) * static_cast<uint32_t>(W));
            WalkRes res = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
            for (uint32_t it = 0; it < p.max_iter; ++it) {
                res = get_center<RI>(src, H, W, pu, pv, p.rint, p.r2);
                if (!walk_move(res, pu, pv, H, W, margin, p.shift_thr)) break;
            }
            res.x = res.x + static_cast<float>(pu);
            res.y = res.y + static_cast<float>(pv);
            keep = passes_minmass(p, res);
            if (keep) {
                fill_row(d, res, seed, frame0 + fr);
                if (p.characterize) characterize_row<T>(src, Rw, H, W, p.circle, d.row);
            }
        }
        const uint32_t pos = wave_append(keep, &ocount[fr * GT_COUNTER_STRIDE]);
        if (keep && pos < out_cap) fout[pos] = d;
    }
}

// ---------------------------------------------------------------------------
// Fused local max + walk + characterize.  One block per FT_TH x FT_TW tile:
//   1. the processed frame over the tile plus an h-pixel halo is staged in LDS
//      (out-of-frame entries -inf);
//   2. row maxima over the k-wide dilation window, then the column test, give
//      the tile's candidates, compacted into an LDS list in raster order with
//      one ballot per wave and a block prefix (no atomics);
//   3. one lane per candidate walks; every disk evaluation that lies inside
//      the staged region reads LDS, the rest read the frame in global memory
//      (same arithmetic, so results do not depend on the tile);
//   4. detections are characterized and appended to the frame's row list with
//      one atomic per block.
constexpr int FT_TW = 64, FT_TH = 32, FT_THREADS = 256;
constexpr int FT_WAVES = FT_THREADS / 64;
constexpr int FT_ITERS = FT_TH * FT_TW / FT_THREADS;

template <typename T, int KT, int RI>
__global__ __launch_bounds__(FT_THREADS) void k_detect_tile(const float *__restrict__ proc,
                                                            const T *__restrict__ raw,
                                                            GtKParams p, int h,
                                                            gt_detection *__restrict__ out,
                                                            uint32_t out_cap,
                                                            uint32_t *__restrict__ ocount,
                                                            uint32_t frame0) {
    extern __shared__ float lds[];
    __shared__ uint32_t s_wave_cnt[FT_ITERS * FT_WAVES];
    __shared__ uint32_t s_wave_off[FT_ITERS * FT_WAVES + 1];
    __shared__ uint16_t s_cand[FT_TH * FT_TW];
    __shared__ uint32_t s_keep_cnt[FT_WAVES];
    __shared__ uint32_t s_base;
    const int k = KT > 0 ? KT : p.k;
    const int a = p.a;
    const int rint = RI > 0 ? RI : p.rint;
    const int RW = FT_TW + 2 * h, RH = FT_TH + 2 * h;
    const int MR = FT_TH + k - 1;  // rows of row-maxima needed by the tile
    float *Sp = lds;               // RH x RW
    float *Sm = Sp + RH * RW;      // MR x FT_TW
    const int H = p.H, W = p.W;
    const size_t HW = static_cast<size_t>(H) * W;
    const int x0 = blockIdx.x * FT_TW, y0 = blockIdx.y * FT_TH;
    const int fr = blockIdx.z;
    const float *P = proc + static_cast<size_t>(fr) * HW;
    const T *Rw = raw + static_cast<size_t>(fr) * HW;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int oy = y0 - h, ox = x0 - h;

    for (int idx = tid; idx < RH * RW; idx += FT_THREADS) {
        const int ry = idx / RW, rx = idx - ry * RW;
        const int gy = oy + ry, gx = ox + rx;
        Sp[idx] = (gy >= 0 && gy < H && gx >= 0 && gx < W) ? P[static_cast<size_t>(gy) * W + gx]
                                                           : -INFINITY;
    }
    __syncthreads();
    // row maxima: Sm[m][cx] = max over j < k of Sp[h - a + m][h + cx - a + j]
    for (int idx = tid; idx < MR * FT_TW; idx += FT_THREADS) {
        const int m = idx / FT_TW, cx = idx - m * FT_TW;
        const float *s = Sp + (h - a + m) * RW + (h + cx - a);
        float mx = s[a];
        if (KT > 0) {
#pragma unroll
            for (int j = 0; j < KT; ++j) mx = maxf_ref(mx, s[j]);
        } else {
            for (int j = 0; j < k; ++j) mx = maxf_ref(mx, s[j]);
        }
        Sm[idx] = mx;
    }
    __syncthreads();
    // candidate test; pixel idx = it * 256 + tid is row it*4 + wave, column lane
    uint64_t masks[FT_ITERS];
    const int margin = p.margin;
    const float thr = p.max_thr;
#pragma unroll
    for (int it = 0; it < FT_ITERS; ++it) {
        const int ty = it * FT_WAVES + wave, cx = lane;
        const int gy = y0 + ty, gx = x0 + cx;
        bool flag = false;
        if (gy < H && gx < W && gy >= margin && gy < H - margin && gx >= margin &&
            gx < W - margin) {
            const float c = Sp[(h + ty) * RW + h + cx];
            if (!(c < thr)) {
                flag = true;
                const float *col = Sm + ty * FT_TW + cx;
                if (KT > 0) {
#pragma unroll
                    for (int i = 0; i < KT; ++i) flag = flag && !(col[i * FT_TW] > c);
                } else {
                    for (int i = 0; i < k; ++i) flag = flag && !(col[i * FT_TW] > c);
                }
            }
        }
        masks[it] = __ballot(flag);
        if (lane == 0) s_wave_cnt[it * FT_WAVES + wave] = __popcll(masks[it]);
    }
    __syncthreads();
    if (tid == 0) {
        uint32_t acc = 0;
        for (int q = 0; q < FT_ITERS * FT_WAVES; ++q) {
            s_wave_off[q] = acc;
            acc += s_wave_cnt[q];
        }
        s_wave_off[FT_ITERS * FT_WAVES] = acc;
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < FT_ITERS; ++it) {
        const uint64_t m = masks[it];
        if ((m >> lane) & 1ull) {
            const uint32_t pos = s_wave_off[it * FT_WAVES + wave] +
                                 __popcll(m & ((1ull << lane) - 1ull));
            s_cand[pos] = static_cast<uint16_t>((it * FT_WAVES + wave) * FT_TW + lane);
        }
    }
    __syncthreads();
    const uint32_t ncand = s_wave_off[FT_ITERS * FT_WAVES];
    const GlobalSrc gsrc{P, W};
    const LdsSrc lsrc{Sp, oy, ox, RW};
    // a disk around (u, v) lies in the staged region when all of it does
    const int ylo = oy + rint, yhi = oy + RH - 1 - rint;
    const int xlo = ox + rint, xhi = ox + RW - 1 - rint;
    const uint32_t frame = frame0 + fr;
    uint32_t *counter = &ocount[fr * GT_COUNTER_STRIDE];
    gt_detection *fout = out + static_cast<size_t>(fr) * out_cap;
    for (uint32_t c0 = 0; c0 < ncand; c0 += FT_THREADS) {
        const uint32_t ci = c0 + tid;
        bool keep = false;
        gt_detection d;
        if (ci < ncand) {
            const int li = s_cand[ci];
            const int ty = li / FT_TW, cx = li - ty * FT_TW;
            int pu = y0 + ty, pv = x0 + cx;
            const uint32_t seed = static_cast<uint32_t>(pu) * static_cast<uint32_t>(W) + pv;
            WalkRes res = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
            for (uint32_t it = 0; it < p.max_iter; ++it) {
                if (pu >= ylo && pu <= yhi && pv >= xlo && pv <= xhi)
                    res = get_center<RI>(lsrc, H, W, pu, pv, p.rint, p.r2);
                else
                    res = get_center<RI>(gsrc, H, W, pu, pv, p.rint, p.r2);
                if (!walk_move(res, pu, pv, H, W, rint, p.shift_thr)) break;
            }
            res.x = res.x + static_cast<float>(pu);
            res.y = res.y + static_cast<float>(pv);
            keep = passes_minmass(p, res);
            if (keep) {
                fill_row(d, res, seed, frame);
                if (p.characterize) {
                    const int cu = f2i(rintf(d.row.x)), cv = f2i(rintf(d.row.y));
                    if (cu >= ylo && cu <= yhi && cv >= xlo && cv <= xhi)
                        characterize_row<T>(lsrc, Rw, H, W, p.circle, d.row);
                    else
                        characterize_row<T>(gsrc, Rw, H, W, p.circle, d.row);
                }
            }
        }
        const uint64_t km = __ballot(keep);
        if (lane == 0) s_keep_cnt[wave] = __popcll(km);
        __syncthreads();
        if (tid == 0) {
            uint32_t tot = 0;
            for (int q = 0; q < FT_WAVES; ++q) tot += s_keep_cnt[q];
            s_base = tot ? atomicAdd(counter, tot) : 0u;
        }
        __syncthreads();
        if (keep) {
            uint32_t pos = s_base + __popcll(km & ((1ull << lane) - 1ull));
            for (int q = 0; q < wave; ++q) pos += s_keep_cnt[q];
            if (pos < out_cap) fout[pos] = d;
        }
        __syncthreads();
    }
}

// characterize_points: rows hold caller positions (x = row, y = column, r);
// mass is recomputed over the disk at the rounded position (get_mass,
// characterize.wgsl:30-58,75) and the characterize columns are filled in place.
template <typename T>
__global__ __launch_bounds__(WK_THREADS) void k_characterize_rows(const float *__restrict__ proc,
                                                                  const T *__restrict__ raw,
                                                                  GtKParams p,
                                                                  gt_detection *__restrict__ rows,
                                                                  uint32_t cap,
                                                                  const uint32_t *__restrict__ counts) {
    const int fr = blockIdx.y;
    const int H = p.H, W = p.W;
    const size_t HW = static_cast<size_t>(H) * W;
    const float *P = proc + static_cast<size_t>(fr) * HW;
    const GlobalSrc gsrc{P, W};
    const T *Rw = raw + static_cast<size_t>(fr) * HW;
    const uint32_t n = min(counts[fr], cap);
    gt_detection *fr_rows = rows + static_cast<size_t>(fr) * cap;
    for (uint32_t i = blockIdx.x * WK_THREADS + threadIdx.x; i < n; i += gridDim.x * WK_THREADS) {
        gt_result_row row = fr_rows[i].row;
        float r = row.r;
        int rint;
        if (r <= 0.0f) {
            rint = (p.circle - 1) / 2;
            r = static_cast<float>(rint);
        } else {
            rint = f2i(ceilf(r));
        }
        const float r2 = r * r;
        const int u = f2i(rintf(row.x)), v = f2i(rintf(row.y));
        float m = 0.0f;
        for (int a = -rint; a <= rint; ++a) {
            const float x = static_cast<float>(a), x2 = x * x;
            for (int b = -rint; b <= rint; ++b) {
                const float y = static_cast<float>(b), y2 = y * y;
                if (x2 + y2 > r2) continue;
                const int pu = u + a, pv = v + b;
                if (pu < 0 || pu >= H || pv < 0 || pv >= W) continue;
                m = m + P[static_cast<size_t>(pu) * W + pv];
            }
        }
        row.mass = m;
        characterize_row<T>(gsrc, Rw, H, W, p.circle, row);
        fr_rows[i].row = row;
    }
}

__global__ void k_zero_u32(uint32_t *p, size_t n) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

template <typename T>
hipError_t launch_bandpass_t(const T *frames, float *proc, uint32_t nframes, const GtKParams &p,
                             hipStream_t s) {
    const dim3 grid((p.W + BP_TW - 1) / BP_TW, (p.H + BP_TH - 1) / BP_TH, nframes);
    const int RH = BP_TH + 2 * p.R, RW = BP_TW + 2 * p.R;
    const size_t lds = sizeof(float) * (static_cast<size_t>(RH) * RW + 2u * RH * BP_TW);
    switch (p.R) {
    case 1: k_bandpass<T, 1><<<grid, BP_THREADS, lds, s>>>(frames, proc, p); break;
    case 2: k_bandpass<T, 2><<<grid, BP_THREADS, lds, s>>>(frames, proc, p); break;
    case 3: k_bandpass<T, 3><<<grid, BP_THREADS, lds, s>>>(frames, proc, p); break;
    case 4: k_bandpass<T, 4><<<grid, BP_THREADS, lds, s>>>(frames, proc, p); break;
    case 5: k_bandpass<T, 5><<<grid, BP_THREADS, lds, s>>>(frames, proc, p); break;
    case 6: k_bandpass<T, 6><<<grid, BP_THREADS, lds, s>>>(frames, proc, p); break;
    case 7: k_bandpass<T, 7><<<grid, BP_THREADS, lds, s>>>(frames, proc, p); break;
    default: k_bandpass<T, 0><<<grid, BP_THREADS, lds, s>>>(frames, proc, p); break;
    }
    return hipGetLastError();
}

template <typename T>
hipError_t launch_walk_t(const float *proc, const T *raw, uint32_t nframes, const GtKParams &p,
                         const uint32_t *cand, uint32_t cap, const uint32_t *ccount,
                         gt_detection *out, uint32_t out_cap, uint32_t *ocount, uint32_t frame0,
                         hipStream_t s) {
    // Enough waves per frame to cover the usual candidate density (about one
    // candidate per 80 pixels) in one sweep; the loop strides over the rest.
    const size_t hw = static_cast<size_t>(p.H) * p.W;
    uint32_t blocks = static_cast<uint32_t>((hw / 80 + WK_THREADS - 1) / WK_THREADS);
    blocks = blocks < 1 ? 1 : (blocks > 1024 ? 1024 : blocks);
    const dim3 grid(blocks, nframes, 1);
#define GT_WALK_CASE(RIV)                                                                        \
    k_walk<T, RIV><<<grid, WK_THREADS, 0, s>>>(proc, raw, p, cand, cap, ccount, out, out_cap, \
                                               ocount, frame0)
    switch (p.rint) {
    case 1: GT_WALK_CASE(1); break;
    case 2: GT_WALK_CASE(2); break;
    case 3: GT_WALK_CASE(3); break;
    case 4: GT_WALK_CASE(4); break;
    case 5: GT_WALK_CASE(5); break;
    case 6: GT_WALK_CASE(6); break;
    case 7: GT_WALK_CASE(7); break;
    default: GT_WALK_CASE(0); break;
    }
#undef GT_WALK_CASE
    return hipGetLastError();
}

}  // namespace

hipError_t gt_launch_zero_u32(uint32_t *p, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const unsigned blocks = static_cast<unsigned>((n + 255) / 256);
    k_zero_u32<<<blocks, 256, 0, s>>>(p, n);
    return hipGetLastError();
}

hipError_t gt_launch_bandpass(const void *frames, int dtype, float *proc, uint32_t nframes,
                              const GtKParams &p, hipStream_t s) {
    switch (dtype) {
    case GT_U16: return launch_bandpass_t(static_cast<const uint16_t *>(frames), proc, nframes, p, s);
    case GT_F32: return launch_bandpass_t(static_cast<const float *>(frames), proc, nframes, p, s);
    case GT_U8: return launch_bandpass_t(static_cast<const uint8_t *>(frames), proc, nframes, p, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t gt_launch_localmax(const float *proc, uint32_t nframes, const GtKParams &p,
                              uint32_t *cand, uint32_t cand_cap, uint32_t *cand_count,
                              hipStream_t s) {
    const dim3 grid((p.W + LM_TW - 1) / LM_TW, (p.H + 3) / 4, nframes);
    k_localmax_direct<<<grid, LM_THREADS, 0, s>>>(proc, p, cand, cand_cap, cand_count);
    return hipGetLastError();
}

int gt_tile_halo(const GtKParams &p) {
    // the staged region must hold the dilation window of every tile pixel and
    // the walk disk of a candidate that moved up to two pixels off the tile
    int h = p.a;
    h = std::max(h, p.k - 1 - p.a);
    h = std::max(h, p.rint + 2);
    return h;
}

size_t gt_tile_lds_bytes(const GtKParams &p) {
    const int h = gt_tile_halo(p);
    const size_t dyn = sizeof(float) * (static_cast<size_t>(FT_TH + 2 * h) * (FT_TW + 2 * h) +
                                        static_cast<size_t>(FT_TH + p.k - 1) * FT_TW);
    const size_t stat = sizeof(uint16_t) * FT_TH * FT_TW + 1024;
    return dyn + stat;
}

namespace {
template <typename T>
hipError_t launch_detect_tile_t(const float *proc, const T *raw, uint32_t nframes,
                                const GtKParams &p, gt_detection *out, uint32_t out_cap,
                                uint32_t *ocount, uint32_t frame0, hipStream_t s) {
    const int h = gt_tile_halo(p);
    const dim3 grid((p.W + FT_TW - 1) / FT_TW, (p.H + FT_TH - 1) / FT_TH, nframes);
    const size_t lds = sizeof(float) * (static_cast<size_t>(FT_TH + 2 * h) * (FT_TW + 2 * h) +
                                        static_cast<size_t>(FT_TH + p.k - 1) * FT_TW);
#define GT_TILE(KV, RV)                                                                      \
    k_detect_tile<T, KV, RV><<<grid, FT_THREADS, lds, s>>>(proc, raw, p, h, out, out_cap, \
                                                            ocount, frame0)
    if (p.rint == 4) {
        switch (p.k) {
        case 9: GT_TILE(9, 4); break;
        case 12: GT_TILE(12, 4); break;
        default: GT_TILE(0, 4); break;
        }
    } else {
        switch (p.k) {
        case 9: GT_TILE(9, 0); break;
        case 12: GT_TILE(12, 0); break;
        default: GT_TILE(0, 0); break;
        }
    }
#undef GT_TILE
    return hipGetLastError();
}

__global__ void k_copy_counts(const uint32_t *__restrict__ strided, uint32_t *__restrict__ dst,
                              uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = strided[i * GT_COUNTER_STRIDE];
}
}  // namespace

hipError_t gt_launch_detect_tile(const float *proc, const void *raw, int dtype, uint32_t nframes,
                                 const GtKParams &p, gt_detection *out, uint32_t out_cap,
                                 uint32_t *ocount, uint32_t frame0, hipStream_t s) {
    switch (dtype) {
    case GT_U16:
        return launch_detect_tile_t(proc, static_cast<const uint16_t *>(raw), nframes, p, out,
                                    out_cap, ocount, frame0, s);
    case GT_F32:
        return launch_detect_tile_t(proc, static_cast<const float *>(raw), nframes, p, out,
                                    out_cap, ocount, frame0, s);
    case GT_U8:
        return launch_detect_tile_t(proc, static_cast<const uint8_t *>(raw), nframes, p, out,
                                    out_cap, ocount, frame0, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t gt_launch_copy_counts(const uint32_t *strided, uint32_t *dst, uint32_t n,
                                 hipStream_t s) {
    if (n == 0) return hipSuccess;
    k_copy_counts<<<(n + 255) / 256, 256, 0, s>>>(strided, dst, n);
    return hipGetLastError();
}

hipError_t gt_launch_walk(const float *proc, const void *raw, int dtype, uint32_t nframes,
                          const GtKParams &p, const uint32_t *cand, uint32_t cand_cap,
                          const uint32_t *cand_count, gt_detection *out, uint32_t out_cap,
                          uint32_t *out_count, uint32_t frame0, hipStream_t s) {
    switch (dtype) {
    case GT_U16:
        return launch_walk_t(proc, static_cast<const uint16_t *>(raw), nframes, p, cand, cand_cap,
                             cand_count, out, out_cap, out_count, frame0, s);
    case GT_F32:
        return launch_walk_t(proc, static_cast<const float *>(raw), nframes, p, cand, cand_cap,
                             cand_count, out, out_cap, out_count, frame0, s);
    case GT_U8:
        return launch_walk_t(proc, static_cast<const uint8_t *>(raw), nframes, p, cand, cand_cap,
                             cand_count, out, out_cap, out_count, frame0, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t gt_launch_characterize_rows(const float *proc, const void *raw, int dtype,
                                       uint32_t nframes, const GtKParams &p, gt_detection *rows,
                                       uint32_t cap, const uint32_t *counts, uint32_t max_count,
                                       hipStream_t s) {
    uint32_t blocks = (max_count + WK_THREADS - 1) / WK_THREADS;
    blocks = blocks < 1 ? 1 : (blocks > 1024 ? 1024 : blocks);
    const dim3 grid(blocks, nframes, 1);
    switch (dtype) {
    case GT_U16:
        k_characterize_rows<uint16_t><<<grid, WK_THREADS, 0, s>>>(
            proc, static_cast<const uint16_t *>(raw), p, rows, cap, counts);
        break;
    case GT_F32:
        k_characterize_rows<float><<<grid, WK_THREADS, 0, s>>>(
            proc, static_cast<const float *>(raw), p, rows, cap, counts);
        break;
    case GT_U8:
        k_characterize_rows<uint8_t><<<grid, WK_THREADS, 0, s>>>(
            proc, static_cast<const uint8_t *>(raw), p, rows, cap, counts);
        break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
