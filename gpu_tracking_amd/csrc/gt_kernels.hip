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

#include "gt_internal.h"

namespace {

// WGSL `1./255.` (preprocess_cols.wgsl:54): f32 bit pattern 0x3B808081.
constexpr float kTrunc = 0.003921568859368562698364257812f;

constexpr int BP_TW = 64, BP_TH = 32, BP_THREADS = 256;
constexpr int LM_TW = 64, LM_TH = 16, LM_THREADS = 256;
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
// below the snr threshold.  Out-of-frame window entries are -inf.
template <int KT>
__global__ __launch_bounds__(LM_THREADS) void k_localmax(const float *__restrict__ proc,
                                                         GtKParams p, uint32_t *__restrict__ cand,
                                                         uint32_t cap,
                                                         uint32_t *__restrict__ count) {
    extern __shared__ float lds[];
    const int k = KT > 0 ? KT : p.k;
    const int a = p.a;
    const int RW = LM_TW + k - 1, RH = LM_TH + k - 1;
    float *Sp = lds;
    float *Sm = Sp + RH * RW;
    const int H = p.H, W = p.W;
    const size_t HW = static_cast<size_t>(H) * W;
    const int x0 = blockIdx.x * LM_TW, y0 = blockIdx.y * LM_TH;
    const int fr = blockIdx.z;
    const float *P = proc + static_cast<size_t>(fr) * HW;
    const int tid = threadIdx.x;
    for (int idx = tid; idx < RH * RW; idx += LM_THREADS) {
        const int ry = idx / RW, rx = idx - ry * RW;
        const int gy = y0 - a + ry, gx = x0 - a + rx;
        Sp[idx] = (gy >= 0 && gy < H && gx >= 0 && gx < W) ? P[static_cast<size_t>(gy) * W + gx]
                                                           : -INFINITY;
    }
    __syncthreads();
    // Row maxima over the k columns of the window.
    for (int idx = tid; idx < RH * LM_TW; idx += LM_THREADS) {
        const int ry = idx / LM_TW, cx = idx - ry * LM_TW;
        const float *s = Sp + ry * RW + cx;
        float m = s[a];
        if (KT > 0) {
#pragma unroll
            for (int j = 0; j < KT; ++j) m = maxf_ref(m, s[j]);
        } else {
            for (int j = 0; j < k; ++j) m = maxf_ref(m, s[j]);
        }
        Sm[idx] = m;
    }
    __syncthreads();
    const int margin = p.margin;
    const float thr = p.max_thr;
    uint32_t *fcand = cand + static_cast<size_t>(fr) * cap;
    for (int base = 0; base < LM_TH * LM_TW; base += LM_THREADS) {
        const int idx = base + tid;
        const int ty = idx / LM_TW, cx = idx - ty * LM_TW;
        const int gy = y0 + ty, gx = x0 + cx;
        bool flag = false;
        if (gy < H && gx < W && gy >= margin && gy < H - margin && gx >= margin &&
            gx < W - margin) {
            const float c = Sp[(ty + a) * RW + cx + a];
            if (!(c < thr)) {
                flag = true;
                const float *col = Sm + ty * LM_TW + cx;
                for (int i = 0; i < k; ++i)
                    if (col[i * LM_TW] > c) { flag = false; break; }
            }
        }
        const uint32_t pos = wave_append(flag, &count[fr]);
        if (flag && pos < cap) fcand[pos] = static_cast<uint32_t>(gy) * static_cast<uint32_t>(W) + gx;
    }
}

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
    const uint32_t pos = wave_append(flag, &count[fr]);
    if (flag && pos < cap)
        cand[static_cast<size_t>(fr) * cap + pos] = static_cast<uint32_t>(gy) * static_cast<uint32_t>(W) + gx;
}

// ---------------------------------------------------------------------------
// Walk + characterize.
struct WalkRes {
    float x, y, mass, max, count;
};

// get_center (walk.wgsl:67-102): centroid offset, mass, count and maximum of
// the processed frame over the disk x^2 + y^2 <= r2 around (u, v).
__device__ __forceinline__ void disk_accumulate(WalkRes &res, const float *__restrict__ P, int H,
                                                int W, int u, int v, int i, int j, float r2) {
    const float x = static_cast<float>(i), x2 = x * x;
    const float y = static_cast<float>(j), y2 = y * y;
    if (x2 + y2 > r2) return;
    const int pu = u + i, pv = v + j;
    if (pu < 0 || pu >= H || pv < 0 || pv >= W) return;
    const float d = P[static_cast<size_t>(pu) * W + pv];
    res.x = fmaf(d, x, res.x);
    res.y = fmaf(d, y, res.y);
    res.mass = res.mass + d;
    res.count = res.count + 1.0f;
    res.max = maxf_ref(res.max, d);
}

template <int RI>
__device__ __forceinline__ WalkRes get_center(const float *__restrict__ P, int H, int W, int u,
                                              int v, int rint_rt, float r2_rt) {
    WalkRes res = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    if constexpr (RI > 0) {
        constexpr float r2 = static_cast<float>(RI) * static_cast<float>(RI);
#pragma unroll
        for (int i = -RI; i <= RI; ++i)
#pragma unroll
            for (int j = -RI; j <= RI; ++j) disk_accumulate(res, P, H, W, u, v, i, j, r2);
    } else {
        for (int i = -rint_rt; i <= rint_rt; ++i)
            for (int j = -rint_rt; j <= rint_rt; ++j) disk_accumulate(res, P, H, W, u, v, i, j, r2_rt);
    }
    res.x = res.x / res.mass;
    res.y = res.y / res.mass;
    return res;
}

// characterize (characterize.wgsl:61-143) at round-half-even of the walked
// centre, with the Rg and eccentricity moment sums accumulated in f64.
template <typename T>
__device__ void characterize_row(const float *__restrict__ P, const T *__restrict__ Rw, int H,
                                 int W, int circle, gt_result_row &row) {
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
            const size_t idx = static_cast<size_t>(pu) * W + pv;
            const float d = P[idx];
            raw_mass = raw_mass + to_f32(Rw[idx]);
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
    const int cu = min(max(u, 0), H - 1), cv = min(max(v, 0), W - 1);
    const float den = row.mass - P[static_cast<size_t>(cu) * W + cv];
    row.ecc = static_cast<float>(sqrt(es * es + ec * ec) / static_cast<double>(den));
    row.Rg = static_cast<float>(sqrt(rg / static_cast<double>(row.mass)));
    row.raw_mass = raw_mass;
    row.signal = signal;
}

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
    const float *P = proc + static_cast<size_t>(fr) * HW;
    const T *Rw = raw + static_cast<size_t>(fr) * HW;
    const uint32_t n = min(ccount[fr], cap);
    const uint32_t *fc = cand + static_cast<size_t>(fr) * cap;
    gt_detection *fout = out + static_cast<size_t>(fr) * out_cap;
    const int rint = RI > 0 ? RI : p.rint;
    const int margin = rint;
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
                res = get_center<RI>(P, H, W, pu, pv, p.rint, p.r2);
                bool changed = false;
                if (res.x > p.shift_thr && pu < H - 1 - margin) { pu += 1; changed = true; }
                else if (res.x < -p.shift_thr && pu > margin) { pu -= 1; changed = true; }
                if (res.y > p.shift_thr && pv < W - 1 - margin) { pv += 1; changed = true; }
                else if (res.y < -p.shift_thr && pv > margin) { pv -= 1; changed = true; }
                if (!changed) break;
            }
            res.x = res.x + static_cast<float>(pu);
            res.y = res.y + static_cast<float>(pv);
            float mm;
            if (p.minmass > 0.0f) mm = p.minmass;
            else mm = p.minmass_snr * res.count * p.image_std * p.rough;
            keep = res.mass >= mm;
            if (keep) {
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
                d.frame = frame0 + fr;
                if (p.characterize) characterize_row<T>(P, Rw, H, W, p.circle, d.row);
            }
        }
        const uint32_t pos = wave_append(keep, &ocount[fr]);
        if (keep && pos < out_cap) fout[pos] = d;
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
    if (p.k > GT_MAX_K_TILED || p.k < 1) {
        const dim3 grid((p.W + LM_TW - 1) / LM_TW, (p.H + 3) / 4, nframes);
        k_localmax_direct<<<grid, LM_THREADS, 0, s>>>(proc, p, cand, cand_cap, cand_count);
        return hipGetLastError();
    }
    const dim3 grid((p.W + LM_TW - 1) / LM_TW, (p.H + LM_TH - 1) / LM_TH, nframes);
    const int RW = LM_TW + p.k - 1, RH = LM_TH + p.k - 1;
    const size_t lds = sizeof(float) * (static_cast<size_t>(RH) * RW + static_cast<size_t>(RH) * LM_TW);
    switch (p.k) {
    case 9: k_localmax<9><<<grid, LM_THREADS, lds, s>>>(proc, p, cand, cand_cap, cand_count); break;
    case 12: k_localmax<12><<<grid, LM_THREADS, lds, s>>>(proc, p, cand, cand_cap, cand_count); break;
    case 14: k_localmax<14><<<grid, LM_THREADS, lds, s>>>(proc, p, cand, cand_cap, cand_count); break;
    default: k_localmax<0><<<grid, LM_THREADS, lds, s>>>(proc, p, cand, cand_cap, cand_count); break;
    }
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
