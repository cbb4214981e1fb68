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
#include <type_traits>

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

// Workgroups are handed to the 8 XCDs (each with its own L2) round-robin in
// launch order.  Tile t of a launch of `ntiles` tiles is given to workgroup
// id w with t = (w % 8) * per + w / 8, per = ceil(ntiles / 8): each XCD works
// through one contiguous range of tiles, so a tile's halo lines are usually
// in the L2 its neighbour filled.  The launch has 8 * per workgroups; ids
// that map past the last tile return at once.
constexpr int kXcds = 8;
// n / d and n % d for n < 2^22 and d >= 1: a float reciprocal estimate (off
// by at most one for quotients below 2^22) and one correction.
__device__ __forceinline__ uint32_t udiv22(uint32_t n, uint32_t d, uint32_t &rem) {
    uint32_t q = static_cast<uint32_t>(static_cast<float>(n) *
                                       __builtin_amdgcn_rcpf(static_cast<float>(d)));
    int32_t r = static_cast<int32_t>(n - q * d);
    if (r < 0) { --q; r += static_cast<int32_t>(d); }
    else if (r >= static_cast<int32_t>(d)) { ++q; r -= static_cast<int32_t>(d); }
    rem = static_cast<uint32_t>(r);
    return q;
}
__device__ __forceinline__ int xcd_tile(int ntiles) {
    const int per = (ntiles + kXcds - 1) / kXcds;
    const int w = blockIdx.x;
    return (w % kXcds) * per + w / kXcds;
}
inline unsigned xcd_grid(size_t ntiles) {
    const size_t per = (ntiles + kXcds - 1) / kXcds;
    return static_cast<unsigned>(per * kXcds);
}

// Appends the lanes with `fa` set to list A (*ca) and those with `fb` set to
// list B (*cb) with one atomic instruction per wavefront (lane 0 adds to A,
// lane 1 to B): one memory round trip for both.  Must be reached by every
// lane of the wave.
__device__ __forceinline__ void wave_append2(bool fa, uint32_t *ca, bool fb, uint32_t *cb,
                                             uint32_t &pa, uint32_t &pb) {
    const unsigned long long ma = __ballot(fa), mb = __ballot(fb);
    const uint32_t lane = __lane_id();
    const unsigned long long below = (1ull << lane) - 1ull;
    uint32_t old = 0u;
    if ((ma | mb) != 0ull) {
        const uint32_t add = static_cast<uint32_t>(__popcll(lane == 0 ? ma : mb));
        uint32_t *ctr = lane == 0 ? ca : cb;
        if (lane < 2 && add != 0u) old = atomicAdd(ctr, add);
    }
    pa = __shfl(old, 0) + static_cast<uint32_t>(__popcll(ma & below));
    pb = __shfl(old, 1) + static_cast<uint32_t>(__popcll(mb & below));
}

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
// Bandpass for integer frames (u8/u16) and radius R <= 7, register blocked.
//
// Integer input makes two things exact that the generic kernel cannot assume:
//   * every boxcar partial sum is an integer below 2^24 ((2R+1)^2 * 65535 <
//     2^24 for R <= 7), so running sums give the sequential sum's bits;
//   * every dividend is 0 or large enough (min(w)^2 / 15 with min(w) >=
//     2^-46) that the reciprocal-and-one-correction quotient below keeps a
//     normal quotient and an exactly representable residual.  Every step then
//     scales exactly with the dividend's binade, so the sequence equals IEEE
//     division for all dividends iff it does for all 2^23 mantissas of one
//     binade; the host checks that for each context's two divisors before it
//     selects this kernel (fast_div), and the CPU test suite checks it for the
//     divisors of common parameter sets.
// The Gaussian taps stay sequential fmaf chains, as in the oracle.
constexpr int B2_TW = 64, B2_TH = 32, B2_THREADS = 256, B2_ROWS = 8;

// a / b correctly rounded, given rb = RN(1/b), for the dividends and
// divisors described above (one correction of the reciprocal quotient).
__device__ __forceinline__ float div_const(float a, float b, float rb) {
    const float q = a * rb;
    const float r = fmaf(-q, b, a);
    return fmaf(r, rb, q);
}
// a / b as RN(a * rh + RN(a * rl)), rh + rl = 1/b to about 48 bits: two
// operations.  The same binade argument applies (a * rl stays normal for
// these dividends), and the host selects it only for divisors whose whole
// binade it has checked against IEEE division.
__device__ __forceinline__ float div_split(float a, float rh, float rl) {
    return fmaf(a, rh, a * rl);
}
// Division by a bandpass norm: the two-operation form when kD2.
template <bool kD2>
__device__ __forceinline__ float div_norm(float a, float b, float rb, float rl) {
    if constexpr (kD2) return div_split(a, rb, rl);
    else return div_const(a, b, rb);
}

template <int R>
struct BpLds {
    static constexpr int RH = B2_TH + 2 * R;
    static constexpr int RWP = (B2_TW + 2 * R + 3) & ~3;   // staged row pitch, 16-byte rows
    // the staged samples, then (once every row pass has read them) the row
    // pass's Gaussian and boxcar sums
    union {
        float Sraw[RH * RWP];   // the frame's samples (u8 / u16), converted when staged
        struct {
            float Sg[RH * B2_TW];
            float Sc[RH * B2_TW];
        };
    };
};

// One 64 x 32 output tile of one frame (f, out: the frame's input and
// processed planes, p.H x p.W).
template <typename T, int R, bool kD2>
__device__ __forceinline__ void bandpass_int_tile(const T *__restrict__ f, float *__restrict__ out,
                                                  const GtKParams &p, int H, int W, int x0,
                                                  int y0, BpLds<R> &L) {
    constexpr int NT = 2 * R + 1;
    constexpr int RH = BpLds<R>::RH;
    constexpr int RWP = BpLds<R>::RWP;
    constexpr int NV = (NT + 3 + 3) / 4;            // float4 reads per row-pass item
    static_assert(B2_TW - 4 + 4 * NV <= RWP, "row-pass reads stay inside the staged row");
    static_assert(B2_TH % B2_ROWS == 0 && (B2_TH / B2_ROWS) * B2_TW == B2_THREADS, "mapping");
    float *Sraw = L.Sraw;
    float *Sg = L.Sg, *Sc = L.Sc;
    const int tid = threadIdx.x;
    constexpr int PAIRS = (B2_TW + 2 * R + 1) / 2;  // u16 pairs per staged row
    static_assert(2 * PAIRS <= RWP, "pairs fit the staged row");
    constexpr int QR = (B2_TW + 2 * R) / 4;         // u16 quads per staged row (R % 4 == 0)
    constexpr int QROWS = B2_THREADS / QR;          // staged rows per pass of quad loads
    constexpr int QPASS = (RH + QROWS - 1) / QROWS;
    if (sizeof(T) == 2 && (R % 4) == 0 && (W % 4) == 0 &&
        (reinterpret_cast<uintptr_t>(f) & 7u) == 0 && y0 - R >= 0 && y0 + B2_TH + R <= H &&
        x0 - R >= 0 && x0 + B2_TW + R <= W) {
        // interior tile of a u16 frame whose staged region starts on a
        // multiple of 4 columns: each thread loads 8-byte quads at one
        // 32-bit offset from a per-pass tile row pointer
        static_assert(R % 4 != 0 || 4 * QR == RWP, "quads fill the staged row");
        const char *tb = reinterpret_cast<const char *>(f + (static_cast<size_t>(y0 - R) * W +
                                                             static_cast<size_t>(x0 - R)));
        const int rq = tid / QR, cq = tid - rq * QR;
        const uint32_t voff = static_cast<uint32_t>(rq) * (2u * static_cast<uint32_t>(W)) +
                              8u * static_cast<uint32_t>(cq);
        const bool active = tid < QROWS * QR;
        uint2 vals[QPASS];
#pragma unroll
        for (int q = 0; q < QPASS; ++q) {
            const char *tq = tb + static_cast<size_t>(q * QROWS) * (2u * static_cast<uint32_t>(W));
            if (active && rq + q * QROWS < RH)
                vals[q] = *reinterpret_cast<const uint2 *>(tq + voff);
        }
#pragma unroll
        for (int q = 0; q < QPASS; ++q) {
            if (active && rq + q * QROWS < RH)
                *reinterpret_cast<float4 *>(Sraw + (rq + q * QROWS) * RWP + 4 * cq) =
                    make_float4(static_cast<float>(vals[q].x & 0xFFFFu),
                                static_cast<float>(vals[q].x >> 16),
                                static_cast<float>(vals[q].y & 0xFFFFu),
                                static_cast<float>(vals[q].y >> 16));
        }
    } else if (sizeof(T) == 2 && (R % 2) == 0 && (W % 2) == 0 &&
        (reinterpret_cast<uintptr_t>(f) & 3u) == 0 && y0 - R >= 0 && y0 + B2_TH + R <= H &&
        x0 - R >= 0 && x0 - R + 2 * PAIRS <= W) {
        // interior tile of a u16 frame: the staged region is whole in the
        // frame and starts on an even column, so rows load as u32 pairs
        constexpr int CN = RH * PAIRS, NL = (CN + B2_THREADS - 1) / B2_THREADS;
        const uint32_t *src = reinterpret_cast<const uint32_t *>(
            f + static_cast<size_t>(y0 - R) * W + (x0 - R));
        const int W2 = W / 2;
        uint32_t vals[NL];
#pragma unroll
        for (int q = 0; q < NL; ++q) {
            const int idx = min(tid + q * B2_THREADS, CN - 1);
            const int ry = idx / PAIRS, rx = idx - ry * PAIRS;
            vals[q] = src[static_cast<size_t>(ry) * W2 + rx];
        }
#pragma unroll
        for (int q = 0; q < NL; ++q) {
            const int idx = tid + q * B2_THREADS;
            if (idx < CN) {
                const int ry = idx / PAIRS, rx = idx - ry * PAIRS;
                *reinterpret_cast<float2 *>(Sraw + ry * RWP + 2 * rx) =
                    make_float2(static_cast<float>(vals[q] & 0xFFFFu),
                                static_cast<float>(vals[q] >> 16));
            }
        }
    } else {
        // issue every global load of the thread before the first LDS store
        constexpr int CN = RH * RWP, NL = (CN + B2_THREADS - 1) / B2_THREADS;
        T vals[NL];
#pragma unroll
        for (int q = 0; q < NL; ++q) {
            const int idx = min(tid + q * B2_THREADS, CN - 1);
            const int ry = idx / RWP, rx = idx - ry * RWP;
            const int gy = min(max(y0 - R + ry, 0), H - 1);
            const int gx = min(max(x0 - R + rx, 0), W - 1);
            vals[q] = f[static_cast<size_t>(gy) * W + gx];
        }
#pragma unroll
        for (int q = 0; q < NL; ++q) {
            const int idx = tid + q * B2_THREADS;
            if (idx < CN) Sraw[idx] = static_cast<float>(vals[q]);
        }
    }
    __syncthreads();
    const float gnorm = p.gnorm, rgnorm = p.rgnorm, rgnorm_lo = p.rgnorm_lo;
    // every column the tile's row pass reads lies in the frame
    const bool xin = x0 - R >= 0 && x0 + B2_TW + R <= W;
    // Row pass: 4 adjacent columns of one staged row per item; item
    // tid + q * B2_THREADS is row ry0 + q * RSTEP, columns cx..cx+3, so a
    // thread's items differ by constant LDS offsets.
    constexpr int IPR = B2_TW / 4;                  // items per staged row
    constexpr int RSTEP = B2_THREADS / IPR;         // staged rows per pass
    static_assert(B2_THREADS % IPR == 0, "whole rows per pass");
    constexpr int RPASS = (RH + RSTEP - 1) / RSTEP;
    const int ry0 = tid / IPR, cxi = (tid % IPR) * 4;
    float4 keep_g[RPASS], keep_c[RPASS];
    // kEdge: the tile touches the frame's left or right edge (its own copy of
    // the pass, so the interior tiles' code carries no edge handling)
    auto row_pass = [&](auto edge) {
        constexpr bool kEdge = decltype(edge)::value;
#pragma unroll
        for (int q = 0; q < RPASS; ++q) {
            const int ry = ry0 + q * RSTEP, cx = cxi;
            if (ry >= RH) continue;
            const float4 *src4 = reinterpret_cast<const float4 *>(Sraw + ry * RWP + cx);
            float v[4 * NV];
#pragma unroll
            for (int k = 0; k < NV; ++k) {
                const float4 t = src4[k];
                v[4 * k + 0] = t.x;
                v[4 * k + 1] = t.y;
                v[4 * k + 2] = t.z;
                v[4 * k + 3] = t.w;
            }
            float gs[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            const int gx0 = x0 + cx - R;   // frame column of v[0]
            if (!kEdge || (gx0 >= 0 && gx0 + 3 + 2 * R < W)) {
#pragma unroll
                for (int t = 0; t < NT; ++t)
#pragma unroll
                    for (int k = 0; k < 4; ++k) gs[k] = fmaf(p.w[t], v[k + t], gs[k]);
            } else {
                // a tap outside the frame is skipped; fmaf(w, 0, s) == s for
                // these non-negative sums, so such samples become zeros
                float vm[NT + 3];
#pragma unroll
                for (int j = 0; j < NT + 3; ++j) {
                    const int px = gx0 + j;
                    vm[j] = px >= 0 && px < W ? v[j] : 0.0f;
                }
#pragma unroll
                for (int t = 0; t < NT; ++t)
#pragma unroll
                    for (int k = 0; k < 4; ++k) gs[k] = fmaf(p.w[t], vm[k + t], gs[k]);
            }
            float cs[4];
            cs[0] = 0.0f;
#pragma unroll
            for (int t = 0; t < NT; ++t) cs[0] = cs[0] + v[t];
#pragma unroll
            for (int k = 1; k < 4; ++k) cs[k] = cs[k - 1] - v[k - 1] + v[k - 1 + NT];
            float4 og, oc;
            og.x = div_norm<kD2>(gs[0], gnorm, rgnorm, rgnorm_lo);
            og.y = div_norm<kD2>(gs[1], gnorm, rgnorm, rgnorm_lo);
            og.z = div_norm<kD2>(gs[2], gnorm, rgnorm, rgnorm_lo);
            og.w = div_norm<kD2>(gs[3], gnorm, rgnorm, rgnorm_lo);
            oc.x = cs[0];
            oc.y = cs[1];
            oc.z = cs[2];
            oc.w = cs[3];
            keep_g[q] = og;
            keep_c[q] = oc;
        }
    };
    if (xin) row_pass(std::false_type{});
    else row_pass(std::true_type{});
    __syncthreads();
#pragma unroll
    for (int q = 0; q < RPASS; ++q) {
        const int ry = ry0 + q * RSTEP;
        if (ry < RH) {
            *reinterpret_cast<float4 *>(Sg + ry * B2_TW + cxi) = keep_g[q];
            *reinterpret_cast<float4 *>(Sc + ry * B2_TW + cxi) = keep_c[q];
        }
    }
    __syncthreads();
    // Column pass: B2_ROWS consecutive rows of one column per thread (the
    // rows of a wavefront are wave-uniform: B2_TW is one wavefront wide).
    static_assert(B2_TW == 64, "one wavefront per row group");
    const int cx = tid % B2_TW;
    const int r0 = __builtin_amdgcn_readfirstlane(tid / B2_TW) * B2_ROWS;
    float g[B2_ROWS + 2 * R], c[B2_ROWS + 2 * R];
#pragma unroll
    for (int t = 0; t < B2_ROWS + 2 * R; ++t) {
        g[t] = Sg[(r0 + t) * B2_TW + cx];
        c[t] = Sc[(r0 + t) * B2_TW + cx];
    }
    const int gx = x0 + cx;
    const int gy0 = y0 + r0;           // frame row of output 0; its taps start at gy0 - R
    const bool inner = gy0 - R >= 0 && gy0 + B2_ROWS - 1 + R < H;
    const float box = p.box_area, rbox = p.rbox, rbox_lo = p.rbox_lo;
    float csum = 0.0f;
#pragma unroll
    for (int t = 0; t < NT; ++t) csum = csum + c[t];
    // the wave's row pointer is uniform; each lane stores at column cx of it
    float *orow = out + static_cast<size_t>(gy0) * W + x0;
    if (inner && x0 + B2_TW <= W) {
        // every tap and every output of the wave's rows lies in the frame
#pragma unroll
        for (int r = 0; r < B2_ROWS; ++r) {
            if (r > 0) csum = csum - c[r - 1] + c[r - 1 + NT];
            float gs = 0.0f;
#pragma unroll
            for (int t = 0; t < NT; ++t) gs = fmaf(p.w[t], g[r + t], gs);
            float res = div_norm<kD2>(gs, gnorm, rgnorm, rgnorm_lo) - div_norm<kD2>(csum, box, rbox, rbox_lo);
            if (p.truncate && res < kTrunc) res = 0.0f;
            orow[cx] = res;
            orow += W;
        }
        return;
    }
    const bool rows_in = gy0 + B2_ROWS <= H && gx < W;
#pragma unroll
    for (int r = 0; r < B2_ROWS; ++r) {
        if (r > 0) csum = csum - c[r - 1] + c[r - 1 + NT];
        float gs = 0.0f;
        if (inner) {
#pragma unroll
            for (int t = 0; t < NT; ++t) gs = fmaf(p.w[t], g[r + t], gs);
        } else {
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const int py = gy0 + r - R + t;
                if (py >= 0 && py < H) gs = fmaf(p.w[t], g[r + t], gs);
            }
        }
        float res = div_norm<kD2>(gs, gnorm, rgnorm, rgnorm_lo) - div_norm<kD2>(csum, box, rbox, rbox_lo);
        if (p.truncate && res < kTrunc) res = 0.0f;
        if (rows_in || (gy0 + r < H && gx < W)) orow[cx] = res;
        orow += W;
    }
}

template <typename T, int R, bool kD2>
__global__ __launch_bounds__(B2_THREADS) void k_bandpass_int(const T *__restrict__ frames,
                                                             float *__restrict__ proc,
                                                             GtKParams p, int nframes,
                                                             GtZero zero) {
    __shared__ __align__(16) BpLds<R> L;
    if (blockIdx.x == 0) {
        // the counters of the chunk's later stages
#pragma unroll
        for (int a = 0; a < 2; ++a)
            for (uint32_t i = threadIdx.x; i < zero.n[a]; i += B2_THREADS) zero.p[a][i] = 0u;
    }
    const int H = p.H, W = p.W;
    const size_t HW = static_cast<size_t>(H) * W;
    const int ntx = (W + B2_TW - 1) / B2_TW, nty = (H + B2_TH - 1) / B2_TH;
    const int tile = xcd_tile(ntx * nty * nframes);
    if (tile >= ntx * nty * nframes) return;
    const int fr = tile / (ntx * nty), trest = tile - fr * (ntx * nty);
    const int x0 = (trest % ntx) * B2_TW, y0 = (trest / ntx) * B2_TH;
    bandpass_int_tile<T, R, kD2>(frames + static_cast<size_t>(fr) * HW,
                            proc + static_cast<size_t>(fr) * HW, p, H, W, x0, y0, L);
}

// The noise estimate of frame fr of a launch (the snr rules' std of the raw
// frame), and the candidate threshold std * snr * rough_snr_factor
// (extract_max.wgsl:37, logspace_max.wgsl:122).
__device__ __forceinline__ float frame_std(const GtKParams &p, int fr) {
    return p.fstd ? p.fstd[fr] : p.image_std;
}
__device__ __forceinline__ float frame_thr(const GtKParams &p, int fr) {
    return p.fstd ? p.fstd[fr] * p.snr * p.rough : p.max_thr;
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
        if (!(c < frame_thr(p, fr))) {
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

}  // namespace

// sin(2 atan2(j, i)) and cos(2 atan2(j, i)) for |i|, |j| <= GT_SC_R in closed
// form, 2ij/(i^2+j^2) and (i^2-j^2)/(i^2+j^2), filled by gt_init_tables().
__constant__ double2 c_sincos2[(2 * GT_SC_R + 1) * (2 * GT_SC_R + 1)];

namespace {

__device__ __forceinline__ double2 sincos2(int i, int j) {
    if (i >= -GT_SC_R && i <= GT_SC_R && j >= -GT_SC_R && j <= GT_SC_R)
        return c_sincos2[(i + GT_SC_R) * (2 * GT_SC_R + 1) + (j + GT_SC_R)];
    const double q = static_cast<double>(i * i + j * j);
    return make_double2(2.0 * static_cast<double>(i) * static_cast<double>(j) / q,
                        (static_cast<double>(i * i) - static_cast<double>(j * j)) / q);
}

// characterize (characterize.wgsl:61-143) at round-half-even of the walked
// centre, with the Rg and eccentricity moment sums accumulated in f64.
struct CharAcc {
    float raw_mass, signal;
    double rg, es, ec;
};

// One disk pixel of characterize; `fold` marks a compile-time (i, j), for
// which the closed-form sin/cos(2 theta) fold to constants.
template <typename T, class S, bool fold>
__device__ __forceinline__ void char_accumulate(CharAcc &acc, const S &src,
                                                const T *__restrict__ Rw, int H, int W, int u,
                                                int v, int i, int j, float r2) {
    const float x = static_cast<float>(i), x2 = x * x;
    const float y = static_cast<float>(j), y2 = y * y;
    const float mask = x2 + y2;
    if (mask > r2) return;
    const int pu = u + i, pv = v + j;
    if (pu < 0 || pu >= H || pv < 0 || pv >= W) return;
    const float d = src.at(pu, pv);
    acc.raw_mass = acc.raw_mass + to_f32(Rw[static_cast<size_t>(pu) * W + pv]);
    acc.signal = maxf_ref(acc.signal, d);
    if (i != 0 || j != 0) {
        double2 sc;
        if constexpr (fold) {
            const double q = static_cast<double>(i * i + j * j);
            sc = make_double2(2.0 * static_cast<double>(i) * static_cast<double>(j) / q,
                              (static_cast<double>(i * i) - static_cast<double>(j * j)) / q);
        } else {
            sc = sincos2(i, j);
        }
        acc.es = fma(static_cast<double>(d), sc.x, acc.es);
        acc.ec = fma(static_cast<double>(d), sc.y, acc.ec);
    } else {
        acc.ec = acc.ec + static_cast<double>(d);
    }
    acc.rg = fma(static_cast<double>(d), static_cast<double>(mask), acc.rg);
}

// characterize (characterize.wgsl:61-143) at round-half-even of the walked
// centre, with the Rg and eccentricity moment sums accumulated in f64.  RI > 0
// unrolls the disk of rows without their own radius, so its loads issue
// together.
template <typename T, int RI = 0, bool kAllRows = false, class S>
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
    CharAcc acc = {0.0f, 0.0f, 0.0, 0.0, 0.0};
    bool done = false;
    if constexpr (RI > 0) {
        if (row.r <= 0.0f && rint == RI) {
            constexpr float r2c = static_cast<float>(RI) * static_cast<float>(RI);
            if constexpr (kAllRows) {
                // whole disk unrolled: every load issues at once
#pragma unroll
                for (int i = -RI; i <= RI; ++i)
#pragma unroll
                    for (int j = -RI; j <= RI; ++j)
                        char_accumulate<T, S, true>(acc, src, Rw, H, W, u, v, i, j, r2c);
            } else {
                // rows one at a time (bounded register use), columns unrolled
                // so a row's loads issue together
#pragma unroll 1
                for (int i = -RI; i <= RI; ++i)
#pragma unroll
                    for (int j = -RI; j <= RI; ++j)
                        char_accumulate<T, S, false>(acc, src, Rw, H, W, u, v, i, j, r2c);
            }
            done = true;
        }
    }
    if (!done) {
        for (int i = -rint; i <= rint; ++i)
            for (int j = -rint; j <= rint; ++j)
                char_accumulate<T, S, false>(acc, src, Rw, H, W, u, v, i, j, r2);
    }
    // the reference reads the centre sample without a bounds check
    const int cu = min(max(u, 0), H - 1), cv = min(max(v, 0), W - 1);
    const float den = row.mass - src.at(cu, cv);
    row.ecc = static_cast<float>(sqrt(acc.es * acc.es + acc.ec * acc.ec) / static_cast<double>(den));
    row.Rg = static_cast<float>(sqrt(acc.rg / static_cast<double>(row.mass)));
    row.raw_mass = acc.raw_mass;
    row.signal = acc.signal;
}

// minmass rule (walk.wgsl:200-213); stdv is the frame's noise estimate
__device__ __forceinline__ bool passes_minmass(const GtKParams &p, const WalkRes &res, float stdv) {
    float mm;
    if (p.minmass > 0.0f) mm = p.minmass;
    else mm = p.minmass_snr * res.count * stdv * p.rough;
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
            keep = passes_minmass(p, res, frame_std(p, fr));
            if (keep) {
                fill_row(d, res, seed, frame0 + fr);
                if (p.characterize) characterize_row<T, RI>(src, Rw, H, W, p.circle, d.row);
            }
        }
        const uint32_t pos = wave_append(keep, &ocount[fr * GT_COUNTER_STRIDE]);
        if (keep && pos < out_cap) fout[pos] = d;
    }
}

// ---------------------------------------------------------------------------
// Fused local max + walk.  One block per FT_TH x FT_TW tile:
//   1. the processed frame over the tile plus halos (TileGeom) is staged in
//      LDS, out-of-frame entries -inf;
//   2. horizontal window maxima: one item = 4 adjacent output columns of one
//      row, read as float4 from the staged row (the left halo is chosen so
//      these reads are 16-byte aligned), written to Hm;
//   3. vertical window maxima and the test: each lane owns a column and
//      FT_WROWS output rows, folding Hm rows in registers; the block's
//      candidates are compacted into an LDS list in raster order (ballots and
//      a 4-entry prefix; the list reuses Hm's storage);
//   4. the first ceil(n/64) waves walk the list, one lane per candidate, for
//      at most FT_WALK_STEPS steps; a disk inside the staged region is read
//      from LDS, any other from the frame in global memory (same arithmetic).
//      Walks still moving go to k_walk_slow through a per-frame list (or walk
//      on here when that list is full);
//   5. finished walks that pass minmass are appended to the frame's rows with
//      one atomic per wave.  Waves without candidates leave early.
// characterize runs afterwards over each frame's compacted rows
// (k_characterize_det), whole wavefronts of detections at a time.
constexpr int FT_TW = 64, FT_TH = 32, FT_THREADS = 256;
constexpr int FT_WAVES = FT_THREADS / 64;
constexpr int FT_WROWS = FT_TH / FT_WAVES;
constexpr uint32_t FT_WALK_STEPS = 2;
// At least FT_MIN_WAVES waves per SIMD: the walk phase is latency bound.
constexpr int FT_MIN_WAVES = 4;

}  // namespace

// Halo layout of a detect tile for dilation side k and walk radius rint.
// The window of output column c starts at staged column hl - a + c, and
// (hl - a) % 4 == 0 keeps the horizontal float4 reads aligned.
struct TileGeom {
    int a, ht, hb, hl, hr, width, pitch, rows;
};
__host__ __device__ constexpr TileGeom tile_geom(int k, int rint) {
    const int a = (k - 1) / 2;  // = i32(k/2 - 0.5) for k >= 1
    const int lo = a > rint + 2 ? a : rint + 2;
    const int hi = (k - 1 - a) > rint + 2 ? (k - 1 - a) : rint + 2;
    const int hl = a + 4 * ((lo - a + 3) / 4);
    const int width = hl + FT_TW + hi;
    return TileGeom{a, lo, hi, hl, hi, width, ((width + 3) & ~3) + 4, lo + FT_TH + hi};
}

namespace {

// Centroid of the disk around (u, v) when the whole disk is inside the frame:
// no per-pixel bounds test, every offset a compile-time constant.  Same
// operations in the same order as get_center, so the same bits.
// kFin: every sample is finite.  Then the zero-offset taps fma(d, 0, acc) are
// skipped: d * 0 is a zero and acc is never -0 (it starts at +0, and an exact
// zero sum rounds to +0), so they leave acc unchanged.
template <int RI, bool kFin = false, class B>
__device__ __forceinline__ WalkRes get_center_inner(const B &base) {
    WalkRes res = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    constexpr float r2 = static_cast<float>(RI) * static_cast<float>(RI);
#pragma unroll
    for (int i = -RI; i <= RI; ++i) {
#pragma unroll
        for (int j = -RI; j <= RI; ++j) {
            const float x = static_cast<float>(i), y = static_cast<float>(j);
            if (x * x + y * y > r2) continue;
            const float d = base(i, j);
            if (!(kFin && i == 0)) res.x = fmaf(d, x, res.x);
            if (!(kFin && j == 0)) res.y = fmaf(d, y, res.y);
            res.mass = res.mass + d;
            res.count = res.count + 1.0f;
            // res.max starts at +0 and is never NaN, so fmaxf agrees with the
            // reference's max() here
            res.max = fmaxf(res.max, d);
        }
    }
    res.x = res.x / res.mass;
    res.y = res.y / res.mass;
    return res;
}

template <int PITCH>
struct LdsRow {
    const float *c;  // LDS address of the disk centre
    __device__ __forceinline__ float operator()(int i, int j) const { return c[i * PITCH + j]; }
};
struct GlobalRow {
    const float *c;
    int W;
    __device__ __forceinline__ float operator()(int i, int j) const {
        return c[static_cast<long long>(i) * W + j];
    }
};

// One walk evaluation at (pu, pv) from the best available source; PITCH is
// the staged region's compile-time row pitch (0: none or not known).
template <int RI, int PITCH, bool kFin = false>
__device__ __forceinline__ WalkRes walk_eval(const GtKParams &p, int H, int W,
                                             const float *__restrict__ P, const float *Sp, int oy,
                                             int ox, int RW, int ylo, int yhi, int xlo, int xhi,
                                             int pu, int pv) {
    if constexpr (RI > 0) {
        const bool in_frame = pu >= RI && pu + RI < H && pv >= RI && pv + RI < W;
        if (in_frame) {
            if constexpr (PITCH > 0) {
                if (pu >= ylo && pu <= yhi && pv >= xlo && pv <= xhi)
                    return get_center_inner<RI, kFin>(LdsRow<PITCH>{Sp + (pu - oy) * PITCH + (pv - ox)});
            }
            return get_center_inner<RI, kFin>(GlobalRow{P + static_cast<size_t>(pu) * W + pv, W});
        }
    }
    if (pu >= ylo && pu <= yhi && pv >= xlo && pv <= xhi)
        return get_center<RI>(LdsSrc{Sp, oy, ox, RW}, H, W, pu, pv, p.rint, p.r2);
    return get_center<RI>(GlobalSrc{P, W}, H, W, pu, pv, p.rint, p.r2);
}

// characterize for a disk of radius RI that lies wholly inside the frame:
// the same operations in the same order as characterize_row, without the
// per-pixel bounds test, so every load is unconditional and they issue
// together (one memory round trip instead of one per pixel).
template <typename T, int RI, class B>
__device__ __forceinline__ void characterize_inner(const B &base, const T *__restrict__ rc, int W,
                                                   gt_result_row &row) {
    constexpr float r2 = static_cast<float>(RI) * static_cast<float>(RI);
    float raw[2 * RI + 1][2 * RI + 1];
#pragma unroll
    for (int i = -RI; i <= RI; ++i) {
        const T *rrow = rc + static_cast<long long>(i) * W;
#pragma unroll
        for (int j = -RI; j <= RI; ++j) {
            const float x = static_cast<float>(i), y = static_cast<float>(j);
            if (x * x + y * y > r2) continue;
            raw[i + RI][j + RI] = to_f32(rrow[j]);
        }
    }
    float raw_mass = 0.0f, signal = 0.0f;
    double rg = 0.0, es = 0.0, ec = 0.0;
#pragma unroll
    for (int i = -RI; i <= RI; ++i) {
#pragma unroll
        for (int j = -RI; j <= RI; ++j) {
            const float x = static_cast<float>(i), y = static_cast<float>(j);
            const float mask = x * x + y * y;
            if (mask > r2) continue;
            const float d = base(i, j);
            raw_mass = raw_mass + raw[i + RI][j + RI];
            // signal starts at +0 and is never NaN: fmaxf matches max() here
            signal = fmaxf(signal, d);
            if (i != 0 || j != 0) {
                const double q = static_cast<double>(i * i + j * j);
                es = fma(static_cast<double>(d), 2.0 * static_cast<double>(i) * static_cast<double>(j) / q, es);
                ec = fma(static_cast<double>(d), (static_cast<double>(i * i) - static_cast<double>(j * j)) / q, ec);
            } else {
                ec = ec + static_cast<double>(d);
            }
            rg = fma(static_cast<double>(d), static_cast<double>(mask), rg);
        }
    }
    const float den = row.mass - base(0, 0);
    row.ecc = static_cast<float>(sqrt(es * es + ec * ec) / static_cast<double>(den));
    row.Rg = static_cast<float>(sqrt(rg / static_cast<double>(row.mass)));
    row.raw_mass = raw_mass;
    row.signal = signal;
}

// Finish a walk at integer position (pu, pv) with last evaluation `res`:
// minmass test, then the row, characterized when asked.
template <typename T, int RI, int PITCH = 0, bool kChar = true>
__device__ __forceinline__ bool finish_walk(const GtKParams &p, int H, int W,
                                            const float *__restrict__ P,
                                            const T *__restrict__ Rw, const float *Sp, int oy,
                                            int ox, int RW, int ylo, int yhi, int xlo, int xhi,
                                            WalkRes res, int pu, int pv, uint32_t seed,
                                            uint32_t frame, gt_detection &d, float stdv) {
    res.x = res.x + static_cast<float>(pu);
    res.y = res.y + static_cast<float>(pv);
    if (!passes_minmass(p, res, stdv)) return false;
    fill_row(d, res, seed, frame);
    if (kChar && p.characterize) {
        const int cu = f2i(rintf(d.row.x)), cv = f2i(rintf(d.row.y));
        const bool in_lds = Sp && cu >= ylo && cu <= yhi && cv >= xlo && cv <= xhi;
        if constexpr (RI > 0) {
            if (cu >= RI && cu + RI < H && cv >= RI && cv + RI < W) {
                const T *rc = Rw + static_cast<size_t>(cu) * W + cv;
                if constexpr (PITCH > 0) {
                    if (in_lds) {
                        characterize_inner<T, RI>(LdsRow<PITCH>{Sp + (cu - oy) * PITCH + (cv - ox)},
                                                  rc, W, d.row);
                        return true;
                    }
                }
                characterize_inner<T, RI>(GlobalRow{P + static_cast<size_t>(cu) * W + cv, W},
                                          rc, W, d.row);
                return true;
            }
        }
        if (in_lds)
            characterize_row<T, RI>(LdsSrc{Sp, oy, ox, RW}, Rw, H, W, p.circle, d.row);
        else
            characterize_row<T, RI>(GlobalSrc{P, W}, Rw, H, W, p.circle, d.row);
    }
    return true;
}

// Workgroup barrier between the phases of a detect tile.
struct SyncBarrier {
    __device__ __forceinline__ static void sync() { __syncthreads(); }
};

template <typename T, int KT, int RI, bool kCand, uint32_t kSteps, class Bar>
__device__ __forceinline__ void detect_compute(
    const float *__restrict__ P, const T *__restrict__ Rw, const GtKParams &p, int H, int W,
    const TileGeom &g, int x0, int y0, uint32_t frame, gt_detection *__restrict__ fout,
    uint32_t out_cap, uint32_t *__restrict__ counter, uint2 *__restrict__ fslow,
    uint32_t slow_cap, uint32_t *__restrict__ slow_counter, float *Sp, float *Hm,
    uint32_t *s_cnt, float thr, float stdv,
    uint32_t *__restrict__ fcand);

// One FT_TH x FT_TW tile (origin x0, y0) of one frame: P, Rw are its
// processed and raw planes (p.H x p.W); its rows go to fout (counted in
// *counter), walks handed over to fslow (counted in *slow_counter); `frame`
// is the rows' frame field.  With kCand the tile only lists its candidate
// pixels (raster index) in fcand, counted in *slow_counter (capacity
// slow_cap), for the learned proposal and its verifying walk.  kSteps walk
// steps run in the tile before a walk still moving is handed over.
template <typename T, int KT, int RI, bool kCand = false, uint32_t kSteps = FT_WALK_STEPS>
__device__ __forceinline__ void detect_tile_body(
    const float *__restrict__ P, const T *__restrict__ Rw, const GtKParams &p, int H, int W,
    const TileGeom &g, int x0, int y0, uint32_t frame, gt_detection *__restrict__ fout,
    uint32_t out_cap,
    uint32_t *__restrict__ counter, uint2 *__restrict__ fslow, uint32_t slow_cap,
    uint32_t *__restrict__ slow_counter, float *lds, uint32_t *s_cnt,
    float thr, float stdv, uint32_t *__restrict__ fcand = nullptr) {
    constexpr bool kFixed = KT > 0 && RI > 0;
    constexpr TileGeom gc = tile_geom(KT > 0 ? KT : 1, RI > 0 ? RI : 1);
    const int RW = g.pitch, RH = g.rows;
    float *Sp = lds;                   // RH x pitch, origin (y0 - ht, x0 - hl)
    float *Hm = lds + RH * RW;         // MR x FT_TW; later the candidate list
    const int tid = threadIdx.x;
    const int oy = y0 - g.ht, ox = x0 - g.hl;

    // 1. staging (the pitch padding is loaded too; it is never compared)
    const bool interior = oy >= 0 && ox >= 0 && oy + RH <= H && ox + RW <= W;
    constexpr int Q4 = gc.pitch / 4;              // float4 per staged row
    constexpr int QROWS = FT_THREADS / Q4;        // staged rows per pass of float4 loads
    constexpr int QP = (gc.rows + QROWS - 1) / QROWS;
    static_assert(gc.pitch % 4 == 0, "staged rows are whole float4s");
    if (kFixed && interior && (W & 3) == 0 && (reinterpret_cast<uintptr_t>(P) & 15u) == 0) {
        // every staged row is 16-byte aligned in the frame: float4 loads at
        // one 32-bit offset per thread from a per-pass tile row pointer
        const float *tb = P + (static_cast<size_t>(oy) * W + static_cast<size_t>(ox));
        const int rq = tid / Q4, cq = tid - rq * Q4;
        const bool active = tid < QROWS * Q4;
        float4 v4[QP];
#pragma unroll
        for (int q = 0; q < QP; ++q) {
            // rows past the staged region (not stored) load its last row:
            // every load stays in the tile and none is predicated
            const int ry = min(rq + q * QROWS, gc.rows - 1);
            const uint32_t off = static_cast<uint32_t>(ry) * static_cast<uint32_t>(W) +
                                 4u * static_cast<uint32_t>(cq);
            v4[q] = *reinterpret_cast<const float4 *>(tb + off);
        }
#pragma unroll
        for (int q = 0; q < QP; ++q)
            if (active && rq + q * QROWS < gc.rows)
                *reinterpret_cast<float4 *>(Sp + (rq + q * QROWS) * gc.pitch + 4 * cq) = v4[q];
    } else if (kFixed && gc.hl % 4 == 0 && (W & 3) == 0 &&
               (reinterpret_cast<uintptr_t>(P) & 15u) == 0) {
        // border tile whose float4 groups start on multiples of 4 frame
        // columns: each group is wholly inside or outside the frame; groups
        // load from clamped coordinates (always in the frame) and the
        // outside ones become -inf
        const int rq = tid / Q4, cq = tid - rq * Q4;
        const bool active = tid < QROWS * Q4;
        const int gx = ox + 4 * cq;
        const bool col_in = gx >= 0 && gx < W;
        const int gxc = min(max(gx, 0), W - 4);
        float4 v4[QP];
#pragma unroll
        for (int q = 0; q < QP; ++q) {
            const int gy = oy + min(rq + q * QROWS, gc.rows - 1);
            const int gyc = min(max(gy, 0), H - 1);
            const float4 t = *reinterpret_cast<const float4 *>(
                P + (static_cast<size_t>(gyc) * static_cast<size_t>(W) + static_cast<size_t>(gxc)));
            const bool in = col_in && gy == gyc;
            v4[q] = in ? t : make_float4(-INFINITY, -INFINITY, -INFINITY, -INFINITY);
        }
#pragma unroll
        for (int q = 0; q < QP; ++q)
            if (active && rq + q * QROWS < gc.rows)
                *reinterpret_cast<float4 *>(Sp + (rq + q * QROWS) * gc.pitch + 4 * cq) = v4[q];
    } else if constexpr (kFixed) {
        constexpr int CN = gc.rows * gc.pitch;
        constexpr int NL = (CN + FT_THREADS - 1) / FT_THREADS;
        float vals[NL];
        if (interior) {
            const float *src = P + static_cast<size_t>(oy) * W + ox;
#pragma unroll
            for (int q = 0; q < NL; ++q) {
                const int idx = min(tid + q * FT_THREADS, CN - 1);
                const int ry = idx / gc.pitch, rx = idx - ry * gc.pitch;
                vals[q] = src[static_cast<size_t>(ry) * W + rx];
            }
        } else {
#pragma unroll
            for (int q = 0; q < NL; ++q) {
                const int idx = tid + q * FT_THREADS;
                const int ry = idx / gc.pitch, rx = idx - ry * gc.pitch;
                const int gy = oy + ry, gx = ox + rx;
                vals[q] = (idx < CN && gy >= 0 && gy < H && gx >= 0 && gx < W)
                              ? P[static_cast<size_t>(gy) * W + gx]
                              : -INFINITY;
            }
        }
#pragma unroll
        for (int q = 0; q < NL; ++q) {
            const int idx = tid + q * FT_THREADS;
            if (idx < CN) Sp[idx] = vals[q];
        }
    } else {
        for (int idx = tid; idx < RH * RW; idx += FT_THREADS) {
            const int ry = idx / RW, rx = idx - ry * RW;
            const int gy = oy + ry, gx = ox + rx;
            Sp[idx] = (gy >= 0 && gy < H && gx >= 0 && gx < W)
                          ? P[static_cast<size_t>(gy) * W + gx]
                          : -INFINITY;
        }
    }
    __syncthreads();
    detect_compute<T, KT, RI, kCand, kSteps, SyncBarrier>(
        P, Rw, p, H, W, g, x0, y0, frame, fout, out_cap, counter, fslow, slow_cap, slow_counter,
        Sp, Hm, s_cnt, thr, stdv, fcand);
}

// Phases 2-5 of a detect tile whose processed region (TileGeom g, origin
// (y0 - g.ht, x0 - g.hl)) is staged in Sp (out-of-frame entries -inf); Hm
// holds the horizontal maxima and then the candidate list.  Bar::sync() is
// the workgroup barrier between the phases.
template <typename T, int KT, int RI, bool kCand, uint32_t kSteps, class Bar>
__device__ __forceinline__ void detect_compute(
    const float *__restrict__ P, const T *__restrict__ Rw, const GtKParams &p, int H, int W,
    const TileGeom &g, int x0, int y0, uint32_t frame, gt_detection *__restrict__ fout,
    uint32_t out_cap, uint32_t *__restrict__ counter, uint2 *__restrict__ fslow,
    uint32_t slow_cap, uint32_t *__restrict__ slow_counter, float *Sp, float *Hm,
    uint32_t *s_cnt, float thr, float stdv,
    uint32_t *__restrict__ fcand) {
    constexpr bool kFixed = KT > 0 && RI > 0;
    constexpr TileGeom gc = tile_geom(KT > 0 ? KT : 1, RI > 0 ? RI : 1);
    constexpr bool kInt = sizeof(T) <= 2;  // integer frames: no NaN in the processed frame
    const int k = KT > 0 ? KT : p.k;
    const int a = g.a;
    const int rint = RI > 0 ? RI : p.rint;
    const int RW = g.pitch, RH = g.rows;
    const int MR = FT_TH + k - 1;      // rows of horizontal maxima the tile needs
    uint16_t *s_cand = reinterpret_cast<uint16_t *>(Hm);
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int oy = y0 - g.ht, ox = x0 - g.hl;

    // integer frames have no NaN: IEEE maximum (one instruction, no operand
    // canonicalization); float frames keep the oracle's comparison
    auto vmax = [](float x, float y) {
        if constexpr (kInt) return __builtin_elementwise_maximum(x, y);
        else return maxf_ref(x, y);
    };
    // 2. horizontal maxima: Hm[m][c] = max_{j<k} Sp[ht - a + m][hl - a + c + j]
    //    (float frames keep the oracle's sequential order: NaN inputs)
    if constexpr (kFixed && kInt) {
        constexpr int NV = (KT + 3 + 3) / 4;  // float4 reads per item
        for (int item = tid; item < MR * (FT_TW / 4); item += FT_THREADS) {
            const int m = item / (FT_TW / 4), c4 = (item % (FT_TW / 4)) * 4;
            const float4 *src =
                reinterpret_cast<const float4 *>(Sp + (gc.ht - gc.a + m) * gc.pitch + (gc.hl - gc.a) + c4);
            float v[4 * NV];
#pragma unroll
            for (int q = 0; q < NV; ++q) {
                const float4 t = src[q];
                v[4 * q] = t.x;
                v[4 * q + 1] = t.y;
                v[4 * q + 2] = t.z;
                v[4 * q + 3] = t.w;
            }
            // the reference starts each row maximum from the centre value
            float core = v[a + 3];
#pragma unroll
            for (int j = 3; j < KT; ++j) core = vmax(core, v[j]);
            float4 o;
            float m0 = vmax(v[a], core), m1 = vmax(v[a + 1], core);
            float m2 = vmax(v[a + 2], core), m3 = vmax(v[a + 3], core);
            m0 = vmax(vmax(vmax(m0, v[0]), v[1]), v[2]);
            m1 = vmax(vmax(vmax(m1, v[1]), v[2]), v[KT]);
            m2 = vmax(vmax(vmax(m2, v[2]), v[KT]), v[KT + 1]);
            m3 = vmax(vmax(vmax(m3, v[KT]), v[KT + 1]), v[KT + 2]);
            o.x = m0;
            o.y = m1;
            o.z = m2;
            o.w = m3;
            *reinterpret_cast<float4 *>(Hm + m * FT_TW + c4) = o;
        }
    } else {
        for (int item = tid; item < MR * FT_TW; item += FT_THREADS) {
            const int m = item / FT_TW, c = item - m * FT_TW;
            const float *s = Sp + (g.ht - a + m) * RW + (g.hl - a) + c;
            float mx = s[a];
            for (int j = 0; j < k; ++j) mx = vmax(mx, s[j]);
            Hm[item] = mx;
        }
    }
    Bar::sync();

    // 3. vertical maxima and the candidate test for rows r0..r0+7 of column
    //    `lane`: Hm row q feeds output row r when 0 <= q - r < k
    const int r0 = wave * FT_WROWS;
    const int margin = p.margin;
    const int gx = x0 + lane;
    uint32_t nwave = 0;  // wave-uniform
    uint64_t mks[FT_WROWS];
    {
        float wmax[FT_WROWS];
#pragma unroll
        for (int r = 0; r < FT_WROWS; ++r) wmax[r] = -INFINITY;
        const float *col = Hm + r0 * FT_TW + lane;
        if constexpr (KT > FT_WROWS && kInt) {
            // integer frames (no NaN: max is order-free): window r..r+KT-1 =
            // suffix max of rows r..7 with prefix max of rows 8..r+KT-1
            float h[FT_WROWS + KT - 1];
#pragma unroll
            for (int q = 0; q < FT_WROWS + KT - 1; ++q) h[q] = col[q * FT_TW];
            float suf[FT_WROWS];
            suf[FT_WROWS - 1] = h[FT_WROWS - 1];
#pragma unroll
            for (int q = FT_WROWS - 2; q >= 0; --q) suf[q] = vmax(h[q], suf[q + 1]);
            float pre = h[FT_WROWS];
#pragma unroll
            for (int q = FT_WROWS + 1; q < FT_WROWS + KT - 1; ++q) {
                pre = vmax(pre, h[q]);
                const int r = q - (KT - 1);   // window r..q ends here
                if (r >= 0 && r < FT_WROWS) wmax[r] = vmax(suf[r], pre);
            }
            // the window ending at row FT_WROWS itself (KT = FT_WROWS + 1)
            if constexpr (KT - 1 == FT_WROWS) wmax[0] = vmax(suf[0], h[FT_WROWS]);
        } else if constexpr (KT > 0) {
#pragma unroll
            for (int q = 0; q < FT_WROWS + KT - 1; ++q) {
                const float hq = col[q * FT_TW];
#pragma unroll
                for (int r = 0; r < FT_WROWS; ++r)
                    if (q - r >= 0 && q - r < KT) wmax[r] = vmax(wmax[r], hq);
            }
        } else {
            for (int q = 0; q < FT_WROWS + k - 1; ++q) {
                const float hq = col[q * FT_TW];
#pragma unroll
                for (int r = 0; r < FT_WROWS; ++r)
                    if (q - r >= 0 && q - r < k) wmax[r] = vmax(wmax[r], hq);
            }
        }
        // the staged region covers every row and column of the tile, so the
        // centre values load unconditionally (all in flight at once); the
        // row test is wave-uniform, the column test per lane
        float cv[FT_WROWS];
#pragma unroll
        for (int r = 0; r < FT_WROWS; ++r) cv[r] = Sp[(g.ht + r0 + r) * RW + g.hl + lane];
        const bool col_ok = gx < W && gx >= margin && gx < W - margin;
        if constexpr (kInt) {
            // finite values: c passes iff c >= max(thr, window maximum); the
            // rows inside the margin are a wave-uniform bit mask, the columns
            // a ballot applied to each row's mask
            const unsigned long long cols = __ballot(col_ok);
            const int ylo = margin - (y0 + r0), yhi = min(H, H - margin) - (y0 + r0);
            const int rlo = min(max(ylo, 0), FT_WROWS), rhi = min(max(yhi, 0), FT_WROWS);
            const uint32_t rows = rhi > rlo ? (1u << rhi) - (1u << rlo) : 0u;
            const float thr_u = __builtin_bit_cast(
                float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, thr)));
#pragma unroll
            for (int r = 0; r < FT_WROWS; ++r) {
                const float m = __builtin_elementwise_maximum(wmax[r], thr_u);
                mks[r] = ((rows >> r) & 1u) ? (__ballot(!(m > cv[r])) & cols) : 0ull;
                nwave += static_cast<uint32_t>(__popcll(mks[r]));
            }
        } else {
#pragma unroll
            for (int r = 0; r < FT_WROWS; ++r) {
                const int gy = y0 + r0 + r;
                const bool row_ok = gy < H && gy >= margin && gy < H - margin;
                const bool flag = row_ok && col_ok && !(cv[r] < thr) && !(wmax[r] > cv[r]);
                mks[r] = __ballot(flag);
                nwave += static_cast<uint32_t>(__popcll(mks[r]));
            }
        }
    }
    if (lane == 0) s_cnt[wave] = nwave;
    Bar::sync();  // Hm is dead from here on: the list takes its place
    uint32_t off = 0, ncand = 0;
#pragma unroll
    for (int w = 0; w < FT_WAVES; ++w) {
        const uint32_t cw = s_cnt[w];
        off += w < wave ? cw : 0u;
        ncand += cw;
    }
#pragma unroll
    for (int r = 0; r < FT_WROWS; ++r) {
        const uint64_t mk = mks[r];
        // the lane's rank among the row's candidates (mbcnt: set bits of mk
        // below the lane); lanes not in mk write nowhere
        const uint32_t below = __builtin_amdgcn_mbcnt_hi(
            static_cast<uint32_t>(mk >> 32), __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mk), 0u));
        const bool mine = (static_cast<uint32_t>(mk >> (lane & 32)) >> (lane & 31)) & 1u;
        if (mine) s_cand[off + below] = static_cast<uint16_t>((r0 + r) * FT_TW + lane);
        off += static_cast<uint32_t>(__popcll(mk));
    }
    Bar::sync();
    if constexpr (kCand) {
        for (uint32_t c0 = static_cast<uint32_t>(wave) * 64u; c0 < ncand; c0 += FT_THREADS) {
            const uint32_t ci = c0 + static_cast<uint32_t>(lane);
            const bool v = ci < ncand;
            uint32_t seed = 0;
            if (v) {
                const int li = s_cand[ci];
                const int ty = li / FT_TW, cx = li - ty * FT_TW;
                seed = static_cast<uint32_t>(y0 + ty) * static_cast<uint32_t>(W) + (x0 + cx);
            }
            const uint32_t pos = wave_append(v, slow_counter);
            if (v && pos < slow_cap) fcand[pos] = seed;
        }
        return;
    }
    if (static_cast<uint32_t>(wave) * 64u >= ncand) return;  // nothing left for this wave

    // 4-5. walks
    const int ylo = oy + rint, yhi = oy + RH - 1 - rint;
    const int xlo = ox + rint, xhi = ox + g.width - 1 - rint;
    constexpr int PITCH = kFixed ? gc.pitch : 0;
    for (uint32_t c0 = static_cast<uint32_t>(wave) * 64u; c0 < ncand; c0 += FT_THREADS) {
        const uint32_t ci = c0 + static_cast<uint32_t>(lane);
        bool keep = false, defer = false;
        gt_detection d;
        int pu = 0, pv = 0;
        uint32_t seed = 0;
        if (ci < ncand) {
            const int li = s_cand[ci];
            const int ty = li / FT_TW, cx = li - ty * FT_TW;
            pu = y0 + ty;
            pv = x0 + cx;
            seed = static_cast<uint32_t>(pu) * static_cast<uint32_t>(W) + pv;
            WalkRes res = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
            bool moving = true;
            uint32_t it = 0;
            for (; it < p.max_iter && it < kSteps; ++it) {
                res = walk_eval<RI, PITCH, kInt>(p, H, W, P, Sp, oy, ox, RW, ylo, yhi, xlo, xhi, pu, pv);
                if (!walk_move(res, pu, pv, H, W, rint, p.shift_thr)) {
                    moving = false;
                    break;
                }
            }
            if (moving && it < p.max_iter) defer = true;
            else keep = finish_walk<T, RI, PITCH, false>(p, H, W, P, Rw, Sp, oy, ox, RW, ylo, yhi, xlo, xhi, res, pu,
                                           pv, seed, frame, d, stdv);
        }
        // long walks go to k_walk_slow, finished ones to the rows: one atomic
        // round trip for both lists
        uint32_t spos, pos;
        wave_append2(defer, slow_counter, keep, counter, spos, pos);
        if (defer && spos < slow_cap)
            fslow[spos] = make_uint2(seed, static_cast<uint32_t>(pu) * static_cast<uint32_t>(W) + pv);
        if (keep && pos < out_cap) fout[pos] = d;
        // the slow list is full: those walks go on here
        const bool over = defer && spos >= slow_cap;
        if (__ballot(over) != 0ull) {
            bool keep2 = false;
            if (over) {
                WalkRes res = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
                for (uint32_t it = kSteps; it < p.max_iter; ++it) {
                    res = walk_eval<RI, PITCH, kInt>(p, H, W, P, Sp, oy, ox, RW, ylo, yhi, xlo, xhi, pu, pv);
                    if (!walk_move(res, pu, pv, H, W, rint, p.shift_thr)) break;
                }
                keep2 = finish_walk<T, RI, PITCH, false>(p, H, W, P, Rw, Sp, oy, ox, RW, ylo, yhi, xlo, xhi, res, pu,
                                           pv, seed, frame, d, stdv);
            }
            const uint32_t pos2 = wave_append(keep2, counter);
            if (keep2 && pos2 < out_cap) fout[pos2] = d;
        }
    }
}

template <typename T, int KT, int RI, uint32_t kSteps = FT_WALK_STEPS>
__global__ __launch_bounds__(FT_THREADS, FT_MIN_WAVES) void k_detect_tile(
    const float *__restrict__ proc, const T *__restrict__ raw, GtKParams p, TileGeom g_rt,
    gt_detection *__restrict__ out, uint32_t out_cap, uint32_t *__restrict__ ocount,
    uint2 *__restrict__ slow, uint32_t slow_cap, uint32_t *__restrict__ scount, uint32_t frame0,
    int nframes) {
    constexpr bool kFixed = KT > 0 && RI > 0;
    constexpr TileGeom gc = tile_geom(KT > 0 ? KT : 1, RI > 0 ? RI : 1);
    const TileGeom g = kFixed ? gc : g_rt;
    extern __shared__ __align__(16) float lds[];
    __shared__ uint32_t s_cnt[FT_WAVES];
    const int H = p.H, W = p.W;
    const size_t HW = static_cast<size_t>(H) * W;
    const int ntx = (W + FT_TW - 1) / FT_TW, nty = (H + FT_TH - 1) / FT_TH;
    const int tile = xcd_tile(ntx * nty * nframes);
    if (tile >= ntx * nty * nframes) return;
    // tile -> (frame, tile row, tile column); the launch keeps ntiles < 2^22
    uint32_t trest, tx;
    const int fr = static_cast<int>(udiv22(static_cast<uint32_t>(tile),
                                           static_cast<uint32_t>(ntx * nty), trest));
    const int ty = static_cast<int>(udiv22(trest, static_cast<uint32_t>(ntx), tx));
    const int x0 = static_cast<int>(tx) * FT_TW, y0 = ty * FT_TH;
    detect_tile_body<T, KT, RI, false, kSteps>(
                                proc + static_cast<size_t>(fr) * HW, raw + static_cast<size_t>(fr) * HW,
                                p, H, W, g, x0, y0, frame0 + fr, out + static_cast<size_t>(fr) * out_cap,
                                out_cap, &ocount[fr * GT_COUNTER_STRIDE],
                                slow + static_cast<size_t>(fr) * slow_cap, slow_cap,
                                &scount[fr * GT_COUNTER_STRIDE], lds, s_cnt,
                                frame_thr(p, fr), frame_std(p, fr));
}

// Candidate pixels of every tile (the local-maximum test of k_detect_tile),
// listed per frame: cand[f * cap + i], count in ccount[f * GT_COUNTER_STRIDE].
template <typename T, int KT, int RI>
__global__ __launch_bounds__(FT_THREADS, FT_MIN_WAVES) void k_detect_cand(
    const float *__restrict__ proc, GtKParams p, TileGeom g_rt, uint32_t *__restrict__ cand,
    uint32_t cap, uint32_t *__restrict__ ccount, int nframes) {
    constexpr bool kFixed = KT > 0 && RI > 0;
    constexpr TileGeom gc = tile_geom(KT > 0 ? KT : 1, RI > 0 ? RI : 1);
    const TileGeom g = kFixed ? gc : g_rt;
    extern __shared__ __align__(16) float lds[];
    __shared__ uint32_t s_cnt[FT_WAVES];
    const int H = p.H, W = p.W;
    const size_t HW = static_cast<size_t>(H) * W;
    const int ntx = (W + FT_TW - 1) / FT_TW, nty = (H + FT_TH - 1) / FT_TH;
    const int tile = xcd_tile(ntx * nty * nframes);
    if (tile >= ntx * nty * nframes) return;
    const int fr = tile / (ntx * nty), trest = tile - fr * (ntx * nty);
    const int x0 = (trest % ntx) * FT_TW, y0 = (trest / ntx) * FT_TH;
    detect_tile_body<T, KT, RI, true>(proc + static_cast<size_t>(fr) * HW, nullptr, p, H, W, g,
                                      x0, y0, 0, nullptr, 0, nullptr, nullptr, cap,
                                      &ccount[fr * GT_COUNTER_STRIDE], lds, s_cnt,
                                      frame_thr(p, fr), frame_std(p, fr),
                                      cand + static_cast<size_t>(fr) * cap);
}

// Continues the walks k_detect_tile handed over: each entry is the seed and
// the position reached after FT_WALK_STEPS steps; the remaining steps read the
// frame in global memory.
template <typename T, int RI>
__global__ __launch_bounds__(WK_THREADS) void k_walk_slow(const float *__restrict__ proc,
                                                          const T *__restrict__ raw, GtKParams p,
                                                          const uint2 *__restrict__ slow,
                                                          uint32_t slow_cap,
                                                          const uint32_t *__restrict__ scount,
                                                          gt_detection *__restrict__ out,
                                                          uint32_t out_cap,
                                                          uint32_t *__restrict__ ocount,
                                                          uint32_t frame0) {
    const int fr = blockIdx.y;
    const int H = p.H, W = p.W;
    const size_t HW = static_cast<size_t>(H) * W;
    const float *P = proc + static_cast<size_t>(fr) * HW;
    const T *Rw = raw + static_cast<size_t>(fr) * HW;
    const uint32_t n = min(scount[fr * GT_COUNTER_STRIDE], slow_cap);
    const uint2 *fs = slow + static_cast<size_t>(fr) * slow_cap;
    gt_detection *fout = out + static_cast<size_t>(fr) * out_cap;
    uint32_t *counter = &ocount[fr * GT_COUNTER_STRIDE];
    const int rint = RI > 0 ? RI : p.rint;
    const uint32_t lane = __lane_id();
    const uint32_t stride = gridDim.x * WK_THREADS;
    for (uint32_t i0 = blockIdx.x * WK_THREADS + (threadIdx.x & ~63u); i0 < n; i0 += stride) {
        const uint32_t i = i0 + lane;
        bool keep = false;
        gt_detection d;
        if (i < n) {
            const uint2 e = fs[i];
            int pu = static_cast<int>(e.y / static_cast<uint32_t>(W));
            int pv = static_cast<int>(e.y - static_cast<uint32_t>(pu) * static_cast<uint32_t>(W));
            WalkRes res = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
            for (uint32_t it = FT_WALK_STEPS; it < p.max_iter; ++it) {
                res = walk_eval<RI, 0, (sizeof(T) <= 2)>(p, H, W, P, nullptr, 0, 0, 0, 1, 0, 1, 0, pu, pv);
                if (!walk_move(res, pu, pv, H, W, rint, p.shift_thr)) break;
            }
            keep = finish_walk<T, RI, 0, false>(p, H, W, P, Rw, nullptr, 0, 0, 0, 1, 0, 1, 0, res, pu, pv,
                                                e.x, frame0 + fr, d, frame_std(p, fr));
        }
        const uint32_t pos = wave_append(keep, counter);
        if (keep && pos < out_cap) fout[pos] = d;
    }
}

// Characterize every detection of the chunk, one lane per row, from the
// frames in global memory (rows of frame f: out[f * out_cap ...], count in
// the strided per-frame counter).
template <typename T, int RI>
__global__ __launch_bounds__(WK_THREADS) void k_characterize_det(const float *__restrict__ proc,
                                                                 const T *__restrict__ raw,
                                                                 GtKParams p,
                                                                 gt_detection *__restrict__ out,
                                                                 uint32_t out_cap,
                                                                 const uint32_t *__restrict__ ocount,
                                                                 uint32_t *__restrict__ dcount) {
    const int fr = blockIdx.y;
    if (dcount && blockIdx.x == 0 && threadIdx.x == 0) dcount[fr] = ocount[fr * GT_COUNTER_STRIDE];
    const int H = p.H, W = p.W;
    const size_t HW = static_cast<size_t>(H) * W;
    const GlobalSrc src{proc + static_cast<size_t>(fr) * HW, W};
    const T *Rw = raw + static_cast<size_t>(fr) * HW;
    const uint32_t n = min(ocount[fr * GT_COUNTER_STRIDE], out_cap);
    gt_detection *rows = out + static_cast<size_t>(fr) * out_cap;
    for (uint32_t i = blockIdx.x * WK_THREADS + threadIdx.x; i < n; i += gridDim.x * WK_THREADS) {
        gt_result_row row = rows[i].row;
        if constexpr (RI > 0) {
            // disk wholly inside the frame: every load issues at once
            const int cu = f2i(rintf(row.x)), cv = f2i(rintf(row.y));
            if (row.r <= 0.0f && p.rint == RI && cu >= RI && cu + RI < H && cv >= RI &&
                cv + RI < W) {
                const size_t c = static_cast<size_t>(cu) * W + cv;
                characterize_inner<T, RI>(GlobalRow{src.p + c, W}, Rw + c, W, row);
                rows[i].row = row;
                continue;
            }
        }
        characterize_row<T, RI, true>(src, Rw, H, W, p.circle, row);
        rows[i].row = row;
    }
}

// ---------------------------------------------------------------------------
// Learned proposal: the walk's end position predicted per candidate (d =
// (dy, dx) from the candidate pixel, confidence logit), checked by one walk
// evaluation.

// Integer offset of the walk's final centre for a predicted end offset d:
// the walk moves one pixel per step while the centroid offset exceeds thr,
// so it stops at the first centre within thr of the end position, the one
// nearest the start: sign(d) * max(0, ceil(|d| - thr)).
__device__ __forceinline__ int proposal_step(float d, float thr) {
    const float a = fabsf(d) - thr;
    const int k = a > 0.0f ? static_cast<int>(ceilf(a)) : 0;
    return d < 0.0f ? -k : k;
}

// One lane per listed candidate: entry i of frame f is list[(f * cap + i) *
// stride] (the candidate pixel) and, when stride is 2, the walk position
// after its first it0 steps in the next word.  A proposal with logit >=
// conf_logit is accepted when one walk evaluation at its centre does not
// move and lands within tol (both axes) of the predicted end position;
// otherwise the walk goes on from that position and step (the full walk).
// Rows as k_walk (minmass; characterize afterwards).  stats[0] += accepted
// proposals, stats[1] += listed candidates.
template <int RI>
__global__ __launch_bounds__(WK_THREADS) void k_walk_learned(
    const float *__restrict__ proc, GtKParams p, const uint32_t *__restrict__ cand,
    uint32_t stride, uint32_t it0, uint32_t cap, const uint32_t *__restrict__ ccount,
    const float4 *__restrict__ props, float conf_logit, float tol,
    gt_detection *__restrict__ out, uint32_t out_cap, uint32_t *__restrict__ ocount,
    unsigned long long *__restrict__ stats, uint32_t frame0) {
    const int fr = blockIdx.y;
    const int H = p.H, W = p.W;
    const size_t HW = static_cast<size_t>(H) * W;
    const float *P = proc + static_cast<size_t>(fr) * HW;
    const uint32_t n = min(ccount[fr * GT_COUNTER_STRIDE], cap);
    const uint32_t *fc = cand + static_cast<size_t>(fr) * cap * stride;
    const float4 *fp = props + static_cast<size_t>(fr) * cap;
    gt_detection *fout = out + static_cast<size_t>(fr) * out_cap;
    uint32_t *counter = &ocount[fr * GT_COUNTER_STRIDE];
    const int rint = RI > 0 ? RI : p.rint;
    const uint32_t lane = __lane_id();
    const uint32_t gstride = gridDim.x * WK_THREADS;
    uint32_t n_acc = 0;
    for (uint32_t i0 = blockIdx.x * WK_THREADS + (threadIdx.x & ~63u); i0 < n; i0 += gstride) {
        const uint32_t i = i0 + lane;
        bool keep = false;
        gt_detection d;
        if (i < n) {
            const uint32_t seed = fc[static_cast<size_t>(i) * stride];
            const uint32_t pos = fc[static_cast<size_t>(i) * stride + stride - 1];
            const int cu = static_cast<int>(seed / static_cast<uint32_t>(W));
            const int cv = static_cast<int>(seed - static_cast<uint32_t>(cu) * static_cast<uint32_t>(W));
            const float4 pr = fp[i];
            WalkRes res = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
            int pu = static_cast<int>(pos / static_cast<uint32_t>(W));
            int pv = static_cast<int>(pos - static_cast<uint32_t>(pu) * static_cast<uint32_t>(W));
            uint32_t it = it0;
            bool done = false, accepted = false;
            if (pr.z >= conf_logit && fabsf(pr.x) < 64.0f && fabsf(pr.y) < 64.0f) {
                // the centres the walk can reach from the candidate
                const int qu = min(max(cu + proposal_step(pr.x, p.shift_thr), min(cu, rint)),
                                   max(cu, H - 1 - rint));
                const int qv = min(max(cv + proposal_step(pr.y, p.shift_thr), min(cv, rint)),
                                   max(cv, W - 1 - rint));
                // a proposal away from a candidate whose walk has not started
                // is taken only if the walk leaves the candidate: a walk
                // that stops there ends there, wherever else is stable
                if (it == 0 && (qu != cu || qv != cv) && p.max_iter > 0) {
                    res = walk_eval<RI, 0>(p, H, W, P, nullptr, 0, 0, 0, 1, 0, 1, 0, pu, pv);
                    it = 1;
                    done = !walk_move(res, pu, pv, H, W, rint, p.shift_thr);
                }
                if (!done) {
                    const WalkRes r =
                        walk_eval<RI, 0>(p, H, W, P, nullptr, 0, 0, 0, 1, 0, 1, 0, qu, qv);
                    int tu = qu, tv = qv;
                    const bool moves = walk_move(r, tu, tv, H, W, rint, p.shift_thr);
                    const float ey = r.x + static_cast<float>(qu) - (static_cast<float>(cu) + pr.x);
                    const float ex = r.y + static_cast<float>(qv) - (static_cast<float>(cv) + pr.y);
                    if (!moves && fabsf(ey) <= tol && fabsf(ex) <= tol) {
                        accepted = done = true;
                        res = r;
                        pu = qu;
                        pv = qv;
                    }
                }
            }
            if (!done) {
                for (; it < p.max_iter; ++it) {
                    res = walk_eval<RI, 0>(p, H, W, P, nullptr, 0, 0, 0, 1, 0, 1, 0, pu, pv);
                    if (!walk_move(res, pu, pv, H, W, rint, p.shift_thr)) break;
                }
            }
            n_acc += accepted ? 1u : 0u;
            keep = finish_walk<uint16_t, RI, 0, false>(p, H, W, P, nullptr, nullptr, 0, 0, 0, 1, 0,
                                                       1, 0, res, pu, pv, seed, frame0 + fr, d,
                                                       frame_std(p, fr));
        }
        const uint32_t pos = wave_append(keep, counter);
        if (keep && pos < out_cap) fout[pos] = d;
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) n_acc += __shfl_xor(n_acc, off);
    if (lane == 0 && n_acc) atomicAdd(&stats[0], static_cast<unsigned long long>(n_acc));
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&stats[1], static_cast<unsigned long long>(n));
}

// Training targets: the full walk from every candidate (the plain method),
// target (end row - candidate row, end column - candidate column, 1 if the
// walk converged, the candidate's raster index as uint32 bits) at flat
// index offs[f] + i (dropped at maxn and beyond).  A walk that does not
// converge (max_iter reached, or no centroid: zero mass) gets offset (0, 0)
// and label 0.  minmass plays no part: it is applied after the verifying
// evaluation, exactly as after the walk.
template <int RI>
__global__ __launch_bounds__(WK_THREADS) void k_cnn_label(
    const float *__restrict__ proc, GtKParams p, const uint32_t *__restrict__ cand, uint32_t cap,
    const uint32_t *__restrict__ ccount, const uint32_t *__restrict__ offs, uint32_t maxn,
    float4 *__restrict__ targets) {
    const int fr = blockIdx.y;
    const int H = p.H, W = p.W;
    const float *P = proc + static_cast<size_t>(fr) * H * W;
    const uint32_t n = min(ccount[fr * GT_COUNTER_STRIDE], cap);
    const uint32_t *fc = cand + static_cast<size_t>(fr) * cap;
    const int rint = RI > 0 ? RI : p.rint;
    for (uint32_t i = blockIdx.x * WK_THREADS + threadIdx.x; i < n; i += gridDim.x * WK_THREADS) {
        const uint32_t dst = offs[fr] + i;
        if (dst >= maxn) break;
        const uint32_t seed = fc[i];
        const int cu = static_cast<int>(seed / static_cast<uint32_t>(W));
        const int cv = static_cast<int>(seed - static_cast<uint32_t>(cu) * static_cast<uint32_t>(W));
        int pu = cu, pv = cv;
        WalkRes res = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        bool stopped = false;
        for (uint32_t it = 0; it < p.max_iter; ++it) {
            res = walk_eval<RI, 0>(p, H, W, P, nullptr, 0, 0, 0, 1, 0, 1, 0, pu, pv);
            if (!walk_move(res, pu, pv, H, W, rint, p.shift_thr)) {
                stopped = true;
                break;
            }
        }
        res.x = res.x + static_cast<float>(pu);
        res.y = res.y + static_cast<float>(pv);
        // label 1: the walk converged (stopped by the move rule within
        // max_iter, with a centroid); otherwise offset (0, 0), label 0
        const float ty = res.x - static_cast<float>(cu), tx = res.y - static_cast<float>(cv);
        const bool ok = stopped && isfinite(ty) && isfinite(tx);
        targets[dst] = make_float4(ok ? ty : 0.0f, ok ? tx : 0.0f, ok ? 1.0f : 0.0f,
                                   __uint_as_float(seed));
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

// Calls f(integral_constant R, bool_constant two-operation division) for
// the integer bandpass's radius 1..7 and division form (fast_div 2: the
// two-operation quotient, checked for both norms).
template <class F>
void with_bandpass_form(const GtKParams &p, F f) {
    auto by_r = [&](auto d2) {
        switch (p.R) {
        case 1: f(std::integral_constant<int, 1>{}, d2); break;
        case 2: f(std::integral_constant<int, 2>{}, d2); break;
        case 3: f(std::integral_constant<int, 3>{}, d2); break;
        case 4: f(std::integral_constant<int, 4>{}, d2); break;
        case 5: f(std::integral_constant<int, 5>{}, d2); break;
        case 6: f(std::integral_constant<int, 6>{}, d2); break;
        default: f(std::integral_constant<int, 7>{}, d2); break;
        }
    };
    if (p.fast_div == 2) by_r(std::true_type{});
    else by_r(std::false_type{});
}

template <typename T>
hipError_t launch_bandpass_t(const T *frames, float *proc, uint32_t nframes, const GtKParams &p,
                             hipStream_t s, const GtZero &z) {
    if constexpr (sizeof(T) <= 2) {
        if (p.fast_div && p.R >= 1 && p.R <= 7) {
            const size_t ntiles = static_cast<size_t>((p.W + B2_TW - 1) / B2_TW) *
                                  ((p.H + B2_TH - 1) / B2_TH) * nframes;
            const dim3 g2(xcd_grid(ntiles));
            const int nf = static_cast<int>(nframes);
            with_bandpass_form(p, [&](auto r, auto d2) {
                k_bandpass_int<T, decltype(r)::value, decltype(d2)::value>
                    <<<g2, B2_THREADS, 0, s>>>(frames, proc, p, nf, z);
            });
            return hipGetLastError();
        }
    }
    for (int a = 0; a < 2; ++a)
        if (z.n[a]) k_zero_u32<<<(z.n[a] + 255) / 256, 256, 0, s>>>(z.p[a], z.n[a]);
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
                              const GtKParams &p, hipStream_t s, const GtZero *zero) {
    const GtZero z = zero ? *zero : GtZero{{nullptr, nullptr}, {0u, 0u}};
    switch (dtype) {
    case GT_U16: return launch_bandpass_t(static_cast<const uint16_t *>(frames), proc, nframes, p, s, z);
    case GT_F32: return launch_bandpass_t(static_cast<const float *>(frames), proc, nframes, p, s, z);
    case GT_U8: return launch_bandpass_t(static_cast<const uint8_t *>(frames), proc, nframes, p, s, z);
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

size_t gt_tile_lds_bytes(const GtKParams &p) {
    const TileGeom g = tile_geom(p.k, p.rint);
    const size_t dyn = sizeof(float) * (static_cast<size_t>(g.rows) * g.pitch +
                                        static_cast<size_t>(FT_TH + p.k - 1) * FT_TW);
    return dyn + 256;
}

namespace {
template <typename T>
hipError_t launch_detect_tile_t(const float *proc, const T *raw, uint32_t nframes,
                                const GtKParams &p, gt_detection *out, uint32_t out_cap,
                                uint32_t *ocount, uint2 *slow, uint32_t slow_cap,
                                uint32_t *scount, uint32_t frame0, uint32_t steps, hipStream_t s) {
    // the kernel decodes tile indices below 2^22: larger launches are split
    // into frame groups
    const size_t tpf = static_cast<size_t>((p.W + FT_TW - 1) / FT_TW) * ((p.H + FT_TH - 1) / FT_TH);
    const size_t fmax = tpf ? std::max<size_t>(1, ((size_t(1) << 22) - 1) / tpf) : nframes;
    if (nframes > fmax) {
        const size_t HW = static_cast<size_t>(p.H) * p.W;
        for (uint32_t f = 0; f < nframes; f += static_cast<uint32_t>(fmax)) {
            const uint32_t nf = std::min<uint32_t>(static_cast<uint32_t>(fmax), nframes - f);
            GtKParams pf = p;
            if (pf.fstd) pf.fstd += f;   // the group's per-frame noise estimates
            const hipError_t e = launch_detect_tile_t(
                proc + f * HW, raw + f * HW, nf, pf, out + static_cast<size_t>(f) * out_cap, out_cap,
                ocount + static_cast<size_t>(f) * GT_COUNTER_STRIDE,
                slow + static_cast<size_t>(f) * slow_cap, slow_cap,
                scount + static_cast<size_t>(f) * GT_COUNTER_STRIDE, frame0 + f, steps, s);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }
    const TileGeom g = tile_geom(p.k, p.rint);
    const size_t ntiles = static_cast<size_t>((p.W + FT_TW - 1) / FT_TW) *
                          ((p.H + FT_TH - 1) / FT_TH) * nframes;
    const dim3 grid(xcd_grid(ntiles));
    const int nf = static_cast<int>(nframes);
    const size_t lds = sizeof(float) * (static_cast<size_t>(g.rows) * g.pitch +
                                        static_cast<size_t>(FT_TH + p.k - 1) * FT_TW);
#define GT_TILE(KV, RV)                                                                          \
    do {                                                                                         \
        if (steps == 1)                                                                          \
            k_detect_tile<T, KV, RV, 1><<<grid, FT_THREADS, lds, s>>>(                           \
                proc, raw, p, g, out, out_cap, ocount, slow, slow_cap, scount, frame0, nf);       \
        else                                                                                     \
            k_detect_tile<T, KV, RV><<<grid, FT_THREADS, lds, s>>>(                              \
                proc, raw, p, g, out, out_cap, ocount, slow, slow_cap, scount, frame0, nf);       \
    } while (0)
    // diameter 9 with separation 7 (k = 9) or 9 (k = 12)
    if (p.rint == 4 && p.k == 9) {
        GT_TILE(9, 4);
    } else if (p.rint == 4 && p.k == 12) {
        GT_TILE(12, 4);
    } else if (p.rint == 4) {
        GT_TILE(0, 4);
    } else {
        GT_TILE(0, 0);
    }
#undef GT_TILE
    return hipGetLastError();
}

template <typename T>
hipError_t launch_walk_slow_t(const float *proc, const T *raw, uint32_t nframes,
                              const GtKParams &p, const uint2 *slow, uint32_t slow_cap,
                              const uint32_t *scount, gt_detection *out, uint32_t out_cap,
                              uint32_t *ocount, uint32_t frame0, hipStream_t s) {
    const size_t hw = static_cast<size_t>(p.H) * p.W;
    uint32_t blocks = static_cast<uint32_t>((hw / 1024 + WK_THREADS - 1) / WK_THREADS);
    blocks = blocks < 1 ? 1 : (blocks > 256 ? 256 : blocks);
    const dim3 grid(blocks, nframes, 1);
    if (p.rint == 4)
        k_walk_slow<T, 4><<<grid, WK_THREADS, 0, s>>>(proc, raw, p, slow, slow_cap, scount, out,
                                                      out_cap, ocount, frame0);
    else
        k_walk_slow<T, 0><<<grid, WK_THREADS, 0, s>>>(proc, raw, p, slow, slow_cap, scount, out,
                                                      out_cap, ocount, frame0);
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
                                 uint32_t *ocount, uint2 *slow, uint32_t slow_cap,
                                 uint32_t *scount, uint32_t frame0, hipStream_t s,
                                 uint32_t walk_steps) {
    if (walk_steps != 1 && walk_steps != FT_WALK_STEPS) return hipErrorInvalidValue;
    switch (dtype) {
    case GT_U16:
        return launch_detect_tile_t(proc, static_cast<const uint16_t *>(raw), nframes, p, out,
                                    out_cap, ocount, slow, slow_cap, scount, frame0, walk_steps, s);
    case GT_F32:
        return launch_detect_tile_t(proc, static_cast<const float *>(raw), nframes, p, out,
                                    out_cap, ocount, slow, slow_cap, scount, frame0, walk_steps, s);
    case GT_U8:
        return launch_detect_tile_t(proc, static_cast<const uint8_t *>(raw), nframes, p, out,
                                    out_cap, ocount, slow, slow_cap, scount, frame0, walk_steps, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t gt_launch_walk_slow(const float *proc, const void *raw, int dtype, uint32_t nframes,
                               const GtKParams &p, const uint2 *slow, uint32_t slow_cap,
                               const uint32_t *scount, gt_detection *out, uint32_t out_cap,
                               uint32_t *ocount, uint32_t frame0, hipStream_t s) {
    switch (dtype) {
    case GT_U16:
        return launch_walk_slow_t(proc, static_cast<const uint16_t *>(raw), nframes, p, slow,
                                  slow_cap, scount, out, out_cap, ocount, frame0, s);
    case GT_F32:
        return launch_walk_slow_t(proc, static_cast<const float *>(raw), nframes, p, slow,
                                  slow_cap, scount, out, out_cap, ocount, frame0, s);
    case GT_U8:
        return launch_walk_slow_t(proc, static_cast<const uint8_t *>(raw), nframes, p, slow,
                                  slow_cap, scount, out, out_cap, ocount, frame0, s);
    default: return hipErrorInvalidValue;
    }
}

namespace {
template <typename T>
hipError_t launch_detect_cand_t(const float *proc, uint32_t nframes, const GtKParams &p,
                                uint32_t *cand, uint32_t cap, uint32_t *ccount, hipStream_t s) {
    const TileGeom g = tile_geom(p.k, p.rint);
    const size_t ntiles = static_cast<size_t>((p.W + FT_TW - 1) / FT_TW) *
                          ((p.H + FT_TH - 1) / FT_TH) * nframes;
    const dim3 grid(xcd_grid(ntiles));
    const int nf = static_cast<int>(nframes);
    const size_t lds = sizeof(float) * (static_cast<size_t>(g.rows) * g.pitch +
                                        static_cast<size_t>(FT_TH + p.k - 1) * FT_TW);
#define GT_CAND(KV, RV) \
    k_detect_cand<T, KV, RV><<<grid, FT_THREADS, lds, s>>>(proc, p, g, cand, cap, ccount, nf)
    if (p.rint == 4 && p.k == 9) GT_CAND(9, 4);
    else if (p.rint == 4 && p.k == 12) GT_CAND(12, 4);
    else if (p.rint == 4) GT_CAND(0, 4);
    else GT_CAND(0, 0);
#undef GT_CAND
    return hipGetLastError();
}

dim3 walk_grid(const GtKParams &p, uint32_t nframes) {
    const size_t hw = static_cast<size_t>(p.H) * p.W;
    uint32_t blocks = static_cast<uint32_t>((hw / 512 + WK_THREADS - 1) / WK_THREADS);
    blocks = blocks < 1 ? 1 : (blocks > 256 ? 256 : blocks);
    return dim3(blocks, nframes, 1);
}
}  // namespace

hipError_t gt_launch_detect_cand(const float *proc, int dtype, uint32_t nframes,
                                 const GtKParams &p, uint32_t *cand, uint32_t cap,
                                 uint32_t *ccount, hipStream_t s) {
    switch (dtype) {
    case GT_U16: return launch_detect_cand_t<uint16_t>(proc, nframes, p, cand, cap, ccount, s);
    case GT_F32: return launch_detect_cand_t<float>(proc, nframes, p, cand, cap, ccount, s);
    case GT_U8: return launch_detect_cand_t<uint8_t>(proc, nframes, p, cand, cap, ccount, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t gt_launch_walk_learned(const float *proc, uint32_t nframes, const GtKParams &p,
                                  const uint32_t *cand, uint32_t stride, uint32_t it0,
                                  uint32_t cap, const uint32_t *ccount, const float4 *props,
                                  float conf_logit, float tol, gt_detection *out,
                                  uint32_t out_cap, uint32_t *ocount, unsigned long long *stats,
                                  uint32_t frame0, hipStream_t s) {
    if (stride != 1 && stride != 2) return hipErrorInvalidValue;
    const dim3 grid = walk_grid(p, nframes);
    if (p.rint == 4)
        k_walk_learned<4><<<grid, WK_THREADS, 0, s>>>(proc, p, cand, stride, it0, cap, ccount,
                                                      props, conf_logit, tol, out, out_cap,
                                                      ocount, stats, frame0);
    else
        k_walk_learned<0><<<grid, WK_THREADS, 0, s>>>(proc, p, cand, stride, it0, cap, ccount,
                                                      props, conf_logit, tol, out, out_cap,
                                                      ocount, stats, frame0);
    return hipGetLastError();
}

hipError_t gt_launch_cnn_label(const float *proc, uint32_t nframes, const GtKParams &p,
                               const uint32_t *cand, uint32_t cap, const uint32_t *ccount,
                               const uint32_t *offs, uint32_t maxn, float4 *targets,
                               hipStream_t s) {
    const dim3 grid = walk_grid(p, nframes);
    if (p.rint == 4)
        k_cnn_label<4><<<grid, WK_THREADS, 0, s>>>(proc, p, cand, cap, ccount, offs, maxn, targets);
    else
        k_cnn_label<0><<<grid, WK_THREADS, 0, s>>>(proc, p, cand, cap, ccount, offs, maxn, targets);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Post-processing of a chunk's detections on the GPU (Trackpy style):
// filter_close and doughnut photometry, one float4 per row: (raw_mass,
// bg_median, raw_mass_corrected, keep).
namespace {

constexpr int FC_THREADS = 256;

// filter_close_trackpy (execute_gpu.rs:556-590): row i of a frame is dropped
// when a row j != i of the same frame lies at d^2 < separation^2 (d^2 =
// (0 + dx^2) + dy^2 in f32, dx = x_j - x_i) with a larger mass, or the same
// mass and a smaller seed (rows are ordered by seed).  Every pair is tested:
// rows of a frame are staged through LDS in tiles.
__global__ __launch_bounds__(FC_THREADS) void k_filter_close_rows(
    const gt_detection *__restrict__ rows, uint32_t cap, const uint32_t *__restrict__ counts,
    float r2, float4 *__restrict__ post) {
    __shared__ float sx[FC_THREADS], sy[FC_THREADS], sm[FC_THREADS];
    __shared__ uint32_t ss[FC_THREADS];
    const uint32_t f = blockIdx.y;
    const uint32_t n = min(counts[f], cap);
    const gt_detection *fr = rows + static_cast<size_t>(f) * cap;
    float4 *fp = post + static_cast<size_t>(f) * cap;
    for (uint32_t i0 = blockIdx.x * FC_THREADS; i0 < n; i0 += gridDim.x * FC_THREADS) {
        const uint32_t i = i0 + threadIdx.x;
        float xi = 0.0f, yi = 0.0f, mi = 0.0f;
        uint32_t si = 0;
        bool fin = false;
        if (i < n) {
            xi = fr[i].row.x;
            yi = fr[i].row.y;
            mi = fr[i].row.mass;
            si = fr[i].seed;
            fin = isfinite(xi) && isfinite(yi);
        }
        bool drop = false;
        for (uint32_t j0 = 0; j0 < n; j0 += FC_THREADS) {
            __syncthreads();
            const uint32_t j = j0 + threadIdx.x;
            if (j < n) {
                sx[threadIdx.x] = fr[j].row.x;
                sy[threadIdx.x] = fr[j].row.y;
                sm[threadIdx.x] = fr[j].row.mass;
                ss[threadIdx.x] = fr[j].seed;
            }
            __syncthreads();
            const uint32_t m = min(static_cast<uint32_t>(FC_THREADS), n - j0);
            if (fin && !drop) {
                for (uint32_t t = 0; t < m; ++t) {
                    if (ss[t] == si) continue;       // the row itself (seeds are unique)
                    const float ddx = sx[t] - xi, ddy = sy[t] - yi;
                    float d2 = 0.0f;
                    d2 = d2 + ddx * ddx;
                    d2 = d2 + ddy * ddy;
                    if (!(d2 < r2)) continue;
                    const float mj = sm[t];
                    if (mi < mj || (mi == mj && si > ss[t])) {
                        drop = true;
                        break;
                    }
                }
            }
        }
        if (i < n) fp[i] = make_float4(0.0f, 0.0f, 0.0f, drop ? 0.0f : 1.0f);
    }
}

// Doughnut photometry (execute_gpu.rs:717-790; stencils kernels.rs:138-171):
// centre (round half away from zero); raw_mass = in-order f32 sum of the raw
// frame over the disk offsets (i-major); bg_median = element len/2 of the
// sorted in-frame ring values; corrected = raw_mass - median * count.  One
// wave per kept row; the ring values are ranked against each other in LDS.
template <typename T>
__global__ __launch_bounds__(64) void k_doughnut_rows(
    const T *__restrict__ raw, int H, int W, const gt_detection *__restrict__ rows, uint32_t cap,
    const uint32_t *__restrict__ counts, const short2 *__restrict__ offs, uint32_t nsig,
    uint32_t nring, float4 *__restrict__ post) {
    extern __shared__ float dl[];
    float *sig = dl;             // nsig
    float *ring = dl + nsig;     // nring (ring values of in-frame offsets, compacted)
    __shared__ uint32_t s_nv;
    __shared__ float s_med;
    const uint32_t f = blockIdx.y;
    const uint32_t n = min(counts[f], cap);
    const T *F = raw + static_cast<size_t>(f) * H * W;
    const gt_detection *fr = rows + static_cast<size_t>(f) * cap;
    float4 *fp = post + static_cast<size_t>(f) * cap;
    const uint32_t lane = threadIdx.x;
    // the largest offset of either stencil (the ring's outer radius)
    int rmax = 0;
    for (uint32_t k = lane; k < nsig + nring; k += 64)
        rmax = max(rmax, max(abs(static_cast<int>(offs[k].x)), abs(static_cast<int>(offs[k].y))));
    for (int o = 32; o >= 1; o >>= 1) rmax = max(rmax, __shfl_xor(rmax, o));
    for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
        if (fp[i].w == 0.0f) continue;       // dropped by filter_close (block-uniform)
        const int64_t mu = f2i(roundf(fr[i].row.x)), mv = f2i(roundf(fr[i].row.y));
        // every offset lands in the frame: no bounds tests, the ring in
        // offset order
        const bool all_in = mu - rmax >= 0 && mu + rmax < H && mv - rmax >= 0 && mv + rmax < W;
        uint32_t nv;
        if (all_in) {
            for (uint32_t k = lane; k < nsig; k += 64)
                sig[k] = to_f32(F[(mu + offs[k].x) * W + (mv + offs[k].y)]);
            for (uint32_t k = lane; k < nring; k += 64)
                ring[k] = to_f32(F[(mu + offs[nsig + k].x) * W + (mv + offs[nsig + k].y)]);
            nv = nring;
        } else {
            for (uint32_t k = lane; k < nsig; k += 64) {
                const int64_t pu = mu + offs[k].x, pv = mv + offs[k].y;
                sig[k] = (pu >= 0 && pu < H && pv >= 0 && pv < W) ? to_f32(F[pu * W + pv]) : 0.0f;
            }
            if (lane == 0) s_nv = 0;
            __syncthreads();
            for (uint32_t k = lane; k < nring; k += 64) {
                const int64_t pu = mu + offs[nsig + k].x, pv = mv + offs[nsig + k].y;
                if (pu >= 0 && pu < H && pv >= 0 && pv < W)
                    ring[atomicAdd(&s_nv, 1u)] = to_f32(F[pu * W + pv]);
            }
            __syncthreads();
            nv = s_nv;
        }
        const uint32_t kth = nv / 2;
        if (lane == 0) s_med = 0.0f;
        __syncthreads();
        // rank each ring value against all of them, up to four values per
        // lane against one pass over the ring
        for (uint32_t a0 = 0; a0 < nv; a0 += 4 * 64) {
            float v[4];
            uint32_t less[4] = {0u, 0u, 0u, 0u}, leq[4] = {0u, 0u, 0u, 0u};
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint32_t a = a0 + lane + 64u * j;
                v[j] = a < nv ? ring[a] : 0.0f;
            }
            for (uint32_t b = 0; b < nv; ++b) {
                const float w = ring[b];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    less[j] += w < v[j] ? 1u : 0u;
                    leq[j] += w <= v[j] ? 1u : 0u;
                }
            }
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (a0 + lane + 64u * j < nv && less[j] <= kth && kth < leq[j])
                    s_med = v[j];   // equal values: the same float
        }
        __syncthreads();
        if (lane == 0) {
            float sm = 0.0f;
            uint32_t ns = 0;
            if (all_in) {
                for (uint32_t k = 0; k < nsig; ++k) sm = sm + sig[k];
                ns = nsig;
            } else {
                for (uint32_t k = 0; k < nsig; ++k) {
                    const int64_t pu = mu + offs[k].x, pv = mv + offs[k].y;
                    if (pu < 0 || pu >= H || pv < 0 || pv >= W) continue;
                    sm = sm + sig[k];
                    ++ns;
                }
            }
            const float med = nv ? s_med : 0.0f;
            fp[i] = make_float4(sm, med, sm - med * static_cast<float>(ns), 1.0f);
        }
        __syncthreads();
    }
}

}  // namespace

hipError_t gt_launch_filter_close_rows(const gt_detection *rows, uint32_t cap,
                                       const uint32_t *counts, uint32_t nframes, float separation,
                                       float4 *post, hipStream_t s) {
    if (nframes == 0) return hipSuccess;
    k_filter_close_rows<<<dim3(8, nframes), FC_THREADS, 0, s>>>(rows, cap, counts,
                                                                separation * separation, post);
    return hipGetLastError();
}

namespace {

// Packs frame f's rows (rows[f * cap ..], min(counts[f], cap) of them) and
// their post-processing values after those of frames 0..f-1, so that a
// chunk's rows read back in one copy.  One workgroup per frame; the frame's
// offset is the sum of the counts before it (a chunk has few frames).
constexpr int PK_THREADS = 256;
static_assert(sizeof(gt_detection) % sizeof(uint32_t) == 0, "rows copy as words");
__global__ __launch_bounds__(PK_THREADS) void k_pack_rows(const gt_detection *__restrict__ rows,
                                                          uint32_t cap,
                                                          const uint32_t *__restrict__ counts,
                                                          const float4 *__restrict__ post,
                                                          gt_detection *__restrict__ orows,
                                                          float4 *__restrict__ opost) {
    const uint32_t f = blockIdx.x;
    uint32_t off = 0;
    for (uint32_t g = threadIdx.x; g < f; g += PK_THREADS) off += min(counts[g], cap);
    for (int o = 32; o >= 1; o >>= 1) off += __shfl_xor(off, o);
    __shared__ uint32_t s_off[PK_THREADS / 64];
    if ((threadIdx.x & 63u) == 0) s_off[threadIdx.x >> 6] = off;
    __syncthreads();
    off = 0;
#pragma unroll
    for (int w = 0; w < PK_THREADS / 64; ++w) off += s_off[w];
    const uint32_t n = min(counts[f], cap);
    constexpr uint32_t WPR = sizeof(gt_detection) / sizeof(uint32_t);
    const uint32_t *src = reinterpret_cast<const uint32_t *>(rows + static_cast<size_t>(f) * cap);
    uint32_t *dst = reinterpret_cast<uint32_t *>(orows + off);
    for (uint32_t i = threadIdx.x; i < n * WPR; i += PK_THREADS) dst[i] = src[i];
    if (post) {
        const float4 *ps = post + static_cast<size_t>(f) * cap;
        for (uint32_t i = threadIdx.x; i < n; i += PK_THREADS) opost[off + i] = ps[i];
    }
}

}  // namespace

hipError_t gt_launch_pack_rows(const gt_detection *rows, uint32_t cap, const uint32_t *counts,
                               uint32_t nframes, const float4 *post, gt_detection *orows,
                               float4 *opost, hipStream_t s) {
    if (nframes == 0) return hipSuccess;
    k_pack_rows<<<nframes, PK_THREADS, 0, s>>>(rows, cap, counts, post, orows, opost);
    return hipGetLastError();
}

size_t gt_doughnut_lds_bytes(uint32_t nsig, uint32_t nring) {
    return sizeof(float) * (static_cast<size_t>(nsig) + nring);
}

hipError_t gt_launch_doughnut_rows(const void *raw, int dtype, int H, int W,
                                   const gt_detection *rows, uint32_t cap, const uint32_t *counts,
                                   uint32_t nframes, const short2 *offs, uint32_t nsig,
                                   uint32_t nring, float4 *post, hipStream_t s) {
    if (nframes == 0) return hipSuccess;
    const size_t lds = gt_doughnut_lds_bytes(nsig, nring);
    const dim3 grid(256, nframes);
#define GT_DOUGH(TT)                                                                          \
    do {                                                                                      \
        if (lds > 65536) {                                                                    \
            hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_doughnut_rows<TT>), \
                                               hipFuncAttributeMaxDynamicSharedMemorySize,     \
                                               static_cast<int>(lds));                         \
            if (e != hipSuccess) return e;                                                    \
        }                                                                                     \
        k_doughnut_rows<TT><<<grid, 64, lds, s>>>(static_cast<const TT *>(raw), H, W, rows,   \
                                                  cap, counts, offs, nsig, nring, post);      \
    } while (0)
    switch (dtype) {
    case GT_U16: GT_DOUGH(uint16_t); break;
    case GT_F32: GT_DOUGH(float); break;
    case GT_U8: GT_DOUGH(uint8_t); break;
    default: return hipErrorInvalidValue;
    }
#undef GT_DOUGH
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Noise estimate of raw frames for the snr / minmass_snr rules
// (src/utils.rs:298-670; sum_image_outplace.wgsl, sum_image_inplace.wgsl,
// std_image.wgsl, last_reduction.wgsl).  Eight-way tree per frame: a level
// of length len has stride s = ceil(len / 8) and out[i] = sum_{j<8} in[i +
// s j] in j order; levels repeat while len > 24; the last values are summed
// in order; mean = sum / f32(N), std = sqrt(sum of squared deviations /
// f32(N - 1)).  NaN-marker pixels (bits 0xFFFFFFFF) add nothing.  Indices
// past a level's length are skipped.
namespace {

__device__ __forceinline__ bool is_marker(float v) { return __float_as_uint(v) == 0xFFFFFFFFu; }

// First level over the frame: the values (kVar false) or the squared
// deviations from mean[f] (kVar true); tmp row f has pitch s.
template <typename T, bool kVar>
__global__ __launch_bounds__(256) void k_std_first(const T *__restrict__ frames, size_t n,
                                                   uint32_t s, const float *__restrict__ mean,
                                                   float *__restrict__ tmp) {
    const uint32_t f = blockIdx.y;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s) return;
    const T *x = frames + static_cast<size_t>(f) * n;
    const float m = kVar ? mean[f] : 0.0f;
    float sum = 0.0f;
#pragma unroll
    for (uint32_t j = 0; j < 8; ++j) {
        const size_t k = i + static_cast<size_t>(s) * j;
        if (k < n) {
            const float d = to_f32(x[k]);
            if constexpr (kVar) {
                float v = 0.0f;
                if (!is_marker(d)) {
                    const float e = d - m;
                    v = e * e;
                }
                sum = sum + v;
            } else {
                if (!is_marker(d)) sum = sum + d;
            }
        }
    }
    tmp[static_cast<size_t>(f) * s + i] = sum;
}

// Next level in place: row f (pitch) of length len -> its first s values.
// Thread i writes index i < s and reads i + s j: no index is written by one
// thread and read by another.
__global__ __launch_bounds__(256) void k_std_level(float *__restrict__ tmp, size_t pitch,
                                                   uint32_t len, uint32_t s) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s) return;
    float *t = tmp + static_cast<size_t>(blockIdx.y) * pitch;
    float sum = 0.0f;
#pragma unroll
    for (uint32_t j = 0; j < 8; ++j) {
        const uint32_t k = i + s * j;
        if (k < len) sum = sum + t[k];
    }
    t[i] = sum;
}

// div[f] (when given) replaces the pixel count as the divisor: mean div,
// variance div - 1 (the adaptive filter's count of unmasked pixels)
__global__ void k_std_last(const float *__restrict__ tmp, size_t pitch, uint32_t len, size_t n,
                           int var, uint32_t nframes, const uint32_t *__restrict__ div,
                           float *__restrict__ out) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nframes) return;
    const float *t = tmp + static_cast<size_t>(f) * pitch;
    float sum = 0.0f;
    for (uint32_t i = 0; i < len; ++i) sum = sum + t[i];
    const uint32_t d = div ? div[f] : static_cast<uint32_t>(n);
    if (var) out[f] = sqrtf(sum / static_cast<float>(d - 1u));
    else out[f] = sum / static_cast<float>(div ? d : n);
}

template <typename T>
hipError_t launch_frame_std_t(const T *frames, uint32_t nframes, size_t n, float *tmp,
                              float *mean, float *std_out, hipStream_t s,
                              const uint32_t *div = nullptr) {
    const uint32_t s0 = static_cast<uint32_t>((n + 7) / 8);
    const dim3 g0((s0 + 255) / 256, nframes);
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 0)
            k_std_first<T, false><<<g0, 256, 0, s>>>(frames, n, s0, nullptr, tmp);
        else
            k_std_first<T, true><<<g0, 256, 0, s>>>(frames, n, s0, mean, tmp);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        uint32_t len = s0;
        while (len > 24) {
            const uint32_t st = (len + 7) / 8;
            k_std_level<<<dim3((st + 255) / 256, nframes), 256, 0, s>>>(tmp, s0, len, st);
            e = hipGetLastError();
            if (e != hipSuccess) return e;
            len = st;
        }
        k_std_last<<<(nframes + 63) / 64, 64, 0, s>>>(tmp, s0, len, n, pass, nframes, div,
                                                       pass == 0 ? mean : std_out);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}
}  // namespace

namespace {
// ---- adaptive_background (adaptive_filter.wgsl, execute_gpu.rs:290-330) ----
template <typename T>
__global__ void k_widen_frames(const T *__restrict__ src, size_t n, float *__restrict__ dst) {
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x)
        dst[i] = to_f32(src[i]);
}

__global__ void k_fill_counter(uint32_t *__restrict__ c, uint32_t n, uint32_t v) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) c[i] = v;
}

// One pass over frame f's rows not yet accepted (pass 0: every row): a row
// is accepted when mass >= (minmass, or minmass_snr * count * std) and
// max_intensity >= snr * std; acceptance marks its walk disk (centre
// rounded half to even) in the f32 copy of the frame with the NaN marker and
// takes u32(count) off the frame's divisor.  state: 1 accepted, 0 not yet.
__global__ __launch_bounds__(256) void k_adapt_filter(
    const gt_detection *__restrict__ rows, uint32_t cap, const uint32_t *__restrict__ ocount,
    GtKParams p, const float *__restrict__ fstd, uint32_t pass, float *__restrict__ masked,
    uint8_t *__restrict__ state, uint32_t *__restrict__ counter) {
    const uint32_t f = blockIdx.y;
    const uint32_t n = min(ocount[f * GT_COUNTER_STRIDE], cap);
    const int H = p.H, W = p.W;
    const float sd = fstd[f];
    float *M = masked + static_cast<size_t>(f) * H * W;
    const gt_detection *fr = rows + static_cast<size_t>(f) * cap;
    uint8_t *st = state + static_cast<size_t>(f) * cap;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (pass > 0 && st[i]) continue;
        const gt_result_row row = fr[i].row;
        float mm;
        if (p.minmass > 0.0f) mm = p.minmass;
        else mm = p.minmass_snr * row.count * sd;
        if (row.mass < mm || row.max_intensity < p.snr * sd) {
            st[i] = 0;
            continue;
        }
        st[i] = 1;
        // LoG rows carry their radius (set_nan, adaptive_filter.wgsl:56-63)
        const int rint = row.r > 0.0f ? f2i(ceilf(row.r)) : p.rint;
        const int u = f2i(rintf(row.x)), v = f2i(rintf(row.y));
        const float r = static_cast<float>(rint), r2 = r * r;
        for (int a = -rint; a <= rint; ++a) {
            const float x = static_cast<float>(a), x2 = x * x;
            for (int b = -rint; b <= rint; ++b) {
                const float y = static_cast<float>(b), y2 = y * y;
                if (x2 + y2 > r2) continue;
                const int pu = u + a, pv = v + b;
                if (pu < 0 || pu >= H || pv < 0 || pv >= W) continue;
                M[static_cast<size_t>(pu) * W + pv] = __uint_as_float(0xFFFFFFFFu);
            }
        }
        atomicSub(&counter[f], static_cast<uint32_t>(row.count));
    }
}

// Keeps frame f's accepted rows, in order, at the front of its slots; one
// block per frame walks the rows in tiles (every read of a tile precedes its
// writes, and a tile's destinations lie below its own end).  The frame's
// count becomes the accepted count (a count above cap is left as it is: the
// capacity error stands).
__global__ __launch_bounds__(256) void k_adapt_compact(gt_detection *__restrict__ rows,
                                                       uint32_t cap,
                                                       uint32_t *__restrict__ ocount,
                                                       const uint8_t *__restrict__ state) {
    __shared__ uint32_t s_scan[256];
    const uint32_t f = blockIdx.x;
    const uint32_t nc = ocount[f * GT_COUNTER_STRIDE];
    if (nc > cap) return;
    gt_detection *fr = rows + static_cast<size_t>(f) * cap;
    const uint8_t *st = state + static_cast<size_t>(f) * cap;
    uint32_t off = 0;
    for (uint32_t t0 = 0; t0 < nc; t0 += 256) {
        const uint32_t i = t0 + threadIdx.x;
        const bool keep = i < nc && st[i];
        gt_detection d;
        if (keep) d = fr[i];
        s_scan[threadIdx.x] = keep ? 1u : 0u;
        __syncthreads();
        for (uint32_t o = 1; o < 256; o <<= 1) {
            const uint32_t v = threadIdx.x >= o ? s_scan[threadIdx.x - o] : 0u;
            __syncthreads();
            s_scan[threadIdx.x] += v;
            __syncthreads();
        }
        if (keep) fr[off + s_scan[threadIdx.x] - 1u] = d;
        const uint32_t tot = s_scan[255];
        __syncthreads();
        off += tot;
    }
    if (threadIdx.x == 0) ocount[f * GT_COUNTER_STRIDE] = off;
}

template <typename T>
hipError_t launch_adaptive_t(const T *frames, uint32_t nf, const GtKParams &p, uint32_t n_iter,
                             gt_detection *out, uint32_t cap, uint32_t *ocount, float *std_tmp,
                             float *mean, float *std2, float *masked, uint8_t *state,
                             uint32_t *counter, hipStream_t s) {
    const size_t HW = static_cast<size_t>(p.H) * p.W;
    k_widen_frames<T><<<2048, 256, 0, s>>>(frames, HW * nf, masked);
    k_fill_counter<<<(nf + 255) / 256, 256, 0, s>>>(counter, nf, static_cast<uint32_t>(HW));
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const dim3 g(16, nf);
    for (uint32_t it = 0; it < n_iter; ++it) {
        const float *sd = p.fstd;
        if (it > 0) {
            e = launch_frame_std_t<float>(masked, nf, HW, std_tmp, mean, std2, s, counter);
            if (e != hipSuccess) return e;
            sd = std2;
        }
        k_adapt_filter<<<g, 256, 0, s>>>(out, cap, ocount, p, sd, it, masked, state, counter);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (n_iter == 0) {
        // no pass: no row is accepted
        k_fill_counter<<<(nf * GT_COUNTER_STRIDE + 255) / 256, 256, 0, s>>>(
            ocount, nf * GT_COUNTER_STRIDE, 0u);
        return hipGetLastError();
    }
    k_adapt_compact<<<nf, 256, 0, s>>>(out, cap, ocount, state);
    return hipGetLastError();
}
}  // namespace

// ---------------------------------------------------------------------------
// Illumination correction.  The profile is the mean frame (mean_from_iter:
// per pixel an f32 running sum in frame order, then / f32(T)) or the frame
// itself, smoothed by smooth_mean_frame.wgsl's two passes; each frame is then
// divided by it (correct_illumination.wgsl).  One lane per pixel; the taps
// come from a device table.

namespace {

template <typename T>
__global__ void k_frames_accumulate(const T *__restrict__ frames, uint32_t nframes, size_t HW,
                                    float *__restrict__ acc) {
    const size_t p = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= HW) return;
    float a = acc[p];
    for (uint32_t t = 0; t < nframes; ++t) a = a + to_f32(frames[static_cast<size_t>(t) * HW + p]);
    acc[p] = a;
}

__global__ void k_scale_div(float *__restrict__ x, size_t n, float d) {
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x)
        x[i] = x[i] / d;
}

// One pass; dim 0 runs along a column (u = row), dim 1 along a row.  Taps
// in order, out-of-frame taps skipped and left out of the norm.
__global__ void k_illum_pass(const float *__restrict__ in, float *__restrict__ out, int H, int W,
                             const float *__restrict__ w, int rint, int dim) {
    const size_t HW = static_cast<size_t>(H) * W;
    const size_t p = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= HW) return;
    const size_t fo = static_cast<size_t>(blockIdx.y) * HW;
    const int r = static_cast<int>(p / W), c = static_cast<int>(p - static_cast<size_t>(r) * W);
    const int u = dim == 0 ? r : c, bound = dim == 0 ? H : W;
    const size_t stride = dim == 0 ? static_cast<size_t>(W) : 1;
    const float *src = in + fo + p;
    float gs = 0.0f, ns = 0.0f;
    const int lo = max(-rint, -u), hi = min(rint, bound - 1 - u);
    for (int i = lo; i <= hi; ++i) {
        const float wi = w[i + rint];
        gs = fmaf(wi, src[static_cast<long long>(i) * static_cast<long long>(stride)], gs);
        ns = ns + wi;
    }
    out[fo + p] = gs / ns;
}

template <typename T>
__global__ void k_illum_divide(const T *__restrict__ frames, uint32_t nframes, size_t HW,
                               const float *__restrict__ profile, size_t pstride,
                               float *__restrict__ out) {
    const size_t n = HW * nframes;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const size_t t = i / HW, pix = i - t * HW;
        out[i] = to_f32(frames[i]) / profile[t * pstride + pix];
    }
}

unsigned grid_for(size_t n, unsigned cap = 65535u * 4u) {
    const size_t b = (n + 255) / 256;
    return static_cast<unsigned>(b < 1 ? 1 : (b > cap ? cap : b));
}

}  // namespace

hipError_t gt_launch_frames_accumulate(const void *frames, int dtype, uint32_t nframes, size_t HW,
                                       float *acc, hipStream_t s) {
    if (nframes == 0 || HW == 0) return hipSuccess;
    const unsigned g = static_cast<unsigned>((HW + 255) / 256);
    switch (dtype) {
    case GT_U16: k_frames_accumulate<<<g, 256, 0, s>>>(static_cast<const uint16_t *>(frames), nframes, HW, acc); break;
    case GT_F32: k_frames_accumulate<<<g, 256, 0, s>>>(static_cast<const float *>(frames), nframes, HW, acc); break;
    case GT_U8: k_frames_accumulate<<<g, 256, 0, s>>>(static_cast<const uint8_t *>(frames), nframes, HW, acc); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t gt_launch_widen(const void *frames, int dtype, size_t n, float *dst, hipStream_t s) {
    if (n == 0) return hipSuccess;
    switch (dtype) {
    case GT_U16: k_widen_frames<<<2048, 256, 0, s>>>(static_cast<const uint16_t *>(frames), n, dst); break;
    case GT_F32: k_widen_frames<<<2048, 256, 0, s>>>(static_cast<const float *>(frames), n, dst); break;
    case GT_U8: k_widen_frames<<<2048, 256, 0, s>>>(static_cast<const uint8_t *>(frames), n, dst); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t gt_launch_scale_div(float *x, size_t n, float d, hipStream_t s) {
    if (n == 0) return hipSuccess;
    k_scale_div<<<grid_for(n), 256, 0, s>>>(x, n, d);
    return hipGetLastError();
}

hipError_t gt_launch_illum_smooth(const float *in, float *tmp, float *out, int H, int W,
                                  uint32_t nframes, const float *w, int rint, hipStream_t s) {
    const size_t HW = static_cast<size_t>(H) * W;
    if (nframes == 0 || HW == 0) return hipSuccess;
    const dim3 g(static_cast<unsigned>((HW + 255) / 256), nframes);
    k_illum_pass<<<g, 256, 0, s>>>(in, tmp, H, W, w, rint, 0);
    k_illum_pass<<<g, 256, 0, s>>>(tmp, out, H, W, w, rint, 1);
    return hipGetLastError();
}

hipError_t gt_launch_illum_divide(const void *frames, int dtype, uint32_t nframes, size_t HW,
                                  const float *profile, size_t pstride, float *out, hipStream_t s) {
    const size_t n = HW * nframes;
    if (n == 0) return hipSuccess;
    const unsigned g = grid_for(n);
    switch (dtype) {
    case GT_U16: k_illum_divide<<<g, 256, 0, s>>>(static_cast<const uint16_t *>(frames), nframes, HW, profile, pstride, out); break;
    case GT_F32: k_illum_divide<<<g, 256, 0, s>>>(static_cast<const float *>(frames), nframes, HW, profile, pstride, out); break;
    case GT_U8: k_illum_divide<<<g, 256, 0, s>>>(static_cast<const uint8_t *>(frames), nframes, HW, profile, pstride, out); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t gt_launch_adaptive(const void *frames, int dtype, uint32_t nframes, const GtKParams &p,
                              uint32_t n_iter, gt_detection *out, uint32_t cap, uint32_t *ocount,
                              float *std_tmp, float *mean, float *std2, float *masked,
                              uint8_t *state, uint32_t *counter, hipStream_t s) {
    if (nframes == 0) return hipSuccess;
    if (!p.fstd) return hipErrorInvalidValue;
    switch (dtype) {
    case GT_U16:
        return launch_adaptive_t(static_cast<const uint16_t *>(frames), nframes, p, n_iter, out,
                                 cap, ocount, std_tmp, mean, std2, masked, state, counter, s);
    case GT_F32:
        return launch_adaptive_t(static_cast<const float *>(frames), nframes, p, n_iter, out, cap,
                                 ocount, std_tmp, mean, std2, masked, state, counter, s);
    case GT_U8:
        return launch_adaptive_t(static_cast<const uint8_t *>(frames), nframes, p, n_iter, out,
                                 cap, ocount, std_tmp, mean, std2, masked, state, counter, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t gt_launch_frame_std(const void *frames, int dtype, uint32_t nframes, size_t n,
                               float *tmp, float *mean, float *std_out, hipStream_t s) {
    if (n <= 24 || nframes == 0) return hipErrorInvalidValue;
    switch (dtype) {
    case GT_U16:
        return launch_frame_std_t(static_cast<const uint16_t *>(frames), nframes, n, tmp, mean,
                                  std_out, s);
    case GT_F32:
        return launch_frame_std_t(static_cast<const float *>(frames), nframes, n, tmp, mean,
                                  std_out, s);
    case GT_U8:
        return launch_frame_std_t(static_cast<const uint8_t *>(frames), nframes, n, tmp, mean,
                                  std_out, s);
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

hipError_t gt_init_tables() {
    // the same expressions as the oracle's closed forms, in f64
    static double2 host[(2 * GT_SC_R + 1) * (2 * GT_SC_R + 1)];
    for (int i = -GT_SC_R; i <= GT_SC_R; ++i)
        for (int j = -GT_SC_R; j <= GT_SC_R; ++j) {
            double2 v = make_double2(0.0, 0.0);
            if (i != 0 || j != 0) {
                const double q = static_cast<double>(i * i + j * j);
                v.x = 2.0 * static_cast<double>(i) * static_cast<double>(j) / q;
                v.y = (static_cast<double>(i * i) - static_cast<double>(j * j)) / q;
            }
            host[(i + GT_SC_R) * (2 * GT_SC_R + 1) + (j + GT_SC_R)] = v;
        }
    return hipMemcpyToSymbol(HIP_SYMBOL(c_sincos2), host, sizeof(host));
}

hipError_t gt_launch_characterize_det(const float *proc, const void *raw, int dtype,
                                      uint32_t nframes, const GtKParams &p, gt_detection *out,
                                      uint32_t out_cap, const uint32_t *ocount, hipStream_t s,
                                      uint32_t *dcount) {
    // about one lane per expected detection (one per ~1000 pixels)
    const size_t hw = static_cast<size_t>(p.H) * p.W;
    uint32_t blocks = static_cast<uint32_t>(hw / 262144);
    blocks = blocks < 1 ? 1 : (blocks > 64 ? 64 : blocks);
    const dim3 grid(blocks, nframes, 1);
#define GT_CHAR(TT)                                                                            \
    do {                                                                                       \
        if (p.rint == 4)                                                                       \
            k_characterize_det<TT, 4><<<grid, WK_THREADS, 0, s>>>(                             \
                proc, static_cast<const TT *>(raw), p, out, out_cap, ocount, dcount);          \
        else                                                                                   \
            k_characterize_det<TT, 0><<<grid, WK_THREADS, 0, s>>>(                             \
                proc, static_cast<const TT *>(raw), p, out, out_cap, ocount, dcount);          \
    } while (0)
    switch (dtype) {
    case GT_U16: GT_CHAR(uint16_t); break;
    case GT_F32: GT_CHAR(float); break;
    case GT_U8: GT_CHAR(uint8_t); break;
    default: return hipErrorInvalidValue;
    }
#undef GT_CHAR
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// LoG style (execute_gpu.rs:237-286).  One block per LG_TH x LG_TW tile of
// one frame computes every scale's LoG plane over the tile plus a one-pixel
// ring (separable_log.wgsl through Laplace<2>, utils.rs:114-296) from one
// LDS copy of the raw tile, keeps the last three planes in LDS, and tests the
// scale-space maxima (logspace_max.wgsl:62-154) as soon as a scale's upper
// neighbour exists.  The planes never go to HBM.  Candidates (seed = scale *
// H * W + raster index) go to a per-frame list; k_walk_log walks them.
namespace {

constexpr int LG_TW = 128, LG_TH = 16;
constexpr int LG_RR = LG_TH + 2, LG_RC = LG_TW + 2;  // plane region: tile + 1-pixel ring
constexpr int LG_MV = 6;                             // rows per item of the vertical pass
constexpr int LG_MH = 5;                             // columns per item of the horizontal pass
constexpr int LG_NG = LG_RC / LG_MH;                 // horizontal items per region row
constexpr int LG_THREADS = 512;
constexpr int LG_WAVES = LG_THREADS / 64;
constexpr int LG_CB = LG_TW / 64;                    // 64-column blocks of the tile
constexpr int LG_TR = LG_TH * LG_CB / LG_WAVES;      // tile rows per lane in the maximum test
constexpr int LG_RU = 24;                            // largest tap radius with unrolled passes
constexpr int LG_BUF = 1024;                         // a tile's buffered candidates
// Two workgroups of LG_THREADS fit a CU's LDS at Rm = 17; their waves need
// at most 128 VGPRs each.
constexpr int LG_MIN_WAVES = 4;
static_assert(LG_RR % LG_MV == 0, "vertical items must tile the region rows");
static_assert(LG_RC % LG_MH == 0 && (LG_MH & 1), "odd horizontal items must tile the region columns");
static_assert(LG_TW % 64 == 0 && LG_WAVES % LG_CB == 0 && LG_TR * (LG_WAVES / LG_CB) == LG_TH,
              "one tile column per lane in the test");

// LDS layout for a largest tap radius Rm:
//   raw   NC columns of CP samples, column-major (NC = LG_RC + 2 Rm, CP =
//         (LG_RR + 2 Rm) | 1): the vertical pass reads down a column at
//         compile-time offsets, and lanes on neighbouring columns fall in
//         different banks.  The last sample of a column is padding.
//   Vd,Vg LG_RR rows of VP >= NC + 1 with VP = LG_RC (mod 64): horizontal
//         item i (row i / LG_NG, columns LG_MH * (i % LG_NG) ...) then starts
//         at LG_MH * i (mod 64), so the 64 items of a wave start in 64
//         different banks (LG_MH odd)
//   L     LG_RR x LG_RC (the current scale's plane)
//   B     LG_BUF candidate seeds, then their LoG values: the tile's
//         candidates, written to the frame's list as one block at the end
struct LogGeom {
    int Rm, NC, CP, VP;
};
inline LogGeom log_geom(int Rm) {
    LogGeom g;
    g.Rm = Rm;
    g.NC = LG_RC + 2 * Rm;
    g.CP = (LG_RR + 2 * Rm) | 1;
    g.VP = g.NC + 1 + ((LG_RC - (g.NC + 1)) % 64 + 64) % 64;
    return g;
}
inline size_t log_lds_bytes(const LogGeom &g) {
    return sizeof(float) * (static_cast<size_t>(g.NC) * g.CP + 2u * LG_RR * g.VP +
                            static_cast<size_t>(LG_RR) * LG_RC + 2u * LG_BUF);
}

// Vertical 1-D pass of one scale for LG_MV consecutive outputs of one column
// (samples colp[0], colp[1], ... down the column):
// ad[m] = sum_i td[i] * x[i + m], ag[m] = sum_i tg[i] * x[i + m], taps in
// order i = 0..ntaps-1 (the reference's -R..R).  The inputs slide through a
// register window; inside a block of LG_MV taps the window index rotates at
// compile time, so no register moves are needed there.  Reads one sample
// past the last one used.
__device__ __forceinline__ void log_vpass(const float *colp, int ntaps,
                                          const float *__restrict__ tg,
                                          const float *__restrict__ td, float (&ag)[LG_MV],
                                          float (&ad)[LG_MV]) {
    float w[LG_MV];
#pragma unroll
    for (int m = 0; m < LG_MV; ++m) {
        w[m] = colp[m];
        ag[m] = 0.0f;
        ad[m] = 0.0f;
    }
    int i = 0;
    for (; i + LG_MV <= ntaps; i += LG_MV) {
#pragma unroll
        for (int k = 0; k < LG_MV; ++k) {
            const float g = tg[i + k], d = td[i + k];
#pragma unroll
            for (int m = 0; m < LG_MV; ++m) {
                const float v = w[(m + k) % LG_MV];
                ad[m] = fmaf(d, v, ad[m]);
                ag[m] = fmaf(g, v, ag[m]);
            }
            w[k] = colp[i + k + LG_MV];
        }
    }
    for (; i < ntaps; ++i) {
        const float g = tg[i], d = td[i];
#pragma unroll
        for (int m = 0; m < LG_MV; ++m) {
            ad[m] = fmaf(d, w[m], ad[m]);
            ag[m] = fmaf(g, w[m], ag[m]);
        }
#pragma unroll
        for (int m = 0; m + 1 < LG_MV; ++m) w[m] = w[m + 1];
        w[LG_MV - 1] = colp[i + LG_MV];
    }
}

// The same sums for a compile-time tap radius R: the whole window is loaded
// first (one wait), then every tap's multiply-adds run back to back.
template <int R>
__device__ __forceinline__ void log_vpass_u(const float *colp, const float *__restrict__ tg,
                                            const float *__restrict__ td, float (&ag)[LG_MV],
                                            float (&ad)[LG_MV]) {
    constexpr int NW = LG_MV + 2 * R;
    float w[NW];
#pragma unroll
    for (int i = 0; i < NW; ++i) w[i] = colp[i];
#pragma unroll
    for (int m = 0; m < LG_MV; ++m) {
        ag[m] = 0.0f;
        ad[m] = 0.0f;
    }
#pragma unroll
    for (int i = 0; i < 2 * R + 1; ++i) {
        const float g = tg[i], d = td[i];
#pragma unroll
        for (int m = 0; m < LG_MV; ++m) {
            ad[m] = fmaf(d, w[i + m], ad[m]);
            ag[m] = fmaf(g, w[i + m], ag[m]);
        }
    }
}

// Horizontal 1-D pass for LG_MH consecutive outputs of one row:
// a0 = sum_j tg[j] * vd[j + m], a1 = sum_j td[j] * vg[j + m].  Reads one
// value past the last one used.
__device__ __forceinline__ void log_hpass(const float *vd, const float *vg, int ntaps,
                                          const float *__restrict__ tg,
                                          const float *__restrict__ td, float (&a0)[LG_MH],
                                          float (&a1)[LG_MH]) {
    float wa[LG_MH], wb[LG_MH];
#pragma unroll
    for (int m = 0; m < LG_MH; ++m) {
        wa[m] = vd[m];
        wb[m] = vg[m];
        a0[m] = 0.0f;
        a1[m] = 0.0f;
    }
    int j = 0;
    for (; j + LG_MH <= ntaps; j += LG_MH) {
#pragma unroll
        for (int k = 0; k < LG_MH; ++k) {
            const float g = tg[j + k], d = td[j + k];
#pragma unroll
            for (int m = 0; m < LG_MH; ++m) {
                const int q = (m + k) % LG_MH;
                a0[m] = fmaf(g, wa[q], a0[m]);
                a1[m] = fmaf(d, wb[q], a1[m]);
            }
            wa[k] = vd[j + k + LG_MH];
            wb[k] = vg[j + k + LG_MH];
        }
    }
    for (; j < ntaps; ++j) {
        const float g = tg[j], d = td[j];
#pragma unroll
        for (int m = 0; m < LG_MH; ++m) {
            a0[m] = fmaf(g, wa[m], a0[m]);
            a1[m] = fmaf(d, wb[m], a1[m]);
        }
#pragma unroll
        for (int m = 0; m + 1 < LG_MH; ++m) {
            wa[m] = wa[m + 1];
            wb[m] = wb[m + 1];
        }
        wa[LG_MH - 1] = vd[j + LG_MH];
        wb[LG_MH - 1] = vg[j + LG_MH];
    }
}

// The same sums for a compile-time tap radius R, one tap set at a time with
// its whole window in registers.
template <int R>
__device__ __forceinline__ void log_hpass_u(const float *vd, const float *vg,
                                            const float *__restrict__ tg,
                                            const float *__restrict__ td, float (&a0)[LG_MH],
                                            float (&a1)[LG_MH]) {
    constexpr int NW = LG_MH + 2 * R;
    {
        float w[NW];
#pragma unroll
        for (int i = 0; i < NW; ++i) w[i] = vd[i];
#pragma unroll
        for (int m = 0; m < LG_MH; ++m) a0[m] = 0.0f;
#pragma unroll
        for (int j = 0; j < 2 * R + 1; ++j) {
            const float g = tg[j];
#pragma unroll
            for (int m = 0; m < LG_MH; ++m) a0[m] = fmaf(g, w[j + m], a0[m]);
        }
    }
    {
        float w[NW];
#pragma unroll
        for (int i = 0; i < NW; ++i) w[i] = vg[i];
#pragma unroll
        for (int m = 0; m < LG_MH; ++m) a1[m] = 0.0f;
#pragma unroll
        for (int j = 0; j < 2 * R + 1; ++j) {
            const float d = td[j];
#pragma unroll
            for (int m = 0; m < LG_MH; ++m) a1[m] = fmaf(d, w[j + m], a1[m]);
        }
    }
}

// Both passes of one scale (tap radius R; RT = R at compile time, or 0):
// vertical pass into Vd/Vg, barrier, horizontal pass into the plane Lk with
// L = (0 - R0) - R1 as the reference subtracts both passes from a zeroed
// plane.  Items go to the waves in blocks of 64, the first block to wave
// `rot`, so that the blocks past a multiple of the wave count land on
// different waves in different workgroups.
template <int RT>
__device__ __forceinline__ void log_scale_passes(const float *Sraw, float *Vd, float *Vg,
                                                 float *Lk, const LogGeom &g, int R,
                                                 const float *__restrict__ tg,
                                                 const float *__restrict__ td, float n1, float n3,
                                                 int wave, int lane_in, int rot) {
    // the item addresses are the same for every scale; the empty asm keeps
    // the compiler from computing them once per tap radius ahead of the
    // scale loop (one live register set per instantiation, spilled)
    int lane = lane_in;
    asm volatile("" : "+v"(lane));
    const int Rr = RT > 0 ? RT : R;
    const int ncv = LG_RC + 2 * Rr;  // vertical-pass columns: region columns -R .. LG_RC + R - 1
    const int nv = (LG_RR / LG_MV) * ncv;
    const int b0 = (wave - rot + LG_WAVES) % LG_WAVES;
    for (int b = b0; b * 64 < nv; b += LG_WAVES) {
        const int item = b * 64 + lane;
        if (item < nv) {
            const int q = item / ncv, vc = item - q * ncv, r0 = q * LG_MV;
            const float *colp = Sraw + (vc + g.Rm - Rr) * g.CP + (r0 + g.Rm - Rr);
            float ag[LG_MV], ad[LG_MV];
            if constexpr (RT > 0) log_vpass_u<RT>(colp, tg, td, ag, ad);
            else log_vpass(colp, 2 * Rr + 1, tg, td, ag, ad);
#pragma unroll
            for (int m = 0; m < LG_MV; ++m) {
                Vd[(r0 + m) * g.VP + vc] = ad[m] * n3;
                Vg[(r0 + m) * g.VP + vc] = ag[m] * n1;
            }
        }
    }
    __syncthreads();
    constexpr int NH = LG_RR * LG_NG;
    for (int b = b0; b * 64 < NH; b += LG_WAVES) {
        const int item = b * 64 + lane;
        if (item < NH) {
            const int r = item / LG_NG, c0 = (item - r * LG_NG) * LG_MH;
            float a0[LG_MH], a1[LG_MH];
            if constexpr (RT > 0)
                log_hpass_u<RT>(Vd + r * g.VP + c0, Vg + r * g.VP + c0, tg, td, a0, a1);
            else
                log_hpass(Vd + r * g.VP + c0, Vg + r * g.VP + c0, 2 * Rr + 1, tg, td, a0, a1);
#pragma unroll
            for (int m = 0; m < LG_MH; ++m) {
                float l = 0.0f - a0[m] * n1;
                l = l - a1[m] * n3;
                Lk[r * LG_RC + c0 + m] = l;
            }
        }
    }
}

// Appends lane flags f[0..N-1] to a list guarded by *counter with one
// atomicAdd per wavefront; pos[s] is flag s's slot (set where f[s]).  Must be
// reached by every lane of the wave.
template <int N>
__device__ __forceinline__ void wave_append_n(const bool (&f)[N], uint32_t *counter,
                                              uint32_t (&pos)[N]) {
    uint64_t mk[N];
    uint32_t tot = 0;
#pragma unroll
    for (int s = 0; s < N; ++s) {
        mk[s] = __ballot(f[s]);
        tot += static_cast<uint32_t>(__popcll(mk[s]));
    }
    if (tot == 0) return;
    const uint32_t lane = __lane_id();
    const uint64_t lt = (1ull << lane) - 1ull;
    uint32_t base = 0u;
    if (lane == 0) base = atomicAdd(counter, tot);
    base = __shfl(base, 0);
#pragma unroll
    for (int s = 0; s < N; ++s) {
        pos[s] = base + static_cast<uint32_t>(__popcll(mk[s] & lt));
        base += static_cast<uint32_t>(__popcll(mk[s]));
    }
}

// Centre value c[s] and largest of the 8 neighbours x[s] in plane Lk of the
// lane's LG_TR tile pixels (tile column tc, tile rows tr0 .. tr0 + LG_TR - 1).
// fmaxf skips NaN, so x is the largest non-NaN neighbour, NaN when all 8
// are: "no neighbour >= c" (is_max_in_plane, logspace_max.wgsl:62-92) is
// then !(x >= c) in every case.
__device__ __forceinline__ void log_plane_stats(const float *Lk, int tr0, int tc,
                                                float (&c)[LG_TR], float (&x)[LG_TR]) {
    const float *b = Lk + tr0 * LG_RC + tc;  // region (tr0, tc): the first pixel's top-left
    float v[LG_TR + 2][3];
#pragma unroll
    for (int i = 0; i < LG_TR + 2; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) v[i][j] = b[i * LG_RC + j];
    float h[LG_TR + 2];
#pragma unroll
    for (int i = 0; i < LG_TR + 2; ++i) h[i] = fmaxf(fmaxf(v[i][0], v[i][1]), v[i][2]);
#pragma unroll
    for (int s = 0; s < LG_TR; ++s) {
        c[s] = v[s + 1][1];
        x[s] = fmaxf(fmaxf(h[s], h[s + 2]), fmaxf(v[s + 1][0], v[s + 1][2]));
    }
}

template <typename T>
__global__ __launch_bounds__(LG_THREADS, LG_MIN_WAVES) void k_log_tile(const T *__restrict__ raw,
                                                         const float *__restrict__ proc,
                                                         const GtLogTable *__restrict__ tab,
                                                         LogGeom g, int H, int W,
                                                         uint32_t *__restrict__ cand, uint32_t cap,
                                                         uint32_t *__restrict__ ccount,
                                                         int nframes,
                                                         const float *__restrict__ fstd, float snr,
                                                         float rough, float *__restrict__ cval) {
    extern __shared__ float lds[];
    const int ntx = (W + LG_TW - 1) / LG_TW, nty = (H + LG_TH - 1) / LG_TH;
    const int per_frame = ntx * nty;
    const int t = xcd_tile(per_frame * nframes);
    if (t >= per_frame * nframes) return;
    const int fr = t / per_frame;
    const int tt = t - fr * per_frame;
    const int ty = tt / ntx, tx = tt - ty * ntx;
    const int y0 = ty * LG_TH, x0 = tx * LG_TW;
    const size_t HW = static_cast<size_t>(H) * W;
    const T *f = raw + static_cast<size_t>(fr) * HW;
    const float *P = proc + static_cast<size_t>(fr) * HW;
    const int Rm = g.Rm;
    float *Sraw = lds;
    float *Vd = Sraw + g.NC * g.CP;
    float *Vg = Vd + LG_RR * g.VP;
    float *Lr = Vg + LG_RR * g.VP;
    uint32_t *Bs = reinterpret_cast<uint32_t *>(Lr + LG_RR * LG_RC);
    float *Bv = reinterpret_cast<float *>(Bs + LG_BUF);
    __shared__ uint32_t s_fill, s_base;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    if (tid == 0) s_fill = 0u;
    // raw tile, column-major: region origin (y0 - 1, x0 - 1) less the largest
    // tap radius; outside the frame 0 (a zero sample adds nothing to a tap
    // sum, as the reference's skipped out-of-frame taps).  Four rows per wave
    // per pass, so that each lane has several loads in flight.
    {
        const int oy = y0 - 1 - Rm, ox = x0 - 1 - Rm;
        const int nrows = LG_RR + 2 * Rm;
        constexpr int U = 4;
        for (int ry0 = wave * U; ry0 < nrows; ry0 += LG_WAVES * U) {
            for (int rx = lane; rx < g.NC; rx += 64) {
                const int gx = ox + rx;
                const bool cok = gx >= 0 && gx < W;
                float v[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int gy = oy + ry0 + u;
                    v[u] = 0.0f;
                    if (cok && ry0 + u < nrows && gy >= 0 && gy < H)
                        v[u] = to_f32(f[static_cast<size_t>(gy) * W + gx]);
                }
#pragma unroll
                for (int u = 0; u < U; ++u)
                    if (ry0 + u < nrows) Sraw[rx * g.CP + ry0 + u] = v[u];
            }
        }
    }
    __syncthreads();
    const int n = tab->n;
    const float snr_thr = fstd ? fstd[fr] * snr * rough : tab->snr_thr;
    uint32_t *cnt = &ccount[fr * GT_COUNTER_STRIDE];
    uint32_t *fc = cand + static_cast<size_t>(fr) * cap;
    const int rot = static_cast<int>(blockIdx.x % LG_WAVES);
    // the maximum test: this lane's pixels are tile column tc, tile rows
    // tr0 .. tr0 + LG_TR - 1; the centre and neighbour maximum of the
    // previous scale (cp, xp) and the one before it (xpp) stay in registers
    const int tr0 = (wave / LG_CB) * LG_TR;
    const int tc = (wave % LG_CB) * 64 + lane;
    const int gx = x0 + tc;
    float cp[LG_TR], xp[LG_TR], xpp[LG_TR];
#pragma unroll
    for (int s = 0; s < LG_TR; ++s) cp[s] = xp[s] = xpp[s] = 0.0f;
    // scale m's candidates: a positive centre with no neighbour >= it in its
    // own plane, in scale m - 1's (m >= 1) and in scale m + 1's (xu, when
    // `up`: m + 1 < n); pixels of the frame's one-pixel border never qualify
    auto test = [&](int m, bool up, const float (&xu)[LG_TR]) {
        bool fl[LG_TR];
#pragma unroll
        for (int s = 0; s < LG_TR; ++s) {
            const int gy = y0 + tr0 + s;
            const float ctr = cp[s];
            bool flag = gy >= 1 && gy <= H - 2 && gx >= 1 && gx <= W - 2 && !(ctr <= 0.0f) &&
                        !(xp[s] >= ctr);
            if (m >= 1) flag = flag && !(xpp[s] >= ctr);
            if (up) flag = flag && !(xu[s] >= ctr);
            if (flag) flag = !(P[static_cast<size_t>(gy) * W + gx] < snr_thr);
            fl[s] = flag;
        }
        // into the tile's buffer; what does not fit goes to the frame's list
        uint32_t pos[LG_TR], gpos[LG_TR];
        bool of[LG_TR];
        wave_append_n<LG_TR>(fl, &s_fill, pos);
#pragma unroll
        for (int s = 0; s < LG_TR; ++s) {
            of[s] = fl[s] && pos[s] >= static_cast<uint32_t>(LG_BUF);
            if (fl[s] && !of[s]) {
                const int gy = y0 + tr0 + s;
                Bs[pos[s]] = static_cast<uint32_t>(m) * static_cast<uint32_t>(HW) +
                             static_cast<uint32_t>(gy) * static_cast<uint32_t>(W) +
                             static_cast<uint32_t>(gx);
                Bv[pos[s]] = cp[s];
            }
        }
        wave_append_n<LG_TR>(of, cnt, gpos);
#pragma unroll
        for (int s = 0; s < LG_TR; ++s) {
            if (of[s] && gpos[s] < cap) {
                const int gy = y0 + tr0 + s;
                fc[gpos[s]] = static_cast<uint32_t>(m) * static_cast<uint32_t>(HW) +
                              static_cast<uint32_t>(gy) * static_cast<uint32_t>(W) +
                              static_cast<uint32_t>(gx);
                if (cval) cval[static_cast<size_t>(fr) * cap + gpos[s]] = cp[s];
            }
        }
    };
    // threshold_rel: the largest LoG value of the tile's own pixels over
    // every scale (each frame pixel is in one tile)
    float lmax = 0.0f;
    for (int k = 0; k < n; ++k) {
        const int R = tab->R[k];
        const float *tg = tab->taps + tab->off[k];
        const float *td = tg + 2 * R + 1;
        const float n1 = tab->n1[k], n3 = tab->n3[k];
        static_assert(LG_RU == 24, "the unrolled cases below are tap radii 1..24");
        switch (R) {
#define GT_LG_CASE(RV)                                                                            \
    case RV:                                                                                      \
        log_scale_passes<RV>(Sraw, Vd, Vg, Lr, g, R, tg, td, n1, n3, wave, lane, rot);           \
        break;
            GT_LG_CASE(1) GT_LG_CASE(2) GT_LG_CASE(3) GT_LG_CASE(4) GT_LG_CASE(5) GT_LG_CASE(6)
            GT_LG_CASE(7) GT_LG_CASE(8) GT_LG_CASE(9) GT_LG_CASE(10) GT_LG_CASE(11)
            GT_LG_CASE(12) GT_LG_CASE(13) GT_LG_CASE(14) GT_LG_CASE(15) GT_LG_CASE(16)
            GT_LG_CASE(17) GT_LG_CASE(18) GT_LG_CASE(19) GT_LG_CASE(20) GT_LG_CASE(21)
            GT_LG_CASE(22) GT_LG_CASE(23) GT_LG_CASE(24)
#undef GT_LG_CASE
        default:
            log_scale_passes<0>(Sraw, Vd, Vg, Lr, g, R, tg, td, n1, n3, wave, lane, rot);
            break;
        }
        __syncthreads();
        float c[LG_TR], x[LG_TR];
        log_plane_stats(Lr, tr0, tc, c, x);
        if (cval) {
#pragma unroll
            for (int s = 0; s < LG_TR; ++s)
                if (y0 + tr0 + s < H && gx < W) lmax = fmaxf(lmax, c[s]);
        }
        if (k >= 1) test(k - 1, true, x);
#pragma unroll
        for (int s = 0; s < LG_TR; ++s) {
            xpp[s] = xp[s];
            xp[s] = x[s];
            cp[s] = c[s];
        }
    }
    if (n > 0) test(n - 1, false, xp);
    // the buffered candidates as one block of the frame's list (the walks
    // then gather from one neighbourhood per block)
    __syncthreads();
    const uint32_t nb = min(s_fill, static_cast<uint32_t>(LG_BUF));
    if (tid == 0) s_base = nb ? atomicAdd(cnt, nb) : 0u;
    __syncthreads();
    for (uint32_t i = tid; i < nb; i += LG_THREADS) {
        const uint32_t gp = s_base + i;
        if (gp < cap) {
            fc[gp] = Bs[i];
            if (cval) cval[static_cast<size_t>(fr) * cap + gp] = Bv[i];
        }
    }
    if (cval) {
        // one atomic per wave: non-negative floats order as their bit patterns
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) lmax = fmaxf(lmax, __shfl_xor(lmax, o));
        if (__lane_id() == 0 && lmax > 0.0f)
            atomicMax(&ccount[fr * GT_COUNTER_STRIDE + 1], __float_as_uint(lmax));
    }
}

// Walk + characterize of the LoG candidates (walk.wgsl with r > 0: rint =
// ceil(r), r2 = r*r, move band rint; characterize.wgsl at the same radius).
// A frame whose candidate list overflowed reports count 0xFFFFFFFF.
//
// Disks wholly inside the frame run over the rows of the scale's disk table
// (row i: columns -jm .. jm), with no per-pixel mask or bounds test: the same
// pixels in the same order as get_center / characterize_row.  A row is read
// LW_C pixels at a time, so that its loads are in flight together; the
// slots of the last group past the row end read the row's last pixel and
// add an exact zero instead (every sum and moment here starts at +0 and
// never becomes -0, and the maxima start at +0), so the bits are the same.
constexpr int LW_C = 4;

// get_center over a table disk around (u, v) inside the frame
__device__ __forceinline__ WalkRes log_center_inner(const float *__restrict__ P, int W, int u,
                                                    int v, int rint, const int8_t *jm) {
    WalkRes res = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    for (int i = -rint; i <= rint; ++i) {
        const int m = jm[i + GT_LOG_WR];
        const float x = static_cast<float>(i);
        const float *row = P + static_cast<size_t>(u + i) * W + v;
        for (int j0 = -m; j0 <= m; j0 += LW_C) {
            float dv[LW_C];
#pragma unroll
            for (int q = 0; q < LW_C; ++q) dv[q] = row[min(j0 + q, m)];
#pragma unroll
            for (int q = 0; q < LW_C; ++q) {
                const bool in = j0 + q <= m;
                const float d = in ? dv[q] : 0.0f;
                res.x = fmaf(d, x, res.x);
                res.y = fmaf(d, static_cast<float>(j0 + q), res.y);
                res.mass = res.mass + d;
                res.count = res.count + (in ? 1.0f : 0.0f);
                res.max = maxf_ref(res.max, d);
            }
        }
    }
    res.x = res.x / res.mass;
    res.y = res.y / res.mass;
    return res;
}

// characterize_row for a row of radius r > 0 (rint = ceil(r)) whose disk
// around the rounded centre (u, v) is inside the frame, over the table disk
template <typename T>
__device__ __forceinline__ void log_characterize_inner(const float *__restrict__ P,
                                                       const T *__restrict__ Rw, int W, int u,
                                                       int v, int rint, const int8_t *jm,
                                                       gt_result_row &row) {
    CharAcc acc = {0.0f, 0.0f, 0.0, 0.0, 0.0};
    for (int i = -rint; i <= rint; ++i) {
        const int m = jm[i + GT_LOG_WR];
        const float x = static_cast<float>(i), x2 = x * x;
        const size_t ro = static_cast<size_t>(u + i) * W + v;
        const float *prow = P + ro;
        const T *rrow = Rw + ro;
        for (int j0 = -m; j0 <= m; j0 += LW_C) {
            float dv[LW_C], rv[LW_C];
            double2 sv[LW_C];
#pragma unroll
            for (int q = 0; q < LW_C; ++q) {
                const int j = min(j0 + q, m);
                dv[q] = prow[j];
                rv[q] = to_f32(rrow[j]);
                sv[q] = sincos2(i, j);
            }
#pragma unroll
            for (int q = 0; q < LW_C; ++q) {
                const int j = j0 + q;
                const bool in = j <= m;
                const float y = static_cast<float>(j), y2 = y * y;
                const float mask = x2 + y2;
                const float d = in ? dv[q] : 0.0f;
                acc.raw_mass = acc.raw_mass + (in ? rv[q] : 0.0f);
                acc.signal = maxf_ref(acc.signal, d);
                if (i != 0 || j != 0) {
                    acc.es = fma(static_cast<double>(d), sv[q].x, acc.es);
                    acc.ec = fma(static_cast<double>(d), sv[q].y, acc.ec);
                } else {
                    acc.ec = acc.ec + static_cast<double>(d);
                }
                acc.rg = fma(static_cast<double>(d), static_cast<double>(mask), acc.rg);
            }
        }
    }
    const float den = row.mass - P[static_cast<size_t>(u) * W + v];
    row.ecc = static_cast<float>(sqrt(acc.es * acc.es + acc.ec * acc.ec) / static_cast<double>(den));
    row.Rg = static_cast<float>(sqrt(acc.rg / static_cast<double>(row.mass)));
    row.raw_mass = acc.raw_mass;
    row.signal = acc.signal;
}

template <typename T>
__global__ __launch_bounds__(WK_THREADS) void k_walk_log(const float *__restrict__ proc,
                                                         const T *__restrict__ raw, GtKParams p,
                                                         const GtLogTable *__restrict__ tab,
                                                         const uint32_t *__restrict__ cand,
                                                         uint32_t cap,
                                                         const uint32_t *__restrict__ ccount,
                                                         gt_detection *__restrict__ out,
                                                         uint32_t out_cap,
                                                         uint32_t *__restrict__ ocount,
                                                         uint32_t frame0,
                                                         const float *__restrict__ cval,
                                                         int nframes, uint32_t wb) {
    constexpr int JW = 2 * GT_LOG_WR + 1;
    __shared__ int8_t s_jm[GT_LOG_MAX_SCALES * JW];
    // block w of the launch serves part b of frame fr's list; consecutive
    // parts of a frame (neighbouring candidates, whose disks overlap) go to
    // one XCD, so its L2 serves the overlaps
    const uint32_t total = static_cast<uint32_t>(nframes) * wb;
    const uint32_t w = static_cast<uint32_t>(xcd_tile(static_cast<int>(total)));
    if (w >= total) return;
    const int fr = static_cast<int>(w / wb);
    const uint32_t b = w - static_cast<uint32_t>(fr) * wb;
    const int H = p.H, W = p.W;
    const uint32_t HW = static_cast<uint32_t>(H) * static_cast<uint32_t>(W);
    const int ns = tab->n;
    for (int q = threadIdx.x; q < ns * JW; q += WK_THREADS)
        s_jm[q] = tab->wjm[q / JW][q - (q / JW) * JW];
    __syncthreads();
    // threshold_rel: the candidate's LoG value must exceed thr_rel times the
    // frame's largest LoG value (k_log_tile's per-frame maximum)
    const float rel_thr =
        cval ? p.thr_rel * __uint_as_float(ccount[fr * GT_COUNTER_STRIDE + 1]) : 0.0f;
    const float *P = proc + static_cast<size_t>(fr) * HW;
    const GlobalSrc src{P, W};
    const T *Rw = raw + static_cast<size_t>(fr) * HW;
    const uint32_t nc = ccount[fr * GT_COUNTER_STRIDE];
    const uint32_t n = min(nc, cap);
    const uint32_t *fc = cand + static_cast<size_t>(fr) * cap;
    gt_detection *fout = out + static_cast<size_t>(fr) * out_cap;
    const float stdv = frame_std(p, fr);
    const uint32_t lane = __lane_id();
    // this block's contiguous share of the list, in whole wavefronts
    const uint32_t per = ((n + wb - 1) / wb + 63u) & ~63u;
    const uint32_t lo = b * per, hi = min(n, lo + per);
    for (uint32_t i0 = lo + (threadIdx.x & ~63u); i0 < hi; i0 += WK_THREADS) {
        const uint32_t i = i0 + lane;
        bool keep = false;
        gt_detection d;
        if (i < hi && (!cval || cval[static_cast<size_t>(fr) * cap + i] > rel_thr)) {
            const uint32_t seed = fc[i];
            const uint32_t m = seed / HW;
            const uint32_t ras = seed - m * HW;
            int pu = static_cast<int>(ras / static_cast<uint32_t>(W));
            int pv = static_cast<int>(ras - static_cast<uint32_t>(pu) * static_cast<uint32_t>(W));
            const float r = tab->radius[m];
            const int rint = f2i(ceilf(r));
            const float r2 = r * r;
            const int8_t *jm = s_jm + m * JW;
            const bool table = rint <= GT_LOG_WR;
            WalkRes res = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
            for (uint32_t it = 0; it < p.max_iter; ++it) {
                if (table && pu >= rint && pu + rint < H && pv >= rint && pv + rint < W)
                    res = log_center_inner(P, W, pu, pv, rint, jm);
                else
                    res = get_center<0>(src, H, W, pu, pv, rint, r2);
                if (!walk_move(res, pu, pv, H, W, rint, p.shift_thr)) break;
            }
            res.x = res.x + static_cast<float>(pu);
            res.y = res.y + static_cast<float>(pv);
            keep = passes_minmass(p, res, stdv);
            if (keep) {
                fill_row(d, res, seed, frame0 + fr);
                d.row.r = r;
                if (p.characterize) {
                    const int u = f2i(rintf(d.row.x)), v = f2i(rintf(d.row.y));
                    if (table && u >= rint && u + rint < H && v >= rint && v + rint < W)
                        log_characterize_inner<T>(P, Rw, W, u, v, rint, jm, d.row);
                    else
                        characterize_row<T>(src, Rw, H, W, p.circle, d.row);
                }
            }
        }
        const uint32_t pos = wave_append(keep, &ocount[fr * GT_COUNTER_STRIDE]);
        if (keep && pos < out_cap) fout[pos] = d;
    }
}

__global__ void k_copy_counts_checked(const uint32_t *__restrict__ ocount,
                                      const uint32_t *__restrict__ ccount, uint32_t ccap,
                                      uint32_t *__restrict__ dst, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        dst[i] = ccount[i * GT_COUNTER_STRIDE] > ccap ? 0xFFFFFFFFu : ocount[i * GT_COUNTER_STRIDE];
}

template <typename T>
hipError_t launch_log_tile_t(const T *raw, const float *proc, uint32_t nframes, const GtKParams &p,
                             const GtLogTable *tab, int Rm, uint32_t *cand, uint32_t cap,
                             uint32_t *ccount, float *cval, hipStream_t s) {
    const LogGeom g = log_geom(Rm);
    const size_t lds = log_lds_bytes(g);
    if (lds > 65536) {
        const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_log_tile<T>),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                                 static_cast<int>(lds));
        if (e != hipSuccess) return e;
    }
    const size_t ntiles = static_cast<size_t>((p.W + LG_TW - 1) / LG_TW) *
                          ((p.H + LG_TH - 1) / LG_TH) * nframes;
    k_log_tile<T><<<xcd_grid(ntiles), LG_THREADS, lds, s>>>(raw, proc, tab, g, p.H, p.W, cand, cap,
                                                            ccount, static_cast<int>(nframes),
                                                            p.fstd, p.snr, p.rough, cval);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_walk_log_t(const float *proc, const T *raw, uint32_t nframes, const GtKParams &p,
                             const GtLogTable *tab, const uint32_t *cand, uint32_t cap,
                             const uint32_t *ccount, gt_detection *out, uint32_t out_cap,
                             uint32_t *ocount, uint32_t frame0, const float *cval, hipStream_t s) {
    // parts per frame: enough blocks to cover a typical frame's list in one
    // pass of each block's wavefronts
    const uint32_t wb = std::min<uint32_t>(1024u, std::max<uint32_t>(1u, cap / WK_THREADS / 4u));
    k_walk_log<T><<<xcd_grid(static_cast<size_t>(wb) * nframes), WK_THREADS, 0, s>>>(
        proc, raw, p, tab, cand, cap, ccount, out, out_cap, ocount, frame0, cval,
        static_cast<int>(nframes), wb);
    return hipGetLastError();
}

}  // namespace

size_t gt_log_lds_bytes(int Rm) { return log_lds_bytes(log_geom(Rm)); }

hipError_t gt_launch_log_tile(const void *frames, int dtype, const float *proc, uint32_t nframes,
                              const GtKParams &p, const GtLogTable *tab, int Rm, uint32_t *cand,
                              uint32_t cap, uint32_t *ccount, float *cval, hipStream_t s) {
    switch (dtype) {
    case GT_U16:
        return launch_log_tile_t(static_cast<const uint16_t *>(frames), proc, nframes, p, tab, Rm,
                                 cand, cap, ccount, cval, s);
    case GT_F32:
        return launch_log_tile_t(static_cast<const float *>(frames), proc, nframes, p, tab, Rm,
                                 cand, cap, ccount, cval, s);
    case GT_U8:
        return launch_log_tile_t(static_cast<const uint8_t *>(frames), proc, nframes, p, tab, Rm,
                                 cand, cap, ccount, cval, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t gt_launch_walk_log(const float *proc, const void *raw, int dtype, uint32_t nframes,
                              const GtKParams &p, const GtLogTable *tab, const uint32_t *cand,
                              uint32_t cap, const uint32_t *ccount, gt_detection *out,
                              uint32_t out_cap, uint32_t *ocount, uint32_t frame0, const float *cval,
                              hipStream_t s) {
    switch (dtype) {
    case GT_U16:
        return launch_walk_log_t(proc, static_cast<const uint16_t *>(raw), nframes, p, tab, cand,
                                 cap, ccount, out, out_cap, ocount, frame0, cval, s);
    case GT_F32:
        return launch_walk_log_t(proc, static_cast<const float *>(raw), nframes, p, tab, cand, cap,
                                 ccount, out, out_cap, ocount, frame0, cval, s);
    case GT_U8:
        return launch_walk_log_t(proc, static_cast<const uint8_t *>(raw), nframes, p, tab, cand,
                                 cap, ccount, out, out_cap, ocount, frame0, cval, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t gt_launch_copy_counts_checked(const uint32_t *ocount, const uint32_t *ccount,
                                         uint32_t ccap, uint32_t *dst, uint32_t n, hipStream_t s) {
    k_copy_counts_checked<<<(n + 255) / 256, 256, 0, s>>>(ocount, ccount, ccap, dst, n);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Batched launches: the frames of a launch may differ in size.  A table of
// GtBatchFrame slots (device memory, read through a pointer so a captured
// graph can be replayed for another launch) gives each slot's frame address,
// size, processed-plane offset and first tile; every stage is one persistent
// launch whose blocks walk the launch's tiles in XCD-contiguous ranges and
// run the same per-tile bodies as the one-size kernels.
namespace {

// Tile of this block in a table launch: blockIdx.y is the frame slot and
// blockIdx.x its tile, remapped so that each XCD (blockIdx.x % 8 for a
// launch whose rows of blocks are multiples of 8 wide) works through one
// contiguous range of the frame's tiles; blocks past the frame's tiles return
// at once.  The grid is sized for the largest frame of a call, so one captured
// graph serves every launch.
// The launch queue's tile of this workgroup (XCD-contiguous ranges of the
// queue, as xcd_tile): its frame slot and the tile's index in that frame.
// False for workgroups past the launch's tiles.
__device__ __forceinline__ bool queue_tile(const GtBatchHeader *__restrict__ hdr, uint32_t &slot,
                                           uint32_t &local, uint32_t &ntx) {
    const uint32_t ntiles = hdr->ntiles;
    const uint32_t per = (gridDim.x / kXcds);
    const uint32_t x = blockIdx.x % kXcds, j = blockIdx.x / kXcds;
    const uint32_t per_used = (ntiles + kXcds - 1) / kXcds;
    // the grid holds every launch's queue (the host sizes it for the largest)
    if (j >= per_used || per_used > per) return false;
    const uint32_t t = x * per_used + j;
    if (t >= ntiles) return false;
    // every segment is read at once (one round of scalar loads); unused
    // segments start past the launch's tiles
    const GtBatchHeader h = *hdr;
    uint32_t first_tile = h.seg[0].first_tile, first_slot = h.seg[0].first_slot;
    uint32_t tpf = h.seg[0].tpf, nx = h.seg[0].ntx;
#pragma unroll
    for (int q = 1; q < GT_BATCH_MAX_SEG; ++q) {
        if (h.seg[q].first_tile <= t) {
            first_tile = h.seg[q].first_tile;
            first_slot = h.seg[q].first_slot;
            tpf = h.seg[q].tpf;
            nx = h.seg[q].ntx;
        }
    }
    uint32_t rem;
    slot = first_slot + udiv22(t - first_tile, tpf, rem);
    local = rem;
    ntx = nx;
    return true;
}

template <typename T, int R, bool kD2>
__global__ __launch_bounds__(B2_THREADS) void k_bandpass_int_batch(
    float *__restrict__ proc, GtKParams p, const GtBatchHeader *__restrict__ hdr,
    const GtBatchFrame *__restrict__ fs) {
    __shared__ __align__(16) BpLds<R> L;
    uint32_t slot, local, ntx;
    if (!queue_tile(hdr, slot, local, ntx)) return;
    const GtBatchFrame e = fs[slot];
    uint32_t tx;
    const int ty = static_cast<int>(udiv22(local, ntx, tx));
    const int x0 = static_cast<int>(tx) * B2_TW, y0 = ty * B2_TH;
    bandpass_int_tile<T, R, kD2>(reinterpret_cast<const T *>(e.in_ptr), proc + e.proc_off, p, e.H, e.W,
                            x0, y0, L);
}

template <typename T, int KT, int RI>
__global__ __launch_bounds__(FT_THREADS, FT_MIN_WAVES) void k_detect_tile_batch(
    const float *__restrict__ proc, GtKParams p, TileGeom g_rt,
    const GtBatchHeader *__restrict__ hdr, const GtBatchFrame *__restrict__ fs,
    gt_detection *__restrict__ out, uint32_t out_cap, uint32_t *__restrict__ ocount,
    uint2 *__restrict__ slow, uint32_t slow_cap, uint32_t *__restrict__ scount) {
    constexpr bool kFixed = KT > 0 && RI > 0;
    constexpr TileGeom gc = tile_geom(KT > 0 ? KT : 1, RI > 0 ? RI : 1);
    const TileGeom g = kFixed ? gc : g_rt;
    extern __shared__ __align__(16) float lds[];
    __shared__ uint32_t s_cnt[FT_WAVES];
    uint32_t slot, local, ntx;
    if (!queue_tile(hdr, slot, local, ntx)) return;
    const GtBatchFrame e = fs[slot];
    uint32_t tx;
    const int ty = static_cast<int>(udiv22(local, ntx, tx));
    const int x0 = static_cast<int>(tx) * FT_TW, y0 = ty * FT_TH;
    // rows go to the output slot; counters and slow lists are per launch slot
    detect_tile_body<T, KT, RI>(proc + e.proc_off, reinterpret_cast<const T *>(e.in_ptr), p, e.H,
                                e.W, g, x0, y0, e.frame, out + static_cast<size_t>(e.out_idx) * out_cap,
                                out_cap, &ocount[slot * GT_COUNTER_STRIDE],
                                slow + static_cast<size_t>(slot) * slow_cap, slow_cap,
                                &scount[slot * GT_COUNTER_STRIDE], lds, s_cnt,
                                p.max_thr, p.image_std);
}

// Slow walks and characterize: slots over blockIdx.y (strided), detections
// over blockIdx.x.
template <typename T, int RI>
__global__ __launch_bounds__(WK_THREADS) void k_walk_slow_batch(
    const float *__restrict__ proc, GtKParams p, const GtBatchHeader *__restrict__ hdr,
    const GtBatchFrame *__restrict__ fs, const uint2 *__restrict__ slow, uint32_t slow_cap,
    const uint32_t *__restrict__ scount, gt_detection *__restrict__ out, uint32_t out_cap,
    uint32_t *__restrict__ ocount) {
    const int nslots = static_cast<int>(hdr->nslots);
    const uint32_t lane = __lane_id();
    const uint32_t stride = gridDim.x * WK_THREADS;
    for (int slot = blockIdx.y; slot < nslots; slot += gridDim.y) {
        const GtBatchFrame e = fs[slot];
        const int H = e.H, W = e.W;
        const float *P = proc + e.proc_off;
        const T *Rw = reinterpret_cast<const T *>(e.in_ptr);
        const uint32_t n = min(scount[slot * GT_COUNTER_STRIDE], slow_cap);
        const uint2 *fsl = slow + static_cast<size_t>(slot) * slow_cap;
        gt_detection *fout = out + static_cast<size_t>(e.out_idx) * out_cap;
        uint32_t *counter = &ocount[slot * GT_COUNTER_STRIDE];
        const int rint = RI > 0 ? RI : p.rint;
        for (uint32_t i0 = blockIdx.x * WK_THREADS + (threadIdx.x & ~63u); i0 < n; i0 += stride) {
            const uint32_t i = i0 + lane;
            bool keep = false;
            gt_detection d;
            if (i < n) {
                const uint2 en = fsl[i];
                int pu = static_cast<int>(en.y / static_cast<uint32_t>(W));
                int pv = static_cast<int>(en.y - static_cast<uint32_t>(pu) * static_cast<uint32_t>(W));
                WalkRes res = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
                for (uint32_t itr = FT_WALK_STEPS; itr < p.max_iter; ++itr) {
                    res = walk_eval<RI, 0, (sizeof(T) <= 2)>(p, H, W, P, nullptr, 0, 0, 0, 1, 0, 1, 0, pu, pv);
                    if (!walk_move(res, pu, pv, H, W, rint, p.shift_thr)) break;
                }
                keep = finish_walk<T, RI, 0, false>(p, H, W, P, Rw, nullptr, 0, 0, 0, 1, 0, 1, 0, res, pu,
                                                    pv, en.x, e.frame, d, p.image_std);
            }
            const uint32_t pos = wave_append(keep, counter);
            if (keep && pos < out_cap) fout[pos] = d;
        }
    }
}

template <typename T, int RI>
__global__ __launch_bounds__(WK_THREADS) void k_characterize_det_batch(
    const float *__restrict__ proc, GtKParams p, const GtBatchHeader *__restrict__ hdr,
    const GtBatchFrame *__restrict__ fs, gt_detection *__restrict__ out, uint32_t out_cap,
    const uint32_t *__restrict__ ocount, uint32_t *__restrict__ counts) {
    const int nslots = static_cast<int>(hdr->nslots);
    for (int slot = blockIdx.y; slot < nslots; slot += gridDim.y) {
        const GtBatchFrame e = fs[slot];
        // the counts are final here: copy them out (k_copy_counts_batch)
        if (counts && blockIdx.x == 0 && threadIdx.x == 0)
            counts[e.out_idx] = ocount[slot * GT_COUNTER_STRIDE];
        const int H = e.H, W = e.W;
        const GlobalSrc src{proc + e.proc_off, W};
        const T *Rw = reinterpret_cast<const T *>(e.in_ptr);
        const uint32_t n = min(ocount[slot * GT_COUNTER_STRIDE], out_cap);
        gt_detection *rows = out + static_cast<size_t>(e.out_idx) * out_cap;
        for (uint32_t i = blockIdx.x * WK_THREADS + threadIdx.x; i < n; i += gridDim.x * WK_THREADS) {
            gt_result_row row = rows[i].row;
            if constexpr (RI > 0) {
                const int cu = f2i(rintf(row.x)), cv = f2i(rintf(row.y));
                if (row.r <= 0.0f && p.rint == RI && cu >= RI && cu + RI < H && cv >= RI &&
                    cv + RI < W) {
                    const size_t c = static_cast<size_t>(cu) * W + cv;
                    characterize_inner<T, RI>(GlobalRow{src.p + c, W}, Rw + c, W, row);
                    rows[i].row = row;
                    continue;
                }
            }
            characterize_row<T, RI, true>(src, Rw, H, W, p.circle, row);
            rows[i].row = row;
        }
    }
}

// First and last kernels of a launch: copy launch *cur's entries into the
// fixed table the stage kernels read, and advance *cur afterwards.
__global__ void k_select_launch(const GtBatchLaunch *__restrict__ launches,
                                const uint32_t *__restrict__ cur,
                                const GtBatchFrame *__restrict__ all,
                                GtBatchHeader *__restrict__ hdr, GtBatchFrame *__restrict__ fs,
                                uint32_t *__restrict__ ocount, uint32_t *__restrict__ scount,
                                uint32_t max_slots) {
    const GtBatchLaunch L = launches[*cur];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < L.nslots) fs[i] = all[L.first + i];
    if (i < max_slots) {
        // the launch's counters (one per slot, GT_COUNTER_STRIDE apart)
        ocount[i * GT_COUNTER_STRIDE] = 0u;
        scount[i * GT_COUNTER_STRIDE] = 0u;
    }
    if (i == 0) {
        GtBatchHeader h = {};
        h.nslots = L.nslots;
        h.nseg = L.nseg;
        h.ntiles = L.ntiles;
        for (uint32_t q = 0; q < GT_BATCH_MAX_SEG; ++q) {
            h.seg[q] = L.seg[q];
            if (q >= L.nseg) h.seg[q].first_tile = 0xFFFFFFFFu;
        }
        *hdr = h;
    }
}

// Advances *cur once k_select_launch has read it (the launch's last kernel).
__global__ void k_next_launch(uint32_t *__restrict__ cur) {
    if (threadIdx.x == 0) *cur = *cur + 1u;
}

__global__ void k_copy_counts_batch(const uint32_t *__restrict__ strided,
                                    const GtBatchHeader *__restrict__ hdr,
                                    const GtBatchFrame *__restrict__ fs, uint32_t *__restrict__ dst) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < hdr->nslots) dst[fs[i].out_idx] = strided[i * GT_COUNTER_STRIDE];
}

template <typename T>
hipError_t launch_batch_t(const GtBatchLaunch *launches, uint32_t *cur, const GtBatchFrame *all,
                          const GtKParams &p, GtBatchHeader *hdr, GtBatchFrame *fs,
                          uint32_t max_slots, float *proc, gt_detection *out, uint32_t out_cap,
                          uint32_t *ocount, uint2 *slow, uint32_t slow_cap, uint32_t *scount,
                          uint32_t *counts, const GtBatchSchedule &sch, hipStream_t s) {
    if (p.R < 1 || p.R > 7 || !p.fast_div) return hipErrorInvalidValue;
    k_select_launch<<<(max_slots + 255) / 256, 256, 0, s>>>(launches, cur, all, hdr, fs, ocount,
                                                            scount, max_slots);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // one queue of the launch's tiles over all its frames (blocks past the
    // launch's tiles return)
    const dim3 gbp(sch.tile_blocks);
    with_bandpass_form(p, [&](auto r, auto d2) {
        k_bandpass_int_batch<T, decltype(r)::value, decltype(d2)::value>
            <<<gbp, B2_THREADS, 0, s>>>(proc, p, hdr, fs);
    });
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    GtKParams kw = p;
    kw.characterize = 0;
    const TileGeom g = tile_geom(p.k, p.rint);
    const size_t lds = sizeof(float) * (static_cast<size_t>(g.rows) * g.pitch +
                                        static_cast<size_t>(FT_TH + p.k - 1) * FT_TW);
    const dim3 gdt(sch.tile_blocks);
#define GT_TILEB(KV, RV)                                                                     \
    k_detect_tile_batch<T, KV, RV><<<gdt, FT_THREADS, lds, s>>>(proc, kw, g, hdr, fs, out,      \
                                                                out_cap, ocount, slow, slow_cap, \
                                                                scount)
    if (p.rint == 4 && p.k == 9) GT_TILEB(9, 4);
    else if (p.rint == 4 && p.k == 12) GT_TILEB(12, 4);
    else if (p.rint == 4) GT_TILEB(0, 4);
    else GT_TILEB(0, 0);
#undef GT_TILEB
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    const dim3 gw(sch.walk_blocks, sch.launch_slots);
    if (p.rint == 4)
        k_walk_slow_batch<T, 4><<<gw, WK_THREADS, 0, s>>>(proc, kw, hdr, fs, slow, slow_cap, scount,
                                                          out, out_cap, ocount);
    else
        k_walk_slow_batch<T, 0><<<gw, WK_THREADS, 0, s>>>(proc, kw, hdr, fs, slow, slow_cap, scount,
                                                          out, out_cap, ocount);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (p.characterize) {
        if (p.rint == 4)
            k_characterize_det_batch<T, 4><<<gw, WK_THREADS, 0, s>>>(proc, p, hdr, fs, out, out_cap,
                                                                     ocount, counts);
        else
            k_characterize_det_batch<T, 0><<<gw, WK_THREADS, 0, s>>>(proc, p, hdr, fs, out, out_cap,
                                                                     ocount, counts);
    } else {
        k_copy_counts_batch<<<(max_slots + 255) / 256, 256, 0, s>>>(ocount, hdr, fs, counts);
    }
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_next_launch<<<1, 64, 0, s>>>(cur);
    return hipGetLastError();
}

}  // namespace

size_t gt_batch_tile_lds_bytes(const GtKParams &p) { return gt_tile_lds_bytes(p); }

hipError_t gt_launch_batch(const GtBatchLaunch *launches, uint32_t *cur,
                           const GtBatchFrame *all, int dtype, const GtKParams &p,
                           GtBatchHeader *hdr, GtBatchFrame *fs, uint32_t max_slots, float *proc,
                           gt_detection *out, uint32_t out_cap, uint32_t *ocount, uint2 *slow,
                           uint32_t slow_cap, uint32_t *scount, uint32_t *counts,
                           const GtBatchSchedule &sch, hipStream_t s) {
    switch (dtype) {
    case GT_U16:
        return launch_batch_t<uint16_t>(launches, cur, all, p, hdr, fs, max_slots, proc, out,
                                        out_cap, ocount, slow, slow_cap, scount, counts, sch, s);
    case GT_U8:
        return launch_batch_t<uint8_t>(launches, cur, all, p, hdr, fs, max_slots, proc, out,
                                       out_cap, ocount, slow, slow_cap, scount, counts, sch, s);
    default: return hipErrorInvalidValue;
    }
}

// Tile counts of one frame in the batched kernels.
static_assert(B2_TW == FT_TW && B2_TH == FT_TH, "one tile queue serves bandpass and detect");
uint32_t gt_batch_tile_rows() { return static_cast<uint32_t>(FT_TH); }
uint32_t gt_batch_tile_cols() { return static_cast<uint32_t>(FT_TW); }
void gt_batch_tiles(int H, int W, uint32_t *bp_tiles, uint32_t *dt_tiles) {
    *bp_tiles = static_cast<uint32_t>(((W + B2_TW - 1) / B2_TW) * ((H + B2_TH - 1) / B2_TH));
    *dt_tiles = static_cast<uint32_t>(((W + FT_TW - 1) / FT_TW) * ((H + FT_TH - 1) / FT_TH));
}

hipError_t gt_plain_head(const void *frames, int dtype, uint32_t nf, const GtKParams &p,
                         const GtPlainScratch &sc, gt_detection *out, uint32_t cap,
                         uint32_t frame0, hipStream_t s, hipEvent_t mid) {
    const uint32_t ncnt = nf * GT_COUNTER_STRIDE;
    const GtZero zero = {{sc.ocount, sc.scount}, {ncnt, ncnt}};
    hipError_t e = gt_launch_bandpass(frames, dtype, sc.proc, nf, p, s, &zero);
    if (e != hipSuccess) return e;
    if (mid && (e = hipEventRecord(mid, s)) != hipSuccess) return e;
    GtKParams kw = p;
    kw.characterize = 0;   // characterize runs in the tail over compacted rows
    return gt_launch_detect_tile(sc.proc, frames, dtype, nf, kw, out, cap, sc.ocount, sc.slow,
                                 sc.slow_cap, sc.scount, frame0, s);
}

hipError_t gt_plain_tail(const void *frames, int dtype, uint32_t nf, const GtKParams &p,
                         const GtPlainScratch &sc, gt_detection *out, uint32_t cap,
                         uint32_t *counts, uint32_t frame0, hipStream_t s) {
    GtKParams kw = p;
    kw.characterize = 0;
    hipError_t e = gt_launch_walk_slow(sc.proc, frames, dtype, nf, kw, sc.slow, sc.slow_cap,
                                       sc.scount, out, cap, sc.ocount, frame0, s);
    if (e != hipSuccess) return e;
    if (p.characterize)   // the counts are final here: characterize copies them too
        return gt_launch_characterize_det(sc.proc, frames, dtype, nf, p, out, cap, sc.ocount, s,
                                          counts);
    return gt_launch_copy_counts(sc.ocount, counts, nf, s);
}
