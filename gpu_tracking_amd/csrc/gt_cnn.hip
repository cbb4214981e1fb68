// gt_cnn.hip -- learned centroid proposal on the gfx950 matrix cores.
//
// One block = four waves; each wave carries one sample (candidate) through
// the network with its activations in its own LDS region:
//   conv1 (M = 81 positions, N = 16, K = 9 taps padded to 32),
//   conv2 (M = 49 positions, N = 32, K = 9 taps x 16 ch) and
//   conv3 (M = 25, N = 32, K = 9 taps x 32 ch) as implicit-im2col GEMMs of
//   v_mfma_f32_16x16x32_bf16 (lane l holds A row l%16, k 8*(l/16)..+8;
//   B column l%16, same k; D column l%16, rows 4*(l/16)..+4),
//   mean pooling and the 32 -> 3 head with wave shuffles.
// Training adds the backward pass: input gradients of conv3 and conv2 per
// wave (transposed-convolution GEMMs on the matrix cores), and weight
// gradients accumulated by the whole block over its four samples (the
// GEMM reduction dimension runs over positions x samples), kept in
// registers across the block's batch share and written once as a per-block
// partial; k_cnn_reduce sums the partials in block order (deterministic)
// and applies Adam.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "gt_cnn.h"
#include "gt_internal.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

constexpr int CT = 256;  // four waves

// float32 -> bfloat16, round to nearest even (finite inputs): the gfx950
// conversion instruction (v_cvt_pk_bf16_f32)
__device__ __forceinline__ uint32_t bf16_bits(float f) {
    return __builtin_bit_cast(uint16_t, static_cast<__bf16>(f));
}
__device__ __forceinline__ float bf16_val(uint32_t h) { return __uint_as_float(h << 16); }
__device__ __forceinline__ float bf16_round(float f) { return bf16_val(bf16_bits(f)); }

__device__ __forceinline__ bf16x8 frag_zero() {
    const u32x4 z = {0u, 0u, 0u, 0u};
    return __builtin_bit_cast(bf16x8, z);
}
// 8 consecutive bf16 (16-byte aligned, LDS or global)
__device__ __forceinline__ bf16x8 frag_ld(const uint16_t *p) {
    return __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4 *>(p));
}
__device__ __forceinline__ bf16x8 frag_pack(const uint32_t (&v)[8]) {
    u32x4 w;
    w[0] = v[0] | (v[1] << 16);
    w[1] = v[2] | (v[3] << 16);
    w[2] = v[4] | (v[5] << 16);
    w[3] = v[6] | (v[7] << 16);
    return __builtin_bit_cast(bf16x8, w);
}
__device__ __forceinline__ f32x4 mma(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
// Orders the wave's LDS accesses across it for the compiler (no instruction).
__device__ __forceinline__ void wave_lds_fence() { asm volatile("" ::: "memory"); }
__device__ __forceinline__ f32x4 zero4() {
    const f32x4 z = {0.0f, 0.0f, 0.0f, 0.0f};
    return z;
}

// Weights staged once per block.
struct alignas(16) BlockW {
    uint16_t w2p[32 * CNN_K2];
    uint16_t w3p[32 * CNN_K3];
    float f[CNN_F_N];
};
// One wave's activations.
struct alignas(16) WaveAct {
    float x[128];            // input patch (bf16 values), 11 x 11
    uint16_t a1[81 * 16];    // conv1 output [position][channel]
    uint16_t a2[64 * 32];    // conv2 output [position][channel] (rows >= 49 unused)
    float a3[25 * 32];       // conv3 output (bf16 values)
    float g[32];             // pooled (bf16 values)
    float o[4];
};
// One wave's gradients (training); padding rows and columns stay zero.
struct alignas(16) WaveGrad {
    uint16_t dz3[32 * 32];   // [m][n], m >= 25 zero
    uint16_t dz3t[32 * 32];  // [n][m]
    uint16_t dz2[64 * 32];   // [q][c], q >= 49 zero
    uint16_t dz2t[32 * 64];  // [c][q]
    float dz1[81 * 16];      // [q][c]
};

// The forward kernel's staged weights: only the f32 tensors (conv1, the
// biases, the head); its conv2 / conv3 B fragments are read from global
// memory straight into registers.  Keeps its LDS at about 43 KB: three
// blocks per CU.
struct alignas(16) FwdW {
    float f[CNN_F_N];
};

constexpr size_t kFwdLds = sizeof(FwdW) + 4 * sizeof(WaveAct);
constexpr size_t kTrainActs = sizeof(BlockW) + 4 * sizeof(WaveAct);   // the train kernel's
constexpr size_t kTrainLds = kTrainActs + 4 * sizeof(WaveGrad) + 32 * sizeof(float);

__device__ __forceinline__ void stage_weights(const uint8_t *__restrict__ pk, BlockW &B, int tid) {
    constexpr int NV = static_cast<int>((CNN_PK_W2T - CNN_PK_W2P) / 16);  // w2p and w3p
    const u32x4 *src = reinterpret_cast<const u32x4 *>(pk + CNN_PK_W2P);
    u32x4 *dst = reinterpret_cast<u32x4 *>(B.w2p);
    for (int i = tid; i < NV; i += CT) dst[i] = src[i];
    const float *fs = reinterpret_cast<const float *>(pk + CNN_PK_F32);
    for (int i = tid; i < CNN_F_N; i += CT) B.f[i] = fs[i];
}

__device__ __forceinline__ void stage_fwd_weights(const uint8_t *__restrict__ pk, FwdW &B,
                                                  int tid) {
    const float *fs = reinterpret_cast<const float *>(pk + CNN_PK_F32);
    for (int i = tid; i < CNN_F_N; i += CT) B.f[i] = fs[i];
}

// The two raw patch values lane `lane` owns (indices lane and lane + 64) of
// the 11 x 11 patch around pixel `pix`; 0 outside the frame.
__device__ __forceinline__ void gather_raw(const float *__restrict__ P, int H, int W, uint32_t pix,
                                           bool valid, int lane, float &v0, float &v1) {
    v0 = 0.0f;
    v1 = 0.0f;
    if (valid) {
        const int cu = static_cast<int>(pix / static_cast<uint32_t>(W));
        const int cv = static_cast<int>(pix - static_cast<uint32_t>(cu) * static_cast<uint32_t>(W));
        {
            const int i = lane, u = cu + i / 11 - 5, v = cv + i % 11 - 5;
            if (u >= 0 && u < H && v >= 0 && v < W) v0 = P[static_cast<size_t>(u) * W + v];
        }
        const int i = lane + 64;
        if (i < 121) {
            const int u = cu + i / 11 - 5, v = cv + i % 11 - 5;
            if (u >= 0 && u < H && v >= 0 && v < W) v1 = P[static_cast<size_t>(u) * W + v];
        }
    }
}
// The raw values normalised by the patch's largest |value| and rounded to
// bf16 (bit patterns).
__device__ __forceinline__ void normalise(float v0, float v1, int lane, uint32_t &h0, uint32_t &h1) {
    float mx = fmaxf(fabsf(v0), fabsf(v1));
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off));
    h0 = mx > 0.0f ? bf16_bits(v0 / mx) : 0u;
    h1 = (mx > 0.0f && lane + 64 < 121) ? bf16_bits(v1 / mx) : 0u;
}
__device__ __forceinline__ void gather_norm(const float *__restrict__ P, int H, int W,
                                            uint32_t pix, bool valid, int lane, uint32_t &h0,
                                            uint32_t &h1) {
    float v0, v1;
    gather_raw(P, H, W, pix, valid, lane, v0, v1);
    normalise(v0, v1, lane, h0, h1);
}

// ---- forward phases (one wave, its own sample) ---------------------------
// conv1 (M = 81 positions in 6 tiles, N = 16 channels, K = 9 taps padded to
// 32) on the matrix cores: lane l's A row is position m = 16 mt + l % 16 with
// taps 8 (l / 16) .. + 8 of its 3 x 3 window, its B column channel l % 16
// with the same taps of w1 (bf16); the bias is added to the f32 sum.
template <class WB>
__device__ __forceinline__ void conv1(const WB &B, WaveAct &A, int lane) {
    const int r16 = lane & 15, kg = lane >> 4;
    uint32_t wv[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int t = kg * 8 + e;
        wv[e] = t < 9 ? (__float_as_uint(B.f[CNN_F_W1 + r16 * 9 + t]) >> 16) : 0u;
    }
    const bf16x8 b = frag_pack(wv);
    const float bias = B.f[CNN_F_B1 + r16];
#pragma unroll
    for (int mt = 0; mt < 6; ++mt) {
        const int m = mt * 16 + r16;
        const int py = m / 9, px = m - py * 9;
        uint32_t xv[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int t = kg * 8 + e;
            xv[e] = (m < 81 && t < 9) ? (__float_as_uint(A.x[(py + t / 3) * 11 + px + t % 3]) >> 16)
                                      : 0u;
        }
        const f32x4 d = mma(frag_pack(xv), b, zero4());
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int row = mt * 16 + kg * 4 + i;
            if (row < 81)
                A.a1[row * 16 + r16] = static_cast<uint16_t>(bf16_bits(fmaxf(d[i] + bias, 0.0f)));
        }
    }
}

__device__ __forceinline__ void conv2(const BlockW &B, WaveAct &A, int lane) {
    const int r16 = lane & 15, kg = lane >> 4;
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        const int m = mt * 16 + r16;
        const bool mv = m < 49;
        const int my = m / 7, mx = m - my * 7;
        f32x4 c0 = zero4(), c1 = zero4();
#pragma unroll
        for (int s = 0; s < 5; ++s) {
            const int k0 = s * 32 + kg * 8, tap = k0 >> 4, ch = k0 & 15;
            bf16x8 a = frag_zero();
            if (mv && tap < 9) a = frag_ld(&A.a1[((my + tap / 3) * 9 + mx + tap % 3) * 16 + ch]);
            c0 = mma(a, frag_ld(&B.w2p[r16 * CNN_K2 + k0]), c0);
            c1 = mma(a, frag_ld(&B.w2p[(16 + r16) * CNN_K2 + k0]), c1);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int row = mt * 16 + kg * 4 + i;
            if (row < 49) {
                A.a2[row * 32 + r16] =
                    static_cast<uint16_t>(bf16_bits(fmaxf(c0[i] + B.f[CNN_F_B2 + r16], 0.0f)));
                A.a2[row * 32 + 16 + r16] =
                    static_cast<uint16_t>(bf16_bits(fmaxf(c1[i] + B.f[CNN_F_B2 + 16 + r16], 0.0f)));
            }
        }
    }
}

__device__ __forceinline__ void conv3(const BlockW &B, WaveAct &A, int lane) {
    const int r16 = lane & 15, kg = lane >> 4;
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
        const int m = mt * 16 + r16;
        const bool mv = m < 25;
        const int my = m / 5, mx = m - my * 5;
        f32x4 c0 = zero4(), c1 = zero4();
#pragma unroll
        for (int s = 0; s < 9; ++s) {
            const int ky = s / 3, kx = s - ky * 3;
            bf16x8 a = frag_zero();
            if (mv) a = frag_ld(&A.a2[((my + ky) * 7 + mx + kx) * 32 + kg * 8]);
            c0 = mma(a, frag_ld(&B.w3p[r16 * CNN_K3 + s * 32 + kg * 8]), c0);
            c1 = mma(a, frag_ld(&B.w3p[(16 + r16) * CNN_K3 + s * 32 + kg * 8]), c1);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int row = mt * 16 + kg * 4 + i;
            if (row < 25) {
                A.a3[row * 32 + r16] = bf16_round(fmaxf(c0[i] + B.f[CNN_F_B3 + r16], 0.0f));
                A.a3[row * 32 + 16 + r16] = bf16_round(fmaxf(c1[i] + B.f[CNN_F_B3 + 16 + r16], 0.0f));
            }
        }
    }
}

// The conv2 and conv3 B operands of this lane, held in registers across the
// samples of the forward kernel (the weights do not change between samples,
// so only the activations are read from LDS per sample).
struct WeightFrags {
    bf16x8 w2[5][2];
    bf16x8 w3[9][2];
};
// The lane's conv2 / conv3 B fragments from the packed weights in global
// memory (the w2p / w3p layouts stage_weights copies to LDS).
__device__ __forceinline__ void load_weight_frags(const uint8_t *__restrict__ pk, int lane,
                                                  WeightFrags &F) {
    const uint16_t *w2p = reinterpret_cast<const uint16_t *>(pk + CNN_PK_W2P);
    const uint16_t *w3p = reinterpret_cast<const uint16_t *>(pk + CNN_PK_W3P);
    const int r16 = lane & 15, kg = lane >> 4;
#pragma unroll
    for (int st = 0; st < 5; ++st) {
        const int k0 = st * 32 + kg * 8;
        F.w2[st][0] = frag_ld(&w2p[r16 * CNN_K2 + k0]);
        F.w2[st][1] = frag_ld(&w2p[(16 + r16) * CNN_K2 + k0]);
    }
#pragma unroll
    for (int st = 0; st < 9; ++st) {
        F.w3[st][0] = frag_ld(&w3p[r16 * CNN_K3 + st * 32 + kg * 8]);
        F.w3[st][1] = frag_ld(&w3p[(16 + r16) * CNN_K3 + st * 32 + kg * 8]);
    }
}

// conv2 and conv3 as above with the B operands from registers (the same
// products in the same order).
template <class WB>
__device__ __forceinline__ void conv2_rf(const WB &B, const WeightFrags &F, WaveAct &A,
                                         int lane) {
    const int r16 = lane & 15, kg = lane >> 4;
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        const int m = mt * 16 + r16;
        const bool mv = m < 49;
        const int my = m / 7, mx = m - my * 7;
        f32x4 c0 = zero4(), c1 = zero4();
#pragma unroll
        for (int st = 0; st < 5; ++st) {
            const int k0 = st * 32 + kg * 8, tap = k0 >> 4, ch = k0 & 15;
            bf16x8 a = frag_zero();
            if (mv && tap < 9) a = frag_ld(&A.a1[((my + tap / 3) * 9 + mx + tap % 3) * 16 + ch]);
            c0 = mma(a, F.w2[st][0], c0);
            c1 = mma(a, F.w2[st][1], c1);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int row = mt * 16 + kg * 4 + i;
            if (row < 49) {
                A.a2[row * 32 + r16] =
                    static_cast<uint16_t>(bf16_bits(fmaxf(c0[i] + B.f[CNN_F_B2 + r16], 0.0f)));
                A.a2[row * 32 + 16 + r16] =
                    static_cast<uint16_t>(bf16_bits(fmaxf(c1[i] + B.f[CNN_F_B2 + 16 + r16], 0.0f)));
            }
        }
    }
}
template <class WB>
__device__ __forceinline__ void conv3_rf(const WB &B, const WeightFrags &F, WaveAct &A,
                                         int lane) {
    const int r16 = lane & 15, kg = lane >> 4;
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
        const int m = mt * 16 + r16;
        const bool mv = m < 25;
        const int my = m / 5, mx = m - my * 5;
        f32x4 c0 = zero4(), c1 = zero4();
#pragma unroll
        for (int st = 0; st < 9; ++st) {
            const int ky = st / 3, kx = st - ky * 3;
            bf16x8 a = frag_zero();
            if (mv) a = frag_ld(&A.a2[((my + ky) * 7 + mx + kx) * 32 + kg * 8]);
            c0 = mma(a, F.w3[st][0], c0);
            c1 = mma(a, F.w3[st][1], c1);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int row = mt * 16 + kg * 4 + i;
            if (row < 25) {
                A.a3[row * 32 + r16] = bf16_round(fmaxf(c0[i] + B.f[CNN_F_B3 + r16], 0.0f));
                A.a3[row * 32 + 16 + r16] = bf16_round(fmaxf(c1[i] + B.f[CNN_F_B3 + 16 + r16], 0.0f));
            }
        }
    }
}

// Mean pooling and the linear head; every lane gets o[0..2].
template <class WB>
__device__ __forceinline__ void head(const WB &B, WaveAct &A, int lane, float (&o)[3]) {
    float g = 0.0f;
    if (lane < 32) {
        for (int m = 0; m < 25; ++m) g = g + A.a3[m * 32 + lane];
        g = bf16_round(g / 25.0f);
        A.g[lane] = g;
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        float v = lane < 32 ? B.f[CNN_F_WFC + j * 32 + lane] * g : 0.0f;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) v = v + __shfl_xor(v, off);
        o[j] = B.f[CNN_F_BFC + j] + v;
    }
}

// ---- backward phases -----------------------------------------------------
// conv3 input gradient of the wave's sample: dz2 = bf16(da2 * (a2 > 0)),
// da2[q][c] = sum_{tap, n} dz3[q - tap][n] * w3[n][c][tap].
__device__ __forceinline__ void da2_own(const uint16_t *__restrict__ w3t, const WaveAct &A,
                                        WaveGrad &G, int lane) {
    const int r16 = lane & 15, kg = lane >> 4;
    for (int mt = 0; mt < 4; ++mt) {
        const int q = mt * 16 + r16;
        const bool qv = q < 49;
        const int qy = q / 7, qx = q - qy * 7;
        f32x4 c0 = zero4(), c1 = zero4();
#pragma unroll
        for (int s = 0; s < 9; ++s) {
            const int ky = s / 3, kx = s - ky * 3, my = qy - ky, mx = qx - kx;
            bf16x8 a = frag_zero();
            if (qv && my >= 0 && my < 5 && mx >= 0 && mx < 5)
                a = frag_ld(&G.dz3[(my * 5 + mx) * 32 + kg * 8]);
            c0 = mma(a, frag_ld(&w3t[r16 * CNN_K3 + s * 32 + kg * 8]), c0);
            c1 = mma(a, frag_ld(&w3t[(16 + r16) * CNN_K3 + s * 32 + kg * 8]), c1);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int row = mt * 16 + kg * 4 + i;
            if (row < 49) {
                const uint32_t v0 = bf16_val(A.a2[row * 32 + r16]) > 0.0f ? bf16_bits(c0[i]) : 0u;
                const uint32_t v1 = bf16_val(A.a2[row * 32 + 16 + r16]) > 0.0f ? bf16_bits(c1[i]) : 0u;
                G.dz2[row * 32 + r16] = static_cast<uint16_t>(v0);
                G.dz2[row * 32 + 16 + r16] = static_cast<uint16_t>(v1);
                G.dz2t[r16 * 64 + row] = static_cast<uint16_t>(v0);
                G.dz2t[(16 + r16) * 64 + row] = static_cast<uint16_t>(v1);
            }
        }
    }
}

// conv2 input gradient: dz1 = da1 * (a1 > 0) (float32),
// da1[q][c] = sum_{tap, n} dz2[q - tap][n] * w2[n][c][tap].
__device__ __forceinline__ void da1_own(const uint16_t *__restrict__ w2t, const WaveAct &A,
                                        WaveGrad &G, int lane) {
    const int r16 = lane & 15, kg = lane >> 4;
    for (int mt = 0; mt < 6; ++mt) {
        const int q = mt * 16 + r16;
        const bool qv = q < 81;
        const int qy = q / 9, qx = q - qy * 9;
        f32x4 c = zero4();
#pragma unroll
        for (int s = 0; s < 9; ++s) {
            const int ky = s / 3, kx = s - ky * 3, my = qy - ky, mx = qx - kx;
            bf16x8 a = frag_zero();
            if (qv && my >= 0 && my < 7 && mx >= 0 && mx < 7)
                a = frag_ld(&G.dz2[(my * 7 + mx) * 32 + kg * 8]);
            c = mma(a, frag_ld(&w2t[r16 * CNN_K3 + s * 32 + kg * 8]), c);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int row = mt * 16 + kg * 4 + i;
            if (row < 81)
                G.dz1[row * 16 + r16] = bf16_val(A.a1[row * 16 + r16]) > 0.0f ? c[i] : 0.0f;
        }
    }
}

// Block-wide conv3 weight gradient over the four waves' samples:
// dW3[k][n] += sum_m a2col[m][k] dz3[m][n], k = tap * 32 + c.  Wave w owns
// row tiles kt = w, w + 4, ... (< 18).
__device__ __forceinline__ void dw3_acc(const WaveAct *Aall, const WaveGrad *Gall, int wave,
                                        int lane, f32x4 (&acc)[5][2]) {
    const int r16 = lane & 15, kg = lane >> 4;
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        const int kt = wave + 4 * i;
        if (kt < 18) {
            const int k = kt * 16 + r16, tap = k >> 5, c = k & 31, ky = tap / 3, kx = tap - ky * 3;
            for (int w = 0; w < 4; ++w) {
                uint32_t v[8];
#pragma unroll
                for (int jj = 0; jj < 8; ++jj) {
                    const int m = kg * 8 + jj;
                    v[jj] = m < 25 ? Aall[w].a2[((m / 5 + ky) * 7 + m % 5 + kx) * 32 + c] : 0u;
                }
                const bf16x8 a = frag_pack(v);
                acc[i][0] = mma(a, frag_ld(&Gall[w].dz3t[r16 * 32 + kg * 8]), acc[i][0]);
                acc[i][1] = mma(a, frag_ld(&Gall[w].dz3t[(16 + r16) * 32 + kg * 8]), acc[i][1]);
            }
        }
    }
}

// Block-wide conv2 weight gradient: dW2[k][n] += sum_q a1col[q][k] dz2[q][n],
// k = tap * 16 + c; wave w owns row tiles kt = w, w + 4, w + 8 (< 9).
__device__ __forceinline__ void dw2_acc(const WaveAct *Aall, const WaveGrad *Gall, int wave,
                                        int lane, f32x4 (&acc)[3][2]) {
    const int r16 = lane & 15, kg = lane >> 4;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const int kt = wave + 4 * i;
        if (kt < 9) {
            const int k = kt * 16 + r16, tap = k >> 4, c = k & 15, ky = tap / 3, kx = tap - ky * 3;
            for (int w = 0; w < 4; ++w) {
#pragma unroll
                for (int st = 0; st < 2; ++st) {
                    uint32_t v[8];
#pragma unroll
                    for (int jj = 0; jj < 8; ++jj) {
                        const int q = st * 32 + kg * 8 + jj;
                        v[jj] = q < 49 ? Aall[w].a1[((q / 7 + ky) * 9 + q % 7 + kx) * 16 + c] : 0u;
                    }
                    const bf16x8 a = frag_pack(v);
                    acc[i][0] = mma(a, frag_ld(&Gall[w].dz2t[r16 * 64 + st * 32 + kg * 8]), acc[i][0]);
                    acc[i][1] = mma(a, frag_ld(&Gall[w].dz2t[(16 + r16) * 64 + st * 32 + kg * 8]), acc[i][1]);
                }
            }
        }
    }
}

// ---- kernels -------------------------------------------------------------
template <bool kFrames>
__global__ __launch_bounds__(CT, 2) void k_cnn_forward(const uint8_t *__restrict__ pk,
                                                       const uint16_t *__restrict__ patches,
                                                       uint32_t n, const float *__restrict__ proc,
                                                       int H, int W,
                                                       const uint32_t *__restrict__ cand,
                                                       uint32_t cstride, uint32_t cap,
                                                       const uint32_t *__restrict__ ccount,
                                                       float4 *__restrict__ out) {
    extern __shared__ __align__(16) uint8_t smem[];
    FwdW &B = *reinterpret_cast<FwdW *>(smem);
    WaveAct *Aall = reinterpret_cast<WaveAct *>(smem + sizeof(FwdW));
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    stage_fwd_weights(pk, B, tid);
    uint32_t count = n;
    const float *P = proc;
    const uint32_t *fc = cand;
    float4 *fo = out;
    if constexpr (kFrames) {
        const uint32_t f = blockIdx.y;
        count = min(ccount[f * GT_COUNTER_STRIDE], cap);
        P = proc + static_cast<size_t>(f) * H * W;
        fc = cand + static_cast<size_t>(f) * cap * cstride;
        fo = out + static_cast<size_t>(f) * cap;
    }
    WaveAct &A = Aall[wave];
    const uint32_t per = 4u * gridDim.x;
    const uint32_t iters = (count + per - 1) / per;   // uniform across the block
    __syncthreads();
    WeightFrags F;
    load_weight_frags(pk, lane, F);
    // sample of iteration it: j(it); its inputs are loaded one iteration
    // ahead (frames: the candidate pixel two ahead), so the loads' latency
    // overlaps the previous sample's network
    auto jof = [&](uint32_t it) {
        return (it * gridDim.x + blockIdx.x) * 4u + static_cast<uint32_t>(wave);
    };
    float v0 = 0.0f, v1 = 0.0f;      // inputs of the current sample
    uint32_t pix_next = 0u;          // frames: candidate pixel of the next sample
    if (iters > 0) {
        const uint32_t j0 = jof(0);
        if constexpr (kFrames) {
            const uint32_t pix0 = j0 < count ? fc[static_cast<size_t>(j0) * cstride] : 0u;
            gather_raw(P, H, W, pix0, j0 < count, lane, v0, v1);
            const uint32_t j1 = jof(1);
            pix_next = j1 < count ? fc[static_cast<size_t>(j1) * cstride] : 0u;
        } else {
            if (j0 < count) {
                v0 = bf16_val(patches[static_cast<size_t>(j0) * CNN_PATCH_PITCH + lane]);
                if (lane + 64 < 121)
                    v1 = bf16_val(patches[static_cast<size_t>(j0) * CNN_PATCH_PITCH + lane + 64]);
            }
        }
    }
    for (uint32_t it = 0; it < iters; ++it) {
        const uint32_t j = jof(it);
        const bool valid = j < count;
        if constexpr (kFrames) {
            uint32_t h0, h1;
            normalise(v0, v1, lane, h0, h1);
            A.x[lane] = bf16_val(h0);
            A.x[lane + 64] = bf16_val(h1);
        } else {
            A.x[lane] = v0;
            A.x[lane + 64] = v1;
        }
        // the next sample's inputs
        const uint32_t jn = jof(it + 1);
        const bool vn = it + 1 < iters && jn < count;
        if constexpr (kFrames) {
            gather_raw(P, H, W, pix_next, vn, lane, v0, v1);
            const uint32_t jnn = jof(it + 2);
            pix_next = (it + 2 < iters && jnn < count) ? fc[static_cast<size_t>(jnn) * cstride] : 0u;
        } else {
            v0 = 0.0f;
            v1 = 0.0f;
            if (vn) {
                v0 = bf16_val(patches[static_cast<size_t>(jn) * CNN_PATCH_PITCH + lane]);
                if (lane + 64 < 121)
                    v1 = bf16_val(patches[static_cast<size_t>(jn) * CNN_PATCH_PITCH + lane + 64]);
            }
        }
        // every phase's activations are the wave's own: the wave's LDS
        // operations complete in order, so a compiler barrier (no
        // workgroup barrier) separates the phases and the four waves run
        // independently
        wave_lds_fence();
        conv1(B, A, lane);
        wave_lds_fence();
        conv2_rf(B, F, A, lane);
        wave_lds_fence();
        conv3_rf(B, F, A, lane);
        wave_lds_fence();
        float o[3];
        head(B, A, lane, o);
        if (valid && lane == 0) fo[j] = make_float4(o[0], o[1], o[2], 0.0f);
        wave_lds_fence();
    }
}

__global__ __launch_bounds__(CT, 1) void k_cnn_train(const uint8_t *__restrict__ pk,
                                                     const uint16_t *__restrict__ patches,
                                                     const float4 *__restrict__ targets,
                                                     const uint32_t *__restrict__ idx, uint32_t nb,
                                                     float *__restrict__ partials,
                                                     float *__restrict__ ploss) {
    extern __shared__ __align__(16) uint8_t smem[];
    BlockW &B = *reinterpret_cast<BlockW *>(smem);
    WaveAct *Aall = reinterpret_cast<WaveAct *>(smem + sizeof(BlockW));
    WaveGrad *Gall = reinterpret_cast<WaveGrad *>(smem + kTrainActs);
    float *s_do = reinterpret_cast<float *>(smem + kTrainActs + 4 * sizeof(WaveGrad));  // [4][4]
    float *s_loss = s_do + 16;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    stage_weights(pk, B, tid);
    {
        u32x4 *g = reinterpret_cast<u32x4 *>(Gall);
        const u32x4 z = {0u, 0u, 0u, 0u};
        for (int i = tid; i < static_cast<int>(4 * sizeof(WaveGrad) / 16); i += CT) g[i] = z;
    }
    const uint16_t *w2t = reinterpret_cast<const uint16_t *>(pk + CNN_PK_W2T);
    const uint16_t *w3t = reinterpret_cast<const uint16_t *>(pk + CNN_PK_W3T);
    WaveAct &A = Aall[wave];
    WaveGrad &G = Gall[wave];
    f32x4 acc3[5][2], acc2[3][2];
#pragma unroll
    for (int i = 0; i < 5; ++i) acc3[i][0] = acc3[i][1] = zero4();
#pragma unroll
    for (int i = 0; i < 3; ++i) acc2[i][0] = acc2[i][1] = zero4();
    float acc_fc = 0.0f, acc_w1 = 0.0f, acc_b = 0.0f, loss_acc = 0.0f;
    const float invB = 1.0f / static_cast<float>(nb);
    const uint32_t per = 4u * gridDim.x;
    const uint32_t iters = (nb + per - 1) / per;
    __syncthreads();
    for (uint32_t it = 0; it < iters; ++it) {
        const uint32_t j = (it * gridDim.x + blockIdx.x) * 4u + static_cast<uint32_t>(wave);
        const bool valid = j < nb;
        const uint32_t si = valid ? idx[j] : 0u;
        for (int i = lane; i < 128; i += 64)
            A.x[i] = (valid && i < 121) ? bf16_val(patches[static_cast<size_t>(si) * CNN_PATCH_PITCH + i]) : 0.0f;
        const float4 tg = valid ? targets[si] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        __syncthreads();
        conv1(B, A, lane);
        __syncthreads();
        conv2(B, A, lane);
        __syncthreads();
        conv3(B, A, lane);
        __syncthreads();
        float o[3];
        head(B, A, lane, o);
        // loss gradient at the outputs, then at conv3's pre-activation
        float d0 = 0.0f, d1 = 0.0f, d2 = 0.0f;
        if (valid) {
            // the position terms count for label 1 only (a label-0 target's
            // offset carries no information)
            const float lab = tg.z;
            const float e0 = lab != 0.0f ? o[0] - tg.x : 0.0f;
            const float e1 = lab != 0.0f ? o[1] - tg.y : 0.0f;
            d0 = lab * 2.0f * e0 * invB;
            d1 = lab * 2.0f * e1 * invB;
            d2 = (1.0f / (1.0f + expf(-o[2])) - lab) * invB;
            if (lane == 0)
                loss_acc += (lab * (e0 * e0 + e1 * e1) + fmaxf(o[2], 0.0f) +
                             log1pf(expf(-fabsf(o[2]))) - lab * o[2]) * invB;
        }
        if (lane == 0) {
            s_do[wave * 4 + 0] = d0;
            s_do[wave * 4 + 1] = d1;
            s_do[wave * 4 + 2] = d2;
        }
        if (lane < 32) {
            const float dg = fmaf(B.f[CNN_F_WFC + 64 + lane], d2,
                                  fmaf(B.f[CNN_F_WFC + 32 + lane], d1, B.f[CNN_F_WFC + lane] * d0));
            const uint32_t hb = bf16_bits(dg / 25.0f);
            for (int m = 0; m < 25; ++m) {
                const uint16_t v = static_cast<uint16_t>(A.a3[m * 32 + lane] > 0.0f ? hb : 0u);
                G.dz3[m * 32 + lane] = v;
                G.dz3t[lane * 32 + m] = v;
            }
        }
        __syncthreads();
        da2_own(w3t, A, G, lane);
        dw3_acc(Aall, Gall, wave, lane, acc3);
        if (tid < 96) {
            const int jj = tid >> 5, c = tid & 31;
            for (int w = 0; w < 4; ++w) acc_fc = fmaf(s_do[w * 4 + jj], Aall[w].g[c], acc_fc);
        } else if (tid < 99) {
            for (int w = 0; w < 4; ++w) acc_fc = acc_fc + s_do[w * 4 + tid - 96];
        }
        if (tid >= 192 && tid < 224) {
            const int c = tid - 192;
            for (int w = 0; w < 4; ++w)
                for (int m = 0; m < 25; ++m) acc_b = acc_b + bf16_val(Gall[w].dz3[m * 32 + c]);
        }
        __syncthreads();
        da1_own(w2t, A, G, lane);
        dw2_acc(Aall, Gall, wave, lane, acc2);
        if (tid >= 224) {
            const int c = tid - 224;
            for (int w = 0; w < 4; ++w)
                for (int q = 0; q < 49; ++q) acc_b = acc_b + bf16_val(Gall[w].dz2[q * 32 + c]);
        }
        __syncthreads();
        if (tid < 144) {
            const int c = tid / 9, t = tid - c * 9, ky = t / 3, kx = t - ky * 3;
            for (int w = 0; w < 4; ++w)
                for (int q = 0; q < 81; ++q) {
                    const int qy = q / 9, qx = q - qy * 9;
                    acc_w1 = fmaf(Aall[w].x[(qy + ky) * 11 + qx + kx], Gall[w].dz1[q * 16 + c], acc_w1);
                }
        } else if (tid < 160) {
            const int c = tid - 144;
            for (int w = 0; w < 4; ++w)
                for (int q = 0; q < 81; ++q) acc_w1 = acc_w1 + Gall[w].dz1[q * 16 + c];
        }
        __syncthreads();
    }
    // this block's partial gradient (every entry written by one thread)
    float *pb = partials + static_cast<size_t>(blockIdx.x) * CNN_NPARAM;
    if (tid < 96) pb[CNN_WFC + tid] = acc_fc;
    else if (tid < 99) pb[CNN_BFC + tid - 96] = acc_fc;
    if (tid < 144) pb[CNN_W1 + tid] = acc_w1;
    else if (tid < 160) pb[CNN_B1 + tid - 144] = acc_w1;
    if (tid >= 192 && tid < 224) pb[CNN_B3 + tid - 192] = acc_b;
    else if (tid >= 224) pb[CNN_B2 + tid - 224] = acc_b;
    const int r16 = lane & 15, kg = lane >> 4;
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        const int kt = wave + 4 * i;
        if (kt < 18) {
#pragma unroll
            for (int nt = 0; nt < 2; ++nt)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int k = kt * 16 + kg * 4 + r, n = nt * 16 + r16;
                    const int tap = k >> 5, c = k & 31;
                    pb[CNN_W3 + n * 288 + c * 9 + tap] = acc3[i][nt][r];
                }
        }
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const int kt = wave + 4 * i;
        if (kt < 9) {
#pragma unroll
            for (int nt = 0; nt < 2; ++nt)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int k = kt * 16 + kg * 4 + r, n = nt * 16 + r16;
                    const int tap = k >> 4, c = k & 15;
                    pb[CNN_W2 + n * 144 + c * 9 + tap] = acc2[i][nt][r];
                }
        }
    }
    if (lane == 0) s_loss[wave] = loss_acc;
    __syncthreads();
    if (tid == 0) ploss[blockIdx.x] = ((s_loss[0] + s_loss[1]) + s_loss[2]) + s_loss[3];
}

// Writes parameter i's packed copies.
__device__ __forceinline__ void pack_one(int i, float w, uint8_t *__restrict__ pk) {
    const uint32_t hb = bf16_bits(w);
    const float wb = bf16_val(hb);
    float *f = reinterpret_cast<float *>(pk + CNN_PK_F32);
    uint16_t *w2p = reinterpret_cast<uint16_t *>(pk + CNN_PK_W2P);
    uint16_t *w3p = reinterpret_cast<uint16_t *>(pk + CNN_PK_W3P);
    uint16_t *w2t = reinterpret_cast<uint16_t *>(pk + CNN_PK_W2T);
    uint16_t *w3t = reinterpret_cast<uint16_t *>(pk + CNN_PK_W3T);
    if (i < CNN_B1) {
        f[CNN_F_W1 + i] = wb;
    } else if (i < CNN_W2) {
        f[CNN_F_B1 + i - CNN_B1] = w;
    } else if (i < CNN_B2) {
        const int j = i - CNN_W2, n = j / 144, c = (j % 144) / 9, tap = j % 9;
        w2p[n * CNN_K2 + tap * 16 + c] = static_cast<uint16_t>(hb);
        w2t[c * CNN_K3 + tap * 32 + n] = static_cast<uint16_t>(hb);
    } else if (i < CNN_W3) {
        f[CNN_F_B2 + i - CNN_B2] = w;
    } else if (i < CNN_B3) {
        const int j = i - CNN_W3, n = j / 288, c = (j % 288) / 9, tap = j % 9;
        w3p[n * CNN_K3 + tap * 32 + c] = static_cast<uint16_t>(hb);
        w3t[c * CNN_K3 + tap * 32 + n] = static_cast<uint16_t>(hb);
    } else if (i < CNN_WFC) {
        f[CNN_F_B3 + i - CNN_B3] = w;
    } else if (i < CNN_BFC) {
        f[CNN_F_WFC + i - CNN_WFC] = wb;
    } else {
        f[CNN_F_BFC + i - CNN_BFC] = w;
    }
}

__global__ void k_cnn_pack(const float *__restrict__ params, uint8_t *__restrict__ pk) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < CNN_NPARAM) pack_one(i, params[i], pk);
}

__global__ void k_cnn_reduce(const float *__restrict__ partials, const float *__restrict__ ploss,
                             uint32_t nblk, float *__restrict__ grad, float *__restrict__ loss,
                             float *__restrict__ params, float *__restrict__ m,
                             float *__restrict__ v, uint8_t *__restrict__ pk, CnnAdam ad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0 && loss) {
        float L = 0.0f;
        for (uint32_t b = 0; b < nblk; ++b) L = L + ploss[b];
        *loss = L;
    }
    if (i >= CNN_NPARAM) return;
    float g = 0.0f;
    for (uint32_t b = 0; b < nblk; ++b) g = g + partials[static_cast<size_t>(b) * CNN_NPARAM + i];
    if (grad) grad[i] = g;
    if (!ad.apply) return;
    // torch.optim.Adam: m, v moving averages; bias-corrected step
    const float mm = ad.b1 * m[i] + ad.omb1 * g;
    const float vv = ad.b2 * v[i] + ad.omb2 * g * g;
    const float denom = sqrtf(vv) / sqrtf(ad.bc2) + ad.eps;
    const float w = params[i] - ad.lr / ad.bc1 * mm / denom;
    m[i] = mm;
    v[i] = vv;
    params[i] = w;
    pack_one(i, w, pk);
}

// Flat sample offsets of a chunk's candidate lists, appended after *total
// (which advances past them).
__global__ void k_sample_offsets(const uint32_t *__restrict__ ccount, uint32_t cap,
                                 uint32_t nframes, uint32_t *__restrict__ total,
                                 uint32_t *__restrict__ offs) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    uint32_t o = *total;
    for (uint32_t f = 0; f < nframes; ++f) {
        offs[f] = o;
        o += min(ccount[f * GT_COUNTER_STRIDE], cap);
    }
    offs[nframes] = o;
    *total = o;
}

// Stored patches (bf16 [.][128], entries 121..127 zero) of every candidate.
__global__ __launch_bounds__(CT) void k_cnn_gather(const float *__restrict__ proc, int H, int W,
                                                   const uint32_t *__restrict__ cand, uint32_t cap,
                                                   const uint32_t *__restrict__ ccount,
                                                   const uint32_t *__restrict__ offs,
                                                   uint32_t maxn, uint16_t *__restrict__ patches) {
    const uint32_t f = blockIdx.y;
    const uint32_t count = min(ccount[f * GT_COUNTER_STRIDE], cap);
    const float *P = proc + static_cast<size_t>(f) * H * W;
    const uint32_t *fc = cand + static_cast<size_t>(f) * cap;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (uint32_t j0 = blockIdx.x * 4u; j0 < count; j0 += 4u * gridDim.x) {
        const uint32_t j = j0 + static_cast<uint32_t>(wave);
        const bool valid = j < count;
        uint32_t h0, h1;
        gather_norm(P, H, W, valid ? fc[j] : 0u, valid, lane, h0, h1);
        const uint32_t dst = offs[f] + j;
        if (valid && dst < maxn) {
            uint16_t *o = patches + static_cast<size_t>(dst) * CNN_PATCH_PITCH;
            o[lane] = static_cast<uint16_t>(h0);
            o[lane + 64] = static_cast<uint16_t>(h1);
        }
    }
}

template <typename K>
hipError_t allow_lds(K kernel, size_t bytes) {
    return hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                               hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes));
}

}  // namespace

hipError_t gt_cnn_launch_pack(const float *params, void *pk, hipStream_t s) {
    k_cnn_pack<<<(CNN_NPARAM + 255) / 256, 256, 0, s>>>(params, static_cast<uint8_t *>(pk));
    return hipGetLastError();
}

hipError_t gt_cnn_launch_forward_patches(const void *pk, const uint16_t *patches, uint32_t n,
                                         float4 *out, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipError_t e = allow_lds(k_cnn_forward<false>, kFwdLds);
    if (e != hipSuccess) return e;
    const uint32_t blocks = std::min<uint32_t>(512u, (n + 3) / 4);
    k_cnn_forward<false><<<blocks, CT, kFwdLds, s>>>(static_cast<const uint8_t *>(pk), patches, n,
                                                     nullptr, 0, 0, nullptr, 1, 0, nullptr, out);
    return hipGetLastError();
}

hipError_t gt_cnn_launch_forward_frames(const void *pk, const float *proc, int H, int W,
                                        uint32_t nframes, const uint32_t *cand, uint32_t cstride,
                                        uint32_t cap, const uint32_t *ccount, float4 *props,
                                        hipStream_t s) {
    if (nframes == 0) return hipSuccess;
    if (cstride != 1 && cstride != 2) return hipErrorInvalidValue;
    hipError_t e = allow_lds(k_cnn_forward<true>, kFwdLds);
    if (e != hipSuccess) return e;
    const uint32_t bx = std::max<uint32_t>(1u, std::min<uint32_t>(32u, 1024u / nframes));
    k_cnn_forward<true><<<dim3(bx, nframes), CT, kFwdLds, s>>>(
        static_cast<const uint8_t *>(pk), nullptr, 0, proc, H, W, cand, cstride, cap, ccount, props);
    return hipGetLastError();
}

hipError_t gt_cnn_launch_gather(const float *proc, int H, int W, uint32_t nframes,
                                const uint32_t *cand, uint32_t cap, const uint32_t *ccount,
                                const uint32_t *offs, uint32_t maxn, uint16_t *patches,
                                hipStream_t s) {
    if (nframes == 0) return hipSuccess;
    k_cnn_gather<<<dim3(64, nframes), CT, 0, s>>>(proc, H, W, cand, cap, ccount, offs, maxn, patches);
    return hipGetLastError();
}

hipError_t gt_cnn_launch_sample_offsets(const uint32_t *ccount, uint32_t cap, uint32_t nframes,
                                        uint32_t *total, uint32_t *offs, hipStream_t s) {
    k_sample_offsets<<<1, 64, 0, s>>>(ccount, cap, nframes, total, offs);
    return hipGetLastError();
}

hipError_t gt_cnn_launch_train(const void *pk, const uint16_t *patches, const float4 *targets,
                               const uint32_t *idx, uint32_t nb, float *partials, float *ploss,
                               hipStream_t s) {
    if (nb == 0) return hipErrorInvalidValue;
    hipError_t e = allow_lds(k_cnn_train, kTrainLds);
    if (e != hipSuccess) return e;
    k_cnn_train<<<CNN_TRAIN_BLOCKS, CT, kTrainLds, s>>>(static_cast<const uint8_t *>(pk), patches,
                                                        targets, idx, nb, partials, ploss);
    return hipGetLastError();
}

hipError_t gt_cnn_launch_reduce(const float *partials, const float *ploss, float *grad,
                                float *loss, float *params, float *m, float *v, void *pk,
                                const CnnAdam &adam, hipStream_t s) {
    k_cnn_reduce<<<(CNN_NPARAM + 255) / 256, 256, 0, s>>>(partials, ploss, CNN_TRAIN_BLOCKS, grad,
                                                          loss, params, m, v,
                                                          static_cast<uint8_t *>(pk), adam);
    return hipGetLastError();
}
