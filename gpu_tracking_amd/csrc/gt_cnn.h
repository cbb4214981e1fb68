// gt_cnn.h -- learned centroid proposal: parameter layout, packed weights and
// kernel launchers (gt_cnn.hip), shared with the host runtime (gt_cnn.cpp).
//
// Network: 11x11 patch of the processed frame around a candidate pixel,
// divided by its largest |value| and rounded to bf16 -> conv3x3 1->16 ->
// conv3x3 16->32 -> conv3x3 32->32 (valid, bias, ReLU, bf16 activations) ->
// mean over the 5x5 positions (bf16) -> linear 32->3 (f32): dy, dx (rows,
// columns, relative to the candidate pixel) and a confidence logit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

// float32 parameter vector, torch layouts: w1 (16,1,3,3) b1 w2 (32,16,3,3)
// b2 w3 (32,32,3,3) b3 wfc (3,32) bfc
constexpr int CNN_W1 = 0, CNN_B1 = 144, CNN_W2 = 160, CNN_B2 = 4768, CNN_W3 = 4800,
              CNN_B3 = 14016, CNN_WFC = 14048, CNN_BFC = 14144, CNN_NPARAM = 14147;
constexpr int CNN_PATCH = 11, CNN_PATCH_PITCH = 128;  // bf16 per stored patch

// Packed copies the kernels read (one device buffer of CNN_PK_BYTES):
//   w2p u16 [32][160]  conv2 B operand, k = tap * 16 + c (k >= 144 zero)
//   w3p u16 [32][288]  conv3 B operand, k = tap * 32 + c
//   w2t u16 [16][288]  conv2 input gradient, [c][tap * 32 + n]
//   w3t u16 [32][288]  conv3 input gradient, [c][tap * 32 + n]
//   f32: w1 [16][9] (bf16 values), b1 [16], b2 [32], b3 [32], wfc [3][32]
//        (bf16 values), bfc [3]
constexpr int CNN_K2 = 160, CNN_K3 = 288;
constexpr size_t CNN_PK_W2P = 0;
constexpr size_t CNN_PK_W3P = CNN_PK_W2P + 32 * CNN_K2 * 2;
constexpr size_t CNN_PK_W2T = CNN_PK_W3P + 32 * CNN_K3 * 2;
constexpr size_t CNN_PK_W3T = CNN_PK_W2T + 16 * CNN_K3 * 2;
constexpr size_t CNN_PK_F32 = CNN_PK_W3T + 32 * CNN_K3 * 2;
constexpr int CNN_F_W1 = 0, CNN_F_B1 = 144, CNN_F_B2 = 160, CNN_F_B3 = 192, CNN_F_WFC = 224,
              CNN_F_BFC = 320, CNN_F_N = 324;
constexpr size_t CNN_PK_BYTES = CNN_PK_F32 + CNN_F_N * 4;

// Adam step arguments (torch.optim.Adam's arithmetic); apply == 0 reduces
// the gradient only.
struct CnnAdam {
    int32_t apply;
    float lr, b1, b2, omb1, omb2, eps, bc1, bc2;
};

// Training kernel blocks (one per CU; each holds four samples at a time).
constexpr uint32_t CNN_TRAIN_BLOCKS = 256;

// Packs the kernels' weight copies from the float32 parameters (pk's
// padding must have been zeroed once).
hipError_t gt_cnn_launch_pack(const float *params, void *pk, hipStream_t s);
// Forward over n stored patches (bf16 [n][128]) -> out[i] = (dy, dx, logit, 0).
hipError_t gt_cnn_launch_forward_patches(const void *pk, const uint16_t *patches, uint32_t n,
                                         float4 *out, hipStream_t s);
// Forward over the candidate lists of a chunk of frames (frame f: proc + f * H * W,
// candidate i's pixel cand[(f * cap + i) * cstride], cstride 1 or 2, count
// min(ccount[f * GT_COUNTER_STRIDE], cap)) -> props[f * cap + i].
hipError_t gt_cnn_launch_forward_frames(const void *pk, const float *proc, int H, int W,
                                        uint32_t nframes, const uint32_t *cand, uint32_t cstride,
                                        uint32_t cap, const uint32_t *ccount, float4 *props,
                                        hipStream_t s);
// Stored patches of every candidate of a chunk, at flat index offs[f] + i
// (offs from gt_cnn_launch_sample_offsets; indices >= maxn are dropped).
hipError_t gt_cnn_launch_gather(const float *proc, int H, int W, uint32_t nframes,
                                const uint32_t *cand, uint32_t cap, const uint32_t *ccount,
                                const uint32_t *offs, uint32_t maxn, uint16_t *patches,
                                hipStream_t s);
// offs[f] = *total + sum of min(ccount[g], cap) over g < f; offs[nframes]
// and *total become the end.
hipError_t gt_cnn_launch_sample_offsets(const uint32_t *ccount, uint32_t cap, uint32_t nframes,
                                        uint32_t *total, uint32_t *offs, hipStream_t s);
// Forward + backward of the batch idx[0..nb) of (patches, targets): partial
// gradients [CNN_TRAIN_BLOCKS][CNN_NPARAM] and partial losses [CNN_TRAIN_BLOCKS].
hipError_t gt_cnn_launch_train(const void *pk, const uint16_t *patches, const float4 *targets,
                               const uint32_t *idx, uint32_t nb, float *partials, float *ploss,
                               hipStream_t s);
// Sums the partials in block order into grad (if not null) and the loss
// into *loss (if not null); with adam.apply, updates params, m, v and pk.
hipError_t gt_cnn_launch_reduce(const float *partials, const float *ploss, float *grad,
                                float *loss, float *params, float *m, float *v, void *pk,
                                const CnnAdam &adam, hipStream_t s);

struct gt_cnn;
// The network's packed weights (device).
const void *gt_cnn_packed(const gt_cnn *n);
