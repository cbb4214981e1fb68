// gt_internal.h -- shared between the HIP kernels and the host runtime.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gtrack.h"

#include "gt_kparams.h"

// Per-frame detection counters live GT_COUNTER_STRIDE words apart so that
// frames processed at the same time do not share a cache line.
#define GT_COUNTER_STRIDE 64

// Kernel launchers (gt_kernels.hip).  All are asynchronous on `s`.
hipError_t gt_launch_zero_u32(uint32_t *p, size_t n, hipStream_t s);
// Up to two u32 arrays a launch clears before any later launch on its stream
// reads them (counters cleared by the bandpass launch instead of launches of
// their own).
struct GtZero {
    uint32_t *p[2];
    uint32_t n[2];
};
hipError_t gt_launch_bandpass(const void *frames, int dtype, float *proc, uint32_t nframes,
                              const GtKParams &p, hipStream_t s, const GtZero *zero = nullptr);
hipError_t gt_launch_localmax(const float *proc, uint32_t nframes, const GtKParams &p,
                              uint32_t *cand, uint32_t cand_cap, uint32_t *cand_count,
                              hipStream_t s);
hipError_t gt_launch_walk(const float *proc, const void *raw, int dtype, uint32_t nframes,
                          const GtKParams &p, const uint32_t *cand, uint32_t cand_cap,
                          const uint32_t *cand_count, gt_detection *out, uint32_t out_cap,
                          uint32_t *out_count, uint32_t frame0, hipStream_t s);
hipError_t gt_launch_characterize_rows(const float *proc, const void *raw, int dtype,
                                       uint32_t nframes, const GtKParams &p, gt_detection *rows,
                                       uint32_t cap, const uint32_t *counts, uint32_t max_count,
                                       hipStream_t s);
// Fused local max + walk + characterize over LDS tiles (the main path).
size_t gt_tile_lds_bytes(const GtKParams &p);
// walk_steps (1 or 2) steps run in the tile; walks still moving after them
// go to the slow list as (seed, position).
hipError_t gt_launch_detect_tile(const float *proc, const void *raw, int dtype, uint32_t nframes,
                                 const GtKParams &p, gt_detection *out, uint32_t out_cap,
                                 uint32_t *ocount, uint2 *slow, uint32_t slow_cap,
                                 uint32_t *scount, uint32_t frame0, hipStream_t s,
                                 uint32_t walk_steps = 2);
// Finishes the walks the tile kernel handed over (slow list: seed, position).
hipError_t gt_launch_walk_slow(const float *proc, const void *raw, int dtype, uint32_t nframes,
                               const GtKParams &p, const uint2 *slow, uint32_t slow_cap,
                               const uint32_t *scount, gt_detection *out, uint32_t out_cap,
                               uint32_t *ocount, uint32_t frame0, hipStream_t s);
hipError_t gt_launch_copy_counts(const uint32_t *strided, uint32_t *dst, uint32_t n,
                                 hipStream_t s);
// Fills the characterize columns of every detection of the chunk (rows of
// frame f at out[f * out_cap], count at ocount[f * GT_COUNTER_STRIDE]).  With
// dcount, also copies the counts there (dcount[f] = ocount[f * stride]), as
// gt_launch_copy_counts does.
hipError_t gt_launch_characterize_det(const float *proc, const void *raw, int dtype,
                                      uint32_t nframes, const GtKParams &p, gt_detection *out,
                                      uint32_t out_cap, const uint32_t *ocount, hipStream_t s,
                                      uint32_t *dcount = nullptr);
// The plain tiled Trackpy path of one chunk (no noise estimate, adaptive
// passes, illumination correction or learned proposal), shared by
// gt_detect_device and the static-scratch entry point.  Scratch: the chunk's
// processed planes, per-frame row and slow-walk counters (GT_COUNTER_STRIDE
// apart, cleared by the head's first launch) and the slow-walk lists.
struct GtPlainScratch {
    float *proc;
    uint32_t *ocount, *scount;
    uint2 *slow;
    uint32_t slow_cap;
};
// Head: bandpass (clearing the counters), then local max + in-tile walks;
// `mid` (may be null) is recorded on s between the two.
hipError_t gt_plain_head(const void *frames, int dtype, uint32_t nf, const GtKParams &p,
                         const GtPlainScratch &sc, gt_detection *out, uint32_t cap,
                         uint32_t frame0, hipStream_t s, hipEvent_t mid);
// Tail: the handed-over walks, then characterize (with p.characterize) or
// the count copy; counts[f] receives frame f's detection count.
hipError_t gt_plain_tail(const void *frames, int dtype, uint32_t nf, const GtKParams &p,
                         const GtPlainScratch &sc, gt_detection *out, uint32_t cap,
                         uint32_t *counts, uint32_t frame0, hipStream_t s);

// LoG style: scale-space maxima of every scale of a chunk (planes kept in
// LDS), then their walks.  Candidate seeds are scale * H * W + raster index.
// With cval (threshold_rel on) the tile kernel also stores each candidate's
// LoG value at cval[f * cap + i] and the frame's largest LoG value (f32 bits)
// at ccount[f * GT_COUNTER_STRIDE + 1], which must start at zero; the walk
// kernel then skips candidates whose value is not above p.thr_rel times it.
size_t gt_log_lds_bytes(int Rm);
hipError_t gt_launch_log_tile(const void *frames, int dtype, const float *proc, uint32_t nframes,
                              const GtKParams &p, const GtLogTable *tab, int Rm, uint32_t *cand,
                              uint32_t cap, uint32_t *ccount, float *cval, hipStream_t s);
hipError_t gt_launch_walk_log(const float *proc, const void *raw, int dtype, uint32_t nframes,
                              const GtKParams &p, const GtLogTable *tab, const uint32_t *cand,
                              uint32_t cap, const uint32_t *ccount, gt_detection *out,
                              uint32_t out_cap, uint32_t *ocount, uint32_t frame0,
                              const float *cval, hipStream_t s);
// Per-frame detection counts, 0xFFFFFFFF where the candidate list overflowed.
hipError_t gt_launch_copy_counts_checked(const uint32_t *ocount, const uint32_t *ccount,
                                         uint32_t ccap, uint32_t *dst, uint32_t n, hipStream_t s);
// Batched launches over frames of different sizes (u8/u16 frames, R <= 7):
// grid y = frame slot of the launch, grid x = tile of that frame.
struct GtBatchSchedule {
    uint32_t tile_blocks;   // queue tiles of the call's largest launch (multiple of 8)
    uint32_t walk_blocks;   // blocks per frame for slow walks and characterize
    uint32_t launch_slots;  // frames per launch at most (grid y of the per-frame stages)
};
// One launch: k_select_launch copies launch *cur's entries of `all` into
// (hdr, fs), the stages run, k_next_launch advances *cur.
hipError_t gt_launch_batch(const GtBatchLaunch *launches, uint32_t *cur,
                           const GtBatchFrame *all, int dtype, const GtKParams &p,
                           GtBatchHeader *hdr, GtBatchFrame *fs, uint32_t max_slots, float *proc,
                           gt_detection *out, uint32_t out_cap, uint32_t *ocount, uint2 *slow,
                           uint32_t slow_cap, uint32_t *scount, uint32_t *counts,
                           const GtBatchSchedule &sch, hipStream_t s);
void gt_batch_tiles(int H, int W, uint32_t *bp_tiles, uint32_t *dt_tiles);
// Rows and columns of the tile the batched bandpass and detect kernels share.
uint32_t gt_batch_tile_rows();
uint32_t gt_batch_tile_cols();
// Learned proposal path: candidate pixels per frame (local-maximum test of
// the tile kernel), the verifying walk over the network's proposals
// (stats[0] += accepted, stats[1] += candidates), and the training targets
// (full walk from every candidate, flat index offs[f] + i).
hipError_t gt_launch_detect_cand(const float *proc, int dtype, uint32_t nframes,
                                 const GtKParams &p, uint32_t *cand, uint32_t cap,
                                 uint32_t *ccount, hipStream_t s);
// List entries: stride 1 = candidate pixel; stride 2 = (candidate pixel,
// position after it0 walk steps), the slow list of the tile kernel.
hipError_t gt_launch_walk_learned(const float *proc, uint32_t nframes, const GtKParams &p,
                                  const uint32_t *cand, uint32_t stride, uint32_t it0,
                                  uint32_t cap, const uint32_t *ccount, const float4 *props,
                                  float conf_logit, float tol, gt_detection *out,
                                  uint32_t out_cap, uint32_t *ocount, unsigned long long *stats,
                                  uint32_t frame0, hipStream_t s);
hipError_t gt_launch_cnn_label(const float *proc, uint32_t nframes, const GtKParams &p,
                               const uint32_t *cand, uint32_t cap, const uint32_t *ccount,
                               const uint32_t *offs, uint32_t maxn, float4 *targets,
                               hipStream_t s);
// Noise estimate (mean and sample std, the snr rules' eight-way tree) of
// nframes raw frames of n > 24 pixels; tmp holds nframes * ceil(n / 8) floats.
hipError_t gt_launch_frame_std(const void *frames, int dtype, uint32_t nframes, size_t n,
                               float *tmp, float *mean, float *std_out, hipStream_t s);
// Post-processing of a chunk's rows (frame f: rows[f * cap ...], counts[f])
// into post[f * cap + i] = (raw_mass, bg_median, corrected, keep):
// filter_close sets keep (and zeroes the rest); the doughnut fills the first
// three for kept rows.  offs: nsig disk offsets, then nring ring offsets
// (i, j), i-major.
hipError_t gt_launch_filter_close_rows(const gt_detection *rows, uint32_t cap,
                                       const uint32_t *counts, uint32_t nframes, float separation,
                                       float4 *post, hipStream_t s);
// Packs each frame's rows (and post values, when post is non-null) of a
// chunk one after another into orows / opost.
hipError_t gt_launch_pack_rows(const gt_detection *rows, uint32_t cap, const uint32_t *counts,
                               uint32_t nframes, const float4 *post, gt_detection *orows,
                               float4 *opost, hipStream_t s);
size_t gt_doughnut_lds_bytes(uint32_t nsig, uint32_t nring);
hipError_t gt_launch_doughnut_rows(const void *raw, int dtype, int H, int W,
                                   const gt_detection *rows, uint32_t cap, const uint32_t *counts,
                                   uint32_t nframes, const short2 *offs, uint32_t nsig,
                                   uint32_t nring, float4 *post, hipStream_t s);
// adaptive_background over a chunk's walked rows (out, cap, strided ocount):
// n_iter passes of the acceptance rule with the noise re-estimated (std2)
// on an f32 copy of the frames (masked, nframes * H * W) with accepted
// particles marked; accepted rows are compacted in place and ocount updated.
// p.fstd holds the first estimate; state: nframes * cap bytes; counter:
// nframes.
hipError_t gt_launch_adaptive(const void *frames, int dtype, uint32_t nframes, const GtKParams &p,
                              uint32_t n_iter, gt_detection *out, uint32_t cap, uint32_t *ocount,
                              float *std_tmp, float *mean, float *std2, float *masked,
                              uint8_t *state, uint32_t *counter, hipStream_t s);
// Illumination correction (correct_illumination/*.wgsl).  acc[p] += f32 of
// each frame's pixel p, frames in order (mean_from_iter's running sum).
hipError_t gt_launch_frames_accumulate(const void *frames, int dtype, uint32_t nframes, size_t HW,
                                       float *acc, hipStream_t s);
// dst[i] = f32(frames[i]) for n pixels.
hipError_t gt_launch_widen(const void *frames, int dtype, size_t n, float *dst, hipStream_t s);
// x[i] = x[i] / d.
hipError_t gt_launch_scale_div(float *x, size_t n, float d, hipStream_t s);
// smooth_mean_frame's two passes over nframes frames: in -> tmp (along
// columns) -> out (along rows); w holds 2 * rint + 1 taps.
hipError_t gt_launch_illum_smooth(const float *in, float *tmp, float *out, int H, int W,
                                  uint32_t nframes, const float *w, int rint, hipStream_t s);
// out[t * HW + p] = f32(frames[t * HW + p]) / profile[t * pstride + p].
hipError_t gt_launch_illum_divide(const void *frames, int dtype, uint32_t nframes, size_t HW,
                                  const float *profile, size_t pstride, float *out, hipStream_t s);
// Fills the constant tables the kernels read (once per device).
hipError_t gt_init_tables();
