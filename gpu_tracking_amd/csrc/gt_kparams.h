// gt_kparams.h -- kernel parameter block shared by the host code and the
// HIP kernels (no HIP headers, so host-only translation units can use it).
#pragma once
#include <stdint.h>

// Largest preprocess radius s/2 the LDS-tiled bandpass handles (s <= 33).
#define GT_MAX_R 16
// Largest dilation side k the LDS-tiled local-max kernel handles; larger
// windows run the untiled kernel.
#define GT_MAX_K_TILED 48
// Radius up to which characterize reads sin/cos(2 theta) from a table.
#define GT_SC_R 24

// Derived per-frame parameters, passed by value as a kernel argument.  The
// first block is the reference's WGSL Params uniform (src/shaders/params.wgsl),
// extended with the f32 tap table the reference recomputes per tap
// (preprocess_rows.wgsl:46) and the walk/characterize constants.
struct GtKParams {
    int32_t H, W;
    int32_t R;                    // preprocess radius s/2; taps i = -R..R
    float w[2 * GT_MAX_R + 1];    // exp(-i^2 / (2 sigma2)) as f32
    float gnorm;                  // sequential f32 sum of w
    float box_area;               // f32(s*s)
    float rgnorm;                 // RN(1 / gnorm)
    float rbox;                   // RN(1 / box_area)
    // k_bandpass_int's division by the norms: 0 IEEE division (general
    // kernel), 1 reciprocal with one fma correction, 2 RN(a*r + RN(a*r_lo))
    // with r + r_lo = 1/norm to ~48 bits (rgnorm, rbox hold r then)
    int32_t fast_div;
    float rgnorm_lo;              // fast_div 2: RN(1/gnorm - rgnorm)
    float rbox_lo;                // fast_div 2: RN(1/box_area - rbox)
    int32_t truncate;             // truncate_preprocessed
    int32_t k;                    // dilation side
    int32_t a;                    // window start offset i32(k/2 - 0.5)
    int32_t margin;
    float max_thr;                // image_std * snr * rough_snr_factor
    int32_t rint;                 // walk radius (circle-1)/2
    float r2;                     // f32(rint)^2
    uint32_t max_iter;
    float shift_thr;
    float minmass;
    float minmass_snr;
    float image_std;
    float rough;
    int32_t characterize;
    int32_t circle;
    float snr;
    float thr_rel;                // LoG threshold_rel (0: off)
    // per-frame noise estimate of the launch's frames (device; null: the
    // constants max_thr and image_std apply to every frame)
    const float *fstd;
};

// LoG style: the radius scan and each scale's tap table (separable_log.wgsl),
// one device copy per context, read by the LoG kernels through a pointer.
#define GT_LOG_MAX_SCALES 64
#define GT_LOG_MAX_R 48
#define GT_LOG_MAX_TAPS (2 * GT_LOG_MAX_SCALES * (2 * GT_LOG_MAX_R + 1))
#define GT_LOG_WR 24  // largest walk radius ceil(r) with a row table
struct GtLogTable {
    int32_t n;                         // scales
    int32_t Rm;                        // largest tap radius over the scales
    float snr_thr;                     // image_std * snr * rough_snr_factor
    int32_t pad;
    float radius[GT_LOG_MAX_SCALES];   // the radius scan
    int32_t R[GT_LOG_MAX_SCALES];      // i32(4 sigma + 0.5), sigma = radius / sqrt(2)
    int32_t off[GT_LOG_MAX_SCALES];    // Gaussian taps at taps[off], second derivative after them
    float n1[GT_LOG_MAX_SCALES];       // 1 / (sqrt(2 pi) sigma)
    float n3[GT_LOG_MAX_SCALES];       // 1 / (sqrt(2 pi) sigma sigma^2)
    float taps[GT_LOG_MAX_TAPS];
    // the walk and characterize disk of scale k, x^2 + y^2 <= radius^2 in
    // float arithmetic: row i (|i| <= ceil(radius)) holds columns
    // -wjm[k][i + GT_LOG_WR] .. wjm[k][i + GT_LOG_WR], none where it is -1;
    // filled where ceil(radius) <= GT_LOG_WR
    int8_t wjm[GT_LOG_MAX_SCALES][2 * GT_LOG_WR + 1];
};

// Batched launches: frames of different sizes in one launch.  One entry per
// frame slot of a launch (grid y of every stage).
struct GtBatchFrame {
    uint64_t in_ptr;      // device address of the frame (dtype pixels, H x W)
    uint64_t proc_off;    // element offset of its processed plane in the scratch pool
    int32_t H, W;
    uint32_t out_idx;     // output slot: rows at out[out_idx * cap], count at counts[out_idx]
    uint32_t frame;       // the frame field of its detections
};
// A launch's tiles form one queue over all its frames: its frames sorted by
// size make runs of equal size (segments), and tile t of the launch is tile
// (t - first_tile) % tpf of slot first_slot + (t - first_tile) / tpf of the
// last segment with first_tile <= t.  Bandpass and detect tiles are both
// 64 x 32, so one queue serves both stages.
#define GT_BATCH_MAX_SEG 8
struct GtBatchSeg {
    uint32_t first_tile;  // queue index of the segment's first tile
    uint32_t first_slot;  // its first frame slot
    uint32_t tpf;         // tiles per frame
    uint32_t ntx;         // tile columns per frame
};
struct GtBatchHeader {
    uint32_t nslots;      // frames of the launch
    uint32_t nseg;        // segments (1 .. GT_BATCH_MAX_SEG)
    uint32_t ntiles;      // tiles of the launch's queue
    uint32_t pad;
    GtBatchSeg seg[GT_BATCH_MAX_SEG];
};
// A call's launches: launch k uses entries [first, first + nslots) of the
// call's entry array with its segments; a device counter says which launch
// runs next.
struct GtBatchLaunch {
    uint32_t first;
    uint32_t nslots;
    uint32_t nseg;
    uint32_t ntiles;
    GtBatchSeg seg[GT_BATCH_MAX_SEG];
};
