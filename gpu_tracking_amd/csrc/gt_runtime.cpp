// gt_runtime.cpp -- host runtime behind include/gtrack.h.
//
// Parameter derivation (src/gpu_setup.rs:528-580), the context (setup_state,
// :611), the device-resident batch path, and execute_gpu's whole-video driver
// with its host post-processing (src/execute_gpu.rs:556-590, 671-817,
// 886-1184).  HIP runtime calls only; kernels live in gt_kernels.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "gt_cnn.h"
#include "gt_internal.h"
#include "gt_host.h"
#include "gt_io.h"

namespace {

thread_local std::string g_last_error = "no error";

gt_status fail(gt_status st, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return st;
}

}  // namespace

// Shared with gt_batch.cpp: records the message for gt_last_error.
gt_status gt_runtime_fail(gt_status st, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return st;
}

namespace {

#define GT_HIP(call)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            return fail(GT_ERR_GPU_DEVICE, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

size_t dtype_size(int32_t dtype) {
    switch (dtype) {
    case GT_U16: return 2;
    case GT_F32: return 4;
    case GT_U8: return 1;
    default: return 0;
    }
}

}  // namespace

struct gt_context {
    gt_params params;
    uint32_t H = 0, W = 0;
    int32_t characterize_points = 0;
    GtKParams kp;
    gt_derived der;
    int device = 0;
    hipStream_t stream = nullptr;

    // device scratch shared by gt_detect_device and gt_execute
    uint32_t chunk = 1;            // frames per pipeline chunk
    uint32_t xchunk = 1;           // frames per staging slot of gt_execute (<= chunk)
    float *d_proc = nullptr;       // chunk * HW
    uint32_t *d_cand = nullptr;    // chunk * cand_cap
    uint32_t *d_ccount = nullptr;  // chunk * GT_COUNTER_STRIDE (candidates, untiled path)
    uint32_t *d_ocount = nullptr;  // chunk * GT_COUNTER_STRIDE (detections)
    uint2 *d_slow = nullptr;       // chunk * slow_cap walks handed to k_walk_slow
    uint32_t *d_scount = nullptr;  // chunk * GT_COUNTER_STRIDE
    uint32_t slow_cap = 0;
    uint32_t cand_cap = 0;
    bool tiled = true;             // fused tile kernel fits in LDS
    bool log = false;              // LoG style
    GtLogTable *d_logtab = nullptr;  // LoG scale table (device copy)
    int32_t log_Rm = 0;
    float *d_cval = nullptr;       // LoG threshold_rel: chunk * cand_cap candidate values

    // learned centroid proposal (tiled Trackpy style only)
    const gt_cnn *cnn = nullptr;   // caller's network; null = plain walk
    float cnn_conf = 0.0f;         // logit threshold
    float cnn_tol = 0.05f;         // verification tolerance (px)
    int32_t cnn_scope = GT_CNN_MOVING;  // which candidates get a proposal
    float4 *d_props = nullptr;     // chunk * cand_cap proposals
    unsigned long long *d_stats = nullptr;  // accepted, candidates
    uint32_t *d_offs = nullptr;    // sample generation: chunk + 1 offsets, then the total

    // noise estimate of each chunk's frames (snr / minmass_snr), kp.fstd = d_std
    float *d_std_tmp = nullptr;    // chunk * ceil(HW / 8)
    float *d_mean = nullptr, *d_std = nullptr;  // chunk
    // adaptive_background: marked f32 copy of the chunk, re-estimates, the
    // divisors, and a state byte per row slot (grown to chunk * cap)
    float *d_adapt_masked = nullptr, *d_std2 = nullptr;
    uint32_t *d_adapt_counter = nullptr;
    uint8_t *d_adapt_state = nullptr;
    size_t adapt_state_cap = 0;

    // illumination correction: the taps, the whole-video profile, the
    // corrected f32 chunk the detection reads, and (per frame) the chunk's
    // smoothing scratch and profiles
    float *d_illum_w = nullptr;
    int32_t illum_rint = 0;
    float *d_profile = nullptr;     // HW (whole-video profile, or the running sum)
    float *d_corr = nullptr;        // chunk * HW
    float *d_illum_tmp = nullptr;   // chunk * HW
    float *d_illum_prof = nullptr;  // chunk * HW (per frame)

    // a second lane for the plain tiled path: odd chunks run on their own
    // stream with their own scratch, so a chunk's slow walks and
    // characterize overlap the next chunk's bandpass
    int32_t lanes = 1;                 // lanes allowed (1 or 2)
    // overlap: every chunk's bandpass and local maxima on the caller's
    // stream, its slow walks and characterize on lane stream 1 (chunk
    // scratch alternating between the two lanes' buffers)
    int32_t overlap = 0;
    hipStream_t lane_stream[2] = {nullptr, nullptr};
    hipEvent_t lane_fork = nullptr, lane_done[2] = {nullptr, nullptr};
    hipEvent_t ovl_det[2] = {nullptr, nullptr};
    float *l1_proc = nullptr;
    uint32_t *l1_ocount = nullptr, *l1_scount = nullptr;
    uint2 *l1_slow = nullptr;

    // gt_execute staging: two slots, so one chunk's detections come back
    // and are post-processed while the next chunk runs on the GPU
    struct Slot {
        void *d_frames = nullptr;      // chunk * HW pixels (largest dtype)
        gt_detection *d_rows = nullptr;  // chunk * det_cap
        uint32_t *d_counts = nullptr;  // chunk
        void *h_stage = nullptr;       // pinned, chunk * HW * 4 bytes
        uint32_t *h_counts = nullptr;  // pinned, chunk
        gt_detection *d_packed = nullptr;  // the chunk's rows one frame after another
        float4 *d_packed_post = nullptr;   // and their post values (gpu_post)
        // pinned copies of the packed rows (and post values): the first
        // pk_copied rows are read back in stream order right after the
        // chunk's kernels, the rest (if any) by drain
        gt_detection *h_pk = nullptr;
        float4 *h_pk_post = nullptr;
        size_t pk_cap = 0, pk_copied = 0;
        hipEvent_t done = nullptr;     // counts of the slot's chunk are on the host
    } slot[2];
    uint32_t det_cap = 0;
    // filter_close and the doughnut on the GPU (Trackpy style, gt_execute):
    // per slot one float4 per row slot (raw_mass, bg_median, corrected, keep)
    bool gpu_post = false;
    short2 *d_dough_offs = nullptr;     // disk offsets, then ring offsets
    uint32_t dough_nsig = 0, dough_nring = 0;
    float4 *d_post[2] = {nullptr, nullptr};
    float *h_post = nullptr;            // pinned, 4 floats per read-back row
    size_t h_post_cap = 0;
    hipStream_t copy_stream = nullptr;  // detection read-back
    gt_detection *h_rows = nullptr;     // pinned read-back buffer
    size_t pk_guess = 65536;            // rows read back with a chunk's kernels
    size_t h_rows_cap = 0;

    // gt_submit_frame / gt_fetch_frame: a ring of frames in flight, each with
    // its pinned frame, device frame, device rows and count; work runs on
    // `stream` in submission order, read-back on ring_copy
    struct RingSlot {
        void *h_frame = nullptr;          // pinned, HW * 4 bytes
        void *d_frame = nullptr;          // HW * 4 bytes
        gt_detection *d_rows = nullptr;   // det_cap rows
        uint32_t *d_count = nullptr;
        uint32_t *h_count = nullptr;      // pinned
        std::vector<gt_detection> h_pos;  // characterize_points positions of the frame
        hipEvent_t done = nullptr;        // the count is on the host
        uint64_t frame_idx = 0;
    };
    std::vector<RingSlot> ring;
    uint32_t ring_depth = 2;
    uint32_t ring_head = 0, ring_count = 0;  // oldest slot in flight, slots in flight
    hipStream_t ring_copy = nullptr;

    // timing
    bool timing = false;
    std::vector<hipEvent_t> events;
    double stage_ms[3] = {0, 0, 0};
    int n_timed_chunks = 0;
    uint32_t timed_chunk_frames = 0;

    std::atomic<int32_t> interrupt{0};
    std::atomic<uint64_t> prog_cur{0}, prog_total{0};
};

namespace {

// Noise-estimate buffers (also used by gt_frame_std on contexts without the
// snr rules).
gt_status alloc_std(gt_context *c) {
    if (c->d_std_tmp) return GT_OK;
    const size_t HW = static_cast<size_t>(c->H) * c->W;
    GT_HIP(hipMalloc(&c->d_std_tmp, sizeof(float) * ((HW + 7) / 8) * c->chunk));
    GT_HIP(hipMalloc(&c->d_mean, sizeof(float) * c->chunk));
    GT_HIP(hipMalloc(&c->d_std, sizeof(float) * c->chunk));
    return GT_OK;
}

gt_status alloc_scratch(gt_context *c) {
    const size_t HW = static_cast<size_t>(c->H) * c->W;
    GT_HIP(hipMalloc(&c->d_proc, sizeof(float) * HW * c->chunk));
    if (c->der.need_std) {
        gt_status st = alloc_std(c);
        if (st != GT_OK) return st;
        c->kp.fstd = c->d_std;
    }
    if (c->der.adaptive) {
        GT_HIP(hipMalloc(&c->d_adapt_masked, sizeof(float) * HW * c->chunk));
        GT_HIP(hipMalloc(&c->d_std2, sizeof(float) * c->chunk));
        GT_HIP(hipMalloc(&c->d_adapt_counter, sizeof(uint32_t) * c->chunk));
    }
    if (c->der.illum_sigma > 0.0f) {
        const float sigma = c->der.illum_sigma;
        const int32_t rint = static_cast<int32_t>(4.0f * sigma + 0.5f);
        std::vector<float> w(2 * static_cast<size_t>(rint) + 1);
        const float sigma2 = sigma * sigma;
        for (int32_t i = -rint; i <= rint; ++i) {
            const float x = static_cast<float>(i);
            const float x2 = x * x;
            w[i + rint] = std::exp(-x2 / (2.0f * sigma2));
        }
        c->illum_rint = rint;
        GT_HIP(hipMalloc(&c->d_illum_w, sizeof(float) * w.size()));
        GT_HIP(hipMemcpy(c->d_illum_w, w.data(), sizeof(float) * w.size(), hipMemcpyHostToDevice));
        GT_HIP(hipMalloc(&c->d_profile, sizeof(float) * HW));
        GT_HIP(hipMalloc(&c->d_corr, sizeof(float) * HW * c->chunk));
        GT_HIP(hipMalloc(&c->d_illum_tmp, sizeof(float) * HW * c->chunk));
        if (c->der.illum_per_frame)
            GT_HIP(hipMalloc(&c->d_illum_prof, sizeof(float) * HW * c->chunk));
    }
    if (!c->tiled)
        GT_HIP(hipMalloc(&c->d_cand, sizeof(uint32_t) * static_cast<size_t>(c->cand_cap) * c->chunk));
    if (c->log && c->der.threshold_rel > 0.0f)
        GT_HIP(hipMalloc(&c->d_cval, sizeof(float) * static_cast<size_t>(c->cand_cap) * c->chunk));
    GT_HIP(hipMalloc(&c->d_ccount, sizeof(uint32_t) * GT_COUNTER_STRIDE * c->chunk));
    GT_HIP(hipMalloc(&c->d_ocount, sizeof(uint32_t) * GT_COUNTER_STRIDE * c->chunk));
    if (c->tiled) {
        // one slot per 8 pixels; a tile whose frame list is full walks on itself
        c->slow_cap = static_cast<uint32_t>(std::max<size_t>(1024, HW / 8));
        GT_HIP(hipMalloc(&c->d_slow, sizeof(uint2) * static_cast<size_t>(c->slow_cap) * c->chunk));
        GT_HIP(hipMalloc(&c->d_scount, sizeof(uint32_t) * GT_COUNTER_STRIDE * c->chunk));
    }
    return GT_OK;
}

gt_status alloc_execute_buffers(gt_context *c) {
    if (c->slot[0].d_frames) return GT_OK;
    const size_t HW = static_cast<size_t>(c->H) * c->W;
    if (c->gpu_post)
        for (auto &dp : c->d_post)
            GT_HIP(hipMalloc(&dp, sizeof(float4) * static_cast<size_t>(c->det_cap) * c->xchunk));
    for (auto &sl : c->slot) {
        GT_HIP(hipMalloc(&sl.d_frames, 4 * HW * c->xchunk));
        GT_HIP(hipMalloc(&sl.d_rows, sizeof(gt_detection) * static_cast<size_t>(c->det_cap) * c->xchunk));
        GT_HIP(hipMalloc(&sl.d_counts, sizeof(uint32_t) * c->xchunk));
        GT_HIP(hipMalloc(&sl.d_packed, sizeof(gt_detection) * static_cast<size_t>(c->det_cap) * c->xchunk));
        if (c->gpu_post)
            GT_HIP(hipMalloc(&sl.d_packed_post, sizeof(float4) * static_cast<size_t>(c->det_cap) * c->xchunk));
        GT_HIP(hipHostMalloc(&sl.h_stage, 4 * HW * c->xchunk, hipHostMallocDefault));
        GT_HIP(hipHostMalloc(reinterpret_cast<void **>(&sl.h_counts), sizeof(uint32_t) * c->xchunk,
                             hipHostMallocDefault));
        GT_HIP(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
    }
    GT_HIP(hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking));
    return GT_OK;
}

// Grows slot sl's pinned read-back buffers to hold n rows, keeping the first
// `keep` rows (no copy into them may be in flight).
gt_status reserve_slot_rows(gt_context *c, gt_context::Slot &sl, size_t n, size_t keep) {
    if (n <= sl.pk_cap) return GT_OK;
    const size_t cap = std::max<size_t>(n, sl.pk_cap + sl.pk_cap / 2);
    gt_detection *r = nullptr;
    float4 *q = nullptr;
    GT_HIP(hipHostMalloc(reinterpret_cast<void **>(&r), sizeof(gt_detection) * cap, hipHostMallocDefault));
    if (c->gpu_post) {
        const hipError_t e = hipHostMalloc(reinterpret_cast<void **>(&q), sizeof(float4) * cap,
                                           hipHostMallocDefault);
        if (e != hipSuccess) {
            (void)hipHostFree(r);
            return fail(GT_ERR_GPU_DEVICE, "pinned allocation failed: %s", hipGetErrorString(e));
        }
    }
    if (keep) {
        std::memcpy(r, sl.h_pk, sizeof(gt_detection) * keep);
        if (q) std::memcpy(q, sl.h_pk_post, sizeof(float4) * keep);
    }
    if (sl.h_pk) (void)hipHostFree(sl.h_pk);
    if (sl.h_pk_post) (void)hipHostFree(sl.h_pk_post);
    sl.h_pk = r;
    sl.h_pk_post = q;
    sl.pk_cap = cap;
    return GT_OK;
}

gt_status reserve_host_rows(gt_context *c, size_t n) {
    if (c->gpu_post && n > c->h_post_cap) {
        if (c->h_post) (void)hipHostFree(c->h_post);
        c->h_post = nullptr;
        c->h_post_cap = 0;
        const size_t cap = std::max<size_t>(n, 2 * c->h_rows_cap + 4096);
        GT_HIP(hipHostMalloc(reinterpret_cast<void **>(&c->h_post), sizeof(float4) * cap,
                             hipHostMallocDefault));
        c->h_post_cap = cap;
    }
    if (n <= c->h_rows_cap) return GT_OK;
    if (c->h_rows) (void)hipHostFree(c->h_rows);
    c->h_rows = nullptr;
    c->h_rows_cap = 0;
    const size_t cap = std::max<size_t>(n, 2 * c->h_rows_cap + 4096);
    GT_HIP(hipHostMalloc(reinterpret_cast<void **>(&c->h_rows), sizeof(gt_detection) * cap,
                         hipHostMallocDefault));
    c->h_rows_cap = cap;
    return GT_OK;
}

void free_ring(gt_context *c) {
    for (auto &r : c->ring) {
        if (r.h_frame) (void)hipHostFree(r.h_frame);
        if (r.d_frame) (void)hipFree(r.d_frame);
        if (r.d_rows) (void)hipFree(r.d_rows);
        if (r.d_count) (void)hipFree(r.d_count);
        if (r.h_count) (void)hipHostFree(r.h_count);
        if (r.done) (void)hipEventDestroy(r.done);
    }
    c->ring.clear();
    if (c->ring_copy) (void)hipStreamDestroy(c->ring_copy);
    c->ring_copy = nullptr;
    c->ring_head = c->ring_count = 0;
}

gt_status alloc_ring(gt_context *c) {
    if (!c->ring.empty()) return GT_OK;
    const size_t HW = static_cast<size_t>(c->H) * c->W;
    c->ring.resize(c->ring_depth);
    for (auto &r : c->ring) {
        GT_HIP(hipHostMalloc(&r.h_frame, 4 * HW, hipHostMallocDefault));
        GT_HIP(hipMalloc(&r.d_frame, 4 * HW));
        GT_HIP(hipMalloc(&r.d_rows, sizeof(gt_detection) * static_cast<size_t>(c->det_cap)));
        GT_HIP(hipMalloc(&r.d_count, sizeof(uint32_t)));
        GT_HIP(hipHostMalloc(reinterpret_cast<void **>(&r.h_count), sizeof(uint32_t),
                             hipHostMallocDefault));
        GT_HIP(hipEventCreateWithFlags(&r.done, hipEventDisableTiming));
    }
    GT_HIP(hipStreamCreateWithFlags(&c->ring_copy, hipStreamNonBlocking));
    return GT_OK;
}

void free_all(gt_context *c) {
    if (c->d_proc) (void)hipFree(c->d_proc);
    if (c->d_cand) (void)hipFree(c->d_cand);
    if (c->d_cval) (void)hipFree(c->d_cval);
    if (c->d_ccount) (void)hipFree(c->d_ccount);
    if (c->d_ocount) (void)hipFree(c->d_ocount);
    if (c->d_slow) (void)hipFree(c->d_slow);
    if (c->d_scount) (void)hipFree(c->d_scount);
    if (c->d_logtab) (void)hipFree(c->d_logtab);
    if (c->d_props) (void)hipFree(c->d_props);
    if (c->d_stats) (void)hipFree(c->d_stats);
    if (c->d_offs) (void)hipFree(c->d_offs);
    if (c->d_std_tmp) (void)hipFree(c->d_std_tmp);
    if (c->d_adapt_masked) (void)hipFree(c->d_adapt_masked);
    if (c->d_std2) (void)hipFree(c->d_std2);
    if (c->d_adapt_counter) (void)hipFree(c->d_adapt_counter);
    if (c->d_adapt_state) (void)hipFree(c->d_adapt_state);
    if (c->d_mean) (void)hipFree(c->d_mean);
    if (c->d_std) (void)hipFree(c->d_std);
    if (c->l1_proc) (void)hipFree(c->l1_proc);
    if (c->l1_ocount) (void)hipFree(c->l1_ocount);
    if (c->l1_scount) (void)hipFree(c->l1_scount);
    if (c->l1_slow) (void)hipFree(c->l1_slow);
    for (int l = 0; l < 2; ++l) {
        if (c->lane_done[l]) (void)hipEventDestroy(c->lane_done[l]);
        if (c->ovl_det[l]) (void)hipEventDestroy(c->ovl_det[l]);
        if (c->lane_stream[l]) (void)hipStreamDestroy(c->lane_stream[l]);
    }
    if (c->lane_fork) (void)hipEventDestroy(c->lane_fork);
    if (c->d_illum_w) (void)hipFree(c->d_illum_w);
    if (c->d_profile) (void)hipFree(c->d_profile);
    if (c->d_corr) (void)hipFree(c->d_corr);
    if (c->d_illum_tmp) (void)hipFree(c->d_illum_tmp);
    if (c->d_illum_prof) (void)hipFree(c->d_illum_prof);
    for (auto &sl : c->slot) {
        if (sl.d_frames) (void)hipFree(sl.d_frames);
        if (sl.d_rows) (void)hipFree(sl.d_rows);
        if (sl.d_counts) (void)hipFree(sl.d_counts);
        if (sl.d_packed) (void)hipFree(sl.d_packed);
        if (sl.d_packed_post) (void)hipFree(sl.d_packed_post);
        if (sl.h_pk) (void)hipHostFree(sl.h_pk);
        if (sl.h_pk_post) (void)hipHostFree(sl.h_pk_post);
        if (sl.h_stage) (void)hipHostFree(sl.h_stage);
        if (sl.h_counts) (void)hipHostFree(sl.h_counts);
        if (sl.done) (void)hipEventDestroy(sl.done);
    }
    if (c->h_rows) (void)hipHostFree(c->h_rows);
    if (c->h_post) (void)hipHostFree(c->h_post);
    for (float4 *dp : c->d_post)
        if (dp) (void)hipFree(dp);
    if (c->d_dough_offs) (void)hipFree(c->d_dough_offs);
    if (c->copy_stream) (void)hipStreamDestroy(c->copy_stream);
    for (hipEvent_t e : c->events) (void)hipEventDestroy(e);
    free_ring(c);
    if (c->stream) (void)hipStreamDestroy(c->stream);
}

// adaptive_background passes over a chunk's walked rows (before characterize).
gt_status run_adaptive(gt_context *c, const void *frames, int32_t dtype, uint32_t nf,
                       gt_detection *out, uint32_t cap, hipStream_t s) {
    const size_t need = static_cast<size_t>(c->chunk) * cap;
    if (need > c->adapt_state_cap) {
        GT_HIP(hipStreamSynchronize(s));
        if (c->d_adapt_state) (void)hipFree(c->d_adapt_state);
        c->d_adapt_state = nullptr;
        c->adapt_state_cap = 0;
        GT_HIP(hipMalloc(&c->d_adapt_state, need));
        c->adapt_state_cap = need;
    }
    GT_HIP(gt_launch_adaptive(frames, dtype, nf, c->kp, c->der.adaptive, out, cap, c->d_ocount,
                              c->d_std_tmp, c->d_mean, c->d_std2, c->d_adapt_masked,
                              c->d_adapt_state, c->d_adapt_counter, s));
    return GT_OK;
}

// The whole-video illumination profile: start (zero the running sum), add
// frames in order, finish (divide by the frame count, then smooth into
// d_profile).  setup_illumination_profile, src/execute_gpu.rs:849-884, over
// mean_from_iter (:819-847).
gt_status illum_begin(gt_context *c, hipStream_t s) {
    GT_HIP(gt_launch_zero_u32(reinterpret_cast<uint32_t *>(c->d_illum_tmp),
                              static_cast<size_t>(c->H) * c->W, s));
    return GT_OK;
}
gt_status illum_add(gt_context *c, const void *frames, int32_t dtype, uint32_t nf, hipStream_t s) {
    GT_HIP(gt_launch_frames_accumulate(frames, dtype, nf, static_cast<size_t>(c->H) * c->W,
                                       c->d_illum_tmp, s));
    return GT_OK;
}
gt_status illum_finish(gt_context *c, uint64_t nframes, hipStream_t s) {
    const size_t HW = static_cast<size_t>(c->H) * c->W;
    GT_HIP(gt_launch_scale_div(c->d_illum_tmp, HW, static_cast<float>(nframes), s));
    // the second pass's input (tmp) and output may not overlap the first's
    GT_HIP(gt_launch_illum_smooth(c->d_illum_tmp, c->d_corr, c->d_profile, static_cast<int>(c->H),
                                  static_cast<int>(c->W), 1, c->d_illum_w, c->illum_rint, s));
    return GT_OK;
}
// The whole-video profile over device frames (gt_detect_device).
gt_status illum_profile_device(gt_context *c, const void *d_frames, int32_t dtype,
                               uint32_t nframes, hipStream_t s) {
    const size_t HW = static_cast<size_t>(c->H) * c->W;
    gt_status st = illum_begin(c, s);
    if (st != GT_OK) return st;
    const size_t esz = dtype_size(dtype);
    for (uint32_t f0 = 0; f0 < nframes; f0 += c->chunk) {
        const uint32_t nf = std::min(c->chunk, nframes - f0);
        st = illum_add(c, static_cast<const char *>(d_frames) + esz * HW * f0, dtype, nf, s);
        if (st != GT_OK) return st;
    }
    return illum_finish(c, nframes, s);
}
// The chunk's frames divided by the profile (per frame: by their own
// smoothed copies) into d_corr, which the detection then reads as f32
// (submit_work, src/execute_gpu.rs:138-166).
gt_status illum_correct(gt_context *c, const void *frames, int32_t dtype, uint32_t nf,
                        hipStream_t s) {
    const size_t HW = static_cast<size_t>(c->H) * c->W;
    if (c->der.illum_per_frame) {
        // f32 copies of the frames, then their smoothed profiles
        GT_HIP(gt_launch_widen(frames, dtype, HW * nf, c->d_corr, s));
        GT_HIP(gt_launch_illum_smooth(c->d_corr, c->d_illum_tmp, c->d_illum_prof,
                                      static_cast<int>(c->H), static_cast<int>(c->W), nf,
                                      c->d_illum_w, c->illum_rint, s));
        GT_HIP(gt_launch_illum_divide(frames, dtype, nf, HW, c->d_illum_prof, HW, c->d_corr, s));
    } else {
        GT_HIP(gt_launch_illum_divide(frames, dtype, nf, HW, c->d_profile, 0, c->d_corr, s));
    }
    return GT_OK;
}

// Whether the second lane applies: the plain tiled path (no LoG, learned
// proposal, noise estimate, adaptive passes or illumination correction,
// whose scratch is single).
bool plain_tiled_only(const gt_context *c) {
    return c->tiled && !c->log && !c->cnn && !c->der.need_std && !c->der.adaptive &&
           !(c->der.illum_sigma > 0.0f);
}

bool lanes_apply(const gt_context *c) { return c->lanes > 1 && plain_tiled_only(c); }

gt_status alloc_lane1(gt_context *c) {
    if (c->l1_proc) return GT_OK;
    const size_t HW = static_cast<size_t>(c->H) * c->W;
    for (int l = 0; l < 2; ++l) {
        GT_HIP(hipStreamCreateWithFlags(&c->lane_stream[l], hipStreamNonBlocking));
        GT_HIP(hipEventCreateWithFlags(&c->lane_done[l], hipEventDisableTiming));
        GT_HIP(hipEventCreateWithFlags(&c->ovl_det[l], hipEventDisableTiming));
    }
    GT_HIP(hipEventCreateWithFlags(&c->lane_fork, hipEventDisableTiming));
    GT_HIP(hipMalloc(&c->l1_ocount, sizeof(uint32_t) * GT_COUNTER_STRIDE * c->chunk));
    GT_HIP(hipMalloc(&c->l1_scount, sizeof(uint32_t) * GT_COUNTER_STRIDE * c->chunk));
    GT_HIP(hipMalloc(&c->l1_slow, sizeof(uint2) * static_cast<size_t>(c->slow_cap) * c->chunk));
    GT_HIP(hipMalloc(&c->l1_proc, sizeof(float) * HW * c->chunk));
    return GT_OK;
}

// Runs the three pipeline stages over `nframes` device frames in chunks.
gt_status run_pipeline(gt_context *c, const void *d_frames, int32_t dtype, uint32_t nframes,
                       gt_detection *d_out, uint32_t cap, uint32_t *d_counts, hipStream_t s) {
    const size_t HW = static_cast<size_t>(c->H) * c->W;
    const size_t esz = dtype_size(dtype);
    GT_HIP(gt_launch_zero_u32(d_counts, nframes, s));
    if (c->timing) {
        for (double &v : c->stage_ms) v = 0.0;
        c->n_timed_chunks = 0;
    }
    const uint32_t nchunks = (nframes + c->chunk - 1) / c->chunk;
    if (c->timing && c->events.size() < 4u * nchunks) {
        while (c->events.size() < 4u * nchunks) {
            hipEvent_t e;
            GT_HIP(hipEventCreate(&e));
            c->events.push_back(e);
        }
    }
    // chunk ci runs on lane ci % 2 when the second lane applies; the lanes
    // start after the work queued on s and s waits for both at the end
    const bool two = lanes_apply(c) && nchunks > 1;
    // overlap: chunk ci's slow walks and characterize run on lane stream 1
    // while s goes on with chunk ci + 1; chunk ci uses lane ci % 2's scratch,
    // and s waits for chunk ci's tail before chunk ci + 2 reuses it
    const bool ovl = !two && c->overlap && plain_tiled_only(c) && nchunks > 1;
    if (two || ovl) {
        gt_status lst = alloc_lane1(c);
        if (lst != GT_OK) return lst;
        GT_HIP(hipEventRecord(c->lane_fork, s));
        for (int l = 0; l < 2; ++l) GT_HIP(hipStreamWaitEvent(c->lane_stream[l], c->lane_fork, 0));
    }
    for (uint32_t ci = 0; ci < nchunks; ++ci) {
        const int lane = (two || ovl) ? static_cast<int>(ci & 1u) : 0;
        hipStream_t cs = two ? c->lane_stream[lane] : s;
        if (ovl && ci >= 2) GT_HIP(hipStreamWaitEvent(s, c->lane_done[lane], 0));
        float *proc = lane ? c->l1_proc : c->d_proc;
        uint32_t *ocount = lane ? c->l1_ocount : c->d_ocount;
        uint32_t *scount = lane ? c->l1_scount : c->d_scount;
        uint2 *slow = lane ? c->l1_slow : c->d_slow;
        const uint32_t f0 = ci * c->chunk;
        const uint32_t nf = std::min(c->chunk, nframes - f0);
        const char *frames = static_cast<const char *>(d_frames) + esz * HW * f0;
        int32_t fdtype = dtype;
        if (c->der.illum_sigma > 0.0f) {
            gt_status ist = illum_correct(c, frames, dtype, nf, cs);
            if (ist != GT_OK) return ist;
            frames = reinterpret_cast<const char *>(c->d_corr);
            fdtype = GT_F32;
        }
        hipEvent_t *ev = c->timing ? &c->events[4 * ci] : nullptr;
        // the plain tiled path's counters (rows, handed-over walks) are
        // cleared by the bandpass launch
        const bool plain_tiled = c->tiled && !c->cnn && !c->log;
        if (!plain_tiled)
            GT_HIP(gt_launch_zero_u32(ocount, static_cast<size_t>(nf) * GT_COUNTER_STRIDE, cs));
        if (!c->tiled || c->log)
            GT_HIP(gt_launch_zero_u32(c->d_ccount, static_cast<size_t>(nf) * GT_COUNTER_STRIDE, cs));
        if (ev) GT_HIP(hipEventRecord(ev[0], cs));
        if (c->der.need_std)
            GT_HIP(gt_launch_frame_std(frames, fdtype, nf, HW, c->d_std_tmp, c->d_mean, c->d_std, cs));
        gt_detection *out = d_out + static_cast<size_t>(f0) * cap;
        if (plain_tiled) {
            // bandpass (clearing the chunk's counters), local max and walk in
            // one kernel; then the walks it handed over and characterize over
            // the frame's compacted detections (whole wavefronts of
            // detections instead of a few lanes per tile)
            const GtPlainScratch sc = {proc, ocount, scount, slow, c->slow_cap};
            GT_HIP(gt_plain_head(frames, fdtype, nf, c->kp, sc, out, cap, f0, cs,
                                 ev ? ev[1] : nullptr));
            if (ev) GT_HIP(hipEventRecord(ev[2], cs));
            hipStream_t ts = cs;   // the chunk's tail: slow walks, characterize, counts
            if (ovl) {
                GT_HIP(hipEventRecord(c->ovl_det[lane], cs));
                ts = c->lane_stream[1];
                GT_HIP(hipStreamWaitEvent(ts, c->ovl_det[lane], 0));
            }
            if (!c->der.adaptive) {
                GT_HIP(gt_plain_tail(frames, fdtype, nf, c->kp, sc, out, cap, d_counts + f0, f0,
                                     ts));
            } else {
                GtKParams kw = c->kp;
                kw.characterize = 0;
                GT_HIP(gt_launch_walk_slow(proc, frames, fdtype, nf, kw, slow, c->slow_cap,
                                           scount, out, cap, ocount, f0, ts));
                gt_status ast = run_adaptive(c, frames, fdtype, nf, out, cap, ts);
                if (ast != GT_OK) return ast;
                if (c->kp.characterize)
                    GT_HIP(gt_launch_characterize_det(proc, frames, fdtype, nf, c->kp, out, cap,
                                                      ocount, ts, d_counts + f0));
                else
                    GT_HIP(gt_launch_copy_counts(ocount, d_counts + f0, nf, ts));
            }
            if (ev) GT_HIP(hipEventRecord(ev[3], ts));
            if (ovl) GT_HIP(hipEventRecord(c->lane_done[lane], ts));
            continue;
        }
        GT_HIP(gt_launch_bandpass(frames, fdtype, proc, nf, c->kp, cs));
        if (ev) GT_HIP(hipEventRecord(ev[1], cs));
        if (c->log) {
            // scale-space maxima of all scales, then their walks (and, with
            // adaptive_background, the acceptance passes before characterize)
            GT_HIP(gt_launch_log_tile(frames, fdtype, proc, nf, c->kp, c->d_logtab, c->log_Rm,
                                      c->d_cand, c->cand_cap, c->d_ccount, c->d_cval, cs));
            if (ev) GT_HIP(hipEventRecord(ev[2], cs));
            GtKParams kw = c->kp;
            if (c->der.adaptive) kw.characterize = 0;
            GT_HIP(gt_launch_walk_log(proc, frames, fdtype, nf, kw, c->d_logtab, c->d_cand,
                                      c->cand_cap, c->d_ccount, out, cap, ocount, f0, c->d_cval,
                                      cs));
            if (c->der.adaptive) {
                gt_status ast = run_adaptive(c, frames, fdtype, nf, out, cap, cs);
                if (ast != GT_OK) return ast;
                if (c->kp.characterize)
                    GT_HIP(gt_launch_characterize_det(proc, frames, fdtype, nf, c->kp, out,
                                                      cap, ocount, cs));
            }
            GT_HIP(gt_launch_copy_counts_checked(ocount, c->d_ccount, c->cand_cap,
                                                 d_counts + f0, nf, cs));
            if (ev) GT_HIP(hipEventRecord(ev[3], cs));
            continue;
        }
        if (c->tiled && c->cnn && c->cnn_scope == GT_CNN_MOVING) {
            // one walk step per candidate in the tile (a walk that stops
            // there is finished: the plain result); the network's proposals
            // for the walks still moving, one verifying evaluation each, the
            // rest of the walk where it fails
            GtKParams kw = c->kp;
            kw.characterize = 0;
            GT_HIP(gt_launch_zero_u32(scount, static_cast<size_t>(nf) * GT_COUNTER_STRIDE, cs));
            GT_HIP(gt_launch_detect_tile(proc, frames, fdtype, nf, kw, out, cap, ocount,
                                         slow, c->slow_cap, scount, f0, cs, 1));
            if (ev) GT_HIP(hipEventRecord(ev[2], cs));
            const uint32_t *list = reinterpret_cast<const uint32_t *>(slow);
            GT_HIP(gt_cnn_launch_forward_frames(gt_cnn_packed(c->cnn), proc,
                                                static_cast<int>(c->H), static_cast<int>(c->W), nf,
                                                list, 2, c->slow_cap, scount, c->d_props, cs));
            GT_HIP(gt_launch_walk_learned(proc, nf, kw, list, 2, 1, c->slow_cap, scount,
                                          c->d_props, c->cnn_conf, c->cnn_tol, out, cap,
                                          ocount, c->d_stats, f0, cs));
            if (c->kp.characterize)
                GT_HIP(gt_launch_characterize_det(proc, frames, fdtype, nf, c->kp, out, cap,
                                                  ocount, cs));
        } else if (c->tiled && c->cnn) {
            // every candidate: the network's proposals for all of them, one
            // verifying walk evaluation each (the full walk where it fails)
            GtKParams kw = c->kp;
            kw.characterize = 0;
            GT_HIP(gt_launch_zero_u32(c->d_ccount, static_cast<size_t>(nf) * GT_COUNTER_STRIDE, cs));
            GT_HIP(gt_launch_detect_cand(proc, fdtype, nf, kw, c->d_cand, c->cand_cap,
                                         c->d_ccount, cs));
            if (ev) GT_HIP(hipEventRecord(ev[2], cs));
            GT_HIP(gt_cnn_launch_forward_frames(gt_cnn_packed(c->cnn), proc,
                                                static_cast<int>(c->H), static_cast<int>(c->W), nf,
                                                c->d_cand, 1, c->cand_cap, c->d_ccount, c->d_props,
                                                cs));
            GT_HIP(gt_launch_walk_learned(proc, nf, kw, c->d_cand, 1, 0, c->cand_cap,
                                          c->d_ccount, c->d_props, c->cnn_conf, c->cnn_tol, out,
                                          cap, ocount, c->d_stats, f0, cs));
            if (c->kp.characterize)
                GT_HIP(gt_launch_characterize_det(proc, frames, fdtype, nf, c->kp, out, cap,
                                                  ocount, cs));
        } else {
            GT_HIP(gt_launch_localmax(proc, nf, c->kp, c->d_cand, c->cand_cap, c->d_ccount, cs));
            if (ev) GT_HIP(hipEventRecord(ev[2], cs));
            GT_HIP(gt_launch_walk(proc, frames, fdtype, nf, c->kp, c->d_cand, c->cand_cap,
                                  c->d_ccount, out, cap, ocount, f0, cs));
            if (c->der.adaptive) {
                gt_status ast = run_adaptive(c, frames, fdtype, nf, out, cap, cs);
                if (ast != GT_OK) return ast;
            }
        }
        GT_HIP(gt_launch_copy_counts(ocount, d_counts + f0, nf, cs));
        if (ev) GT_HIP(hipEventRecord(ev[3], cs));
    }
    if (two) {
        for (int l = 0; l < 2; ++l) {
            GT_HIP(hipEventRecord(c->lane_done[l], c->lane_stream[l]));
            GT_HIP(hipStreamWaitEvent(s, c->lane_done[l], 0));
        }
    }
    if (ovl)
        for (int l = 0; l < 2; ++l) GT_HIP(hipStreamWaitEvent(s, c->lane_done[l], 0));
    if (c->timing) {
        c->n_timed_chunks = static_cast<int>(nchunks);
        c->timed_chunk_frames = std::min(c->chunk, nframes);
    }
    return GT_OK;
}

}  // namespace

extern "C" {

int32_t gt_abi_version(void) { return GT_ABI_VERSION; }

const char *gt_last_error(void) { return g_last_error.c_str(); }

void gt_params_default(gt_params *p) { gt_host_params_default(p); }

uint32_t gt_column_names(const gt_params *params, const char **names, const char **types) {
    return gt_host_column_names(params, names, types);
}

gt_status gt_context_create(const gt_params *params, uint32_t nrows, uint32_t ncols,
                            int32_t characterize_points, gt_context **out) {
    if (!params || !out) return fail(GT_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (nrows == 0 || ncols == 0)
        return fail(GT_ERR_INVALID_ARGUMENT, "frame dimensions must be positive");
    if (static_cast<uint64_t>(nrows) * ncols > 0xFFFFFFFFull)
        return fail(GT_ERR_INVALID_ARGUMENT, "frames above 2^32 pixels are not supported");
    gt_derived der;
    std::string why;
    if (!gt_host_derive(params, nrows, ncols, &der, &why))
        return fail(GT_ERR_INVALID_ARGUMENT, "%s", why.c_str());
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(GT_ERR_GPU_ADAPTER, "no HIP device is available");
    if (der.adaptive && characterize_points)
        return fail(GT_ERR_INVALID_ARGUMENT,
                    "adaptive_background does not apply to characterize_points");
    gt_context *c = new (std::nothrow) gt_context();
    if (!c) return fail(GT_ERR_GPU_DEVICE, "out of host memory");
    c->params = *params;
    c->H = nrows;
    c->W = ncols;
    c->characterize_points = characterize_points;
    c->der = der;
    gt_host_fill_kparams(&der, nrows, ncols, &c->kp);
    // the fused tile kernel needs its staged region in 64 KiB of LDS
    c->tiled = der.style == GT_STYLE_TRACKPY && c->kp.k >= 1 && gt_tile_lds_bytes(c->kp) <= 65536;
    c->log = der.style == GT_STYLE_LOG;
    if (c->log && static_cast<uint64_t>(nrows) * ncols * der.radii.size() > 0xFFFFFFFFull) {
        delete c;
        return fail(GT_ERR_INVALID_ARGUMENT,
                    "LoG candidates are keyed by scale * nrows * ncols + pixel, which must stay "
                    "below 2^32");
    }
    if (hipGetDevice(&c->device) != hipSuccess) c->device = 0;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(GT_ERR_GPU_DEVICE, "hipStreamCreate failed");
    }
    if (gt_init_tables() != hipSuccess) {
        free_all(c);
        delete c;
        return fail(GT_ERR_GPU_DEVICE, "could not initialise the device tables");
    }
    const size_t HW = static_cast<size_t>(nrows) * ncols;
    // Frames per chunk: bounded by a 4 GiB scratch budget and by 512 frames
    // (longer chunks spread the slow walks' and characterize's tails over
    // more frames).  Per frame of device scratch, in bytes per pixel: the
    // processed plane (4); the slow-walk list (1: HW/8 entries of 8 B) on the
    // tiled path, else the candidate list (4) and, with threshold_rel, its
    // values (4); adaptive_background's masked plane (4); illumination
    // correction's planes (8, 12 per frame).  The two gt_execute staging
    // slots, frame + detection buffer each, are allocated on first use.
    size_t bpp = 4 + (c->tiled && !c->log ? 1 : 4);
    if (c->log && der.threshold_rel > 0.0f) bpp += 4;
    if (der.adaptive) bpp += 4;
    if (der.illum_sigma > 0.0f) bpp += der.illum_per_frame ? 12 : 8;
    const size_t per_frame = HW * bpp;
    size_t chunk = (size_t(4) << 30) / (per_frame ? per_frame : 1);
    c->chunk = static_cast<uint32_t>(std::max<size_t>(1, std::min<size_t>(chunk, 512)));
    // host frames are staged 64 at a time: smaller pinned slots, and the
    // first slot of a video starts on the GPU sooner
    c->xchunk = std::min<uint32_t>(c->chunk, 64);
    c->cand_cap = static_cast<uint32_t>(HW);
    c->det_cap = static_cast<uint32_t>(HW);
    gt_status st = alloc_scratch(c);
    if (st == GT_OK && c->log) {
        GtLogTable *tab = new (std::nothrow) GtLogTable;
        if (!tab) {
            st = fail(GT_ERR_GPU_DEVICE, "out of host memory");
        } else {
            gt_host_fill_logtable(&der, tab);
            c->log_Rm = tab->Rm;
            hipError_t e = hipMalloc(&c->d_logtab, sizeof(GtLogTable));
            if (e == hipSuccess) e = hipMemcpy(c->d_logtab, tab, sizeof(GtLogTable), hipMemcpyHostToDevice);
            delete tab;
            if (e != hipSuccess) st = fail(GT_ERR_GPU_DEVICE, "LoG table upload failed: %s", hipGetErrorString(e));
        }
    }
    if (st == GT_OK && der.style == GT_STYLE_TRACKPY && !characterize_points &&
        (der.filter_close || der.doughnut)) {
        // filter_close and the doughnut run on the GPU when the doughnut's
        // stencils fit one wave's LDS
        c->gpu_post = true;
        if (der.doughnut) {
            std::vector<std::pair<int32_t, int32_t>> sig, ring;
            gt_host_doughnut_offsets(static_cast<float>(der.circle) / 2.0f, der.bg_radius,
                                     der.gap_radius, &sig, &ring);
            const int32_t lim = 32767;
            bool fits = gt_doughnut_lds_bytes(static_cast<uint32_t>(sig.size()),
                                              static_cast<uint32_t>(ring.size())) <= 65536;
            for (const auto &o : sig) fits = fits && std::abs(o.first) <= lim && std::abs(o.second) <= lim;
            for (const auto &o : ring) fits = fits && std::abs(o.first) <= lim && std::abs(o.second) <= lim;
            if (!fits) {
                c->gpu_post = false;
            } else {
                std::vector<short2> tab;
                for (const auto &o : sig) tab.push_back(make_short2(static_cast<short>(o.first), static_cast<short>(o.second)));
                for (const auto &o : ring) tab.push_back(make_short2(static_cast<short>(o.first), static_cast<short>(o.second)));
                c->dough_nsig = static_cast<uint32_t>(sig.size());
                c->dough_nring = static_cast<uint32_t>(ring.size());
                hipError_t e = hipMalloc(&c->d_dough_offs, sizeof(short2) * std::max<size_t>(1, tab.size()));
                if (e == hipSuccess && !tab.empty())
                    e = hipMemcpy(c->d_dough_offs, tab.data(), sizeof(short2) * tab.size(),
                                  hipMemcpyHostToDevice);
                if (e != hipSuccess) st = fail(GT_ERR_GPU_DEVICE, "doughnut table upload failed: %s", hipGetErrorString(e));
            }
        }
    }
    if (st != GT_OK) {
        std::string msg = g_last_error;
        free_all(c);
        delete c;
        return fail(st, "%s", msg.c_str());
    }
    *out = c;
    return GT_OK;
}

void gt_context_destroy(gt_context *c) {
    if (!c) return;
    (void)hipStreamSynchronize(c->stream);
    free_all(c);
    delete c;
}

gt_status gt_detect_device(gt_context *c, const void *d_frames, int32_t dtype, uint32_t nframes,
                           gt_detection *d_out, uint32_t cap_per_frame, uint32_t *d_counts,
                           void *stream) {
    if (!c) return fail(GT_ERR_INVALID_ARGUMENT, "null context");
    if (!dtype_size(dtype)) return fail(GT_ERR_INVALID_ARGUMENT, "unknown dtype %d", dtype);
    if (nframes == 0) return fail(GT_ERR_EMPTY_ITERATOR, "the provided video doesn't contain any frames.");
    if (!d_frames || !d_out || !d_counts || cap_per_frame == 0)
        return fail(GT_ERR_INVALID_ARGUMENT, "null device buffer or zero capacity");
    if (c->ring_count)
        return fail(GT_ERR_INVALID_ARGUMENT, "frames submitted with gt_submit_frame are in flight");
    hipStream_t s = static_cast<hipStream_t>(stream);  // NULL: the default stream
    if (c->der.illum_sigma > 0.0f && !c->der.illum_per_frame) {
        // the profile of this call's frames (the video it is given)
        gt_status ist = illum_profile_device(c, d_frames, dtype, nframes, s);
        if (ist != GT_OK) return ist;
    }
    return run_pipeline(c, d_frames, dtype, nframes, d_out, cap_per_frame, d_counts, s);
}

// submit_work (src/execute_gpu.rs:95-422) for one host frame: staged into the
// ring's pinned slot, then the copy, the pass sequence and the count's
// read-back are queued on the context's stream.
gt_status gt_submit_frame(gt_context *c, const void *px, int32_t dtype, uint64_t n_pixels,
                          uint64_t frame_idx, const gt_result_row *pos, uint32_t npos) {
    if (!c) return fail(GT_ERR_INVALID_ARGUMENT, "null context");
    const size_t esz = dtype_size(dtype);
    if (!esz) return fail(GT_ERR_INVALID_ARGUMENT, "unknown dtype %d", dtype);
    if (!px) return fail(GT_ERR_INVALID_ARGUMENT, "null frame");
    const size_t HW = static_cast<size_t>(c->H) * c->W;
    if (n_pixels != HW)
        return fail(GT_ERR_DIMENSION_MISMATCH,
                    "dimension mismatch. provided dimensions are (%u, %u), but frame %llu contains "
                    "%llu pixels",
                    c->H, c->W, static_cast<unsigned long long>(frame_idx),
                    static_cast<unsigned long long>(n_pixels));
    if (c->characterize_points && (npos > 0 && !pos))
        return fail(GT_ERR_INVALID_ARGUMENT, "null positions");
    if (!c->characterize_points && (pos || npos))
        return fail(GT_ERR_INVALID_ARGUMENT,
                    "positions given to a context created without characterize_points");
    if (c->characterize_points && npos > c->det_cap)
        return fail(GT_ERR_CAPACITY, "%u positions for one frame exceed the capacity of %u", npos,
                    c->det_cap);
    if (c->der.illum_sigma > 0.0f && !c->der.illum_per_frame)
        return fail(GT_ERR_INVALID_ARGUMENT,
                    "illumination correction over the whole video needs every frame first "
                    "(gt_execute); per-frame correction streams");
    if (c->ring_count == c->ring_depth)
        return fail(GT_ERR_INVALID_ARGUMENT, "%u frames are in flight: fetch one first",
                    c->ring_count);
    gt_status st = alloc_ring(c);
    if (st != GT_OK) {
        std::string msg = g_last_error;
        free_ring(c);
        return fail(st, "%s", msg.c_str());
    }
    gt_context::RingSlot &r = c->ring[(c->ring_head + c->ring_count) % c->ring_depth];
    std::memcpy(r.h_frame, px, esz * HW);
    GT_HIP(hipMemcpyAsync(r.d_frame, r.h_frame, esz * HW, hipMemcpyHostToDevice, c->stream));
    if (c->characterize_points) {
        // characterize at the given positions (submit_work's characterization
        // branch, src/execute_gpu.rs:190-229): seed = position index
        r.h_pos.resize(npos);
        for (uint32_t i = 0; i < npos; ++i) {
            r.h_pos[i].row = pos[i];
            r.h_pos[i].seed = i;
            r.h_pos[i].frame = static_cast<uint32_t>(frame_idx);
        }
        if (npos)
            GT_HIP(hipMemcpyAsync(r.d_rows, r.h_pos.data(), sizeof(gt_detection) * npos,
                                  hipMemcpyHostToDevice, c->stream));
        *r.h_count = npos;
        GT_HIP(hipMemcpyAsync(r.d_count, r.h_count, sizeof(uint32_t), hipMemcpyHostToDevice,
                              c->stream));
        const void *fr = r.d_frame;
        int32_t fdt = dtype;
        if (c->der.illum_sigma > 0.0f) {
            gt_status ist = illum_correct(c, fr, dtype, 1, c->stream);
            if (ist != GT_OK) return ist;
            fr = c->d_corr;
            fdt = GT_F32;
        }
        GT_HIP(gt_launch_bandpass(fr, fdt, c->d_proc, 1, c->kp, c->stream));
        GT_HIP(gt_launch_characterize_rows(c->d_proc, fr, fdt, 1, c->kp, r.d_rows, c->det_cap,
                                           r.d_count, npos, c->stream));
    } else {
        st = run_pipeline(c, r.d_frame, dtype, 1, r.d_rows, c->det_cap, r.d_count, c->stream);
        if (st != GT_OK) return st;
    }
    GT_HIP(hipMemcpyAsync(r.h_count, r.d_count, sizeof(uint32_t), hipMemcpyDeviceToHost,
                          c->stream));
    GT_HIP(hipEventRecord(r.done, c->stream));
    r.frame_idx = frame_idx;
    ++c->ring_count;
    return GT_OK;
}

// get_work (src/execute_gpu.rs:424-467): waits for the oldest frame in
// flight and copies its rows out in ascending seed order.
gt_status gt_fetch_frame(gt_context *c, uint64_t *frame_idx, gt_result_row *out, uint32_t cap,
                         uint32_t *n) {
    if (!c || !frame_idx || !n || (cap && !out)) return fail(GT_ERR_INVALID_ARGUMENT, "null argument");
    if (c->ring_count == 0) return fail(GT_ERR_INVALID_ARGUMENT, "no frame is in flight");
    gt_context::RingSlot &r = c->ring[c->ring_head];
    GT_HIP(hipEventSynchronize(r.done));
    const uint32_t cnt = *r.h_count;
    *frame_idx = r.frame_idx;
    auto pop = [&]() {
        c->ring_head = (c->ring_head + 1) % c->ring_depth;
        --c->ring_count;
    };
    if (cnt == 0xFFFFFFFFu || cnt > c->det_cap) {
        pop();
        *n = 0;
        return fail(GT_ERR_CAPACITY, "frame %llu exceeded the candidate or detection capacity",
                    static_cast<unsigned long long>(r.frame_idx));
    }
    *n = cnt;
    if (cnt > cap)   // the frame stays in flight: fetch again with room for *n rows
        return fail(GT_ERR_CAPACITY, "frame %llu has %u rows; the buffer holds %u",
                    static_cast<unsigned long long>(r.frame_idx), cnt, cap);
    gt_status st = reserve_host_rows(c, cnt);
    if (st != GT_OK) return st;
    if (cnt) {
        GT_HIP(hipMemcpyAsync(c->h_rows, r.d_rows, sizeof(gt_detection) * cnt,
                              hipMemcpyDeviceToHost, c->ring_copy));
        GT_HIP(hipStreamSynchronize(c->ring_copy));
        std::sort(c->h_rows, c->h_rows + cnt,
                  [](const gt_detection &a, const gt_detection &b) { return a.seed < b.seed; });
        for (uint32_t i = 0; i < cnt; ++i) out[i] = c->h_rows[i].row;
    }
    pop();
    return GT_OK;
}

gt_status gt_context_set_inflight(gt_context *c, uint32_t depth) {
    if (!c) return fail(GT_ERR_INVALID_ARGUMENT, "null context");
    if (depth < 1 || depth > 16) return fail(GT_ERR_INVALID_ARGUMENT, "depth must be 1..16");
    if (c->ring_count) return fail(GT_ERR_INVALID_ARGUMENT, "frames are in flight");
    if (depth != c->ring_depth) {
        (void)hipStreamSynchronize(c->stream);
        free_ring(c);
        c->ring_depth = depth;
    }
    return GT_OK;
}

gt_status gt_context_set_lanes(gt_context *c, int32_t lanes) {
    if (!c) return fail(GT_ERR_INVALID_ARGUMENT, "null context");
    if (lanes != 1 && lanes != 2) return fail(GT_ERR_INVALID_ARGUMENT, "lanes must be 1 or 2");
    c->lanes = lanes;
    return GT_OK;
}

gt_status gt_context_set_overlap(gt_context *c, int32_t enable) {
    if (!c) return fail(GT_ERR_INVALID_ARGUMENT, "null context");
    c->overlap = enable != 0;
    return GT_OK;
}

gt_status gt_set_timing(gt_context *c, int32_t enable) {
    if (!c) return fail(GT_ERR_INVALID_ARGUMENT, "null context");
    c->timing = enable != 0;
    return GT_OK;
}

int32_t gt_stage_times(gt_context *c, double *stage_ms, int32_t n, uint32_t *launches,
                       uint32_t *frames_per_launch) {
    if (!c || !c->timing || c->n_timed_chunks == 0) return 0;
    if (launches) *launches = static_cast<uint32_t>(c->n_timed_chunks);
    if (frames_per_launch) *frames_per_launch = c->timed_chunk_frames;
    double acc[3] = {0, 0, 0};
    for (int ci = 0; ci < c->n_timed_chunks; ++ci) {
        hipEvent_t *ev = &c->events[4 * ci];
        if (hipEventSynchronize(ev[3]) != hipSuccess) return 0;
        for (int k = 0; k < 3; ++k) {
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, ev[k], ev[k + 1]) != hipSuccess) return 0;
            acc[k] += ms;
        }
    }
    const int m = std::min<int>(n, 3);
    for (int k = 0; k < m; ++k) stage_ms[k] = acc[k];
    return m;
}

gt_status gt_set_interrupt(gt_context *c, int32_t value) {
    if (!c) return fail(GT_ERR_INVALID_ARGUMENT, "null context");
    c->interrupt.store(value);
    return GT_OK;
}

gt_status gt_progress(gt_context *c, uint64_t *cur, uint64_t *total) {
    if (!c) return fail(GT_ERR_INVALID_ARGUMENT, "null context");
    if (cur) *cur = c->prog_cur.load();
    if (total) *total = c->prog_total.load();
    return GT_OK;
}

void gt_output_free(gt_output *out) {
    if (!out) return;
    free(out->data);
    out->data = nullptr;
    out->nrows = 0;
}

}  // extern "C"

namespace {

// Where gt_execute's frames come from: a host array or a video file.
struct FrameSrc {
    int32_t dtype = GT_U16;
    uint64_t nframes = 0;
    const char *host = nullptr;
    const gt_file *file = nullptr;
};

// Threads that stage a chunk's frames (copies or file reads).
unsigned io_threads() {
    const unsigned hw = std::thread::hardware_concurrency();
    return std::max(1u, std::min(8u, hw ? hw : 4u));
}

gt_status execute_frames(gt_context *c, const FrameSrc &src, const uint64_t *keys, uint64_t nkeys,
                         const float *points, uint64_t npoints, uint32_t point_cols,
                         int32_t points_has_frames, int32_t points_has_r,
                         const uint64_t *linker_reset_points, uint64_t nreset, gt_output *out) {
    const int32_t dtype = src.dtype;
    const uint64_t nframes = src.nframes;
    if (c->ring_count)
        return fail(GT_ERR_INVALID_ARGUMENT, "frames submitted with gt_submit_frame are in flight");
    c->interrupt.store(0);
    gt_status st = alloc_execute_buffers(c);
    if (st != GT_OK) return st;

    gt_host_plan plan;
    std::string why;
    gt_status pst = gt_host_plan_frames(c->characterize_points, nframes, keys, nkeys, points,
                                        npoints, point_cols, points_has_frames, points_has_r,
                                        &plan, &why);
    if (pst != GT_OK) return fail(pst, "%s", why.c_str());
    if (plan.frames.empty()) return fail(GT_ERR_EMPTY_ITERATOR, "the provided video doesn't contain any frames.");
    if (c->characterize_points && !points)
        return fail(GT_ERR_INVALID_ARGUMENT, "context was created for characterize_points but no points were given");
    if (!c->characterize_points && points)
        return fail(GT_ERR_INVALID_ARGUMENT, "points given to a context created without characterize_points");

    const size_t HW = static_cast<size_t>(c->H) * c->W;
    const size_t esz = dtype_size(dtype);
    gt_host_post post(&c->params, &c->der, c->H, c->W, linker_reset_points, nreset);
    if (!post.ok()) return fail(GT_ERR_INVALID_ARGUMENT, "%s", post.error().c_str());
    c->prog_total.store(plan.frames.size());
    c->prog_cur.store(0);

    // Host post-processing (filter_close / pruning, doughnut, linking) runs
    // on its own thread, frame by frame in order, as the reference's
    // post-process thread (execute_gpu.rs:914-987), while this thread keeps
    // the GPU busy: chunk k+1 is submitted before chunk k's detections are
    // read back and handed over.
    struct Item {
        uint64_t frame;
        std::vector<gt_detection> rows;
        std::vector<float> post;   // GPU filter_close / doughnut results, 4 per row
        const char *raw;
        std::vector<char> rawbuf;  // a file frame's copy when the host needs it
    };
    const bool keep_raw = post.needs_raw() && !c->gpu_post;
    std::deque<Item> queue;
    std::mutex mu;
    std::condition_variable cv_item, cv_space;
    bool producer_done = false;
    gt_status post_status = GT_OK;
    std::string post_err;
    const size_t max_queued = 4 * static_cast<size_t>(c->xchunk) + 8;
    std::thread worker([&]() {
        std::vector<float> rawf;
        for (;;) {
            Item it;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_item.wait(lk, [&] { return !queue.empty() || producer_done; });
                if (queue.empty()) return;
                it = std::move(queue.front());
                queue.pop_front();
            }
            cv_space.notify_one();
            bool failed;
            {
                std::lock_guard<std::mutex> lk(mu);
                failed = post_status != GT_OK;
            }
            if (failed) continue;   // drain without work after an error
            const float *rawp = nullptr;
            if (keep_raw) {
                rawf.resize(HW);
                gt_host_widen(it.raw, dtype, HW, rawf.data());
                rawp = rawf.data();
            }
            std::string perr;
            gt_status ps = post.process(it.rows, it.frame, rawp, &c->interrupt, &perr,
                                        c->gpu_post ? it.post.data() : nullptr);
            if (ps != GT_OK) {
                std::lock_guard<std::mutex> lk(mu);
                post_status = ps;
                post_err = perr;
            }
            c->prog_cur.fetch_add(1);
        }
    });
    auto post_failed = [&]() {
        std::lock_guard<std::mutex> lk(mu);
        return post_status != GT_OK;
    };
    auto finish_worker = [&]() {
        {
            std::lock_guard<std::mutex> lk(mu);
            producer_done = true;
        }
        cv_item.notify_all();
        worker.join();
    };

    const size_t nplan = plan.frames.size();
    const size_t nchunks = (nplan + c->xchunk - 1) / c->xchunk;
    // stage frames idx[0..nf) into slot sl's pinned buffer: copies (host
    // array) or decodes (file), frame f by thread f % nt
    auto stage = [&](gt_context::Slot &sl, const uint64_t *idx, uint32_t nf) -> gt_status {
        const unsigned nt = std::min<unsigned>(io_threads(), nf);
        std::vector<gt_status> sts(nt, GT_OK);
        std::vector<std::string> whys(nt);
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nt; ++t)
            th.emplace_back([&, t]() {
                for (uint32_t f = t; f < nf; f += nt) {
                    char *dst = static_cast<char *>(sl.h_stage) + esz * HW * f;
                    if (src.host) {
                        std::memcpy(dst, src.host + esz * HW * idx[f], esz * HW);
                        continue;
                    }
                    const gt_status rs = gt_io_read(src.file, idx[f], dst, &whys[t]);
                    if (rs != GT_OK) {
                        sts[t] = rs;
                        return;
                    }
                }
            });
        for (auto &t : th) t.join();
        for (unsigned t = 0; t < nt; ++t)
            if (sts[t] != GT_OK) return fail(sts[t], "%s", whys[t].c_str());
        return GT_OK;
    };
    if (c->der.illum_sigma > 0.0f && !c->der.illum_per_frame) {
        // the whole video's mean frame, smoothed, before any detection
        // (execute_provider, src/execute_gpu.rs:1324-1330): every frame of the
        // source, whatever the keys
        auto profile = [&]() -> gt_status {
            gt_status ist = illum_begin(c, c->stream);
            std::vector<uint64_t> idx(c->xchunk);
            for (uint64_t f0 = 0; f0 < nframes && ist == GT_OK; f0 += c->xchunk) {
                const uint32_t nf = static_cast<uint32_t>(std::min<uint64_t>(c->xchunk, nframes - f0));
                for (uint32_t f = 0; f < nf; ++f) idx[f] = f0 + f;
                auto &sl = c->slot[0];
                GT_HIP(hipStreamSynchronize(c->stream));   // the slot's last copy is done
                ist = stage(sl, idx.data(), nf);
                if (ist != GT_OK) return ist;
                GT_HIP(hipMemcpyAsync(sl.d_frames, sl.h_stage, esz * HW * nf,
                                      hipMemcpyHostToDevice, c->stream));
                ist = illum_add(c, sl.d_frames, dtype, nf, c->stream);
            }
            if (ist == GT_OK) ist = illum_finish(c, nframes, c->stream);
            GT_HIP(hipStreamSynchronize(c->stream));
            return ist;
        };
        const gt_status ist = profile();
        if (ist != GT_OK) {
            finish_worker();
            return ist;
        }
    }
    // stage chunk k's frames in slot k % 2's pinned buffer (host work)
    auto stage_chunk = [&](size_t k) -> gt_status {
        auto &sl = c->slot[k % 2];
        const size_t base = k * c->xchunk;
        const uint32_t nf = static_cast<uint32_t>(std::min<size_t>(c->xchunk, nplan - base));
        for (uint32_t f = 0; f < nf; ++f) {
            const uint64_t idx = plan.frames[base + f];
            if (idx >= nframes)
                return fail(GT_ERR_FRAME_OUT_OF_BOUNDS,
                            "Requested a frame outside the length\n of the video. Video length is %llu, requested frame %llu",
                            static_cast<unsigned long long>(nframes), static_cast<unsigned long long>(idx));
        }
        // the slot's previous chunk has been drained
        return stage(sl, plan.frames.data() + base, nf);
    };
    // queue chunk k's copy to the device and its GPU work
    auto submit = [&](size_t k) -> gt_status {
        auto &sl = c->slot[k % 2];
        const size_t base = k * c->xchunk;
        const uint32_t nf = static_cast<uint32_t>(std::min<size_t>(c->xchunk, nplan - base));
        GT_HIP(hipMemcpyAsync(sl.d_frames, sl.h_stage, esz * HW * nf, hipMemcpyHostToDevice, c->stream));
        gt_status st2;
        if (c->characterize_points)
            st2 = gt_host_characterize_points(c, k % 2, base, nf, dtype, plan);
        else
            st2 = run_pipeline(c, sl.d_frames, dtype, nf, sl.d_rows, c->det_cap, sl.d_counts, c->stream);
        if (st2 != GT_OK) return st2;
        if (c->gpu_post) {
            float4 *dp = c->d_post[k % 2];
            GT_HIP(gt_launch_filter_close_rows(sl.d_rows, c->det_cap, sl.d_counts, nf,
                                               c->der.filter_close ? static_cast<float>(c->der.separation) : 0.0f,
                                               dp, c->stream));
            if (c->der.doughnut)
                GT_HIP(gt_launch_doughnut_rows(sl.d_frames, dtype, static_cast<int>(c->H),
                                               static_cast<int>(c->W), sl.d_rows, c->det_cap,
                                               sl.d_counts, nf, c->d_dough_offs, c->dough_nsig,
                                               c->dough_nring, dp, c->stream));
        }
        // the chunk's rows (and post values) one frame after another, read
        // back in one copy each by drain
        GT_HIP(gt_launch_pack_rows(sl.d_rows, c->det_cap, sl.d_counts, nf,
                                   c->gpu_post ? c->d_post[k % 2] : nullptr, sl.d_packed,
                                   sl.d_packed_post, c->stream));
        // read back as many rows as recent chunks held, in stream order (a
        // copy on another stream would queue behind the next chunk's upload)
        sl.pk_copied = std::min<size_t>(c->pk_guess, static_cast<size_t>(c->det_cap) * nf);
        {
            gt_status rs = reserve_slot_rows(c, sl, sl.pk_copied, 0);
            if (rs != GT_OK) return rs;
        }
        GT_HIP(hipMemcpyAsync(sl.h_pk, sl.d_packed, sizeof(gt_detection) * sl.pk_copied,
                              hipMemcpyDeviceToHost, c->stream));
        if (c->gpu_post)
            GT_HIP(hipMemcpyAsync(sl.h_pk_post, sl.d_packed_post, sizeof(float4) * sl.pk_copied,
                                  hipMemcpyDeviceToHost, c->stream));
        GT_HIP(hipMemcpyAsync(sl.h_counts, sl.d_counts, sizeof(uint32_t) * nf, hipMemcpyDeviceToHost, c->stream));
        GT_HIP(hipEventRecord(sl.done, c->stream));
        return GT_OK;
    };
    // read chunk k's detections back and hand them to the post thread
    auto drain = [&](size_t k) -> gt_status {
        auto &sl = c->slot[k % 2];
        const size_t base = k * c->xchunk;
        const uint32_t nf = static_cast<uint32_t>(std::min<size_t>(c->xchunk, nplan - base));
        GT_HIP(hipEventSynchronize(sl.done));
        size_t total = 0;
        for (uint32_t f = 0; f < nf; ++f) {
            const uint32_t n = sl.h_counts[f];
            if (n > c->det_cap)
                return fail(GT_ERR_CAPACITY, "frame %llu produced %u detections, above the capacity of %u",
                            static_cast<unsigned long long>(plan.frames[base + f]), n, c->det_cap);
            total += n;
        }
        if (total > sl.pk_copied) {
            // more rows than the copy queued with the chunk: the rest now
            gt_status rs = reserve_slot_rows(c, sl, total, sl.pk_copied);
            if (rs != GT_OK) return rs;
            const size_t rest = total - sl.pk_copied;
            GT_HIP(hipMemcpyAsync(sl.h_pk + sl.pk_copied, sl.d_packed + sl.pk_copied,
                                  sizeof(gt_detection) * rest, hipMemcpyDeviceToHost, c->copy_stream));
            if (c->gpu_post)
                GT_HIP(hipMemcpyAsync(sl.h_pk_post + sl.pk_copied, sl.d_packed_post + sl.pk_copied,
                                      sizeof(float4) * rest, hipMemcpyDeviceToHost, c->copy_stream));
            GT_HIP(hipStreamSynchronize(c->copy_stream));
        }
        c->pk_guess = std::max(c->pk_guess, total + total / 4);
        const float *pk_post = reinterpret_cast<const float *>(sl.h_pk_post);
        size_t off = 0;
        for (uint32_t f = 0; f < nf; ++f) {
            const uint32_t n = sl.h_counts[f];
            Item it;
            it.frame = plan.frames[base + f];
            it.rows.assign(sl.h_pk + off, sl.h_pk + off + n);
            if (c->gpu_post) it.post.assign(pk_post + 4 * off, pk_post + 4 * (off + n));
            if (src.host) {
                it.raw = src.host + esz * HW * it.frame;
            } else if (keep_raw) {
                const char *st = static_cast<const char *>(sl.h_stage) + esz * HW * f;
                it.rawbuf.assign(st, st + esz * HW);
                it.raw = it.rawbuf.data();
            } else {
                it.raw = nullptr;
            }
            off += n;
            std::unique_lock<std::mutex> lk(mu);
            cv_space.wait(lk, [&] { return queue.size() < max_queued || post_status != GT_OK; });
            queue.push_back(std::move(it));
            lk.unlock();
            cv_item.notify_one();
        }
        return GT_OK;
    };

    gt_status lst = GT_OK;
    if (nchunks) lst = stage_chunk(0);
    if (nchunks && lst == GT_OK) lst = submit(0);
    for (size_t k = 0; k < nchunks && lst == GT_OK; ++k) {
        if (c->interrupt.load()) {
            lst = fail(GT_ERR_INTERRUPTED, "Keyboard Interrupt, my custom version");
            break;
        }
        if (post_failed()) break;
        // chunk k+1 is staged on the host and queued while the GPU works on
        // chunk k (slot (k+1) % 2 last held chunk k-1, drained in the
        // previous pass)
        if (k + 1 < nchunks) {
            lst = stage_chunk(k + 1);
            if (lst == GT_OK) lst = submit(k + 1);
            if (lst != GT_OK) break;
        }
        lst = drain(k);
    }
    // leave no work in flight on the slots before returning
    (void)hipStreamSynchronize(c->stream);
    finish_worker();
    if (lst != GT_OK) return lst;
    if (post_status != GT_OK) return fail(post_status, "%s", post_err.c_str());
    std::string ferr;
    gt_status fst = post.finish(out, &ferr);
    if (fst != GT_OK) return fail(fst, "%s", ferr.c_str());
    return GT_OK;
}

}  // namespace

extern "C" {

gt_status gt_execute(gt_context *c, const void *frames, int32_t dtype, uint64_t nframes,
                     const uint64_t *keys, uint64_t nkeys, const float *points, uint64_t npoints,
                     uint32_t point_cols, int32_t points_has_frames, int32_t points_has_r,
                     const uint64_t *linker_reset_points, uint64_t nreset, gt_output *out) {
    if (!c || !out) return fail(GT_ERR_INVALID_ARGUMENT, "null argument");
    out->data = nullptr;
    out->nrows = 0;
    out->ncols = 0;
    if (!dtype_size(dtype)) return fail(GT_ERR_INVALID_ARGUMENT, "unknown dtype %d", dtype);
    if (nframes > 0 && !frames) return fail(GT_ERR_INVALID_ARGUMENT, "null frames");
    FrameSrc src;
    src.dtype = dtype;
    src.nframes = nframes;
    src.host = static_cast<const char *>(frames);
    return execute_frames(c, src, keys, nkeys, points, npoints, point_cols, points_has_frames,
                          points_has_r, linker_reset_points, nreset, out);
}

gt_status gt_execute_file(gt_context *c, const gt_file *f, const uint64_t *keys, uint64_t nkeys,
                          const float *points, uint64_t npoints, uint32_t point_cols,
                          int32_t points_has_frames, int32_t points_has_r,
                          const uint64_t *linker_reset_points, uint64_t nreset, gt_output *out) {
    if (!c || !f || !out) return fail(GT_ERR_INVALID_ARGUMENT, "null argument");
    out->data = nullptr;
    out->nrows = 0;
    out->ncols = 0;
    uint32_t H, W;
    gt_io_dims(f, &H, &W);
    if (H != c->H || W != c->W)
        return fail(GT_ERR_DIMENSION_MISMATCH,
                    "the file's frames are %ux%u but the context was created for %ux%u", H, W,
                    c->H, c->W);
    FrameSrc src;
    src.dtype = gt_io_dtype(f);
    src.nframes = gt_io_nframes(f);
    src.file = f;
    return execute_frames(c, src, keys, nkeys, points, npoints, point_cols, points_has_frames,
                          points_has_r, linker_reset_points, nreset, out);
}

gt_status gt_file_open(const char *path, int32_t has_channel, uint32_t channel, gt_file **out) {
    if (!path || !out) return fail(GT_ERR_INVALID_ARGUMENT, "null argument");
    std::string why;
    const gt_status st = gt_io_open(path, has_channel, channel, out, &why);
    if (st != GT_OK) return fail(st, "%s", why.c_str());
    return GT_OK;
}

void gt_file_close(gt_file *f) { gt_io_close(f); }

gt_status gt_file_info(const gt_file *f, uint32_t *nrows, uint32_t *ncols, uint64_t *nframes,
                       int32_t *dtype) {
    if (!f) return fail(GT_ERR_INVALID_ARGUMENT, "null file");
    uint32_t H, W;
    gt_io_dims(f, &H, &W);
    if (nrows) *nrows = H;
    if (ncols) *ncols = W;
    if (nframes) *nframes = gt_io_nframes(f);
    if (dtype) *dtype = gt_io_dtype(f);
    return GT_OK;
}

gt_status gt_file_read(const gt_file *f, uint64_t first, uint64_t n, void *dst) {
    if (!f || (n && !dst)) return fail(GT_ERR_INVALID_ARGUMENT, "null argument");
    uint32_t H, W;
    gt_io_dims(f, &H, &W);
    const size_t bytes = dtype_size(gt_io_dtype(f)) * static_cast<size_t>(H) * W;
    std::string why;
    for (uint64_t i = 0; i < n; ++i) {
        const gt_status st = gt_io_read(f, first + i, static_cast<char *>(dst) + bytes * i, &why);
        if (st != GT_OK) return fail(st, "%s", why.c_str());
    }
    return GT_OK;
}

uint32_t gt_file_channels(const gt_file *f, uint32_t *channels, uint32_t cap) {
    if (!f) return 0;
    const std::vector<uint32_t> v = gt_io_channels(f);
    for (size_t i = 0; i < v.size() && i < cap; ++i) channels[i] = v[i];
    return static_cast<uint32_t>(v.size());
}

gt_status gt_file_mean(const gt_file *f, float *out) {
    if (!f || !out) return fail(GT_ERR_INVALID_ARGUMENT, "null argument");
    uint32_t H, W;
    gt_io_dims(f, &H, &W);
    const size_t HW = static_cast<size_t>(H) * W;
    const int32_t dt = gt_io_dtype(f);
    const size_t esz = dtype_size(dt);
    const uint64_t n = gt_io_nframes(f);
    std::vector<char> fr(esz * HW);
    std::vector<float> wide(HW);
    std::fill(out, out + HW, 0.0f);
    std::string why;
    for (uint64_t t = 0; t < n; ++t) {
        const gt_status st = gt_io_read(f, t, fr.data(), &why);
        if (st != GT_OK) return fail(st, "%s", why.c_str());
        gt_host_widen(fr.data(), dt, HW, wide.data());
        for (size_t i = 0; i < HW; ++i) out[i] += wide[i];
    }
    const float nf = static_cast<float>(n);
    for (size_t i = 0; i < HW; ++i) out[i] /= nf;
    return GT_OK;
}

gt_status gt_write_tiff(const char *path, const float *frames, uint64_t nframes, uint32_t nrows,
                        uint32_t ncols) {
    if (!path || (nframes && !frames)) return fail(GT_ERR_INVALID_ARGUMENT, "null argument");
    std::string why;
    const gt_status st = gt_io_write_tiff_u16(path, frames, nframes, nrows, ncols, &why);
    if (st != GT_OK) return fail(st, "%s", why.c_str());
    return GT_OK;
}

gt_status gt_frame_std(gt_context *c, const void *d_frames, int32_t dtype, uint32_t nframes,
                       float *d_std, void *stream) {
    if (!c) return fail(GT_ERR_INVALID_ARGUMENT, "null context");
    if (!dtype_size(dtype)) return fail(GT_ERR_INVALID_ARGUMENT, "unknown dtype %d", dtype);
    if (nframes == 0) return fail(GT_ERR_EMPTY_ITERATOR, "the provided video doesn't contain any frames.");
    if (!d_frames || !d_std) return fail(GT_ERR_INVALID_ARGUMENT, "null argument");
    const size_t HW = static_cast<size_t>(c->H) * c->W;
    if (HW <= 24) return fail(GT_ERR_INVALID_ARGUMENT, "the noise estimate needs more than 24 pixels");
    gt_status st = alloc_std(c);
    if (st != GT_OK) return st;
    hipStream_t s = static_cast<hipStream_t>(stream);  // NULL: the default stream
    const size_t esz = dtype_size(dtype);
    for (uint32_t f0 = 0; f0 < nframes; f0 += c->chunk) {
        const uint32_t nf = std::min(c->chunk, nframes - f0);
        GT_HIP(gt_launch_frame_std(static_cast<const char *>(d_frames) + esz * HW * f0, dtype, nf,
                                   HW, c->d_std_tmp, c->d_mean, d_std + f0, s));
    }
    return GT_OK;
}

gt_status gt_context_set_cnn(gt_context *c, const gt_cnn *net, float conf_logit, float tol,
                             int32_t scope) {
    if (!c) return fail(GT_ERR_INVALID_ARGUMENT, "null context");
    if (!net) {
        c->cnn = nullptr;
        return GT_OK;
    }
    if (scope != GT_CNN_MOVING && scope != GT_CNN_ALL)
        return fail(GT_ERR_INVALID_ARGUMENT, "unknown proposal scope %d", scope);
    if (!c->tiled || c->log || c->der.adaptive)
        return fail(GT_ERR_INVALID_ARGUMENT,
                    "the learned proposal runs in the tiled Trackpy-style path only "
                    "(without adaptive_background)");
    if (!std::isfinite(conf_logit) && conf_logit != INFINITY)
        return fail(GT_ERR_INVALID_ARGUMENT, "confidence threshold must not be NaN or -inf");
    if (!(tol >= 0.0f) || !std::isfinite(tol))
        return fail(GT_ERR_INVALID_ARGUMENT, "tolerance must be finite and >= 0");
    GT_HIP(hipStreamSynchronize(c->stream));
    const size_t n = static_cast<size_t>(c->chunk) * c->cand_cap;
    if (!c->d_cand) GT_HIP(hipMalloc(&c->d_cand, sizeof(uint32_t) * n));
    if (!c->d_props) GT_HIP(hipMalloc(&c->d_props, sizeof(float4) * n));
    if (!c->d_stats) {
        GT_HIP(hipMalloc(&c->d_stats, 2 * sizeof(unsigned long long)));
        GT_HIP(hipMemset(c->d_stats, 0, 2 * sizeof(unsigned long long)));
    }
    c->cnn = net;
    c->cnn_conf = conf_logit;
    c->cnn_tol = tol;
    c->cnn_scope = scope;
    return GT_OK;
}

gt_status gt_context_cnn_stats(gt_context *c, uint64_t *accepted, uint64_t *candidates,
                               int32_t reset) {
    if (!c || !accepted || !candidates) return fail(GT_ERR_INVALID_ARGUMENT, "null argument");
    unsigned long long h[2] = {0, 0};
    if (c->d_stats) {
        GT_HIP(hipDeviceSynchronize());
        GT_HIP(hipMemcpy(h, c->d_stats, sizeof(h), hipMemcpyDeviceToHost));
        if (reset) GT_HIP(hipMemset(c->d_stats, 0, sizeof(h)));
    }
    *accepted = h[0];
    *candidates = h[1];
    return GT_OK;
}

gt_status gt_cnn_make_samples(gt_context *c, const void *d_frames, int32_t dtype,
                              uint32_t nframes, uint16_t *d_patches, float *d_targets,
                              uint64_t cap, uint64_t *n_written, uint64_t *n_total, void *stream) {
    if (!c) return fail(GT_ERR_INVALID_ARGUMENT, "null context");
    if (!dtype_size(dtype)) return fail(GT_ERR_INVALID_ARGUMENT, "unknown dtype %d", dtype);
    if (nframes == 0) return fail(GT_ERR_EMPTY_ITERATOR, "the provided video doesn't contain any frames.");
    if (!d_frames || !d_patches || !d_targets || !n_written)
        return fail(GT_ERR_INVALID_ARGUMENT, "null argument");
    if (!c->tiled || c->log)
        return fail(GT_ERR_INVALID_ARGUMENT,
                    "training samples come from the tiled Trackpy-style path only");
    hipStream_t s = static_cast<hipStream_t>(stream);  // NULL: the default stream
    const size_t HW = static_cast<size_t>(c->H) * c->W;
    if (!c->d_cand) GT_HIP(hipMalloc(&c->d_cand, sizeof(uint32_t) * c->chunk * static_cast<size_t>(c->cand_cap)));
    if (!c->d_offs) GT_HIP(hipMalloc(&c->d_offs, sizeof(uint32_t) * (c->chunk + 2)));
    uint32_t *d_total = c->d_offs + c->chunk + 1;
    GT_HIP(gt_launch_zero_u32(d_total, 1, s));
    const uint32_t maxn = static_cast<uint32_t>(std::min<uint64_t>(cap, 0xFFFFFFFFull));
    const size_t esz = dtype_size(dtype);
    GtKParams kw = c->kp;
    kw.characterize = 0;
    for (uint32_t f0 = 0; f0 < nframes; f0 += c->chunk) {
        const uint32_t nf = std::min(c->chunk, nframes - f0);
        const char *frames = static_cast<const char *>(d_frames) + esz * HW * f0;
        GT_HIP(gt_launch_zero_u32(c->d_ccount, static_cast<size_t>(nf) * GT_COUNTER_STRIDE, s));
        if (c->der.need_std)
            GT_HIP(gt_launch_frame_std(frames, dtype, nf, HW, c->d_std_tmp, c->d_mean, c->d_std, s));
        GT_HIP(gt_launch_bandpass(frames, dtype, c->d_proc, nf, c->kp, s));
        GT_HIP(gt_launch_detect_cand(c->d_proc, dtype, nf, kw, c->d_cand, c->cand_cap, c->d_ccount, s));
        GT_HIP(gt_cnn_launch_sample_offsets(c->d_ccount, c->cand_cap, nf, d_total, c->d_offs, s));
        GT_HIP(gt_cnn_launch_gather(c->d_proc, static_cast<int>(c->H), static_cast<int>(c->W), nf,
                                    c->d_cand, c->cand_cap, c->d_ccount, c->d_offs, maxn,
                                    d_patches, s));
        GT_HIP(gt_launch_cnn_label(c->d_proc, nf, kw, c->d_cand, c->cand_cap, c->d_ccount,
                                   c->d_offs, maxn, reinterpret_cast<float4 *>(d_targets), s));
    }
    uint32_t total = 0;
    GT_HIP(hipMemcpyAsync(&total, d_total, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    GT_HIP(hipStreamSynchronize(s));
    *n_written = std::min<uint64_t>(total, maxn);
    if (n_total) *n_total = total;
    return GT_OK;
}

}  // extern "C"

// characterize_points for one chunk (submit_work's characterization branch,
// src/execute_gpu.rs:190-229, which drops max_rows/extract_max/walk from the
// pass order, src/gpu_setup.rs:877-891): the chunk's frames are already on the
// device; upload each frame's positions as rows, bandpass, then characterize
// the rows in place.  Row seeds are the positions' order within the frame.
gt_status gt_host_characterize_points(gt_context *c, int slot, size_t base, uint32_t nf,
                                      int32_t dtype, const gt_host_plan &plan) {
    auto &sl = c->slot[slot];
    std::vector<gt_detection> staged;
    std::vector<uint32_t> counts(nf, 0);
    uint32_t max_count = 0;
    for (uint32_t f = 0; f < nf; ++f) {
        const size_t b = plan.pt_begin[base + f], e = plan.pt_end[base + f];
        const size_t n = e - b;
        if (n > c->det_cap)
            return fail(GT_ERR_CAPACITY, "%zu positions for one frame exceed the capacity of %u", n,
                        c->det_cap);
        counts[f] = static_cast<uint32_t>(n);
        max_count = std::max<uint32_t>(max_count, static_cast<uint32_t>(n));
        staged.resize(n);
        for (size_t i = 0; i < n; ++i) {
            staged[i].row = plan.pts[b + i];
            staged[i].seed = static_cast<uint32_t>(i);
            staged[i].frame = static_cast<uint32_t>(plan.frames[base + f]);
        }
        if (n)
            GT_HIP(hipMemcpyAsync(sl.d_rows + static_cast<size_t>(f) * c->det_cap, staged.data(),
                                  sizeof(gt_detection) * n, hipMemcpyHostToDevice, c->stream));
        // the staging vector is reused: wait for this copy before refilling it
        GT_HIP(hipStreamSynchronize(c->stream));
    }
    std::memcpy(sl.h_counts, counts.data(), sizeof(uint32_t) * nf);
    GT_HIP(hipMemcpyAsync(sl.d_counts, sl.h_counts, sizeof(uint32_t) * nf, hipMemcpyHostToDevice,
                          c->stream));
    const void *frames = sl.d_frames;
    if (c->der.illum_sigma > 0.0f) {
        gt_status ist = illum_correct(c, frames, dtype, nf, c->stream);
        if (ist != GT_OK) return ist;
        frames = c->d_corr;
        dtype = GT_F32;
    }
    GT_HIP(gt_launch_bandpass(frames, dtype, c->d_proc, nf, c->kp, c->stream));
    GT_HIP(gt_launch_characterize_rows(c->d_proc, frames, dtype, nf, c->kp, sl.d_rows,
                                       c->det_cap, sl.d_counts, max_count, c->stream));
    return GT_OK;
}
