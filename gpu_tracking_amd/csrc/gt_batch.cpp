// gt_batch.cpp -- batched detection over frames of different sizes.
//
// The reference's GpuState serves one frame size (src/execute_gpu.rs:1074-1080);
// a batch context instead takes any mix of frame sizes (same detection
// parameters).  A call's frames are split into launches (a pixel budget and a
// frame-count limit); every launch is one kernel per stage over the launch's
// frame table (grid y = frame, grid x = tile of that frame).  All launch
// tables of a call go to the device in one copy; the stage sequence, framed by
// a kernel that selects the next launch's table on the device and one that
// advances a device launch counter, is captured once into a HIP graph (again
// only when the output buffers or grid sizes change) and replayed once per
// launch with no host work in between.  gt_batch_tune picks the launch size
// and the slow-walk / characterize grid by timing candidates.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "gt_host.h"
#include "gt_internal.h"

// from gt_runtime.cpp
gt_status gt_runtime_fail(gt_status st, const char *fmt, ...);

#define GB_HIP(call)                                                                             \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return gt_runtime_fail(GT_ERR_GPU_DEVICE, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

struct gt_batch {
    gt_params params;
    gt_derived der;
    GtKParams kp;
    hipStream_t stream = nullptr;
    uint64_t max_frame_pixels = 0, max_launch_pixels = 0;
    uint32_t max_slots = 0;
    float *d_proc = nullptr;          // max_launch_pixels
    uint32_t *d_ocount = nullptr;     // max_slots * GT_COUNTER_STRIDE
    uint32_t *d_scount = nullptr;
    uint2 *d_slow = nullptr;          // max_slots * slow_cap
    uint32_t slow_cap = 0;
    GtBatchHeader *d_hdr = nullptr;   // the running launch's table ...
    GtBatchFrame *d_fs = nullptr;     // ... max_slots entries
    // a call's launch tables: all entries, the launches, the launch counter
    GtBatchFrame *d_all = nullptr;
    GtBatchLaunch *d_launches = nullptr;
    uint32_t *d_cur = nullptr;
    size_t all_cap = 0, launch_cap = 0;
    char *h_stage = nullptr;          // pinned staging of the tables (+ the counter's zero)
    size_t h_cap = 0;
    hipEvent_t stage_free = nullptr;  // the last upload from h_stage has completed
    hipEvent_t ev_in = nullptr, ev_out = nullptr;  // ordering against the caller's stream
    uint32_t walk_blocks = 4;         // blocks per frame for slow walks and characterize
    uint64_t launch_px = 0;           // pixels per launch in use (<= max_launch_pixels)
    int32_t use_graph = 1;
    uint64_t launches = 0;            // launches of the last call
    // the captured stage sequence, kept while the call's arguments repeat
    hipGraphExec_t gexec = nullptr;
    struct GraphKey {
        gt_detection *out;
        uint32_t cap;
        uint32_t *counts;
        int32_t dtype;
        GtBatchFrame *all;
        GtBatchLaunch *launches;
        uint32_t bp_blocks, dt_blocks, walk_blocks, launch_slots;
        bool operator==(const GraphKey &o) const {
            return out == o.out && cap == o.cap && counts == o.counts && dtype == o.dtype &&
                   all == o.all && launches == o.launches && bp_blocks == o.bp_blocks &&
                   dt_blocks == o.dt_blocks && walk_blocks == o.walk_blocks &&
                   launch_slots == o.launch_slots;
        }
    } gkey{};
};

namespace {

void free_batch(gt_batch *b) {
    if (b->d_proc) (void)hipFree(b->d_proc);
    if (b->d_ocount) (void)hipFree(b->d_ocount);
    if (b->d_scount) (void)hipFree(b->d_scount);
    if (b->d_slow) (void)hipFree(b->d_slow);
    if (b->d_hdr) (void)hipFree(b->d_hdr);
    if (b->d_fs) (void)hipFree(b->d_fs);
    if (b->d_all) (void)hipFree(b->d_all);
    if (b->d_launches) (void)hipFree(b->d_launches);
    if (b->d_cur) (void)hipFree(b->d_cur);
    if (b->h_stage) (void)hipHostFree(b->h_stage);
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->stage_free) (void)hipEventDestroy(b->stage_free);
    if (b->ev_in) (void)hipEventDestroy(b->ev_in);
    if (b->ev_out) (void)hipEventDestroy(b->ev_out);
    if (b->stream) (void)hipStreamDestroy(b->stream);
}

// Splits frames[0..n) into launches: consecutive frames while the frame
// limit and the pixel budget last.  Returns the end index of the launch
// starting at i.
uint32_t launch_end(const gt_batch *b, const gt_frame_desc *frames, uint32_t n, uint32_t i,
                    uint64_t launch_px) {
    uint64_t px = 0;
    uint32_t j = i;
    while (j < n && j - i < b->max_slots) {
        const uint64_t hw = static_cast<uint64_t>(frames[j].nrows) * frames[j].ncols;
        if (j > i && px + hw > launch_px) break;
        px += hw;
        ++j;
    }
    return j;
}

// Grows a device buffer to hold `need` elements (contents are not kept).
template <typename T>
gt_status grow(T **p, size_t *cap, size_t need) {
    if (need <= *cap) return GT_OK;
    const size_t c = std::max(need, 2 * *cap);
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    GB_HIP(hipMalloc(p, sizeof(T) * c));
    *cap = c;
    return GT_OK;
}

gt_status run_batch(gt_batch *b, const gt_frame_desc *frames, uint32_t n, int32_t dtype,
                    gt_detection *d_out, uint32_t cap, uint32_t *d_counts, hipStream_t s,
                    uint32_t walk_blocks, uint64_t launch_px, bool graph) {
    // the call's launches and their tables
    std::vector<GtBatchLaunch> launches;
    for (uint32_t i = 0; i < n;) {
        const uint32_t j = launch_end(b, frames, n, i, launch_px);
        launches.push_back(GtBatchLaunch{i, j - i});
        i = j;
    }
    const size_t nl = launches.size();
    // the pinned staging buffer and the device tables may still serve the
    // previous call's upload and launches
    GB_HIP(hipEventSynchronize(b->stage_free));
    const size_t bytes = sizeof(GtBatchFrame) * n + sizeof(GtBatchLaunch) * nl + sizeof(uint32_t);
    if (bytes > b->h_cap) {
        const size_t c = std::max(bytes, 2 * b->h_cap);
        if (b->h_stage) (void)hipHostFree(b->h_stage);
        b->h_stage = nullptr;
        b->h_cap = 0;
        GB_HIP(hipHostMalloc(&b->h_stage, c, hipHostMallocDefault));
        b->h_cap = c;
    }
    if (n > b->all_cap || nl > b->launch_cap) GB_HIP(hipStreamSynchronize(s));
    gt_status st = grow(&b->d_all, &b->all_cap, n);
    if (st == GT_OK) st = grow(&b->d_launches, &b->launch_cap, nl);
    if (st != GT_OK) return st;
    GtBatchFrame *e = reinterpret_cast<GtBatchFrame *>(b->h_stage);
    GtBatchLaunch *hl = reinterpret_cast<GtBatchLaunch *>(b->h_stage + sizeof(GtBatchFrame) * n);
    uint32_t *zero = reinterpret_cast<uint32_t *>(b->h_stage + sizeof(GtBatchFrame) * n +
                                                  sizeof(GtBatchLaunch) * nl);
    uint32_t maxbp = 1, maxdt = 1, maxslots = 1;
    for (const GtBatchLaunch &L : launches) {
        uint64_t off = 0;
        for (uint32_t f = L.first; f < L.first + L.nslots; ++f) {
            GtBatchFrame &x = e[f];
            x.in_ptr = reinterpret_cast<uint64_t>(frames[f].data);
            x.proc_off = off;
            x.H = static_cast<int32_t>(frames[f].nrows);
            x.W = static_cast<int32_t>(frames[f].ncols);
            x.out_idx = f;
            x.frame = frames[f].frame;
            off += static_cast<uint64_t>(x.H) * x.W;
            uint32_t tb, td;
            gt_batch_tiles(x.H, x.W, &tb, &td);
            maxbp = std::max(maxbp, tb);
            maxdt = std::max(maxdt, td);
        }
        maxslots = std::max(maxslots, L.nslots);
    }
    std::copy(launches.begin(), launches.end(), hl);
    *zero = 0;
    GB_HIP(hipMemcpyAsync(b->d_all, e, sizeof(GtBatchFrame) * n, hipMemcpyHostToDevice, s));
    GB_HIP(hipMemcpyAsync(b->d_launches, hl, sizeof(GtBatchLaunch) * nl, hipMemcpyHostToDevice, s));
    GB_HIP(hipMemcpyAsync(b->d_cur, zero, sizeof(uint32_t), hipMemcpyHostToDevice, s));
    // grids: x = tiles of the call's largest frame, y = frames of its largest launch
    GtBatchSchedule sch;
    sch.bp_blocks = (maxbp + 7) / 8 * 8;
    sch.dt_blocks = (maxdt + 7) / 8 * 8;
    sch.walk_blocks = walk_blocks;
    sch.launch_slots = maxslots;
    auto enqueue = [&]() {
        return gt_launch_batch(b->d_launches, b->d_cur, b->d_all, dtype, b->kp, b->d_hdr, b->d_fs,
                               b->max_slots, b->d_proc, d_out, cap, b->d_ocount, b->d_slow,
                               b->slow_cap, b->d_scount, d_counts, sch, s);
    };
    hipGraphExec_t gexec = nullptr;
    const gt_batch::GraphKey key{d_out,           cap,           d_counts,        dtype,
                                 b->d_all,        b->d_launches, sch.bp_blocks,   sch.dt_blocks,
                                 sch.walk_blocks, sch.launch_slots};
    if (graph && b->gexec && b->gkey == key) {
        gexec = b->gexec;
    } else if (graph) {
        if (b->gexec) {
            GB_HIP(hipStreamSynchronize(s));   // the old graph may still run
            (void)hipGraphExecDestroy(b->gexec);
            b->gexec = nullptr;
        }
        hipGraph_t g = nullptr;
        GB_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        hipError_t le = enqueue();
        hipError_t ce = hipStreamEndCapture(s, &g);
        if (le != hipSuccess || ce != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return gt_runtime_fail(GT_ERR_GPU_DEVICE, "graph capture of the batch stages failed: %s",
                                   hipGetErrorString(le != hipSuccess ? le : ce));
        }
        hipError_t ie = hipGraphInstantiate(&gexec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ie != hipSuccess)
            return gt_runtime_fail(GT_ERR_GPU_DEVICE, "hipGraphInstantiate failed: %s",
                                   hipGetErrorString(ie));
        b->gexec = gexec;
        b->gkey = key;
    }
    for (size_t k = 0; k < nl; ++k) {
        const hipError_t e1 = gexec ? hipGraphLaunch(gexec, s) : enqueue();
        if (e1 != hipSuccess)
            return gt_runtime_fail(GT_ERR_GPU_DEVICE, "batch launch failed: %s", hipGetErrorString(e1));
    }
    // the staging buffer and tables are free again once this call's launches are done
    GB_HIP(hipEventRecord(b->stage_free, s));
    b->launches = nl;
    return GT_OK;
}

gt_status check_frames(const gt_batch *b, const gt_frame_desc *frames, uint32_t n, int32_t dtype) {
    if (dtype != GT_U16 && dtype != GT_U8)
        return gt_runtime_fail(GT_ERR_INVALID_ARGUMENT, "batched detection takes u8 or u16 frames");
    for (uint32_t i = 0; i < n; ++i) {
        const uint64_t hw = static_cast<uint64_t>(frames[i].nrows) * frames[i].ncols;
        if (!frames[i].data || hw == 0)
            return gt_runtime_fail(GT_ERR_INVALID_ARGUMENT, "frame %u: null or empty", i);
        if (hw > b->max_frame_pixels || hw > b->max_launch_pixels)
            return gt_runtime_fail(GT_ERR_INVALID_ARGUMENT,
                                   "frame %u has %llu pixels, above the batch context's limit", i,
                                   static_cast<unsigned long long>(hw));
    }
    return GT_OK;
}

}  // namespace

extern "C" {

gt_status gt_batch_create(const gt_params *params, uint64_t max_frame_pixels,
                          uint64_t max_launch_pixels, uint32_t max_slots, gt_batch **out) {
    if (!params || !out) return gt_runtime_fail(GT_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (max_frame_pixels == 0 || max_launch_pixels < max_frame_pixels || max_slots == 0 ||
        max_slots > 65535 || max_frame_pixels > 0xFFFFFFFFull)
        return gt_runtime_fail(GT_ERR_INVALID_ARGUMENT,
                               "need 0 < max_frame_pixels <= max_launch_pixels and "
                               "0 < max_slots <= 65535");
    if (params->style != GT_STYLE_TRACKPY)
        return gt_runtime_fail(GT_ERR_INVALID_ARGUMENT, "batched detection runs the Trackpy style");
    if (params->has_illumination_sigma)
        return gt_runtime_fail(GT_ERR_INVALID_ARGUMENT,
                               "batched detection does not apply illumination correction (it "
                               "is defined per video)");
    if (params->has_snr || params->has_minmass_snr || params->has_adaptive_background)
        return gt_runtime_fail(GT_ERR_INVALID_ARGUMENT,
                               "batched detection does not apply the snr / minmass_snr rules "
                               "or adaptive_background");
    gt_derived der;
    std::string why;
    if (!gt_host_derive(params, 1, 1, &der, &why))
        return gt_runtime_fail(GT_ERR_INVALID_ARGUMENT, "%s", why.c_str());
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return gt_runtime_fail(GT_ERR_GPU_ADAPTER, "no HIP device is available");
    gt_batch *b = new (std::nothrow) gt_batch();
    if (!b) return gt_runtime_fail(GT_ERR_GPU_DEVICE, "out of host memory");
    b->params = *params;
    b->der = der;
    gt_host_fill_kparams(&der, 1, 1, &b->kp);
    if (b->kp.R < 1 || b->kp.R > 7 || !b->kp.fast_div || b->kp.k < 1 ||
        gt_tile_lds_bytes(b->kp) > 65536) {
        delete b;
        return gt_runtime_fail(GT_ERR_INVALID_ARGUMENT,
                               "batched detection needs smoothing_size <= 15 and a dilation the "
                               "tile kernel holds in LDS");
    }
    b->max_frame_pixels = max_frame_pixels;
    b->max_launch_pixels = max_launch_pixels;
    b->max_slots = max_slots;
    b->slow_cap = static_cast<uint32_t>(std::max<uint64_t>(1024, max_frame_pixels / 8));
    b->launch_px = max_launch_pixels;
    gt_status st = GT_OK;
    auto ok = [&](hipError_t e, const char *what) {
        if (e != hipSuccess && st == GT_OK)
            st = gt_runtime_fail(GT_ERR_GPU_DEVICE, "%s failed: %s", what, hipGetErrorString(e));
        return st == GT_OK;
    };
    (void)(ok(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "hipStreamCreate") &&
           ok(gt_init_tables(), "table init") &&
           ok(hipMalloc(&b->d_proc, sizeof(float) * max_launch_pixels), "hipMalloc") &&
           ok(hipMalloc(&b->d_ocount, sizeof(uint32_t) * GT_COUNTER_STRIDE * max_slots), "hipMalloc") &&
           ok(hipMalloc(&b->d_scount, sizeof(uint32_t) * GT_COUNTER_STRIDE * max_slots), "hipMalloc") &&
           ok(hipMalloc(&b->d_slow, sizeof(uint2) * static_cast<size_t>(b->slow_cap) * max_slots), "hipMalloc") &&
           ok(hipMalloc(&b->d_hdr, sizeof(GtBatchHeader)), "hipMalloc") &&
           ok(hipMalloc(&b->d_fs, sizeof(GtBatchFrame) * max_slots), "hipMalloc") &&
           ok(hipMalloc(&b->d_cur, sizeof(uint32_t)), "hipMalloc") &&
           ok(hipEventCreateWithFlags(&b->stage_free, hipEventDisableTiming), "hipEventCreate") &&
           ok(hipEventRecord(b->stage_free, b->stream), "hipEventRecord") &&
           ok(hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming), "hipEventCreate") &&
           ok(hipEventCreateWithFlags(&b->ev_out, hipEventDisableTiming), "hipEventCreate"));
    if (st != GT_OK) {
        free_batch(b);
        delete b;
        return st;
    }
    *out = b;
    return GT_OK;
}

void gt_batch_destroy(gt_batch *b) {
    if (!b) return;
    (void)hipStreamSynchronize(b->stream);
    free_batch(b);
    delete b;
}

gt_status gt_batch_detect(gt_batch *b, const gt_frame_desc *frames, uint32_t n, int32_t dtype,
                          gt_detection *d_out, uint32_t cap_per_frame, uint32_t *d_counts,
                          void *stream) {
    if (!b) return gt_runtime_fail(GT_ERR_INVALID_ARGUMENT, "null batch context");
    if (n == 0) return gt_runtime_fail(GT_ERR_EMPTY_ITERATOR, "the provided video doesn't contain any frames.");
    if (!frames || !d_out || !d_counts || cap_per_frame == 0)
        return gt_runtime_fail(GT_ERR_INVALID_ARGUMENT, "null argument or zero capacity");
    gt_status st = check_frames(b, frames, n, dtype);
    if (st != GT_OK) return st;
    // the work runs on the context's own stream (a graph cannot be captured on
    // a legacy default stream), ordered after and before the caller's stream
    // (NULL: the default stream)
    hipStream_t user = static_cast<hipStream_t>(stream);
    if (user != b->stream) {
        GB_HIP(hipEventRecord(b->ev_in, user));
        GB_HIP(hipStreamWaitEvent(b->stream, b->ev_in, 0));
    }
    st = run_batch(b, frames, n, dtype, d_out, cap_per_frame, d_counts, b->stream, b->walk_blocks,
                   b->launch_px, b->use_graph != 0);
    if (st != GT_OK) return st;
    if (user != b->stream) {
        GB_HIP(hipEventRecord(b->ev_out, b->stream));
        GB_HIP(hipStreamWaitEvent(user, b->ev_out, 0));
    }
    return GT_OK;
}

gt_status gt_batch_set_graph(gt_batch *b, int32_t enable) {
    if (!b) return gt_runtime_fail(GT_ERR_INVALID_ARGUMENT, "null batch context");
    b->use_graph = enable != 0;
    return GT_OK;
}

gt_status gt_batch_get_schedule(gt_batch *b, uint64_t *sched4, uint64_t *launches) {
    if (!b || !sched4) return gt_runtime_fail(GT_ERR_INVALID_ARGUMENT, "null argument");
    sched4[0] = b->launch_px;
    sched4[1] = b->walk_blocks;
    sched4[2] = b->max_slots;
    sched4[3] = b->max_launch_pixels;
    if (launches) *launches = b->launches;
    return GT_OK;
}

gt_status gt_batch_tune(gt_batch *b, const gt_frame_desc *frames, uint32_t n, int32_t dtype,
                        gt_detection *d_out, uint32_t cap_per_frame, uint32_t *d_counts,
                        float *best_ms) {
    // Coordinate search: first the pixels per launch, then the slow-walk /
    // characterize blocks per frame; each candidate runs the sample once to
    // warm up and once timed with events; the fastest value of each knob is
    // kept.
    if (!b) return gt_runtime_fail(GT_ERR_INVALID_ARGUMENT, "null batch context");
    if (n == 0 || !frames || !d_out || !d_counts || cap_per_frame == 0)
        return gt_runtime_fail(GT_ERR_INVALID_ARGUMENT, "empty sample or null argument");
    gt_status st = check_frames(b, frames, n, dtype);
    if (st != GT_OK) return st;
    uint64_t maxfr = 1;
    for (uint32_t i = 0; i < n; ++i)
        maxfr = std::max<uint64_t>(maxfr, static_cast<uint64_t>(frames[i].nrows) * frames[i].ncols);
    hipEvent_t e0, e1;
    GB_HIP(hipEventCreate(&e0));
    GB_HIP(hipEventCreate(&e1));
    auto time_it = [&](uint32_t wb, uint64_t lpx, float *ms) -> gt_status {
        gt_status s1 = run_batch(b, frames, n, dtype, d_out, cap_per_frame, d_counts, b->stream, wb,
                                 lpx, false);
        if (s1 != GT_OK) return s1;
        GB_HIP(hipEventRecord(e0, b->stream));
        s1 = run_batch(b, frames, n, dtype, d_out, cap_per_frame, d_counts, b->stream, wb, lpx,
                       false);
        if (s1 != GT_OK) return s1;
        GB_HIP(hipEventRecord(e1, b->stream));
        GB_HIP(hipEventSynchronize(e1));
        GB_HIP(hipEventElapsedTime(ms, e0, e1));
        return GT_OK;
    };
    uint32_t wb = 4;
    uint64_t lpx = b->max_launch_pixels;
    float best = 0.0f;
    st = time_it(wb, lpx, &best);
    for (int d = 1; d <= 3 && st == GT_OK; ++d) {
        const uint64_t c = b->max_launch_pixels >> d;
        if (c < maxfr) break;
        float ms = 0.0f;
        st = time_it(wb, c, &ms);
        if (st == GT_OK && ms < best) {
            best = ms;
            lpx = c;
        }
    }
    const uint32_t walks[] = {1, 2, 8, 16};
    for (int o = 0; o < 4 && st == GT_OK; ++o) {
        float ms = 0.0f;
        st = time_it(walks[o], lpx, &ms);
        if (st == GT_OK && ms < best) {
            best = ms;
            wb = walks[o];
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st != GT_OK) return st;
    b->walk_blocks = wb;
    b->launch_px = lpx;
    if (best_ms) *best_ms = best;
    return GT_OK;
}

}  // extern "C"
