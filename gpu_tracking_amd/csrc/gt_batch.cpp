// gt_batch.cpp -- batched detection over frames of different sizes.
//
// The reference's GpuState serves one frame size (src/execute_gpu.rs:1074-1080);
// a batch context instead takes any mix of frame sizes (same detection
// parameters) and runs each launch's frames through one persistent launch
// per stage, reading the launch's frame table from device memory.  The
// stage sequence of a launch is captured once into a HIP graph per call and
// replayed for every launch of the call; only the table changes between
// replays.  Grid sizes come from gt_batch_tune (timing candidates on sample
// frames).
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

namespace {

#define GB_HIP(call)                                                                             \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return gt_runtime_fail(GT_ERR_GPU_DEVICE, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

}  // namespace

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
    char *d_table = nullptr;          // header + max_slots entries
    char *h_table[2] = {nullptr, nullptr};  // pinned staging of the table
    hipEvent_t table_free[2] = {nullptr, nullptr};
    hipEvent_t ev_in = nullptr, ev_out = nullptr;  // ordering against the caller's stream
    GtBatchSchedule sched{};
    int32_t use_graph = 1;
    uint64_t launches = 0;            // launches of the last call
};

namespace {

size_t table_bytes(uint32_t slots) {
    return sizeof(GtBatchHeader) + sizeof(GtBatchFrame) * static_cast<size_t>(slots);
}

void free_batch(gt_batch *b) {
    if (b->d_proc) (void)hipFree(b->d_proc);
    if (b->d_ocount) (void)hipFree(b->d_ocount);
    if (b->d_scount) (void)hipFree(b->d_scount);
    if (b->d_slow) (void)hipFree(b->d_slow);
    if (b->d_table) (void)hipFree(b->d_table);
    for (int i = 0; i < 2; ++i) {
        if (b->h_table[i]) (void)hipHostFree(b->h_table[i]);
        if (b->table_free[i]) (void)hipEventDestroy(b->table_free[i]);
    }
    if (b->ev_in) (void)hipEventDestroy(b->ev_in);
    if (b->ev_out) (void)hipEventDestroy(b->ev_out);
    if (b->stream) (void)hipStreamDestroy(b->stream);
}

GtBatchSchedule default_schedule() {
    int cus = 256;
    int dev = 0;
    if (hipGetDevice(&dev) == hipSuccess) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0)
            cus = prop.multiProcessorCount;
    }
    const uint32_t c8 = static_cast<uint32_t>((cus + 7) / 8 * 8);
    return GtBatchSchedule{4 * c8, 4 * c8, 4, 256};
}

// Splits frames[0..n) into launches: consecutive frames while the slots and
// the pixel budget last.  Returns the end index of the launch starting at i.
uint32_t launch_end(const gt_batch *b, const gt_frame_desc *frames, uint32_t n, uint32_t i) {
    uint64_t px = 0;
    uint32_t j = i;
    while (j < n && j - i < b->max_slots) {
        const uint64_t hw = static_cast<uint64_t>(frames[j].nrows) * frames[j].ncols;
        if (j > i && px + hw > b->max_launch_pixels) break;
        px += hw;
        ++j;
    }
    return j;
}

gt_status run_batch(gt_batch *b, const gt_frame_desc *frames, uint32_t n, int32_t dtype,
                    gt_detection *d_out, uint32_t cap, uint32_t *d_counts, hipStream_t s,
                    const GtBatchSchedule &sch, bool graph) {
    const GtBatchHeader *hdr = reinterpret_cast<const GtBatchHeader *>(b->d_table);
    const GtBatchFrame *fs = reinterpret_cast<const GtBatchFrame *>(b->d_table + sizeof(GtBatchHeader));
    hipGraphExec_t gexec = nullptr;
    if (graph) {
        // record the stage sequence once; replays read the current table
        hipGraph_t g = nullptr;
        GB_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        hipError_t le = gt_launch_batch(nullptr, dtype, b->kp, hdr, fs, b->max_slots, b->d_proc,
                                        d_out, cap, b->d_ocount, b->d_slow, b->slow_cap,
                                        b->d_scount, d_counts, sch, s);
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
    }
    gt_status st = GT_OK;
    uint32_t k = 0;
    b->launches = 0;
    for (uint32_t i = 0; i < n && st == GT_OK;) {
        const uint32_t j = launch_end(b, frames, n, i);
        const int buf = static_cast<int>(k & 1u);
        // the pinned table buffer is reused two launches later
        if (hipEventSynchronize(b->table_free[buf]) != hipSuccess) {
            st = gt_runtime_fail(GT_ERR_GPU_DEVICE, "hipEventSynchronize failed");
            break;
        }
        GtBatchHeader *h = reinterpret_cast<GtBatchHeader *>(b->h_table[buf]);
        GtBatchFrame *e = reinterpret_cast<GtBatchFrame *>(b->h_table[buf] + sizeof(GtBatchHeader));
        uint32_t bp = 0, dt = 0;
        uint64_t off = 0;
        for (uint32_t f = i; f < j; ++f) {
            GtBatchFrame &x = e[f - i];
            x.in_ptr = reinterpret_cast<uint64_t>(frames[f].data);
            x.proc_off = off;
            x.H = static_cast<int32_t>(frames[f].nrows);
            x.W = static_cast<int32_t>(frames[f].ncols);
            x.bp_tile0 = bp;
            x.dt_tile0 = dt;
            x.out_idx = f;
            x.frame = frames[f].frame;
            uint32_t tb, td;
            gt_batch_tiles(x.H, x.W, &tb, &td);
            bp += tb;
            dt += td;
            off += static_cast<uint64_t>(x.H) * x.W;
        }
        h->nslots = j - i;
        h->bp_tiles = bp;
        h->dt_tiles = dt;
        h->pad = 0;
        hipError_t e1 = hipMemcpyAsync(b->d_table, b->h_table[buf], table_bytes(j - i),
                                       hipMemcpyHostToDevice, s);
        if (e1 == hipSuccess) e1 = hipEventRecord(b->table_free[buf], s);
        if (e1 == hipSuccess) {
            if (gexec) e1 = hipGraphLaunch(gexec, s);
            else
                e1 = gt_launch_batch(nullptr, dtype, b->kp, hdr, fs, b->max_slots, b->d_proc, d_out,
                                     cap, b->d_ocount, b->d_slow, b->slow_cap, b->d_scount, d_counts,
                                     sch, s);
        }
        if (e1 != hipSuccess)
            st = gt_runtime_fail(GT_ERR_GPU_DEVICE, "batch launch failed: %s", hipGetErrorString(e1));
        ++k;
        ++b->launches;
        i = j;
    }
    if (gexec) {
        // the executable graph may be destroyed once its launches are done
        (void)hipStreamSynchronize(s);
        (void)hipGraphExecDestroy(gexec);
    }
    return st;
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
        max_frame_pixels > 0xFFFFFFFFull)
        return gt_runtime_fail(GT_ERR_INVALID_ARGUMENT,
                               "need 0 < max_frame_pixels <= max_launch_pixels and max_slots > 0");
    if (params->style != GT_STYLE_TRACKPY)
        return gt_runtime_fail(GT_ERR_INVALID_ARGUMENT, "batched detection runs the Trackpy style");
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
    b->sched = default_schedule();
    gt_status st = GT_OK;
    auto ok = [&](hipError_t e, const char *what) {
        if (e != hipSuccess && st == GT_OK)
            st = gt_runtime_fail(GT_ERR_GPU_DEVICE, "%s failed: %s", what, hipGetErrorString(e));
        return st == GT_OK;
    };
    if (ok(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "hipStreamCreate") &&
        ok(gt_init_tables(), "table init") &&
        ok(hipMalloc(&b->d_proc, sizeof(float) * max_launch_pixels), "hipMalloc") &&
        ok(hipMalloc(&b->d_ocount, sizeof(uint32_t) * GT_COUNTER_STRIDE * max_slots), "hipMalloc") &&
        ok(hipMalloc(&b->d_scount, sizeof(uint32_t) * GT_COUNTER_STRIDE * max_slots), "hipMalloc") &&
        ok(hipMalloc(&b->d_slow, sizeof(uint2) * static_cast<size_t>(b->slow_cap) * max_slots), "hipMalloc") &&
        ok(hipMalloc(&b->d_table, table_bytes(max_slots)), "hipMalloc") &&
        ok(hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming), "hipEventCreate") &&
        ok(hipEventCreateWithFlags(&b->ev_out, hipEventDisableTiming), "hipEventCreate")) {
        for (int i = 0; i < 2 && st == GT_OK; ++i) {
            ok(hipHostMalloc(&b->h_table[i], table_bytes(max_slots), hipHostMallocDefault), "hipHostMalloc");
            if (st == GT_OK) ok(hipEventCreateWithFlags(&b->table_free[i], hipEventDisableTiming), "hipEventCreate");
            if (st == GT_OK) ok(hipEventRecord(b->table_free[i], b->stream), "hipEventRecord");
        }
    }
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
    hipStream_t user = static_cast<hipStream_t>(stream);
    if (user && user != b->stream) {
        GB_HIP(hipEventRecord(b->ev_in, user));
        GB_HIP(hipStreamWaitEvent(b->stream, b->ev_in, 0));
    }
    st = run_batch(b, frames, n, dtype, d_out, cap_per_frame, d_counts, b->stream, b->sched,
                   b->use_graph != 0);
    if (st != GT_OK) return st;
    if (user && user != b->stream) {
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

gt_status gt_batch_get_schedule(gt_batch *b, uint32_t *sched4, uint64_t *launches) {
    if (!b || !sched4) return gt_runtime_fail(GT_ERR_INVALID_ARGUMENT, "null argument");
    sched4[0] = b->sched.bp_blocks;
    sched4[1] = b->sched.dt_blocks;
    sched4[2] = b->sched.walk_blocks;
    sched4[3] = b->sched.slot_blocks;
    if (launches) *launches = b->launches;
    return GT_OK;
}

gt_status gt_batch_tune(gt_batch *b, const gt_frame_desc *frames, uint32_t n, int32_t dtype,
                        gt_detection *d_out, uint32_t cap_per_frame, uint32_t *d_counts,
                        float *best_ms) {
    // Coordinate search over the persistent grid sizes: each candidate runs
    // the sample once to warm up and once timed with events; the fastest
    // value of each knob is kept before the next knob is searched.
    if (!b) return gt_runtime_fail(GT_ERR_INVALID_ARGUMENT, "null batch context");
    if (n == 0 || !frames || !d_out || !d_counts || cap_per_frame == 0)
        return gt_runtime_fail(GT_ERR_INVALID_ARGUMENT, "empty sample or null argument");
    gt_status st = check_frames(b, frames, n, dtype);
    if (st != GT_OK) return st;
    hipEvent_t e0, e1;
    GB_HIP(hipEventCreate(&e0));
    GB_HIP(hipEventCreate(&e1));
    auto time_it = [&](const GtBatchSchedule &sch, float *ms) -> gt_status {
        gt_status s1 = run_batch(b, frames, n, dtype, d_out, cap_per_frame, d_counts, b->stream, sch, false);
        if (s1 != GT_OK) return s1;
        GB_HIP(hipEventRecord(e0, b->stream));
        s1 = run_batch(b, frames, n, dtype, d_out, cap_per_frame, d_counts, b->stream, sch, false);
        if (s1 != GT_OK) return s1;
        GB_HIP(hipEventRecord(e1, b->stream));
        GB_HIP(hipEventSynchronize(e1));
        GB_HIP(hipEventElapsedTime(ms, e0, e1));
        return GT_OK;
    };
    const GtBatchSchedule base = default_schedule();
    const uint32_t c8 = base.bp_blocks / 4;
    GtBatchSchedule cur = base;
    float best = 0.0f;
    st = time_it(cur, &best);
    const uint32_t mults[] = {2, 3, 4, 6, 8};
    const uint32_t walks[] = {1, 2, 4, 8};
    for (int knob = 0; knob < 3 && st == GT_OK; ++knob) {
        const int nopt = knob == 2 ? 4 : 5;
        for (int o = 0; o < nopt && st == GT_OK; ++o) {
            GtBatchSchedule c = cur;
            if (knob == 0) c.bp_blocks = mults[o] * c8;
            else if (knob == 1) c.dt_blocks = mults[o] * c8;
            else c.walk_blocks = walks[o];
            if (c.bp_blocks == cur.bp_blocks && c.dt_blocks == cur.dt_blocks &&
                c.walk_blocks == cur.walk_blocks)
                continue;
            float ms = 0.0f;
            st = time_it(c, &ms);
            if (st == GT_OK && ms < best) {
                best = ms;
                cur = c;
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st != GT_OK) return st;
    b->sched = cur;
    if (best_ms) *best_ms = best;
    return GT_OK;
}

}  // extern "C"
