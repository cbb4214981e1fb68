// gt_batch.cpp -- batched detection over frames of different sizes.
//
// The reference's GpuState serves one frame size (src/execute_gpu.rs:1074-1080);
// a batch context instead takes any mix of frame sizes (same detection
// parameters).  A call's frames, sorted by size, are split into launches (a
// pixel budget, a frame-count limit and at most GT_BATCH_MAX_SEG runs of
// equal size); every launch is one kernel per stage: the bandpass and detect
// kernels take one queue of all the launch's tiles (frame by frame through
// the size runs), the per-frame stages take grid y = frame.  A call whose
// pixels fit the budget is one launch per stage.  All launch
// tables of a call go to the device in one copy; the stage sequence, framed by
// a kernel that selects the next launch's table on the device and one that
// advances a device launch counter, is captured once into a HIP graph (again
// only when the output buffers or grid sizes change) and replayed once per
// launch with no host work in between.  Launches are dealt to up to four
// lanes (stream, scratch and graph each) so that consecutive launches
// overlap.  gt_batch_tune picks the launch size, the slow-walk /
// characterize grid and the lane count by timing candidates.
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

// One lane: a stream with its own scratch, launch list and captured graph.
// A call's launches are dealt to the lanes in turn (launch k to lane k % n),
// so consecutive launches overlap as separate contexts on separate streams
// do; each frame's rows and count still go to its own output slots.
struct BatchLane {
    hipStream_t stream = nullptr;
    float *d_proc = nullptr;          // max_launch_pixels
    uint32_t *d_ocount = nullptr;     // max_slots * GT_COUNTER_STRIDE
    uint32_t *d_scount = nullptr;
    uint2 *d_slow = nullptr;          // max_slots * slow_cap
    GtBatchHeader *d_hdr = nullptr;   // the running launch's table ...
    GtBatchFrame *d_fs = nullptr;     // ... max_slots entries
    GtBatchLaunch *d_launches = nullptr;  // this lane's launches of the call
    size_t launch_cap = 0;
    uint32_t *d_cur = nullptr;        // device launch counter
    hipEvent_t done = nullptr;
    // captured stage sequences, one per grid shape, kept while the call's
    // arguments repeat
    struct GraphKey {
        gt_detection *out;
        uint32_t cap;
        uint32_t *counts;
        int32_t dtype;
        GtBatchFrame *all;
        GtBatchLaunch *launches;
        uint32_t tile_blocks, walk_blocks, launch_slots;
        bool operator==(const GraphKey &o) const {
            return out == o.out && cap == o.cap && counts == o.counts && dtype == o.dtype &&
                   all == o.all && launches == o.launches && tile_blocks == o.tile_blocks &&
                   walk_blocks == o.walk_blocks && launch_slots == o.launch_slots;
        }
    };
    std::vector<std::pair<GraphKey, hipGraphExec_t>> graphs;
};

constexpr uint32_t kMaxLanes = 4;

struct gt_batch {
    gt_params params;
    gt_derived der;
    GtKParams kp;
    hipStream_t stream = nullptr;     // uploads, fork and join of the lanes
    uint64_t max_frame_pixels = 0, max_launch_pixels = 0;
    uint32_t max_slots = 0;
    uint32_t slow_cap = 0;
    BatchLane lanes[kMaxLanes];
    uint32_t nlanes = 1;              // lanes in use
    // a call's launch tables: all entries (shared by the lanes)
    GtBatchFrame *d_all = nullptr;
    size_t all_cap = 0;
    // pinned staging of the tables (+ the counters' zeros), two buffers used
    // by alternate calls: a call fills one while the previous call's launches
    // (uploaded from the other) still run
    char *h_stage[2] = {nullptr, nullptr};
    size_t h_cap[2] = {0, 0};
    hipEvent_t stage_free[2] = {nullptr, nullptr};  // the last call using buffer i is done
    uint32_t stage_next = 0;          // buffer of the next call
    hipEvent_t uploaded = nullptr;    // the call's tables are on the device
    hipEvent_t ev_in = nullptr, ev_out = nullptr;  // ordering against the caller's stream
    uint32_t walk_blocks = 4;         // blocks per frame for slow walks and characterize
    uint64_t launch_px = 0;           // pixels per launch in use (<= max_launch_pixels)
    int32_t use_graph = 1;
    uint64_t launches = 0;            // launches of the last call
    GtBatchSchedule last_sch = {0u, 0u, 0u};  // grid shape of the last call
    uint32_t last_max_seg = 0;        // most size runs in one launch of the last call
};

namespace {

void free_lane(BatchLane &L) {
    if (L.d_proc) (void)hipFree(L.d_proc);
    if (L.d_ocount) (void)hipFree(L.d_ocount);
    if (L.d_scount) (void)hipFree(L.d_scount);
    if (L.d_slow) (void)hipFree(L.d_slow);
    if (L.d_hdr) (void)hipFree(L.d_hdr);
    if (L.d_fs) (void)hipFree(L.d_fs);
    if (L.d_launches) (void)hipFree(L.d_launches);
    if (L.d_cur) (void)hipFree(L.d_cur);
    for (auto &kg : L.graphs) (void)hipGraphExecDestroy(kg.second);
    if (L.done) (void)hipEventDestroy(L.done);
    if (L.stream) (void)hipStreamDestroy(L.stream);
    L = BatchLane();
}

void free_batch(gt_batch *b) {
    for (BatchLane &L : b->lanes) free_lane(L);
    if (b->d_all) (void)hipFree(b->d_all);
    for (int i = 0; i < 2; ++i) {
        if (b->h_stage[i]) (void)hipHostFree(b->h_stage[i]);
        if (b->stage_free[i]) (void)hipEventDestroy(b->stage_free[i]);
    }
    if (b->uploaded) (void)hipEventDestroy(b->uploaded);
    if (b->ev_in) (void)hipEventDestroy(b->ev_in);
    if (b->ev_out) (void)hipEventDestroy(b->ev_out);
    if (b->stream) (void)hipStreamDestroy(b->stream);
}

// Allocates lane l's stream and scratch on first use.
gt_status alloc_lane(gt_batch *b, uint32_t l) {
    BatchLane &L = b->lanes[l];
    if (L.stream) return GT_OK;
    GB_HIP(hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking));
    GB_HIP(hipEventCreateWithFlags(&L.done, hipEventDisableTiming));
    GB_HIP(hipMalloc(&L.d_proc, sizeof(float) * b->max_launch_pixels));
    GB_HIP(hipMalloc(&L.d_ocount, sizeof(uint32_t) * GT_COUNTER_STRIDE * b->max_slots));
    GB_HIP(hipMalloc(&L.d_scount, sizeof(uint32_t) * GT_COUNTER_STRIDE * b->max_slots));
    GB_HIP(hipMalloc(&L.d_slow, sizeof(uint2) * static_cast<size_t>(b->slow_cap) * b->max_slots));
    GB_HIP(hipMalloc(&L.d_hdr, sizeof(GtBatchHeader)));
    GB_HIP(hipMalloc(&L.d_fs, sizeof(GtBatchFrame) * b->max_slots));
    GB_HIP(hipMalloc(&L.d_cur, sizeof(uint32_t)));
    return GT_OK;
}

// Splits frames[order[0..n)] (sorted by size) into launches: consecutive
// frames while the frame limit, the pixel budget and the segment limit (runs
// of equal frame size) last.  Returns the end index of the launch starting at
// i.
uint32_t launch_end(const gt_batch *b, const gt_frame_desc *frames, const uint32_t *order,
                    uint32_t n, uint32_t i, uint64_t launch_px) {
    uint64_t px = 0;
    uint32_t j = i, nseg = 0;
    while (j < n && j - i < b->max_slots) {
        const gt_frame_desc &fj = frames[order[j]];
        const uint64_t hw = static_cast<uint64_t>(fj.nrows) * fj.ncols;
        if (j > i && px + hw > launch_px) break;
        const bool new_seg = j == i || fj.nrows != frames[order[j - 1]].nrows ||
                             fj.ncols != frames[order[j - 1]].ncols;
        if (new_seg && nseg == GT_BATCH_MAX_SEG) break;
        nseg += new_seg ? 1u : 0u;
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
                    uint32_t walk_blocks, uint64_t launch_px, bool graph, uint32_t nlanes) {
    nlanes = std::max(1u, std::min(nlanes, kMaxLanes));
    // the call's frames by size (stable), so that a launch holds frames of
    // one size where it can and its grids fit them; results go to each
    // frame's own output slot, so the order does not show
    std::vector<uint32_t> order(n);
    for (uint32_t i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t c) {
        const uint64_t pa = static_cast<uint64_t>(frames[a].nrows) * frames[a].ncols;
        const uint64_t pc = static_cast<uint64_t>(frames[c].nrows) * frames[c].ncols;
        if (pa != pc) return pa < pc;
        return frames[a].ncols < frames[c].ncols;
    });
    // the call's launches and their tables
    std::vector<GtBatchLaunch> launches;
    for (uint32_t i = 0; i < n;) {
        const uint32_t j = launch_end(b, frames, order.data(), n, i, launch_px);
        GtBatchLaunch L = {};
        L.first = i;
        L.nslots = j - i;
        // the launch's tile queue: one segment per run of equal frame size
        for (uint32_t q = i; q < j; ++q) {
            const gt_frame_desc &fq = frames[order[q]];
            if (q == i || fq.nrows != frames[order[q - 1]].nrows ||
                fq.ncols != frames[order[q - 1]].ncols) {
                uint32_t tb, td;
                gt_batch_tiles(static_cast<int>(fq.nrows), static_cast<int>(fq.ncols), &tb, &td);
                GtBatchSeg &sg = L.seg[L.nseg++];
                sg.first_tile = L.ntiles;
                sg.first_slot = q - i;
                sg.tpf = td;
                sg.ntx = static_cast<uint32_t>((fq.ncols + 63) / 64);
            }
            L.ntiles += L.seg[L.nseg - 1].tpf;
        }
        launches.push_back(L);
        i = j;
    }
    const size_t nl = launches.size();
    nlanes = static_cast<uint32_t>(std::min<size_t>(nlanes, nl));
    for (uint32_t l = 0; l < nlanes; ++l) {
        gt_status ast = alloc_lane(b, l);
        if (ast != GT_OK) return ast;
    }
    // launch k goes to lane k % nlanes; lane l's list starts at lane_off[l]
    std::vector<size_t> lane_n(nlanes, 0), lane_off(nlanes + 1, 0);
    for (size_t k = 0; k < nl; ++k) ++lane_n[k % nlanes];
    for (uint32_t l = 0; l < nlanes; ++l) lane_off[l + 1] = lane_off[l] + lane_n[l];
    // this call's staging buffer may still serve the upload of the call
    // before the previous one; the device tables are written by uploads
    // ordered on s after the previous call's launches
    const uint32_t sb = b->stage_next;
    b->stage_next ^= 1u;
    GB_HIP(hipEventSynchronize(b->stage_free[sb]));
    const size_t bytes = sizeof(GtBatchFrame) * n + sizeof(GtBatchLaunch) * nl +
                         sizeof(uint32_t) * kMaxLanes;
    if (bytes > b->h_cap[sb]) {
        const size_t c = std::max(bytes, 2 * b->h_cap[sb]);
        if (b->h_stage[sb]) (void)hipHostFree(b->h_stage[sb]);
        b->h_stage[sb] = nullptr;
        b->h_cap[sb] = 0;
        GB_HIP(hipHostMalloc(&b->h_stage[sb], c, hipHostMallocDefault));
        b->h_cap[sb] = c;
    }
    char *stage = b->h_stage[sb];
    bool regrow = n > b->all_cap;
    for (uint32_t l = 0; l < nlanes; ++l) regrow = regrow || lane_n[l] > b->lanes[l].launch_cap;
    // tables grown while the previous call may still read them: wait for it
    if (regrow) GB_HIP(hipEventSynchronize(b->stage_free[sb ^ 1u]));
    gt_status st = grow(&b->d_all, &b->all_cap, n);
    for (uint32_t l = 0; l < nlanes && st == GT_OK; ++l)
        st = grow(&b->lanes[l].d_launches, &b->lanes[l].launch_cap, lane_n[l]);
    if (st != GT_OK) return st;
    GtBatchFrame *e = reinterpret_cast<GtBatchFrame *>(stage);
    GtBatchLaunch *hl = reinterpret_cast<GtBatchLaunch *>(stage + sizeof(GtBatchFrame) * n);
    uint32_t *zero = reinterpret_cast<uint32_t *>(stage + sizeof(GtBatchFrame) * n +
                                                  sizeof(GtBatchLaunch) * nl);
    // the frame tables; one grid shape for every launch of the call: the tile
    // queue of its largest launch (multiple of 8) and its most frames
    GtBatchSchedule sch_all = {8u, walk_blocks, 1u};
    uint32_t max_seg = 0;
    for (size_t k = 0; k < nl; ++k) {
        const GtBatchLaunch &L = launches[k];
        uint64_t off = 0;
        for (uint32_t q = L.first; q < L.first + L.nslots; ++q) {
            const uint32_t f = order[q];
            GtBatchFrame &x = e[q];
            x.in_ptr = reinterpret_cast<uint64_t>(frames[f].data);
            x.proc_off = off;
            x.H = static_cast<int32_t>(frames[f].nrows);
            x.W = static_cast<int32_t>(frames[f].ncols);
            x.out_idx = f;
            x.frame = frames[f].frame;
            off += static_cast<uint64_t>(x.H) * x.W;
        }
        sch_all.tile_blocks = std::max(sch_all.tile_blocks, (L.ntiles + 7) / 8 * 8);
        sch_all.launch_slots = std::max(sch_all.launch_slots, L.nslots);
        max_seg = std::max(max_seg, L.nseg);
    }
    std::vector<size_t> fill(lane_off.begin(), lane_off.end() - 1);
    for (size_t k = 0; k < nl; ++k) hl[fill[k % nlanes]++] = launches[k];
    for (uint32_t l = 0; l < kMaxLanes; ++l) zero[l] = 0;
    // the device tables are rewritten after the previous call's launches,
    // whichever stream it ran on
    GB_HIP(hipStreamWaitEvent(s, b->stage_free[sb ^ 1u], 0));
    GB_HIP(hipMemcpyAsync(b->d_all, e, sizeof(GtBatchFrame) * n, hipMemcpyHostToDevice, s));
    for (uint32_t l = 0; l < nlanes; ++l) {
        GB_HIP(hipMemcpyAsync(b->lanes[l].d_launches, hl + lane_off[l],
                              sizeof(GtBatchLaunch) * lane_n[l], hipMemcpyHostToDevice, s));
        GB_HIP(hipMemcpyAsync(b->lanes[l].d_cur, zero + l, sizeof(uint32_t), hipMemcpyHostToDevice,
                              s));
    }
    GB_HIP(hipEventRecord(b->uploaded, s));
    for (uint32_t l = 0; l < nlanes; ++l) {
        BatchLane &L = b->lanes[l];
        hipStream_t ls = L.stream;
        GB_HIP(hipStreamWaitEvent(ls, b->uploaded, 0));
        // graphs of a previous call with other arguments are dropped (after
        // the lane's earlier work, which may still replay them)
        const BatchLane::GraphKey probe{d_out, cap, d_counts, dtype, b->d_all, L.d_launches,
                                        0, 0, 0};
        bool stale = false;
        for (auto &kg : L.graphs)
            stale = stale || kg.first.out != probe.out || kg.first.cap != probe.cap ||
                    kg.first.counts != probe.counts || kg.first.dtype != probe.dtype ||
                    kg.first.all != probe.all || kg.first.launches != probe.launches ||
                    kg.first.walk_blocks != walk_blocks;
        if (stale) {
            GB_HIP(hipStreamSynchronize(ls));
            for (auto &kg : L.graphs) (void)hipGraphExecDestroy(kg.second);
            L.graphs.clear();
        }
        for (size_t q = lane_off[l]; q < lane_off[l + 1]; ++q) {
            // every launch of the call replays the same grid shape
            const GtBatchSchedule &sch = sch_all;
            auto enqueue = [&]() {
                return gt_launch_batch(L.d_launches, L.d_cur, b->d_all, dtype, b->kp, L.d_hdr,
                                       L.d_fs, b->max_slots, L.d_proc, d_out, cap, L.d_ocount,
                                       L.d_slow, b->slow_cap, L.d_scount, d_counts, sch, ls);
            };
            hipGraphExec_t gexec = nullptr;
            if (graph) {
                const BatchLane::GraphKey key{d_out,           cap,           d_counts,
                                              dtype,           b->d_all,      L.d_launches,
                                              sch.tile_blocks, sch.walk_blocks, sch.launch_slots};
                for (auto &kg : L.graphs)
                    if (kg.first == key) gexec = kg.second;
                if (!gexec) {
                    hipGraph_t g = nullptr;
                    GB_HIP(hipStreamBeginCapture(ls, hipStreamCaptureModeThreadLocal));
                    hipError_t le = enqueue();
                    hipError_t ce = hipStreamEndCapture(ls, &g);
                    if (le != hipSuccess || ce != hipSuccess) {
                        if (g) (void)hipGraphDestroy(g);
                        return gt_runtime_fail(GT_ERR_GPU_DEVICE,
                                               "graph capture of the batch stages failed: %s",
                                               hipGetErrorString(le != hipSuccess ? le : ce));
                    }
                    hipError_t ie = hipGraphInstantiate(&gexec, g, nullptr, nullptr, 0);
                    (void)hipGraphDestroy(g);
                    if (ie != hipSuccess)
                        return gt_runtime_fail(GT_ERR_GPU_DEVICE, "hipGraphInstantiate failed: %s",
                                               hipGetErrorString(ie));
                    L.graphs.emplace_back(key, gexec);
                }
            }
            const hipError_t e1 = gexec ? hipGraphLaunch(gexec, ls) : enqueue();
            if (e1 != hipSuccess)
                return gt_runtime_fail(GT_ERR_GPU_DEVICE, "batch launch failed: %s",
                                       hipGetErrorString(e1));
        }
        GB_HIP(hipEventRecord(L.done, ls));
        GB_HIP(hipStreamWaitEvent(s, L.done, 0));
    }
    // this call's staging buffer is free again, and the next call may rewrite
    // the tables, once its launches are done
    GB_HIP(hipEventRecord(b->stage_free[sb], s));
    b->launches = nl;
    b->last_sch = sch_all;
    b->last_max_seg = max_seg;
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
           ok(hipEventCreateWithFlags(&b->uploaded, hipEventDisableTiming), "hipEventCreate") &&
           ok(hipEventCreateWithFlags(&b->stage_free[0], hipEventDisableTiming), "hipEventCreate") &&
           ok(hipEventCreateWithFlags(&b->stage_free[1], hipEventDisableTiming), "hipEventCreate") &&
           ok(hipEventRecord(b->stage_free[0], b->stream), "hipEventRecord") &&
           ok(hipEventRecord(b->stage_free[1], b->stream), "hipEventRecord") &&
           ok(hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming), "hipEventCreate") &&
           ok(hipEventCreateWithFlags(&b->ev_out, hipEventDisableTiming), "hipEventCreate"));
    if (st == GT_OK) st = alloc_lane(b, 0);
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
    for (BatchLane &L : b->lanes)
        if (L.stream) (void)hipStreamSynchronize(L.stream);
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
                   b->launch_px, b->use_graph != 0, b->nlanes);
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

gt_status gt_batch_set_lanes(gt_batch *b, uint32_t nlanes) {
    if (!b) return gt_runtime_fail(GT_ERR_INVALID_ARGUMENT, "null batch context");
    if (nlanes < 1 || nlanes > kMaxLanes)
        return gt_runtime_fail(GT_ERR_INVALID_ARGUMENT, "lanes must be between 1 and %u", kMaxLanes);
    b->nlanes = nlanes;
    return GT_OK;
}

uint32_t gt_batch_get_lanes(const gt_batch *b) { return b ? b->nlanes : 0; }

gt_status gt_batch_get_schedule(gt_batch *b, uint64_t *sched4, uint64_t *launches) {
    if (!b || !sched4) return gt_runtime_fail(GT_ERR_INVALID_ARGUMENT, "null argument");
    sched4[0] = b->launch_px;
    sched4[1] = b->walk_blocks;
    sched4[2] = b->max_slots;
    sched4[3] = b->max_launch_pixels;
    if (launches) *launches = b->launches;
    return GT_OK;
}

gt_status gt_batch_get_schedule_info(gt_batch *b, gt_batch_schedule_info *out) {
    if (!b || !out) return gt_runtime_fail(GT_ERR_INVALID_ARGUMENT, "null argument");
    gt_batch_schedule_info r = {};
    r.pixels_per_launch = b->launch_px;
    r.max_pixels_per_launch = b->max_launch_pixels;
    r.walk_blocks = b->walk_blocks;
    r.max_frames_per_launch = b->max_slots;
    r.lanes = b->nlanes;
    r.tile_rows = gt_batch_tile_rows();
    r.tile_cols = gt_batch_tile_cols();
    r.launches = static_cast<uint32_t>(b->launches);
    r.tile_blocks = b->last_sch.tile_blocks;
    r.launch_slots = b->last_sch.launch_slots;
    r.max_segments = b->last_max_seg;
    r.graph = b->use_graph;
    *out = r;
    return GT_OK;
}

gt_status gt_batch_tune(gt_batch *b, const gt_frame_desc *frames, uint32_t n, int32_t dtype,
                        gt_detection *d_out, uint32_t cap_per_frame, uint32_t *d_counts,
                        float *best_ms) {
    // Coordinate search: first the pixels per launch, then the slow-walk /
    // characterize blocks per frame, then the lanes; each candidate runs the
    // sample once to warm up and twice timed with events (the faster run
    // counts); the fastest value of each knob is kept.
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
    uint32_t lanes = 1;
    auto time_it = [&](uint32_t wb, uint64_t lpx, float *ms) -> gt_status {
        gt_status s1 = run_batch(b, frames, n, dtype, d_out, cap_per_frame, d_counts, b->stream, wb,
                                 lpx, false, lanes);
        if (s1 != GT_OK) return s1;
        for (int rep = 0; rep < 2; ++rep) {
            GB_HIP(hipEventRecord(e0, b->stream));
            s1 = run_batch(b, frames, n, dtype, d_out, cap_per_frame, d_counts, b->stream, wb,
                           lpx, false, lanes);
            if (s1 != GT_OK) return s1;
            GB_HIP(hipEventRecord(e1, b->stream));
            GB_HIP(hipEventSynchronize(e1));
            float t = 0.0f;
            GB_HIP(hipEventElapsedTime(&t, e0, e1));
            if (rep == 0 || t < *ms) *ms = t;
        }
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
    const uint32_t lane_opts[] = {2, 4};
    uint32_t best_lanes = 1;
    for (int o = 0; o < 2 && st == GT_OK; ++o) {
        lanes = lane_opts[o];
        float ms = 0.0f;
        st = time_it(wb, lpx, &ms);
        if (st == GT_OK && ms < best) {
            best = ms;
            best_lanes = lanes;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st != GT_OK) return st;
    b->walk_blocks = wb;
    b->launch_px = lpx;
    b->nlanes = best_lanes;
    if (best_ms) *best_ms = best;
    return GT_OK;
}

}  // extern "C"
