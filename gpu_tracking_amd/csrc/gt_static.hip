// gt_static.hip -- gt_detect_device_static: the plain Trackpy-style path of
// gt_detect_device with its scratch (processed planes, per-frame counters,
// slow-walk lists) in a file-scope device array instead of memory the
// library allocates, for callers that cannot allocate device memory.
//
// Host code never passes the array's name to a kernel: its device address is
// looked up once per device with hipGetSymbolAddress (each device holds its
// own copy of the array) and kept per device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "gt_host.h"
#include "gt_internal.h"

namespace {

// 64 MiB per device, held until the process ends.  One 1024 x 1024 frame
// takes about 5.25 MB of it (4 B/px processed plane, 1 B/px slow-walk list,
// two 256-byte counters), so a call runs in chunks of 12 such frames.
constexpr size_t kScratchBytes = size_t(64) << 20;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_scratch[kMaxDevices] = {};

constexpr size_t align256(size_t n) { return (n + 255) & ~size_t(255); }

// The current device's copy of the array (and, on the device's first call,
// the constant tables the characterize kernels read).
gt_status device_scratch(void **out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return gt_runtime_fail(GT_ERR_GPU_ADAPTER, "no HIP device");
    if (dev < 0 || dev >= kMaxDevices)
        return gt_runtime_fail(GT_ERR_GPU_DEVICE, "device %d is beyond the %d this entry point serves",
                               dev, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return gt_runtime_fail(GT_ERR_GPU_DEVICE, "hipGetSymbolAddress(scratch) failed: %s",
                                   hipGetErrorString(e));
        e = gt_init_tables();
        if (e != hipSuccess)
            return gt_runtime_fail(GT_ERR_GPU_DEVICE, "could not initialise the device tables: %s",
                                   hipGetErrorString(e));
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return GT_OK;
}

size_t dtype_bytes(int32_t dtype) {
    return dtype == GT_U16 ? 2 : dtype == GT_F32 ? 4 : dtype == GT_U8 ? 1 : 0;
}

}  // namespace

extern "C" gt_status gt_detect_device_static(const gt_params *params, uint32_t nrows,
                                             uint32_t ncols, const void *d_frames, int32_t dtype,
                                             uint32_t nframes, gt_detection *d_out,
                                             uint32_t cap_per_frame, uint32_t *d_counts,
                                             void *stream) {
    if (!params) return gt_runtime_fail(GT_ERR_INVALID_ARGUMENT, "null parameters");
    const size_t esz = dtype_bytes(dtype);
    if (!esz) return gt_runtime_fail(GT_ERR_INVALID_ARGUMENT, "unknown dtype %d", dtype);
    if (nframes == 0)
        return gt_runtime_fail(GT_ERR_EMPTY_ITERATOR, "the provided video doesn't contain any frames.");
    if (!d_frames || !d_out || !d_counts || cap_per_frame == 0)
        return gt_runtime_fail(GT_ERR_INVALID_ARGUMENT, "null device buffer or zero capacity");
    if (nrows == 0 || ncols == 0 || static_cast<uint64_t>(nrows) * ncols > 0xFFFFFFFFull)
        return gt_runtime_fail(GT_ERR_INVALID_ARGUMENT, "frame dimensions must be in 1 .. 2^32 pixels");
    gt_derived der;
    std::string why;
    if (!gt_host_derive(params, nrows, ncols, &der, &why))
        return gt_runtime_fail(GT_ERR_INVALID_ARGUMENT, "%s", why.c_str());
    if (der.style != GT_STYLE_TRACKPY || der.need_std || der.adaptive || der.illum_sigma > 0.0f)
        return gt_runtime_fail(GT_ERR_INVALID_ARGUMENT,
                               "the static-scratch entry point runs the plain Trackpy-style path "
                               "(no LoG, snr, minmass_snr, adaptive_background or illumination "
                               "correction)");
    GtKParams kp;
    gt_host_fill_kparams(&der, nrows, ncols, &kp);
    if (kp.k < 1 || gt_tile_lds_bytes(kp) > 65536)
        return gt_runtime_fail(GT_ERR_INVALID_ARGUMENT,
                               "separation %u gives a dilation window beyond the tiled kernels",
                               der.separation);
    // per frame: processed plane, row and slow-walk counters, slow-walk list
    // (one slot per 8 pixels, as a context's)
    const size_t HW = static_cast<size_t>(nrows) * ncols;
    const uint32_t slow_cap = static_cast<uint32_t>(std::max<size_t>(1024, HW / 8));
    const size_t per_frame = align256(sizeof(float) * HW) + 2 * 256 +
                             align256(sizeof(uint2) * static_cast<size_t>(slow_cap));
    if (per_frame > kScratchBytes)
        return gt_runtime_fail(GT_ERR_CAPACITY,
                               "a %u x %u frame needs %zu bytes of scratch; the static scratch "
                               "holds %zu",
                               nrows, ncols, per_frame, kScratchBytes);
    void *base = nullptr;
    gt_status st = device_scratch(&base);
    if (st != GT_OK) return st;
    const uint32_t chunk = static_cast<uint32_t>(
        std::min<size_t>({kScratchBytes / per_frame, static_cast<size_t>(nframes), size_t(256)}));
    char *b = static_cast<char *>(base);
    GtPlainScratch sc;
    sc.proc = reinterpret_cast<float *>(b);
    b += align256(sizeof(float) * HW * chunk);
    sc.ocount = reinterpret_cast<uint32_t *>(b);
    b += align256(sizeof(uint32_t) * GT_COUNTER_STRIDE * chunk);
    sc.scount = reinterpret_cast<uint32_t *>(b);
    b += align256(sizeof(uint32_t) * GT_COUNTER_STRIDE * chunk);
    sc.slow = reinterpret_cast<uint2 *>(b);
    sc.slow_cap = slow_cap;
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipError_t e = gt_launch_zero_u32(d_counts, nframes, s);
    for (uint32_t f0 = 0; e == hipSuccess && f0 < nframes; f0 += chunk) {
        const uint32_t nf = std::min(chunk, nframes - f0);
        const char *frames = static_cast<const char *>(d_frames) + esz * HW * f0;
        gt_detection *out = d_out + static_cast<size_t>(f0) * cap_per_frame;
        e = gt_plain_head(frames, dtype, nf, kp, sc, out, cap_per_frame, f0, s, nullptr);
        if (e == hipSuccess)
            e = gt_plain_tail(frames, dtype, nf, kp, sc, out, cap_per_frame, d_counts + f0, f0, s);
    }
    if (e != hipSuccess)
        return gt_runtime_fail(GT_ERR_GPU_DEVICE, "launch failed: %s", hipGetErrorString(e));
    return GT_OK;
}
