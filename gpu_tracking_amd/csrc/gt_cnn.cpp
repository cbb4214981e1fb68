// gt_cnn.cpp -- host side of the learned centroid proposal (include/gtrack.h):
// the network object (float32 parameters, Adam moments, packed weight
// copies), its initialiser, forward, gradient, Adam step and training loop.
// Kernels live in gt_cnn.hip; sample generation and the detection path that
// uses the network live in gt_runtime.cpp (they need the context).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "gt_cnn.h"
#include "gt_internal.h"

gt_status gt_runtime_fail(gt_status st, const char *fmt, ...);

struct gt_cnn {
    hipStream_t stream = nullptr;
    float *d_params = nullptr, *d_m = nullptr, *d_v = nullptr, *d_grad = nullptr;
    float *d_partials = nullptr, *d_ploss = nullptr, *d_losses = nullptr;
    uint8_t *d_pk = nullptr;
    uint32_t *d_idx = nullptr;  // a training epoch's sample order
    size_t idx_cap = 0, losses_cap = 0;
    uint64_t steps = 0;         // Adam steps since the parameters were set
};

// Packed weights the kernels read (gt_runtime.cpp's detection path).
const void *gt_cnn_packed(const gt_cnn *n) { return n->d_pk; }

namespace {

#define GC_HIP(call)                                                                             \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return gt_runtime_fail(GT_ERR_GPU_DEVICE, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

uint64_t splitmix(uint64_t &s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
double uniform01(uint64_t &s) { return static_cast<double>(splitmix(s) >> 11) * 0x1.0p-53; }

// He-uniform weights (bound sqrt(6 / fan_in)), zero biases.
void he_init(uint64_t seed, std::vector<float> &p) {
    p.assign(CNN_NPARAM, 0.0f);
    uint64_t s = seed;
    auto fill = [&](int off, int n, int fan_in) {
        const double b = std::sqrt(6.0 / fan_in);
        for (int i = 0; i < n; ++i) p[off + i] = static_cast<float>((2.0 * uniform01(s) - 1.0) * b);
    };
    fill(CNN_W1, CNN_B1 - CNN_W1, 9);
    fill(CNN_W2, CNN_B2 - CNN_W2, 144);
    fill(CNN_W3, CNN_B3 - CNN_W3, 288);
    fill(CNN_WFC, CNN_BFC - CNN_WFC, 32);
}

void free_cnn(gt_cnn *n) {
    for (float *p : {n->d_params, n->d_m, n->d_v, n->d_grad, n->d_partials, n->d_ploss, n->d_losses})
        if (p) (void)hipFree(p);
    if (n->d_pk) (void)hipFree(n->d_pk);
    if (n->d_idx) (void)hipFree(n->d_idx);
    if (n->stream) (void)hipStreamDestroy(n->stream);
}

gt_status upload_params(gt_cnn *n, const float *h) {
    GC_HIP(hipMemcpyAsync(n->d_params, h, sizeof(float) * CNN_NPARAM, hipMemcpyHostToDevice, n->stream));
    GC_HIP(hipMemsetAsync(n->d_m, 0, sizeof(float) * CNN_NPARAM, n->stream));
    GC_HIP(hipMemsetAsync(n->d_v, 0, sizeof(float) * CNN_NPARAM, n->stream));
    GC_HIP(gt_cnn_launch_pack(n->d_params, n->d_pk, n->stream));
    GC_HIP(hipStreamSynchronize(n->stream));
    n->steps = 0;
    return GT_OK;
}

// torch.optim.Adam's defaults (betas 0.9, 0.999; eps 1e-8) at step t
CnnAdam adam_args(uint64_t t, float lr, bool apply) {
    CnnAdam a;
    a.apply = apply ? 1 : 0;
    a.lr = lr;
    a.b1 = 0.9f;
    a.b2 = 0.999f;
    a.omb1 = static_cast<float>(1.0 - 0.9);
    a.omb2 = static_cast<float>(1.0 - 0.999);
    a.eps = 1e-8f;
    a.bc1 = static_cast<float>(1.0 - std::pow(0.9, static_cast<double>(t)));
    a.bc2 = static_cast<float>(1.0 - std::pow(0.999, static_cast<double>(t)));
    return a;
}

gt_status check_batch(const gt_cnn *n, const void *patches, const void *targets, const void *idx,
                      uint32_t nb) {
    if (!n) return gt_runtime_fail(GT_ERR_INVALID_ARGUMENT, "null network");
    if (!patches || !targets || !idx || nb == 0)
        return gt_runtime_fail(GT_ERR_INVALID_ARGUMENT, "null buffer or empty batch");
    return GT_OK;
}

gt_status grow_losses(gt_cnn *n, size_t need) {
    if (need <= n->losses_cap) return GT_OK;
    if (n->d_losses) (void)hipFree(n->d_losses);
    n->d_losses = nullptr;
    n->losses_cap = 0;
    GC_HIP(hipMalloc(&n->d_losses, sizeof(float) * need));
    n->losses_cap = need;
    return GT_OK;
}

// One Adam step on the batch idx[0..nb); the batch loss lands in *d_loss.
gt_status adam_step(gt_cnn *n, const uint16_t *patches, const float *targets,
                    const uint32_t *d_idx, uint32_t nb, float lr, float *d_loss) {
    GC_HIP(gt_cnn_launch_train(n->d_pk, patches, reinterpret_cast<const float4 *>(targets), d_idx,
                               nb, n->d_partials, n->d_ploss, n->stream));
    ++n->steps;
    const CnnAdam ad = adam_args(n->steps, lr, true);
    GC_HIP(gt_cnn_launch_reduce(n->d_partials, n->d_ploss, nullptr, d_loss, n->d_params, n->d_m,
                                n->d_v, n->d_pk, ad, n->stream));
    return GT_OK;
}

}  // namespace

extern "C" {

gt_status gt_cnn_create(uint64_t seed, gt_cnn **out) {
    if (!out) return gt_runtime_fail(GT_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return gt_runtime_fail(GT_ERR_GPU_ADAPTER, "no HIP device is available");
    gt_cnn *n = new (std::nothrow) gt_cnn();
    if (!n) return gt_runtime_fail(GT_ERR_GPU_DEVICE, "out of host memory");
    hipError_t e = hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking);
    const size_t np = sizeof(float) * CNN_NPARAM;
    if (e == hipSuccess) e = hipMalloc(&n->d_params, np);
    if (e == hipSuccess) e = hipMalloc(&n->d_m, np);
    if (e == hipSuccess) e = hipMalloc(&n->d_v, np);
    if (e == hipSuccess) e = hipMalloc(&n->d_grad, np);
    if (e == hipSuccess) e = hipMalloc(&n->d_partials, np * CNN_TRAIN_BLOCKS);
    if (e == hipSuccess) e = hipMalloc(&n->d_ploss, sizeof(float) * CNN_TRAIN_BLOCKS);
    if (e == hipSuccess) e = hipMalloc(&n->d_pk, CNN_PK_BYTES);
    if (e == hipSuccess) e = hipMemset(n->d_pk, 0, CNN_PK_BYTES);  // w2p's padding columns
    if (e != hipSuccess) {
        free_cnn(n);
        delete n;
        return gt_runtime_fail(GT_ERR_GPU_DEVICE, "network allocation failed: %s", hipGetErrorString(e));
    }
    gt_status st = grow_losses(n, 1);
    std::vector<float> p;
    he_init(seed, p);
    if (st == GT_OK) st = upload_params(n, p.data());
    if (st != GT_OK) {
        free_cnn(n);
        delete n;
        return st;
    }
    *out = n;
    return GT_OK;
}

void gt_cnn_destroy(gt_cnn *n) {
    if (!n) return;
    (void)hipStreamSynchronize(n->stream);
    free_cnn(n);
    delete n;
}

gt_status gt_cnn_set_params(gt_cnn *n, const float *params) {
    if (!n || !params) return gt_runtime_fail(GT_ERR_INVALID_ARGUMENT, "null argument");
    for (int i = 0; i < CNN_NPARAM; ++i)
        if (!std::isfinite(params[i]))
            return gt_runtime_fail(GT_ERR_INVALID_ARGUMENT, "parameter %d is not finite", i);
    return upload_params(n, params);
}

gt_status gt_cnn_get_params(gt_cnn *n, float *params) {
    if (!n || !params) return gt_runtime_fail(GT_ERR_INVALID_ARGUMENT, "null argument");
    GC_HIP(hipMemcpyAsync(params, n->d_params, sizeof(float) * CNN_NPARAM, hipMemcpyDeviceToHost,
                          n->stream));
    GC_HIP(hipStreamSynchronize(n->stream));
    return GT_OK;
}

gt_status gt_cnn_adam_steps(gt_cnn *n, uint64_t *steps) {
    if (!n || !steps) return gt_runtime_fail(GT_ERR_INVALID_ARGUMENT, "null argument");
    *steps = n->steps;
    return GT_OK;
}

gt_status gt_cnn_forward(gt_cnn *n, const uint16_t *d_patches, uint64_t count, float *d_out,
                         void *stream) {
    if (!n || !d_patches || !d_out) return gt_runtime_fail(GT_ERR_INVALID_ARGUMENT, "null argument");
    if (count > 0xFFFFFFFFull) return gt_runtime_fail(GT_ERR_INVALID_ARGUMENT, "at most 2^32 - 1 patches");
    hipStream_t s = static_cast<hipStream_t>(stream);  // NULL: the default stream
    GC_HIP(gt_cnn_launch_forward_patches(n->d_pk, d_patches, static_cast<uint32_t>(count),
                                         reinterpret_cast<float4 *>(d_out), s));
    return GT_OK;
}

gt_status gt_cnn_gradient(gt_cnn *n, const uint16_t *d_patches, const float *d_targets,
                          const uint32_t *d_idx, uint32_t nb, float *grad, float *loss) {
    gt_status st = check_batch(n, d_patches, d_targets, d_idx, nb);
    if (st != GT_OK) return st;
    if (!grad) return gt_runtime_fail(GT_ERR_INVALID_ARGUMENT, "null gradient buffer");
    GC_HIP(gt_cnn_launch_train(n->d_pk, d_patches, reinterpret_cast<const float4 *>(d_targets),
                               d_idx, nb, n->d_partials, n->d_ploss, n->stream));
    const CnnAdam ad = adam_args(1, 0.0f, false);
    GC_HIP(gt_cnn_launch_reduce(n->d_partials, n->d_ploss, n->d_grad, n->d_losses, n->d_params,
                                n->d_m, n->d_v, n->d_pk, ad, n->stream));
    GC_HIP(hipMemcpyAsync(grad, n->d_grad, sizeof(float) * CNN_NPARAM, hipMemcpyDeviceToHost, n->stream));
    float L = 0.0f;
    GC_HIP(hipMemcpyAsync(&L, n->d_losses, sizeof(float), hipMemcpyDeviceToHost, n->stream));
    GC_HIP(hipStreamSynchronize(n->stream));
    if (loss) *loss = L;
    return GT_OK;
}

gt_status gt_cnn_step(gt_cnn *n, const uint16_t *d_patches, const float *d_targets,
                      const uint32_t *d_idx, uint32_t nb, float lr, float *loss) {
    gt_status st = check_batch(n, d_patches, d_targets, d_idx, nb);
    if (st != GT_OK) return st;
    if (!(lr > 0.0f) || !std::isfinite(lr))
        return gt_runtime_fail(GT_ERR_INVALID_ARGUMENT, "learning rate must be positive");
    st = adam_step(n, d_patches, d_targets, d_idx, nb, lr, n->d_losses);
    if (st != GT_OK) return st;
    float L = 0.0f;
    GC_HIP(hipMemcpyAsync(&L, n->d_losses, sizeof(float), hipMemcpyDeviceToHost, n->stream));
    GC_HIP(hipStreamSynchronize(n->stream));
    if (loss) *loss = L;
    return GT_OK;
}

gt_status gt_cnn_train(gt_cnn *n, const uint16_t *d_patches, const float *d_targets,
                       uint64_t count, uint32_t epochs, uint32_t batch, float lr, uint64_t seed,
                       float *loss_per_epoch) {
    if (!n || !d_patches || !d_targets) return gt_runtime_fail(GT_ERR_INVALID_ARGUMENT, "null argument");
    if (count == 0 || count > 0xFFFFFFFFull || batch == 0 || epochs == 0)
        return gt_runtime_fail(GT_ERR_INVALID_ARGUMENT,
                               "need 0 < samples < 2^32, a positive batch and epoch count");
    if (!(lr > 0.0f) || !std::isfinite(lr))
        return gt_runtime_fail(GT_ERR_INVALID_ARGUMENT, "learning rate must be positive");
    const uint32_t N = static_cast<uint32_t>(count);
    if (N > n->idx_cap) {
        if (n->d_idx) (void)hipFree(n->d_idx);
        n->d_idx = nullptr;
        n->idx_cap = 0;
        GC_HIP(hipMalloc(&n->d_idx, sizeof(uint32_t) * N));
        n->idx_cap = N;
    }
    const uint32_t nsteps = (N + batch - 1) / batch;
    gt_status st = grow_losses(n, nsteps);
    if (st != GT_OK) return st;
    std::vector<uint32_t> order(N);
    std::vector<float> losses(nsteps);
    uint64_t rs = seed;
    for (uint32_t ep = 0; ep < epochs; ++ep) {
        // a fresh shuffle per epoch (Fisher-Yates)
        for (uint32_t i = 0; i < N; ++i) order[i] = i;
        for (uint32_t i = N - 1; i > 0; --i) {
            const uint32_t j = static_cast<uint32_t>(splitmix(rs) % (static_cast<uint64_t>(i) + 1));
            std::swap(order[i], order[j]);
        }
        GC_HIP(hipMemcpyAsync(n->d_idx, order.data(), sizeof(uint32_t) * N, hipMemcpyHostToDevice,
                              n->stream));
        for (uint32_t k = 0; k < nsteps; ++k) {
            const uint32_t b0 = k * batch, nb = std::min(batch, N - b0);
            st = adam_step(n, d_patches, d_targets, n->d_idx + b0, nb, lr, n->d_losses + k);
            if (st != GT_OK) return st;
        }
        GC_HIP(hipMemcpyAsync(losses.data(), n->d_losses, sizeof(float) * nsteps,
                              hipMemcpyDeviceToHost, n->stream));
        GC_HIP(hipStreamSynchronize(n->stream));
        if (loss_per_epoch) {
            double s = 0.0;
            for (float L : losses) s += L;
            loss_per_epoch[ep] = static_cast<float>(s / nsteps);
        }
    }
    return GT_OK;
}

}  // extern "C"
