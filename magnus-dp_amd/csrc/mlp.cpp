// mlp.cpp -- C ABI of the resolution controller: feature extraction and a
// 3-layer MLP (in -> hidden -> hidden -> out, ReLU) trained with Adam on the
// device.  FP32 master parameters, bfloat16 weight copies read by the layer
// kernels, FP32 activations, gradients and accumulation.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "host.hpp"
#include "kernels_ctrl.hpp"
#include "magnus_capi.h"

using namespace magnus_host;

struct magnus_mlp {
    int in = 0, hid = 0, out = 0;
    uint64_t max_batch = 0;
    uint64_t step = 0;
    size_t np = 0;  // parameters
    size_t oW1, ob1, oW2, ob2, oW3, ob3;
    float* p = nullptr;      // FP32 masters
    uint16_t* p16 = nullptr; // bfloat16 copies
    float* g = nullptr;      // gradients of the last step
    float* m = nullptr;      // Adam moments
    float* v = nullptr;
    float *h1 = nullptr, *h2 = nullptr, *logits = nullptr, *dlogits = nullptr, *dh1 = nullptr, *dh2 = nullptr;
    void* block = nullptr;   // one device allocation for all of the above
};

namespace {

void layout(magnus_mlp* n) {
    size_t o = 0;
    n->oW1 = o;
    o += (size_t)n->hid * n->in;
    n->ob1 = o;
    o += n->hid;
    n->oW2 = o;
    o += (size_t)n->hid * n->hid;
    n->ob2 = o;
    o += n->hid;
    n->oW3 = o;
    o += (size_t)n->out * n->hid;
    n->ob3 = o;
    o += n->out;
    n->np = o;
}

magnus::LinearArgs layer(const magnus_mlp* n, int which, int B) {
    magnus::LinearArgs a{};
    a.B = B;
    if (which == 1) {
        a.in = n->in, a.out = n->hid, a.relu = 1;
        a.w16 = n->p16 + n->oW1, a.bias = n->p + n->ob1, a.dw = n->g + n->oW1, a.db = n->g + n->ob1;
    } else if (which == 2) {
        a.in = n->hid, a.out = n->hid, a.relu = 1;
        a.w16 = n->p16 + n->oW2, a.bias = n->p + n->ob2, a.dw = n->g + n->oW2, a.db = n->g + n->ob2;
    } else {
        a.in = n->hid, a.out = n->out, a.relu = 0;
        a.w16 = n->p16 + n->oW3, a.bias = n->p + n->ob3, a.dw = n->g + n->oW3, a.db = n->g + n->ob3;
    }
    return a;
}

int forward3(magnus_mlp* n, int B, const float* x, float* logits, hipStream_t s) {
    magnus::LinearArgs l1 = layer(n, 1, B), l2 = layer(n, 2, B), l3 = layer(n, 3, B);
    l1.x = x, l1.y = n->h1;
    l2.x = n->h1, l2.y = n->h2;
    l3.x = n->h2, l3.y = logits;
    hipError_t e = magnus::launch_linear_fwd(l1, s);
    if (e == hipSuccess) e = magnus::launch_linear_fwd(l2, s);
    if (e == hipSuccess) e = magnus::launch_linear_fwd(l3, s);
    if (e != hipSuccess) return device_error(e, "mlp forward");
    return MAGNUS_OK;
}

int check_batch(const magnus_mlp* n, uint64_t batch) {
    if (batch > n->max_batch) return fail(MAGNUS_EINVAL, "batch exceeds the network's max_batch");
    return MAGNUS_OK;
}

}  // namespace

extern "C" {

int magnus_mlp_create(uint32_t in_dim, uint32_t hidden, uint32_t out_dim, uint64_t max_batch, uint64_t seed,
                      magnus_mlp** out) {
    if (!out) return fail(MAGNUS_EINVAL, "null argument");
    *out = nullptr;
    if (in_dim == 0 || hidden == 0 || out_dim == 0 || max_batch == 0)
        return fail(MAGNUS_EINVAL, "layer widths and max_batch must be positive");
    if (in_dim > 4096 || hidden > 4096 || out_dim > 4096) return fail(MAGNUS_EINVAL, "layer width exceeds 4096");
    auto* n = new magnus_mlp();
    n->in = (int)in_dim, n->hid = (int)hidden, n->out = (int)out_dim, n->max_batch = max_batch;
    layout(n);
    const size_t act = max_batch * (4 * (size_t)hidden + 2 * (size_t)out_dim);
    const size_t bytes = n->np * (4 * sizeof(float) + sizeof(uint16_t)) + act * sizeof(float) + 1024;
    if (hipMalloc(&n->block, bytes) != hipSuccess) {
        (void)hipGetLastError();
        delete n;
        return fail(MAGNUS_ENOMEM, "mlp: device allocation failed");
    }
    char* at = static_cast<char*>(n->block);
    auto take = [&](size_t b) {
        char* r = at;
        at += (b + 255) / 256 * 256;
        return r;
    };
    n->p = (float*)take(n->np * 4);
    n->g = (float*)take(n->np * 4);
    n->m = (float*)take(n->np * 4);
    n->v = (float*)take(n->np * 4);
    n->p16 = (uint16_t*)take(n->np * 2);
    n->h1 = (float*)take(max_batch * hidden * 4);
    n->h2 = (float*)take(max_batch * hidden * 4);
    n->dh1 = (float*)take(max_batch * hidden * 4);
    n->dh2 = (float*)take(max_batch * hidden * 4);
    n->logits = (float*)take(max_batch * out_dim * 4);
    n->dlogits = (float*)take(max_batch * out_dim * 4);
    // He-normal weights, zero biases
    std::vector<float> host(n->np, 0.0f);
    std::mt19937_64 rng(seed);
    auto fill = [&](size_t off, int rows, int cols) {
        std::normal_distribution<double> nd(0.0, std::sqrt(2.0 / cols));
        for (size_t i = 0; i < (size_t)rows * cols; ++i) host[off + i] = (float)nd(rng);
    };
    fill(n->oW1, n->hid, n->in);
    fill(n->oW2, n->hid, n->hid);
    fill(n->oW3, n->out, n->hid);
    if (int rc = magnus_mlp_set_params(n, host.data(), n->np)) {
        magnus_mlp_destroy(n);
        return rc;
    }
    *out = n;
    return MAGNUS_OK;
}

void magnus_mlp_destroy(magnus_mlp* n) {
    if (!n) return;
    (void)hipDeviceSynchronize();
    if (n->block) (void)hipFree(n->block);
    delete n;
}

uint64_t magnus_mlp_param_count(const magnus_mlp* n) { return n ? n->np : 0; }

int magnus_mlp_get_params(const magnus_mlp* n, float* host, uint64_t len) {
    if (!n || !host) return fail(MAGNUS_EINVAL, "null argument");
    if (len < n->np) return fail(MAGNUS_EINVAL, "buffer too small");
    if (hipDeviceSynchronize() != hipSuccess || hipMemcpy(host, n->p, n->np * 4, hipMemcpyDeviceToHost) != hipSuccess)
        return device_error(hipGetLastError(), "mlp params");
    return MAGNUS_OK;
}

int magnus_mlp_get_grads(const magnus_mlp* n, float* host, uint64_t len) {
    if (!n || !host) return fail(MAGNUS_EINVAL, "null argument");
    if (len < n->np) return fail(MAGNUS_EINVAL, "buffer too small");
    if (hipDeviceSynchronize() != hipSuccess || hipMemcpy(host, n->g, n->np * 4, hipMemcpyDeviceToHost) != hipSuccess)
        return device_error(hipGetLastError(), "mlp grads");
    return MAGNUS_OK;
}

int magnus_mlp_set_params(magnus_mlp* n, const float* host, uint64_t len) {
    if (!n || !host) return fail(MAGNUS_EINVAL, "null argument");
    if (len < n->np) return fail(MAGNUS_EINVAL, "buffer too small");
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(n->p, host, n->np * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(n->m, 0, n->np * 4);
    if (e == hipSuccess) e = hipMemset(n->v, 0, n->np * 4);
    if (e == hipSuccess) e = hipMemset(n->g, 0, n->np * 4);
    if (e == hipSuccess) e = magnus::launch_to_bf16(n->p, n->p16, (long long)n->np, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) return device_error(e, "mlp set params");
    n->step = 0;
    return MAGNUS_OK;
}

int magnus_mlp_forward(magnus_mlp* n, uint64_t batch, const float* x, float* logits, void* stream) {
    if (!n || !x || !logits) return fail(MAGNUS_EINVAL, "null argument");
    if (int rc = check_batch(n, batch)) return rc;
    return forward3(n, (int)batch, x, logits, (hipStream_t)stream);
}

int magnus_mlp_train_step(magnus_mlp* n, uint64_t batch, const float* x, const int32_t* labels, float lr,
                          float* loss_rows, void* stream) {
    if (!n || !x || !labels || !loss_rows) return fail(MAGNUS_EINVAL, "null argument");
    if (int rc = check_batch(n, batch)) return rc;
    if (batch == 0) return MAGNUS_OK;
    hipStream_t s = (hipStream_t)stream;
    const int B = (int)batch;
    if (int rc = forward3(n, B, x, n->logits, s)) return rc;
    hipError_t e = magnus::launch_softmax_xent(n->logits, labels, n->dlogits, loss_rows, B, n->out, s);
    magnus::LinearArgs l3 = layer(n, 3, B), l2 = layer(n, 2, B), l1 = layer(n, 1, B);
    l3.x = n->h2, l3.dy = n->dlogits, l3.dx = n->dh2, l3.xpost = n->h2;
    l2.x = n->h1, l2.dy = n->dh2, l2.dx = n->dh1, l2.xpost = n->h1;
    l1.x = x, l1.dy = n->dh1, l1.dx = nullptr;
    if (e == hipSuccess) e = magnus::launch_linear_bwd(l3, s);
    if (e == hipSuccess) e = magnus::launch_linear_bwd(l2, s);
    if (e == hipSuccess) e = magnus::launch_linear_bwd(l1, s);
    if (e != hipSuccess) return device_error(e, "mlp backward");
    n->step += 1;
    magnus::AdamArgs a{};
    a.p = n->p, a.g = n->g, a.m = n->m, a.v = n->v, a.p16 = n->p16;
    a.n = (long long)n->np;
    a.lr = lr, a.beta1 = 0.9f, a.beta2 = 0.999f, a.eps = 1e-8f;
    a.c1 = (float)(1.0 - std::pow(0.9, (double)n->step));
    a.c2 = (float)(1.0 - std::pow(0.999, (double)n->step));
    e = magnus::launch_adam(a, s);
    if (e != hipSuccess) return device_error(e, "mlp adam");
    return MAGNUS_OK;
}

uint64_t magnus_ctrl_feature_count(void) { return magnus::kCtrlFeatures; }

int magnus_ctrl_features_f64(uint64_t count, uint64_t grid, uint64_t dim, double t0, double tf, const double* in,
                             float* out, void* stream) {
    if (!in || !out) return fail(MAGNUS_EINVAL, "null argument");
    if (dim == 0 || dim > 16) return fail(MAGNUS_EINVAL, "controller features take 1 <= dim <= 16");
    if (grid < 9 || grid > 129 || ((grid - 1) & (grid - 2)) != 0)
        return fail(MAGNUS_EINVAL, "controller features take a grid of 2^k + 1 points, 9..129");
    if (count == 0) return MAGNUS_OK;
    magnus::CtrlArgs a{};
    a.in = in, a.out = out, a.B = (long long)count, a.d = (int)dim, a.G = (int)grid, a.T = tf - t0;
    hipError_t e = magnus::launch_ctrl_features(a, (hipStream_t)stream);
    if (e != hipSuccess) return device_error(e, "feature kernel launch");
    return MAGNUS_OK;
}

}  // extern "C"
