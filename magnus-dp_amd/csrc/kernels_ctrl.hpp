// kernels_ctrl.hpp -- resolution-controller feature kernel and the MLP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace magnus {

constexpr int kCtrlFeatures = 24;

struct CtrlArgs {
    const double* in;  // [B][G][d][d]
    float* out;        // [B][kCtrlFeatures]
    long long B;
    int d, G;
    double T;  // tf - t0
};
hipError_t launch_ctrl_features(const CtrlArgs& a, hipStream_t s);

// Dense layer kernels of the controller MLP (kernels_mlp.hip).  Weights are
// stored [out][in]; W16 is their bfloat16 copy, read by the forward and the
// data backward; accumulation is FP32.
struct LinearArgs {
    const float* x;         // [B][in]
    const uint16_t* w16;    // [out][in] bfloat16
    const float* bias;      // [out]
    float* y;               // [B][out]
    const float* dy;        // [B][out]     (backward)
    float* dx;              // [B][in]      (backward data), or nullptr
    const float* xpost;     // [B][in] activation whose ReLU mask gates dx, or nullptr
    float* dw;              // [out][in]    (backward weights)
    float* db;              // [out]
    int B, in, out;
    int relu;               // forward: apply ReLU
};
hipError_t launch_linear_fwd(const LinearArgs& a, hipStream_t s);
hipError_t launch_linear_bwd(const LinearArgs& a, hipStream_t s);

// softmax cross-entropy over rows of logits; dlogits = (softmax - onehot) / B;
// loss[b] per row
hipError_t launch_softmax_xent(const float* logits, const int32_t* labels, float* dlogits, float* loss, int B,
                               int C, hipStream_t s);

// Adam on FP32 master parameters, refreshing their bfloat16 copy (w16 may be
// nullptr for parameters without one)
struct AdamArgs {
    float* p;
    const float* g;
    float* m;
    float* v;
    uint16_t* p16;
    long long n;
    float lr, beta1, beta2, eps, c1, c2;  // c1 = 1 - beta1^t, c2 = 1 - beta2^t
};
hipError_t launch_adam(const AdamArgs& a, hipStream_t s);
hipError_t launch_to_bf16(const float* p, uint16_t* p16, long long n, hipStream_t s);

}  // namespace magnus
