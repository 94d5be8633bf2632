// kernels_mlp.hip -- dense layers, loss and optimizer of the resolution
// controller's MLP.  Weights live as FP32 masters (the optimizer's copy) and
// bfloat16 copies (what the forward and the data backward read); activations,
// gradients and every accumulation are FP32.  The layers are small (widths up
// to a few hundred, batches of ~1k), so these are plain VALU kernels: one
// thread per output element with the shared operand staged in LDS.
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "kernels_ctrl.hpp"

namespace magnus {

namespace {

__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// round to nearest even; NaN stays a quiet NaN
__device__ __forceinline__ uint16_t f32_to_bf16(float f) {
    const uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (uint16_t)((u >> 16) | 0x0040u);
    const uint32_t bias = 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)((u + bias) >> 16);
}

constexpr int kRows = 4;  // batch rows per workgroup in the forward / data backward

// y[b][o] = act(bias[o] + sum_i x[b][i] w[o][i]); workgroup = kRows rows x all outputs
__global__ void __launch_bounds__(256) linear_fwd(LinearArgs a) {
    extern __shared__ float xs[];  // [kRows][in]
    const int b0 = blockIdx.x * kRows;
    for (int t = threadIdx.x; t < kRows * a.in; t += blockDim.x) {
        const int r = t / a.in, i = t - r * a.in;
        xs[t] = (b0 + r < a.B) ? a.x[(long long)(b0 + r) * a.in + i] : 0.0f;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < kRows * a.out; t += blockDim.x) {
        const int r = t / a.out, o = t - r * a.out;
        if (b0 + r >= a.B) continue;
        const uint16_t* w = a.w16 + (long long)o * a.in;
        const float* x = xs + r * a.in;
        float acc = 0.0f;
        for (int i = 0; i < a.in; ++i) acc = fmaf(x[i], bf16_to_f32(w[i]), acc);
        acc += a.bias[o];
        if (a.relu) acc = fmaxf(acc, 0.0f);
        a.y[(long long)(b0 + r) * a.out + o] = acc;
    }
}

// dx[b][i] = (sum_o dy[b][o] w[o][i]) * [xpost[b][i] > 0]
__global__ void __launch_bounds__(256) linear_bwd_data(LinearArgs a) {
    extern __shared__ float ds[];  // [kRows][out]
    const int b0 = blockIdx.x * kRows;
    for (int t = threadIdx.x; t < kRows * a.out; t += blockDim.x) {
        const int r = t / a.out, o = t - r * a.out;
        ds[t] = (b0 + r < a.B) ? a.dy[(long long)(b0 + r) * a.out + o] : 0.0f;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < kRows * a.in; t += blockDim.x) {
        const int r = t / a.in, i = t - r * a.in;
        if (b0 + r >= a.B) continue;
        const float* dy = ds + r * a.out;
        float acc = 0.0f;
        for (int o = 0; o < a.out; ++o) acc = fmaf(dy[o], bf16_to_f32(a.w16[(long long)o * a.in + i]), acc);
        const long long idx = (long long)(b0 + r) * a.in + i;
        if (a.xpost && !(a.xpost[idx] > 0.0f)) acc = 0.0f;
        a.dx[idx] = acc;
    }
}

// dw[o][i] = sum_b dy[b][o] x[b][i], db[o] = sum_b dy[b][o]; one thread per
// (o, i) (i = in for the bias), rows b in order
__global__ void __launch_bounds__(256) linear_bwd_weight(LinearArgs a) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long per = (long long)a.in + 1;
    if (t >= (long long)a.out * per) return;
    const int o = (int)(t / per), i = (int)(t - (long long)o * per);
    float acc = 0.0f;
    if (i < a.in) {
        for (int b = 0; b < a.B; ++b) acc = fmaf(a.dy[(long long)b * a.out + o], a.x[(long long)b * a.in + i], acc);
        a.dw[(long long)o * a.in + i] = acc;
    } else {
        for (int b = 0; b < a.B; ++b) acc += a.dy[(long long)b * a.out + o];
        a.db[o] = acc;
    }
}

__global__ void __launch_bounds__(256) softmax_xent(const float* logits, const int32_t* labels, float* dlogits,
                                                   float* loss, int B, int C) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const float* l = logits + (long long)b * C;
    float mx = l[0];
    for (int c = 1; c < C; ++c) mx = fmaxf(mx, l[c]);
    float se = 0.0f;
    for (int c = 0; c < C; ++c) se += expf(l[c] - mx);
    const int y = labels[b];
    loss[b] = logf(se) - (l[y] - mx);
    const float inv = 1.0f / ((float)B * se);
    float* g = dlogits + (long long)b * C;
    for (int c = 0; c < C; ++c) g[c] = expf(l[c] - mx) * inv - (c == y ? 1.0f / (float)B : 0.0f);
}

__global__ void __launch_bounds__(256) adam_kernel(AdamArgs a) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const float g = a.g[i];
    const float m = a.beta1 * a.m[i] + (1.0f - a.beta1) * g;
    const float v = a.beta2 * a.v[i] + (1.0f - a.beta2) * g * g;
    a.m[i] = m;
    a.v[i] = v;
    const float p = a.p[i] - a.lr * (m / a.c1) / (sqrtf(v / a.c2) + a.eps);
    a.p[i] = p;
    if (a.p16) a.p16[i] = f32_to_bf16(p);
}

__global__ void __launch_bounds__(256) to_bf16(const float* p, uint16_t* p16, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p16[i] = f32_to_bf16(p[i]);
}

}  // namespace

hipError_t launch_linear_fwd(const LinearArgs& a, hipStream_t s) {
    if (a.B == 0) return hipSuccess;
    const size_t lds = (size_t)kRows * a.in * sizeof(float);
    if (lds > 64 * 1024) return hipErrorInvalidValue;
    hipLaunchKernelGGL(linear_fwd, dim3((a.B + kRows - 1) / kRows), dim3(256), lds, s, a);
    return hipGetLastError();
}

hipError_t launch_linear_bwd(const LinearArgs& a, hipStream_t s) {
    if (a.B == 0) return hipSuccess;
    if (a.dx) {
        const size_t lds = (size_t)kRows * a.out * sizeof(float);
        if (lds > 64 * 1024) return hipErrorInvalidValue;
        hipLaunchKernelGGL(linear_bwd_data, dim3((a.B + kRows - 1) / kRows), dim3(256), lds, s, a);
    }
    const long long nt = (long long)a.out * (a.in + 1);
    hipLaunchKernelGGL(linear_bwd_weight, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_softmax_xent(const float* logits, const int32_t* labels, float* dlogits, float* loss, int B, int C,
                               hipStream_t s) {
    if (B == 0) return hipSuccess;
    hipLaunchKernelGGL(softmax_xent, dim3((B + 255) / 256), dim3(256), 0, s, logits, labels, dlogits, loss, B, C);
    return hipGetLastError();
}

hipError_t launch_adam(const AdamArgs& a, hipStream_t s) {
    if (a.n == 0) return hipSuccess;
    hipLaunchKernelGGL(adam_kernel, dim3((unsigned)((a.n + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_to_bf16(const float* p, uint16_t* p16, long long n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(to_bf16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, p16, n);
    return hipGetLastError();
}

}  // namespace magnus
