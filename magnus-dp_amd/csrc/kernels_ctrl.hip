// kernels_ctrl.hip -- resolution-controller features from a coarse sampling.
//
// Input: A(t_g) on G = 2^k + 1 uniform points of [t0, tf] (the reference's
// grid, np.linspace), [B][G][d][d] FP64, d <= 16.  Output [B][kCtrlFeatures]
// FP32, all log-scaled:
//   0  max_g ||A_g||_2 T          (spectral norm, power iteration on A^T A)
//   1  mean_g ||A_g||_2 T
//   2  max_g ||A_g||_F T
//   3  trapezoid mean of ||A_g||_F times T
//   4..19  energy fraction of harmonics 1..16 of the periodic DFT over the
//          first G-1 points (sum over entries of |A_hat_k|^2 / total energy)
//   20 energy-weighted mean harmonic (k >= 1)
//   21 max over 4 sample pairs of ||[A_i, A_j]||_F T^2
//   22 ||mean_g A_g||_F T
//   23 feature 21 / (feature 2)^2  (non-commutativity relative to size)
// One 256-thread workgroup per instance.
#include <hip/hip_runtime.h>

#include "kernels_ctrl.hpp"

namespace magnus {

namespace {

constexpr int kThreads = 256;
constexpr int kMaxG = 129;
constexpr int kHarm = 16;

// The same iteration with the dimension known at compile time, on the Gram
// matrix B = A^T A formed once (its upper triangle from the sample in memory,
// mirrored): one D x D matrix-vector product per iteration instead of two,
// with ||A v|| = sqrt(v . B v).  B and the vectors stay in registers (a
// runtime dimension would index them in scratch).
template <int D>
__device__ double spectral_norm_d(const double* __restrict__ Ag) {
    double B[D * D], v[D], u[D];
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int k = i; k < D; ++k) {
            double acc = 0.0;
#pragma unroll
            for (int r = 0; r < D; ++r) acc = fma(Ag[r * D + i], Ag[r * D + k], acc);
            B[i * D + k] = acc;
            B[k * D + i] = acc;
        }
#pragma unroll
    for (int i = 0; i < D; ++i) v[i] = 1.0 + 0.01 * i;
    double lam = 0.0;
    for (int it = 0; it < 60; ++it) {
        double nv = 0.0, wn = 0.0;
#pragma unroll
        for (int i = 0; i < D; ++i) {
            double acc = 0.0;
#pragma unroll
            for (int k = 0; k < D; ++k) acc = fma(B[i * D + k], v[k], acc);
            u[i] = acc;  // u = A^T A v
            nv = fma(acc, acc, nv);
            wn = fma(v[i], acc, wn);  // v . A^T A v = ||A v||^2
        }
        nv = sqrt(nv);
        if (nv == 0.0) return 0.0;
        lam = sqrt(fmax(wn, 0.0));  // ||A v_prev|| with ||v_prev|| = 1 after the first step
        const double inv = 1.0 / nv;  // one division per iteration, not D
#pragma unroll
        for (int k = 0; k < D; ++k) v[k] = u[k] * inv;
    }
    return lam;
}

__device__ double spectral_norm(const double* A, int d) {
    // power iteration on A^T A from a fixed start with all components nonzero
    double v[16], w[16];
    for (int i = 0; i < d; ++i) v[i] = 1.0 + 0.01 * i;
    double lam = 0.0;
    for (int it = 0; it < 60; ++it) {
        double nv = 0.0;
        for (int i = 0; i < d; ++i) {
            double acc = 0.0;
            for (int k = 0; k < d; ++k) acc = fma(A[i * d + k], v[k], acc);
            w[i] = acc;  // w = A v
        }
        for (int k = 0; k < d; ++k) {
            double acc = 0.0;
            for (int i = 0; i < d; ++i) acc = fma(A[i * d + k], w[i], acc);
            v[k] = acc;  // v = A^T A v
            nv = fma(acc, acc, nv);
        }
        nv = sqrt(nv);
        if (nv == 0.0) return 0.0;
        double wn = 0.0;
        for (int i = 0; i < d; ++i) wn = fma(w[i], w[i], wn);
        lam = sqrt(wn);  // ||A v_prev|| with ||v_prev|| = 1 after the first step
        const double inv = 1.0 / nv;
        for (int k = 0; k < d; ++k) v[k] *= inv;
    }
    return lam;
}

// D > 0: the dimension as a compile-time constant (the controller's d = 8)
template <int D>
__global__ void __launch_bounds__(kThreads) ctrl_features(CtrlArgs a) {
    const long long b = blockIdx.x;
    const int d = a.d, m = d * d, G = a.G, tid = threadIdx.x;
    const double T = a.T;
    const double* A = a.in + b * (long long)G * m;
    __shared__ double s_spec[kMaxG], s_fro[kMaxG];
    __shared__ double s_energy[kHarm + 1];
    __shared__ double s_total;
    __shared__ double s_comm[4];
    __shared__ double s_mean[256];
    __shared__ double s_cs[kMaxG], s_sn[kMaxG];  // twiddles of the periodic DFT
    const int N = G - 1;
    // per-sample norms
    for (int g = tid; g < G; g += kThreads) {
        const double* Ag = A + (long long)g * m;
        double f = 0.0;
        for (int e = 0; e < m; ++e) f = fma(Ag[e], Ag[e], f);
        s_fro[g] = sqrt(f);
        s_spec[g] = D > 0 ? spectral_norm_d<(D > 0 ? D : 1)>(Ag) : spectral_norm(Ag, d);
    }
    for (int j = tid; j < N; j += kThreads) sincospi(2.0 * (double)j / N, &s_sn[j], &s_cs[j]);
    // harmonic energies of the periodic DFT over N = G - 1 points
    if (tid <= kHarm) s_energy[tid] = 0.0;
    if (tid == 0) s_total = 0.0;
    __syncthreads();
    for (int p = tid; p < kHarm * m; p += kThreads) {
        const int k = 1 + p / m, e = p - (k - 1) * m;
        double re = 0.0, im = 0.0;
        for (int i = 0; i < N; ++i) {
            const int j = (k * i) & (N - 1);  // k i mod N, N a power of two
            const double sn = s_sn[j], cs = s_cs[j];
            const double x = A[(long long)i * m + e];
            re = fma(x, cs, re);
            im = fma(-x, sn, im);
        }
        atomicAdd(&s_energy[k], re * re + im * im);
    }
    for (int e = tid; e < m; e += kThreads) {
        double mean = 0.0, tot = 0.0;
        for (int i = 0; i < N; ++i) {
            const double x = A[(long long)i * m + e];
            mean += x;
            tot = fma(x, x, tot);
        }
        s_mean[e] = mean / N;
        atomicAdd(&s_total, tot * N);  // Parseval: sum_k |X_k|^2 = N sum_i |x_i|^2
    }
    // commutators of 4 sample pairs
    if (tid < 4) {
        const int i0 = tid < 2 ? 0 : (tid == 2 ? N / 4 : N / 8);
        const int j0 = tid == 0 ? N / 4 : (tid == 1 ? N / 2 : (tid == 2 ? 3 * N / 4 : 5 * N / 8));
        const double* X = A + (long long)i0 * m;
        const double* Y = A + (long long)j0 * m;
        double c = 0.0;
        for (int r = 0; r < d; ++r)
            for (int col = 0; col < d; ++col) {
                double xy = 0.0, yx = 0.0;
                for (int k = 0; k < d; ++k) {
                    xy = fma(X[r * d + k], Y[k * d + col], xy);
                    yx = fma(Y[r * d + k], X[k * d + col], yx);
                }
                const double df = xy - yx;
                c = fma(df, df, c);
            }
        s_comm[tid] = sqrt(c);
    }
    __syncthreads();
    if (tid != 0) return;
    double smax = 0.0, ssum = 0.0, frob_max = 0.0, ftrap = 0.0;
    for (int g = 0; g < G; ++g) {
        smax = fmax(smax, s_spec[g]);
        ssum += s_spec[g];
        frob_max = fmax(frob_max, s_fro[g]);
        ftrap += (g == 0 || g == G - 1) ? 0.5 * s_fro[g] : s_fro[g];
    }
    ftrap /= (G - 1);
    double eh = 0.0, kmean = 0.0;
    for (int k = 1; k <= kHarm; ++k) {
        eh += s_energy[k];
        kmean += k * s_energy[k];
    }
    double cmax = 0.0;
    for (int i = 0; i < 4; ++i) cmax = fmax(cmax, s_comm[i]);
    double mfro = 0.0;
    for (int e = 0; e < m; ++e) mfro = fma(s_mean[e], s_mean[e], mfro);
    mfro = sqrt(mfro);
    const double tiny = 1e-30;
    float* out = a.out + b * kCtrlFeatures;
    out[0] = (float)log(tiny + smax * T);
    out[1] = (float)log(tiny + ssum / G * T);
    out[2] = (float)log(tiny + frob_max * T);
    out[3] = (float)log(tiny + ftrap * T);
    for (int k = 1; k <= kHarm; ++k) out[3 + k] = (float)log(1e-12 + s_energy[k] / (tiny + s_total));
    out[20] = (float)log(1e-12 + kmean / (tiny + eh));
    out[21] = (float)log(tiny + cmax * T * T);
    out[22] = (float)log(tiny + mfro * T);
    out[23] = (float)log(tiny + cmax / (tiny + frob_max * frob_max));
}

}  // namespace

hipError_t launch_ctrl_features(const CtrlArgs& a, hipStream_t s) {
    if (a.B == 0) return hipSuccess;
    if (a.d < 1 || a.d > 16 || a.G < 9 || a.G > kMaxG || ((a.G - 1) & (a.G - 2)) != 0) return hipErrorInvalidValue;
    if (a.d == 8)
        hipLaunchKernelGGL(ctrl_features<8>, dim3((unsigned)a.B), dim3(kThreads), 0, s, a);
    else
        hipLaunchKernelGGL(ctrl_features<0>, dim3((unsigned)a.B), dim3(kThreads), 0, s, a);
    return hipGetLastError();
}

}  // namespace magnus
