// kernels_sample.hip -- A(t) sampled on the device for the Fourier forms of the
// benchmark workloads, on the reference's grid t_i = np.linspace(t0, tf, S)
// (src/api/magnus/__init__.py:58-85 samples a Python callable there).
//
//   matrix form:      A(t) = C_0 + sum_{k=1..M} ( sin(2 pi k t) C_{2k-1} + cos(2 pi k t) C_{2k} )
//   SpaceCurve form:  a_c(t) = sum_{k=1..M} amp_{k,c} sin(2 pi k t + phase_{k,c})
//
// The grid and every product and sum are rounded separately in numpy's order
// (t_i = i*step + t0, the last point set to tf; ((2 pi) k) t; terms added in k
// order), so the samples differ from a numpy evaluation only by the sin/cos
// implementations (a few ulp).
#include "kernels_sample.hpp"

namespace magnus {

namespace {

constexpr double kTwoPi = 6.283185307179586;  // 2 * np.pi

__device__ __forceinline__ double grid_t(long long i, long long S, double t0, double tf, double step) {
    if (i == S - 1) return tf;
    return __dadd_rn(__dmul_rn((double)i, step), t0);
}

// one workgroup handles SB consecutive samples of one instance: the sin/cos
// table of its samples is built once in LDS, then the entries are filled
constexpr int kSampleMaxModes = 32;
constexpr int kSampleThreads = 256;

// STAGE: the instance's (2M+1) coefficient matrices fit kSampleStage doubles
// and are copied to LDS once per workgroup (every entry reads 2M+1 of them per
// sample); otherwise they are read from global memory.
constexpr int kSampleStage = 2048;

template <bool STAGE>
__global__ void __launch_bounds__(kSampleThreads) sample_fourier_kernel(SampleArgs a) {
    __shared__ double tab[kSampleThreads * 2];  // [sample][mode][sin, cos], SB * modes <= 256
    __shared__ double cst[STAGE ? kSampleStage : 1];
    const int m = a.m, modes = a.modes;
    const int SB = a.sb;
    const long long blocks_per_inst = (a.S + SB - 1) / SB;
    const long long b = blockIdx.x / blocks_per_inst;
    const long long i0 = (blockIdx.x - b * blocks_per_inst) * SB;
    const double* Cg = a.coeffs + b * (long long)(2 * modes + 1) * m;
    if (STAGE)
        for (int x = threadIdx.x; x < (2 * modes + 1) * m; x += blockDim.x) cst[x] = Cg[x];
    for (int x = threadIdx.x; x < SB * modes; x += blockDim.x) {
        const int si = x / modes, k = x - si * modes + 1;
        const long long i = i0 + si;
        double sv = 0.0, cv = 0.0;
        if (i < a.S) {
            const double t = grid_t(i, a.S, a.t0, a.tf, a.step);
            const double arg = __dmul_rn(__dmul_rn(kTwoPi, (double)k), t);
            sv = sin(arg);
            cv = cos(arg);
        }
        tab[2 * x] = sv;
        tab[2 * x + 1] = cv;
    }
    __syncthreads();
    const double* C = STAGE ? cst : Cg;
    for (int x = threadIdx.x; x < SB * m; x += blockDim.x) {
        const int si = x / m, e = x - si * m;
        const long long i = i0 + si;
        if (i >= a.S) continue;
        double v = C[e];
        for (int k = 1; k <= modes; ++k) {
            const double* tk = tab + 2 * (si * modes + k - 1);
            v = __dadd_rn(v, __dmul_rn(tk[0], C[(2 * k - 1) * m + e]));
            v = __dadd_rn(v, __dmul_rn(tk[1], C[(2 * k) * m + e]));
        }
        a.out[(b * a.S + i) * m + e] = v;
    }
}

// one thread per (instance, sample, component)
__global__ void __launch_bounds__(256) sample_sc_fourier_kernel(SampleArgs a) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= a.B * a.S * 3) return;
    const long long b = idx / (a.S * 3);
    const long long rem = idx - b * a.S * 3;
    const long long i = rem / 3;
    const int c = (int)(rem - i * 3);
    const double t = grid_t(i, a.S, a.t0, a.tf, a.step);
    const double* amp = a.amp + b * (long long)a.modes * 3;
    const double* ph = a.phase + b * (long long)a.modes * 3;
    double v = 0.0;
    for (int k = 1; k <= a.modes; ++k) {
        const double arg = __dadd_rn(__dmul_rn(__dmul_rn(kTwoPi, (double)k), t), ph[(k - 1) * 3 + c]);
        v = __dadd_rn(v, __dmul_rn(amp[(k - 1) * 3 + c], sin(arg)));
    }
    a.out[idx] = v;
}

}  // namespace

int sample_max_modes() { return kSampleMaxModes; }

hipError_t launch_sample_fourier(SampleArgs a, hipStream_t s) {
    // samples per workgroup: about eight entries per thread (a workgroup of only
    // 256 entries spent most of its time in the sin/cos table, built by a few
    // lanes), keeping the table within 256 (sample, mode) pairs
    int sb = a.m >= kSampleThreads ? 1 : (8 * kSampleThreads) / a.m;
    const int sb_tab = kSampleThreads / (a.modes > 0 ? a.modes : 1);
    if (sb > sb_tab) sb = sb_tab;
    if (sb < 1) sb = 1;
    a.sb = sb;
    const long long blocks = a.B * ((a.S + sb - 1) / sb);
    void* h = timer_begin("sample_fourier", s);
    if ((long long)(2 * a.modes + 1) * a.m <= kSampleStage)
        hipLaunchKernelGGL(sample_fourier_kernel<true>, dim3(blocks), dim3(kSampleThreads), 0, s, a);
    else
        hipLaunchKernelGGL(sample_fourier_kernel<false>, dim3(blocks), dim3(kSampleThreads), 0, s, a);
    timer_end(h);
    return hipGetLastError();
}

hipError_t launch_sample_sc_fourier(const SampleArgs& a, hipStream_t s) {
    const long long total = a.B * a.S * 3;
    void* h = timer_begin("sample_sc_fourier", s);
    hipLaunchKernelGGL(sample_sc_fourier_kernel, dim3((total + 255) / 256), dim3(256), 0, s, a);
    timer_end(h);
    return hipGetLastError();
}

}  // namespace magnus
