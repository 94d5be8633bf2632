// kernels_sample.hip -- A(t) sampled on the device for the Fourier forms of the
// benchmark workloads, on the reference's grid t_i = np.linspace(t0, tf, S)
// (src/api/magnus/__init__.py:58-85 samples a Python callable there).
//
//   matrix form:      A(t) = C_0 + sum_{k=1..M} ( sin(2 pi k t) C_{2k-1} + cos(2 pi k t) C_{2k} )
//   SpaceCurve form:  a_c(t) = sum_{k=1..M} amp_{k,c} sin(2 pi k t + phase_{k,c})
//
// The grid and the sin/cos arguments are rounded separately in numpy's order
// (t_i = i*step + t0, the last point set to tf; ((2 pi) k) t), and the
// matrix form's terms are accumulated in k order, each with one fused
// multiply-add (C_0, then + sin_k C_{2k-1}, then + cos_k C_{2k}), so the
// samples differ from a numpy evaluation by the sin/cos implementations and
// the fused roundings (a few ulp).  The d = 8 sweeps that form samples from
// the coefficients (kernels_wbd8.hip) use this table and accumulation order.
//
// No other contraction: the pragma precedes the includes so that it also
// covers the inline arithmetic helpers the headers define.
#pragma clang fp contract(off)

#include "kernels_sample.hpp"

namespace magnus {

namespace {

constexpr double kTwoPi = 6.283185307179586;  // 2 * np.pi

__device__ __forceinline__ double grid_t(long long i, long long S, double t0, double tf, double step) {
    if (i == S - 1) return tf;
    return __dadd_rn(__dmul_rn((double)i, step), t0);
}

// One workgroup handles SB consecutive samples of IB consecutive instances:
// the sin/cos table of its samples (which depends on the grid and the modes
// only) is built once in LDS, then the entries of each instance are filled in
// turn.  The table is the expensive part (two FP64 transcendentals per sample
// and mode), so sharing it across instances leaves the kernel bound by its
// output writes.
constexpr int kSampleMaxModes = 32;
constexpr int kSampleThreads = 256;

// STAGE: the instance's (2M+1) coefficient matrices fit kSampleStage doubles
// and are copied to LDS once per instance (every entry reads 2M+1 of them per
// sample); otherwise they are read from global memory.
constexpr int kSampleStage = 2048;

template <bool STAGE>
__global__ void __launch_bounds__(kSampleThreads) sample_fourier_kernel(SampleArgs a) {
    __shared__ double tab[kSampleThreads * 2];  // [sample][mode][sin, cos], SB * modes <= 256
    __shared__ double cst[STAGE ? kSampleStage : 1];
    const int m = a.m, modes = a.modes;
    const int SB = a.sb, IB = a.ib;
    const long long blocks_per_group = (a.S + SB - 1) / SB;
    const long long g = blockIdx.x / blocks_per_group;
    const long long i0 = (blockIdx.x - g * blocks_per_group) * SB;
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
    for (int bi = 0; bi < IB; ++bi) {
        const long long b = g * IB + bi;
        if (b >= a.B) break;
        const double* Cg = a.coeffs + b * (long long)(2 * modes + 1) * m;
        if (STAGE) {
            __syncthreads();  // the previous instance's coefficients are read
            for (int x = threadIdx.x; x < (2 * modes + 1) * m; x += blockDim.x) cst[x] = Cg[x];
        }
        __syncthreads();  // the table (and the staged coefficients) are in LDS
        const double* C = STAGE ? cst : Cg;
        for (int x = threadIdx.x; x < SB * m; x += blockDim.x) {
            const int si = x / m, e = x - si * m;
            const long long i = i0 + si;
            if (i >= a.S) continue;
            double v = C[e];
            for (int k = 1; k <= modes; ++k) {
                const double* tk = tab + 2 * (si * modes + k - 1);
                v = __fma_rn(tk[0], C[(2 * k - 1) * m + e], v);
                v = __fma_rn(tk[1], C[(2 * k) * m + e], v);
            }
            a.out[(b * a.S + i) * m + e] = v;
        }
    }
}

// The same for m dividing the workgroup size and at most kRegModes modes:
// thread x keeps entry e = x % m of the instance's 2M + 1 coefficient
// matrices in registers and fills that entry of samples x / m, x / m + 256 / m,
// ..., so that only the table is read from LDS (a broadcast within each
// wavefront); the per-entry arithmetic is the generic kernel's.
constexpr int kRegModes = 8;

__global__ void __launch_bounds__(kSampleThreads) sample_fourier_reg(SampleArgs a) {
    __shared__ double tab[kSampleThreads * 2];
    const int m = a.m, modes = a.modes;
    const int SB = a.sb, IB = a.ib;
    const long long blocks_per_group = (a.S + SB - 1) / SB;
    const long long g = blockIdx.x / blocks_per_group;
    const long long i0 = (blockIdx.x - g * blocks_per_group) * SB;
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
    const int e = threadIdx.x % m, s0 = threadIdx.x / m, sstride = kSampleThreads / m;
    for (int bi = 0; bi < IB; ++bi) {
        const long long b = g * IB + bi;
        if (b >= a.B) break;
        const double* Cg = a.coeffs + b * (long long)(2 * modes + 1) * m + e;
        double c[2 * kRegModes + 1];
#pragma unroll
        for (int q = 0; q < 2 * kRegModes + 1; ++q) c[q] = q < 2 * modes + 1 ? Cg[(long long)q * m] : 0.0;
        for (int si = s0; si < SB; si += sstride) {
            const long long i = i0 + si;
            if (i >= a.S) break;
            double v = c[0];
#pragma unroll
            for (int k = 1; k <= kRegModes; ++k) {
                if (k > modes) break;
                const double* tk = tab + 2 * (si * modes + k - 1);
                v = __fma_rn(tk[0], c[2 * k - 1], v);
                v = __fma_rn(tk[1], c[2 * k], v);
            }
            a.out[(b * a.S + i) * m + e] = v;
        }
    }
}

// one thread per (sample, mode)
__global__ void __launch_bounds__(256) fourier_table_kernel(SampleArgs a, double* out) {
    const long long x = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= a.S * a.modes) return;
    const long long i = x / a.modes;
    const int k = (int)(x - i * a.modes) + 1;
    const double t = grid_t(i, a.S, a.t0, a.tf, a.step);
    const double arg = __dmul_rn(__dmul_rn(kTwoPi, (double)k), t);
    out[2 * x] = sin(arg);
    out[2 * x + 1] = cos(arg);
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
    // instances per workgroup: eight share a table once the batch still gives
    // thousands of workgroups
    const long long sblocks = (a.S + sb - 1) / sb;
    a.ib = a.B * sblocks >= 8 * 4096 ? 8 : 1;
    const long long blocks = ((a.B + a.ib - 1) / a.ib) * sblocks;
    void* h = timer_begin("sample_fourier", s);
    if (a.m <= kSampleThreads && kSampleThreads % a.m == 0 && a.modes <= kRegModes)
        hipLaunchKernelGGL(sample_fourier_reg, dim3(blocks), dim3(kSampleThreads), 0, s, a);
    else if ((long long)(2 * a.modes + 1) * a.m <= kSampleStage)
        hipLaunchKernelGGL(sample_fourier_kernel<true>, dim3(blocks), dim3(kSampleThreads), 0, s, a);
    else
        hipLaunchKernelGGL(sample_fourier_kernel<false>, dim3(blocks), dim3(kSampleThreads), 0, s, a);
    timer_end(h);
    return hipGetLastError();
}

hipError_t launch_fourier_table(const SampleArgs& a, double* out, hipStream_t s) {
    const long long total = a.S * a.modes;
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(fourier_table_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, a, out);
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
