// kernels_wbd.hip -- GL fold of the dense W-basis path and its adjoint.
//
// The level-parallel kernels, run with one job per instance, shift 0 and
// weight 1, produce the basis streams W_j = Prefix(A W_{j-1}) (W_1 = Prefix(A)),
// their totals What_j and the sums Z_l = Sum(A W_l) (Z_0 = Sum(A)).  The
// reference's per-node prefixes are P^(j)(q) = W_j + s_q sum_{l<j} W_l T_{j-l}(q)
// (linearity of Prefix/Sum, associativity of the product; src/magnus.hpp:230-254
// computes them one level and node at a time), so all Q nodes come from one
// sweep, and what depends on q is d x d algebra per instance:
//   T_1 = What_1,  T_j = What_j + s_q sum_{l=1}^{j-1} What_l T_{j-l},
//   Omega_1 = Z_0, Omega_j = sum_q w_q (Z_{j-1} + s_q sum_{l=0}^{j-2} Z_l T_{j-1-l}).
// One 256-thread workgroup per instance; every thread owns the entries
// e = tid + 256 t of every matrix, so accumulations need no atomics and a
// barrier separates a matrix's last write from its first read as an operand.
#include "kernels.hpp"
#include "rules.hpp"

namespace magnus {

namespace {

constexpr int kFoldThreads = 256;

// C[e] += alpha * (X' Y')[e] for the entries this thread owns;
// X' = X or X^T, Y' = Y or Y^T (d x d, row-major)
__device__ void mat_acc(double* __restrict__ C, double alpha, const double* __restrict__ X, bool tx,
                        const double* __restrict__ Y, bool ty, int d) {
    const int m = d * d;
    for (int e = threadIdx.x; e < m; e += kFoldThreads) {
        const int r = e / d, c = e - r * d;
        double acc = 0.0;
        for (int k = 0; k < d; ++k) {
            const double xv = tx ? X[k * d + r] : X[r * d + k];
            const double yv = ty ? Y[c * d + k] : Y[k * d + c];
            acc = fma(xv, yv, acc);
        }
        C[e] = fma(alpha, acc, C[e]);
    }
}

__device__ void mat_set(double* __restrict__ C, double alpha, const double* __restrict__ X, int m) {
    for (int e = threadIdx.x; e < m; e += kFoldThreads) C[e] = X ? alpha * X[e] : 0.0;
}

__device__ void mat_axpy(double* __restrict__ C, double alpha, const double* __restrict__ X, int m) {
    for (int e = threadIdx.x; e < m; e += kFoldThreads) C[e] = fma(alpha, X[e], C[e]);
}

// T_1..T_K of node q; T_j at T + j m (T_0 unused)
__device__ void fold_totals(const WbdArgs& a, const double* What, double* T, double s) {
    const int K = a.n - 1, m = a.m, d = a.d;
    const long long M = m;
    for (int j = 1; j <= K; ++j) {
        mat_set(T + j * M, 1.0, What + (j - 1) * M, m);
        for (int l = 1; l < j; ++l) mat_acc(T + j * M, s, What + (l - 1) * M, false, T + (j - l) * M, false, d);
        __syncthreads();
    }
}

// Z_l of the instance: row l of z for l >= 1, z0 for l = 0
__device__ __forceinline__ const double* zrow(const WbdArgs& a, long long b, int l) {
    return l == 0 ? a.z0 + b * a.m : a.z + (b * a.n + l) * (long long)a.m;
}

// cotangent of Omega_j (j = 1..n) from the output cotangent, or nullptr
__device__ __forceinline__ const double* obar(const WbdArgs& a, long long b, int j) {
    const double* c = a.cot + b * (long long)a.n_out * a.m;
    if (a.op == OP_MANY) return c + (j - 1) * (long long)a.m;
    if (a.op == OP_SUM) return c;
    return j == a.n ? c : nullptr;
}

__global__ void __launch_bounds__(kFoldThreads) wbd_fold(WbdArgs a) {
    const long long b = blockIdx.x;
    const int n = a.n, K = n - 1, m = a.m, d = a.d;
    const long long M = m;
    double* work = a.work + b * (long long)(4 * K + 1) * M;
    double* T = work;                 // T_j at T + j M, j = 1..K
    double* Om = work + (K + 1) * M;  // Omega_j at Om + (j-1) M, j = 1..n
    const double* What = a.totals + b * (long long)a.n * M;
    mat_set(Om, 1.0, zrow(a, b, 0), m);
    for (int j = 2; j <= n; ++j) mat_set(Om + (j - 1) * M, 0.0, nullptr, m);
    __syncthreads();
    for (int q = 0; q < a.Q; ++q) {
        const double s = a.x[q] - 1.0, wq = a.w[q];
        fold_totals(a, What, T, s);
        for (int j = 2; j <= n; ++j) {
            double* O = Om + (j - 1) * M;
            mat_axpy(O, wq, zrow(a, b, j - 1), m);
            for (int l = 0; l <= j - 2; ++l) mat_acc(O, wq * s, zrow(a, b, l), false, T + (j - 1 - l) * M, false, d);
        }
        __syncthreads();
    }
    double* out = a.out + b * (long long)a.n_out * m;
    for (int e = threadIdx.x; e < m; e += kFoldThreads) {
        if (a.op == OP_MANY) {
            for (int j = 0; j < n; ++j) out[j * (long long)m + e] = Om[j * (long long)m + e];
        } else if (a.op == OP_SUM) {
            double v = Om[e];
            for (int j = 1; j < n; ++j) v += Om[j * (long long)m + e];
            out[e] = v;
        } else {
            out[e] = Om[(n - 1) * (long long)m + e];
        }
    }
}

// Adjoint of the fold: Zbar_0..Zbar_K and What_bar_1..What_bar_K from the
// output cotangent (reverse of wbd_fold, node by node).
__global__ void __launch_bounds__(kFoldThreads) wbd_fold_vjp(WbdArgs a) {
    const long long b = blockIdx.x;
    const int n = a.n, K = n - 1, m = a.m, d = a.d;
    const long long M = m;
    double* work = a.work + b * (long long)(4 * K + 1) * M;
    double* T = work;                  // T_j at T + j M
    double* Tb = work + (K + 1) * M;   // T_bar_j at Tb + j M
    const double* What = a.totals + b * (long long)a.n * M;
    double* Zb = a.zbar + b * (long long)a.n * M;  // row l = Zbar_l
    double* Wb = a.wbar + b * (long long)a.n * M;  // row j-1 = What_bar_j
    mat_set(Zb, 1.0, obar(a, b, 1), m);  // Omega_1 = Z_0
    for (int l = 1; l <= K; ++l) mat_set(Zb + l * M, 0.0, nullptr, m);
    for (int j = 1; j <= n; ++j) mat_set(Wb + (j - 1) * M, 0.0, nullptr, m);
    __syncthreads();
    for (int q = 0; q < a.Q; ++q) {
        const double s = a.x[q] - 1.0, wq = a.w[q];
        fold_totals(a, What, T, s);
        for (int j = 1; j <= K; ++j) mat_set(Tb + j * M, 0.0, nullptr, m);
        __syncthreads();
        for (int j = 2; j <= n; ++j) {
            const double* Ob = obar(a, b, j);
            if (!Ob) continue;
            mat_axpy(Zb + (j - 1) * M, wq, Ob, m);
            for (int l = 0; l <= j - 2; ++l) {
                mat_acc(Zb + l * M, wq * s, Ob, false, T + (j - 1 - l) * M, true, d);
                mat_acc(Tb + (j - 1 - l) * M, wq * s, zrow(a, b, l), true, Ob, false, d);
            }
        }
        __syncthreads();
        for (int j = K; j >= 1; --j) {
            mat_axpy(Wb + (j - 1) * M, 1.0, Tb + j * M, m);
            for (int l = 1; l < j; ++l) {
                mat_acc(Wb + (l - 1) * M, s, Tb + j * M, false, T + (j - l) * M, true, d);
                mat_acc(Tb + (j - l) * M, s, What + (l - 1) * M, true, Tb + j * M, false, d);
            }
            __syncthreads();
        }
    }
}

}  // namespace

hipError_t launch_wbd_fold(const WbdArgs& a, hipStream_t s) {
    if (a.B == 0) return hipSuccess;
    void* h = timer_begin("wbd_fold", s);
    hipLaunchKernelGGL(wbd_fold, dim3((unsigned)a.B), dim3(kFoldThreads), 0, s, a);
    timer_end(h);
    return hipGetLastError();
}

hipError_t launch_wbd_fold_vjp(const WbdArgs& a, hipStream_t s) {
    if (a.B == 0) return hipSuccess;
    void* h = timer_begin("wbd_fold_vjp", s);
    hipLaunchKernelGGL(wbd_fold_vjp, dim3((unsigned)a.B), dim3(kFoldThreads), 0, s, a);
    timer_end(h);
    return hipGetLastError();
}

}  // namespace magnus
