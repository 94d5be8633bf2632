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
// One 256-thread workgroup per (instance, node q) writes node q's partial
// results; a reduce kernel adds the nodes in fixed q order.  Products stage
// both operands in LDS (transposed on the way in) and, for d in {16, 32, 48,
// 64}, each thread computes a (d/16) x (d/16) register tile; other d take
// one entry per thread, and d > 64 reads the operands from global memory.
// d <= 16 runs an instantiation with 16 x 16 staging (4 KB of LDS), so that
// eight of these latency-bound workgroups share a CU instead of two.
// Every thread owns the same entries of every matrix it accumulates, so the
// accumulations need no atomics; a barrier separates a matrix's last write
// from its first read as an operand.
#include "kernels.hpp"
#include "rules.hpp"

namespace magnus {

namespace {

constexpr int kFoldThreads = 256;
constexpr int kFoldLds = 64;       // largest d staged in LDS
constexpr int kFoldLdsSmall = 16;  // staging capacity of the d <= 16 instantiation (4 KB instead of 64 KB)

// CAP: largest d staged in LDS.  A workgroup holds two CAP x CAP operands, so
// the d <= 16 instantiation fits eight workgroups per CU where the d <= 64 one
// fits two; the arithmetic is the same for every d <= CAP.
template <int CAP>
struct FoldLds {
    double x[CAP * CAP];
    double y[CAP * CAP];
};

// C[e] += alpha * (X' Y')[e] for the entries this thread owns (X' = X or X^T,
// Y' = Y or Y^T; d x d row-major).  Ends with a barrier: C and the staging
// buffers are free for the next call.
template <int CAP>
__device__ void mat_acc(FoldLds<CAP>& L, double* __restrict__ C, double alpha, const double* __restrict__ X, bool tx,
                        const double* __restrict__ Y, bool ty, int d) {
    const int m = d * d, tid = threadIdx.x;
    if (d > CAP) {
        for (int e = tid; e < m; e += kFoldThreads) {
            const int r = e / d, c = e - r * d;
            double acc = 0.0;
            for (int k = 0; k < d; ++k)
                acc = fma(tx ? X[k * d + r] : X[r * d + k], ty ? Y[c * d + k] : Y[k * d + c], acc);
            C[e] = fma(alpha, acc, C[e]);
        }
        __syncthreads();
        return;
    }
    // stage X' and Y' row-major
    for (int e = tid; e < m; e += kFoldThreads) {
        const int r = e / d, c = e - r * d;
        L.x[e] = tx ? X[c * d + r] : X[e];
        L.y[e] = ty ? Y[c * d + r] : Y[e];
    }
    __syncthreads();
    if ((d & 15) == 0) {
        const int t = d >> 4;  // register tile t x t, threads on a 16 x 16 grid
        const int r0 = (tid >> 4) * t, c0 = (tid & 15) * t;
        double acc[4][4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = 0.0;
        for (int k = 0; k < d; ++k) {
            double xv[4], yv[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                xv[i] = i < t ? L.x[(r0 + i) * d + k] : 0.0;
                yv[i] = i < t ? L.y[k * d + c0 + i] : 0.0;
            }
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = fma(xv[i], yv[j], acc[i][j]);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (i < t && j < t) {
                    double* c = C + (r0 + i) * d + c0 + j;
                    *c = fma(alpha, acc[i][j], *c);
                }
    } else {
        for (int e = tid; e < m; e += kFoldThreads) {
            const int r = e / d, c = e - r * d;
            double acc = 0.0;
            for (int k = 0; k < d; ++k) acc = fma(L.x[r * d + k], L.y[k * d + c], acc);
            C[e] = fma(alpha, acc, C[e]);
        }
    }
    __syncthreads();
}

// Entry ownership of mat_set / mat_axpy must match mat_acc's: the tiled
// layout gives thread tid the entries (r0 + i, c0 + j).
template <int CAP>
__device__ __forceinline__ int owned_count(int d) { return ((d & 15) == 0 && d <= CAP) ? (d >> 4) * (d >> 4) : 0; }
__device__ __forceinline__ int owned_entry(int d, int u) {
    const int t = d >> 4, tid = threadIdx.x;
    return ((tid >> 4) * t + u / t) * d + (tid & 15) * t + u % t;
}

template <int CAP>
__device__ void mat_set(double* __restrict__ C, double alpha, const double* __restrict__ X, int d) {
    const int m = d * d, oc = owned_count<CAP>(d);
    if (oc) {
        for (int u = 0; u < oc; ++u) {
            const int e = owned_entry(d, u);
            C[e] = X ? alpha * X[e] : 0.0;
        }
        return;
    }
    for (int e = threadIdx.x; e < m; e += kFoldThreads) C[e] = X ? alpha * X[e] : 0.0;
}

template <int CAP>
__device__ void mat_axpy(double* __restrict__ C, double alpha, const double* __restrict__ X, int d) {
    const int m = d * d, oc = owned_count<CAP>(d);
    if (oc) {
        for (int u = 0; u < oc; ++u) {
            const int e = owned_entry(d, u);
            C[e] = fma(alpha, X[e], C[e]);
        }
        return;
    }
    for (int e = threadIdx.x; e < m; e += kFoldThreads) C[e] = fma(alpha, X[e], C[e]);
}

// T_1..T_K of node q; T_j at T + j m (T_0 unused)
template <int CAP>
__device__ void fold_totals(FoldLds<CAP>& L, const WbdArgs& a, const double* What, double* T, double s) {
    const int K = a.n - 1, d = a.d;
    const long long M = a.m;
    for (int j = 1; j <= K; ++j) {
        mat_set<CAP>(T + j * M, 1.0, What + (j - 1) * M, d);
        for (int l = 1; l < j; ++l) mat_acc(L, T + j * M, s, What + (l - 1) * M, false, T + (j - l) * M, false, d);
        __syncthreads();
    }
}

// Z_l of the instance: row l of z for l >= 1, z0 for l = 0
__device__ __forceinline__ int row_stride(const WbdArgs& a) { return a.rows ? a.rows : a.n; }

__device__ __forceinline__ const double* zrow(const WbdArgs& a, long long b, int l) {
    return l == 0 ? a.z0 + b * a.m : a.z + (b * row_stride(a) + l) * (long long)a.m;
}

// cotangent of Omega_j (j = 1..n) from the output cotangent, or nullptr
__device__ __forceinline__ const double* obar(const WbdArgs& a, long long b, int j) {
    const double* c = a.cot + b * (long long)a.n_out * a.m;
    if (a.op == OP_MANY) return c + (j - 1) * (long long)a.m;
    if (a.op == OP_SUM) return c;
    return j == a.n ? c : nullptr;
}

// node q's contribution to Omega_2..Omega_n: work rows [b][q][n-1][m]
template <int CAP>
__global__ void __launch_bounds__(kFoldThreads) wbd_fold(WbdArgs a) {
    __shared__ FoldLds<CAP> L;
    const long long b = blockIdx.x / a.Q;
    const int q = (int)(blockIdx.x - b * a.Q);
    const int n = a.n, K = n - 1, d = a.d;
    const long long M = a.m;
    double* T = a.work + (long long)blockIdx.x * (2 * K + 1) * M;  // T_j at T + j M
    double* Om = a.part + ((long long)blockIdx.x * n) * M;          // Omega_j at Om + (j-1) M, j = 2..n
    const double* What = a.totals + b * (long long)row_stride(a) * M;
    const double s = a.x[q] - 1.0, wq = a.w[q];
    fold_totals(L, a, What, T, s);
    for (int j = 2; j <= n; ++j) {
        double* O = Om + (j - 1) * M;
        mat_set<CAP>(O, wq, zrow(a, b, j - 1), d);
        for (int l = 0; l <= j - 2; ++l) mat_acc(L, O, wq * s, zrow(a, b, l), false, T + (j - 1 - l) * M, false, d);
    }
}

// All orders lo..n at once: one workgroup per (instance, flattened node f),
// node f of order korder[f] folding that order's Omega_2..Omega_k shares;
// the same arithmetic as wbd_fold for each order
template <int CAP>
__global__ void __launch_bounds__(kFoldThreads) wbd_fold_orders(WbdArgs a) {
    __shared__ FoldLds<CAP> L;
    const long long b = blockIdx.x / a.nflat;
    const int f = (int)(blockIdx.x - b * a.nflat);
    const int k = a.korder[f], nmax = a.n, Kmax = nmax - 1, d = a.d;
    const long long M = a.m;
    double* T = a.work + (long long)blockIdx.x * (2 * Kmax + 1) * M;  // T_j at T + j M
    double* Om = a.part + ((long long)blockIdx.x * nmax) * M;          // Omega_j at Om + (j-1) M, j = 2..k
    const double* What = a.totals + b * (long long)row_stride(a) * M;
    const double s = a.x[f] - 1.0, wq = a.w[f];
    WbdArgs ak = a;
    ak.n = k;  // fold_totals runs T_1..T_{k-1}
    fold_totals(L, ak, What, T, s);
    for (int j = 2; j <= k; ++j) {
        double* O = Om + (j - 1) * M;
        mat_set<CAP>(O, wq, zrow(a, b, j - 1), d);
        for (int l = 0; l <= j - 2; ++l) mat_acc(L, O, wq * s, zrow(a, b, l), false, T + (j - 1 - l) * M, false, d);
    }
}

// out row k - kout (op sum of order k, lo <= k <= n) from Z_0 and the node partials of
// order k, Omega_j's nodes added in q order, then the Omega_j in j order
__global__ void __launch_bounds__(256) wbd_fold_orders_reduce(WbdArgs a) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long M = a.m;
    const int nk = a.n - a.lo + 1;
    if (idx >= a.B * nk * M) return;
    const long long b = idx / (nk * M);
    const long long r = idx - b * nk * M;
    const int k = a.lo + (int)(r / M);
    const long long e = r - (long long)(k - a.lo) * M;
    int f0 = 0;
    while (f0 < a.nflat && a.korder[f0] != k) ++f0;
    const int Q = (k + 3) / 2;
    double v = a.z0[b * M + e];
    for (int j = 2; j <= k; ++j) {
        double om = 0.0;
        for (int q = 0; q < Q; ++q) om += a.part[(((b * a.nflat) + f0 + q) * a.n + (j - 1)) * M + e];
        v += om;
    }
    a.out[b * a.ostride + (long long)(k - a.kout) * M + e] = v;
}

// out from Omega_1 = Z_0 and the node partials, nodes added in q order
__global__ void __launch_bounds__(256) wbd_fold_reduce(WbdArgs a) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long M = a.m;
    if (idx >= a.B * M) return;
    const long long b = idx / M;
    const long long e = idx - b * M;
    const int n = a.n;
    auto omega = [&](int j) {  // Omega_j, j >= 2
        double v = 0.0;
        for (int q = 0; q < a.Q; ++q) v += a.part[((b * a.Q + q) * n + (j - 1)) * M + e];
        return v;
    };
    double* out = a.out + b * (a.ostride ? a.ostride : (long long)a.n_out * M) + e;
    const double o1 = a.z0[b * M + e];
    if (a.op == OP_MANY) {
        out[0] = o1;
        for (int j = 2; j <= n; ++j) out[(j - 1) * M] = omega(j);
    } else if (a.op == OP_SUM) {
        double v = o1;
        for (int j = 2; j <= n; ++j) v += omega(j);
        out[0] = v;
    } else {
        out[0] = omega(n);
    }
}

// Adjoint of node q's fold: its Zbar_l and What_bar_j contributions, rows
// [b][q][n][m] of part (Zbar) and part2 (What_bar).
template <int CAP>
__global__ void __launch_bounds__(kFoldThreads) wbd_fold_vjp(WbdArgs a) {
    __shared__ FoldLds<CAP> L;
    const long long b = blockIdx.x / a.Q;
    const int q = (int)(blockIdx.x - b * a.Q);
    const int n = a.n, K = n - 1, d = a.d;
    const long long M = a.m;
    double* T = a.work + (long long)blockIdx.x * (2 * K + 1) * M;  // T_j at T + j M
    double* Tb = T + K * M;                                        // T_bar_j at Tb + j M
    const double* What = a.totals + b * (long long)row_stride(a) * M;
    double* Zb = a.part + (long long)blockIdx.x * n * M;   // row l = Zbar_l (node q)
    double* Wb = a.part2 + (long long)blockIdx.x * n * M;  // row j-1 = What_bar_j (node q)
    const double s = a.x[q] - 1.0, wq = a.w[q];
    for (int l = 0; l < n; ++l) {
        mat_set<CAP>(Zb + l * M, 0.0, nullptr, d);
        mat_set<CAP>(Wb + l * M, 0.0, nullptr, d);
    }
    for (int j = 1; j <= K; ++j) mat_set<CAP>(Tb + j * M, 0.0, nullptr, d);
    fold_totals(L, a, What, T, s);
    for (int j = 2; j <= n; ++j) {
        const double* Ob = obar(a, b, j);
        if (!Ob) continue;
        mat_axpy<CAP>(Zb + (j - 1) * M, wq, Ob, d);
        for (int l = 0; l <= j - 2; ++l) {
            mat_acc(L, Zb + l * M, wq * s, Ob, false, T + (j - 1 - l) * M, true, d);
            mat_acc(L, Tb + (j - 1 - l) * M, wq * s, zrow(a, b, l), true, Ob, false, d);
        }
    }
    __syncthreads();
    for (int j = K; j >= 1; --j) {
        mat_axpy<CAP>(Wb + (j - 1) * M, 1.0, Tb + j * M, d);
        for (int l = 1; l < j; ++l) {
            mat_acc(L, Wb + (l - 1) * M, s, Tb + j * M, false, T + (j - l) * M, true, d);
            mat_acc(L, Tb + (j - l) * M, s, What + (l - 1) * M, true, Tb + j * M, false, d);
        }
        __syncthreads();
    }
}

// zbar / wbar = sum of the node partials in q order (+ Omega_1's cotangent on Zbar_0)
__global__ void __launch_bounds__(256) wbd_fold_vjp_reduce(WbdArgs a) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long M = a.m;
    const int n = a.n;
    if (idx >= a.B * n * M) return;
    const long long b = idx / (n * M);
    const long long r = idx - b * n * M;  // row * M + e
    double z = 0.0, w = 0.0;
    if (r < M) {
        const double* o1 = obar(a, b, 1);
        z = o1 ? o1[r] : 0.0;
    }
    for (int q = 0; q < a.Q; ++q) {
        const long long off = ((b * a.Q + q) * n) * M + r;
        z += a.part[off];
        w += a.part2[off];
    }
    a.zbar[idx] = z;
    a.wbar[idx] = w;
}

// ---------------------------------------------------------------------------
// Small problems (d <= 8, n <= NM): one wavefront per (instance, node).  Each
// lane owns one entry of every matrix; the instance's totals and sums, the
// node's T_j and (VJP) its cotangents sit in LDS (row-major), so there is no
// global-memory round trip between the dependent products.  For d < 8 lane e
// owns entry e and a product is 2d LDS reads and d FMAs per lane, the
// operations of mat_acc's d <= 8 branch in the same order.  For d = 8 a
// product is two v_mfma_f64_4x4x4f64 (four 4 x 4 output blocks over two
// halves of the contraction) from 4 LDS reads per lane, and lanes own the
// entries in the instruction's output layout, so each result lands in its
// owner.  The workgroup is one wavefront, so its barriers cost nothing.
// ---------------------------------------------------------------------------
constexpr int kSmallD = 8;
constexpr int kSmallM = kSmallD * kSmallD;

// (X' Y')[r][c] summed over k ascending from 0.0, X' = X or X^T, Y' = Y or Y^T;
// unrolled to kSmallD steps so that all operand reads are issued before the
// first FMA waits on one (a loop over the runtime d waited on LDS every step)
template <bool TX, bool TY>
__device__ __forceinline__ double sdot(const double* X, const double* Y, int d, int r, int c) {
    double x[kSmallD], y[kSmallD];
#pragma unroll
    for (int k = 0; k < kSmallD; ++k) {
        x[k] = k < d ? (TX ? X[k * d + r] : X[r * d + k]) : 0.0;
        y[k] = k < d ? (TY ? Y[c * d + k] : Y[k * d + c]) : 0.0;
    }
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < kSmallD; ++k)
        if (k < d) acc = fma(x[k], y[k], acc);
    return acc;
}

// The entry lane L owns: e = L for d < 8; for d = 8 the output layout of
// v_mfma_f64_4x4x4f64 (block b's D[i][j] in lane 16i + 4b + j) with block
// b = 2I + J holding rows 4I.., columns 4J.. of the product
__device__ __forceinline__ int small_owned(int lane, int d) {
    if (d != kSmallD) return lane;
    return 8 * (4 * ((lane >> 3) & 1) + (lane >> 4)) + 4 * ((lane >> 2) & 1) + (lane & 3);
}

// this lane's entry of X' Y' (X' = X or X^T, Y' = Y or Y^T), entry (r, c) =
// small_owned(lane, d) as (row, column)
template <bool TX, bool TY>
__device__ __forceinline__ double sprod(const double* X, const double* Y, int d, int r, int c) {
    if (d != kSmallD) return sdot<TX, TY>(X, Y, d, r, c);
    // block b = 2I + J of instruction h: A[i][k] = X'[4I + i][4h + k] in lane
    // 16k + 4b + i, B[k][j] = Y'[4h + k][4J + j] in lane 16k + 4b + j
    const int lane = (int)threadIdx.x, k = lane >> 4, q = lane & 3, I = (lane >> 3) & 1, J = (lane >> 2) & 1;
    double acc = 0.0;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int ar = 4 * I + q, ac = 4 * h + k, br = 4 * h + k, bc = 4 * J + q;
        const double x = TX ? X[ac * kSmallD + ar] : X[ar * kSmallD + ac];
        const double y = TY ? Y[bc * kSmallD + br] : Y[br * kSmallD + bc];
        acc = __builtin_amdgcn_mfma_f64_4x4x4f64(x, y, acc, 0, 0, 0);
    }
    return acc;
}

template <int NM>
struct SmallFold {
    double wh[NM][kSmallM];  // What_1..What_K at rows 0..K-1
    double z[NM][kSmallM];   // Z_0..Z_{n-1}
    double t[NM][kSmallM];   // T_1..T_K at rows 1..K
};

// load What_1..What_K and Z_0..Z_{n-1} of instance b (orders up to n), then T_1..T_K
template <int NM>
__device__ void small_totals(SmallFold<NM>& L, const WbdArgs& a, long long b, int n, double s) {
    const int K = n - 1, d = a.d, e = small_owned((int)threadIdx.x, d), r = e / d, c = e - r * d;
    const long long M = a.m;
    const bool live = e < a.m;
    const double* What = a.totals + b * (long long)row_stride(a) * M;
    if (live) {
        for (int j = 0; j < K; ++j) L.wh[j][e] = What[j * M + e];
        for (int l = 0; l < n; ++l) L.z[l][e] = zrow(a, b, l)[e];
    }
    __syncthreads();
    for (int j = 1; j <= K; ++j) {
        if (live) {
            double v = 1.0 * L.wh[j - 1][e];
            for (int l = 1; l < j; ++l) v = fma(s, sprod<false, false>(L.wh[l - 1], L.t[j - l], d, r, c), v);
            L.t[j][e] = v;
        }
        __syncthreads();
    }
}

// node (weight wq, shift s) shares of Omega_2..Omega_n into part rows Om + (j-1) M
template <int NM>
__device__ void small_omegas(SmallFold<NM>& L, const WbdArgs& a, int n, double s, double wq, double* Om) {
    const int d = a.d, e = small_owned((int)threadIdx.x, d), r = e / d, c = e - r * d;
    if (e >= a.m) return;
    const long long M = a.m;
    for (int j = 2; j <= n; ++j) {
        double v = wq * L.z[j - 1][e];
        for (int l = 0; l <= j - 2; ++l) v = fma(wq * s, sprod<false, false>(L.z[l], L.t[j - 1 - l], d, r, c), v);
        Om[(j - 1) * M + e] = v;
    }
}

template <int NM>
__global__ void __launch_bounds__(64) wbd_fold_small(WbdArgs a) {
    __shared__ SmallFold<NM> L;
    const long long b = blockIdx.x / a.Q;
    const int q = (int)(blockIdx.x - b * a.Q);
    const double s = a.x[q] - 1.0;
    small_totals(L, a, b, a.n, s);
    small_omegas(L, a, a.n, s, a.w[q], a.part + ((long long)blockIdx.x * a.n) * a.m);
}

template <int NM>
__global__ void __launch_bounds__(64) wbd_fold_orders_small(WbdArgs a) {
    __shared__ SmallFold<NM> L;
    const long long b = blockIdx.x / a.nflat;
    const int f = (int)(blockIdx.x - b * a.nflat);
    const int k = a.korder[f];
    const double s = a.x[f] - 1.0;
    small_totals(L, a, b, k, s);
    small_omegas(L, a, k, s, a.w[f], a.part + ((long long)blockIdx.x * a.n) * a.m);
}

template <int NM>
struct SmallFoldVjp {
    SmallFold<NM> f;
    double tb[NM][kSmallM];  // T_bar_1..T_bar_K at rows 1..K
    double ob[NM][kSmallM];  // cotangent of Omega_j at row j-1 (rows with obar(j) set)
};

// Zbar_l and What_bar_j are only ever accumulated entry by entry, so each lane
// keeps its entries of them in registers (loops unrolled to NM, so that the
// row indices are compile-time); the LDS holds only matrices read as product
// operands, which lets more of these one-wavefront workgroups share a CU.
template <int NM>
__global__ void __launch_bounds__(64) wbd_fold_vjp_small(WbdArgs a) {
    __shared__ SmallFoldVjp<NM> L;
    const long long b = blockIdx.x / a.Q;
    const int q = (int)(blockIdx.x - b * a.Q);
    const int n = a.n, K = n - 1, d = a.d, e = small_owned((int)threadIdx.x, d), r = e / d, c = e - r * d;
    const long long M = a.m;
    const bool live = e < a.m;
    const double s = a.x[q] - 1.0, wq = a.w[q];
    double zb[NM], wb[NM];  // Zbar_0..Zbar_{n-1}, What_bar_1..What_bar_n (rows 0..n-1)
#pragma unroll
    for (int l = 0; l < NM; ++l) {
        zb[l] = 0.0;
        wb[l] = 0.0;
    }
    if (live) {
        for (int j = 1; j <= K; ++j) L.tb[j][e] = 0.0;
        for (int j = 2; j <= n; ++j) {
            const double* Ob = obar(a, b, j);
            if (Ob) L.ob[j - 1][e] = Ob[e];
        }
    }
    small_totals(L.f, a, b, n, s);  // ends with a barrier: the loads above are visible
    auto& T = L.f.t;
    auto& Z = L.f.z;
    if (live) {
#pragma unroll
        for (int j = 2; j <= NM; ++j) {
            if (j > n) break;
            if (!obar(a, b, j)) continue;
            const double* Ob = L.ob[j - 1];
            zb[j - 1] = fma(wq, Ob[e], zb[j - 1]);
#pragma unroll
            for (int l = 0; l <= j - 2; ++l) {
                zb[l] = fma(wq * s, sprod<false, true>(Ob, T[j - 1 - l], d, r, c), zb[l]);
                L.tb[j - 1 - l][e] = fma(wq * s, sprod<true, false>(Z[l], Ob, d, r, c), L.tb[j - 1 - l][e]);
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = NM - 1; j >= 1; --j) {
        if (j > K) continue;
        if (live) {
            wb[j - 1] = fma(1.0, L.tb[j][e], wb[j - 1]);
#pragma unroll
            for (int l = 1; l < j; ++l) {
                wb[l - 1] = fma(s, sprod<false, true>(L.tb[j], T[j - l], d, r, c), wb[l - 1]);
                L.tb[j - l][e] = fma(s, sprod<true, false>(L.f.wh[l - 1], L.tb[j], d, r, c), L.tb[j - l][e]);
            }
        }
        __syncthreads();
    }
    if (live) {
        double* Zb = a.part + (long long)blockIdx.x * n * M;
        double* Wb = a.part2 + (long long)blockIdx.x * n * M;
#pragma unroll
        for (int l = 0; l < NM; ++l) {
            if (l >= n) break;
            Zb[l * M + e] = zb[l];
            Wb[l * M + e] = wb[l];
        }
    }
}

// the small kernels apply at d <= 8 and n <= 16 (NM = 8 or 16)
__device__ __host__ inline int small_fold_nm(int d, int n) { return d > kSmallD ? 0 : n <= 8 ? 8 : n <= 16 ? 16 : 0; }

}  // namespace

hipError_t launch_wbd_fold(const WbdArgs& a, hipStream_t s) {
    if (a.B == 0) return hipSuccess;
    void* h = timer_begin("wbd_fold", s);
    const int nm = small_fold_nm(a.d, a.n);
    if (nm == 8)
        hipLaunchKernelGGL(wbd_fold_small<8>, dim3((unsigned)(a.B * a.Q)), dim3(64), 0, s, a);
    else if (nm == 16)
        hipLaunchKernelGGL(wbd_fold_small<16>, dim3((unsigned)(a.B * a.Q)), dim3(64), 0, s, a);
    else if (a.d <= kFoldLdsSmall)
        hipLaunchKernelGGL(wbd_fold<kFoldLdsSmall>, dim3((unsigned)(a.B * a.Q)), dim3(kFoldThreads), 0, s, a);
    else
        hipLaunchKernelGGL(wbd_fold<kFoldLds>, dim3((unsigned)(a.B * a.Q)), dim3(kFoldThreads), 0, s, a);
    hipLaunchKernelGGL(wbd_fold_reduce, dim3((unsigned)((a.B * a.m + 255) / 256)), dim3(256), 0, s, a);
    timer_end(h);
    return hipGetLastError();
}

hipError_t launch_wbd_fold_orders(const WbdArgs& a, hipStream_t s) {
    if (a.B == 0) return hipSuccess;
    void* h = timer_begin("wbd_fold", s);
    const unsigned grid = (unsigned)(a.B * a.nflat);
    const int nm = small_fold_nm(a.d, a.n);
    if (nm == 8)
        hipLaunchKernelGGL(wbd_fold_orders_small<8>, dim3(grid), dim3(64), 0, s, a);
    else if (nm == 16)
        hipLaunchKernelGGL(wbd_fold_orders_small<16>, dim3(grid), dim3(64), 0, s, a);
    else if (a.d <= kFoldLdsSmall)
        hipLaunchKernelGGL(wbd_fold_orders<kFoldLdsSmall>, dim3(grid), dim3(kFoldThreads), 0, s, a);
    else
        hipLaunchKernelGGL(wbd_fold_orders<kFoldLds>, dim3(grid), dim3(kFoldThreads), 0, s, a);
    const long long outs = a.B * (a.n - a.lo + 1) * a.m;
    hipLaunchKernelGGL(wbd_fold_orders_reduce, dim3((unsigned)((outs + 255) / 256)), dim3(256), 0, s, a);
    timer_end(h);
    return hipGetLastError();
}

hipError_t launch_wbd_fold_vjp(const WbdArgs& a, hipStream_t s) {
    if (a.B == 0) return hipSuccess;
    void* h = timer_begin("wbd_fold_vjp", s);
    const int nm = small_fold_nm(a.d, a.n);
    if (nm == 8)
        hipLaunchKernelGGL(wbd_fold_vjp_small<8>, dim3((unsigned)(a.B * a.Q)), dim3(64), 0, s, a);
    else if (nm == 16)
        hipLaunchKernelGGL(wbd_fold_vjp_small<16>, dim3((unsigned)(a.B * a.Q)), dim3(64), 0, s, a);
    else if (a.d <= kFoldLdsSmall)
        hipLaunchKernelGGL(wbd_fold_vjp<kFoldLdsSmall>, dim3((unsigned)(a.B * a.Q)), dim3(kFoldThreads), 0, s, a);
    else
        hipLaunchKernelGGL(wbd_fold_vjp<kFoldLds>, dim3((unsigned)(a.B * a.Q)), dim3(kFoldThreads), 0, s, a);
    hipLaunchKernelGGL(wbd_fold_vjp_reduce, dim3((unsigned)((a.B * a.n * a.m + 255) / 256)), dim3(256), 0, s, a);
    timer_end(h);
    return hipGetLastError();
}

}  // namespace magnus
