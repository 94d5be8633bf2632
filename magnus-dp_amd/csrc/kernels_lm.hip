// kernels_lm.hip -- level-parallel Magnus kernels: the reference's per-node
// level recursion (src/magnus.hpp:12-261) and its VJP (src/grad.hpp:14-349)
// for any matrix policy and integrator.
//
// Work decomposition.  One job = (instance b, Gauss-Legendre node q).  A level
// of the recursion is one product per sample followed by one prefix scan per
// matrix entry.  The products of a level depend only on the previous level's
// prefix array, so they run for all (job, sample, entry) at once
// (lm_fwd_prod, lm_vjp_prod); the scans and sums, which are sequential in the
// sample index, run one thread per (job, entry) with the reference's
// per-entry recurrences and order of operations (rules.hpp): lm_fwd_scan,
// lm_vjp_scan.  Products keep the reference's summation order too, so every
// scalar is computed as the reference computes it; only the q-lane reduction
// is done in fixed q order by the reduce kernels.
#include "kernels.hpp"
#include "products.hpp"
#include "rules.hpp"

namespace magnus {

namespace {

// P^{(j)} of job (b, q): carry[b][q][j-1] or a ping-pong buffer
__device__ __forceinline__ double* lm_slot(const LmArgs& a, long long job, long long b, int q, int j) {
    if (a.carry) return a.carry + ((b * a.Q + q) * (long long)(a.n - 1) + (j - 1)) * a.S * a.m;
    return a.pingpong + (job * 2 + (j & 1)) * a.S * a.m;
}

// ---------------------------------------------------------------------------
// forward
// ---------------------------------------------------------------------------

// Y_i = A_i (P^{(j-1)}_i + s_q T), T = P^{(j-1)}_{S-1}  (magnus.hpp:230-254)
__global__ void __launch_bounds__(256) lm_fwd_prod(LmArgs a) {
    const long long per = a.S * a.m;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= a.jobs * per) return;
    const long long job = idx / per;
    const long long rem = idx - job * per;
    const long long i = rem / a.m;
    const int e = (int)(rem - i * a.m);
    const long long b = job / a.Qj;
    const int q = (int)(job - b * a.Qj);
    const double shift = a.x[q] - 1;  // magnus.hpp:65
    const double* P = lm_slot(a, job, b, q, a.level - 1);
    const double* T = P + (a.S - 1) * a.m;
    const double* Ai = a.in + (b * a.S + i) * a.m_in;
    double f;
    if (a.sc) f = prod_sc(Ai + (a.sc == 2 ? 1 : 0), P + i * a.m, T, e, shift);
    else f = prod_generic(Ai, P + i * a.m, T, a.d, e / a.d, e % a.d, shift);
    a.Y[idx] = f;
}

// ---------------------------------------------------------------------------
// Row products for small d (<= 16, generic matrices): one thread per (sample,
// output row), the sample's operands staged once in LDS, so a product costs
// its operands' reads once instead of once per output entry.  Sums run in the
// same index order as prod_generic / prod_vjp.
// ---------------------------------------------------------------------------
constexpr int kRowMaxD = 16;

// Where a workgroup's samples live (forward row products): one thread per
// sample splits the flat sample index p = job * S + i (two 64-bit divisions)
// once, so the staging loop, which touches every entry, does no division by S
// or Q.  The table sits behind the staged matrices in the dynamic LDS, sized by
// the workgroup's sample count.
struct RowSample {
    const double* P;  // previous level (or carry slot) of the sample's job, row 0
    const double* A;  // the sample's input matrix
    long long job, i;
    double shift;
};

__host__ __device__ __forceinline__ int rows_spb(int d) { return 256 / d; }
// LDS stride of one sample's d x d staging: d^2 + 1 doubles.  The inner loops
// read one address per sample (broadcast over its d row threads) for the 64 / d
// samples of a wavefront at once; at a stride of d^2 doubles (a multiple of
// 32 dwords for even d) all of them fall in one bank pair, and the odd stride
// spreads up to 16 samples over distinct pairs.
__host__ __device__ __forceinline__ int rows_stride(int d) { return d * d + 1; }
// the VJP stages three matrices per sample; stay within 64 KB of dynamic LDS
__host__ __device__ __forceinline__ int vjp_rows_spb(int d) {
    const int s = 256 / d, lim = 65536 / (3 * rows_stride(d) * 8);
    return s < lim ? s : lim;
}
// dynamic LDS of the forward rows launch: the staged matrices, then the table
__host__ __device__ __forceinline__ size_t rows_lds_bytes(int d, int spb) {
    return (size_t)spb * (rows_stride(d) * sizeof(double) + sizeof(RowSample));
}

__device__ __forceinline__ void row_samples(const LmArgs& a, long long p0, int spb, RowSample* tab) {
    const int g = threadIdx.x;
    if (g >= spb) return;
    const long long p = p0 + g;
    if (p >= a.jobs * a.S) return;
    RowSample r;
    r.job = p / a.S;
    r.i = p - r.job * a.S;
    const long long b = r.job / a.Qj;
    const int q = (int)(r.job - b * a.Qj);
    r.shift = a.x[q] - 1;
    r.P = lm_slot(a, r.job, b, q, a.level - 1);
    r.A = a.in + (b * a.S + r.i) * (long long)a.m_in;
    tab[g] = r;
}

// Y_i = A_i (P_i + s T): LDS holds b = P_i + s T per sample
__global__ void __launch_bounds__(256) lm_fwd_prod_rows(LmArgs a) {
    extern __shared__ double lds[];
    const int d = a.d, m = a.m, spb = rows_spb(d), sm = rows_stride(d);
    RowSample* tab = reinterpret_cast<RowSample*>(lds + spb * sm);
    const long long total = a.jobs * a.S;
    const long long p0 = (long long)blockIdx.x * spb;
    const int live = (int)(total - p0 < spb ? total - p0 : spb);
    row_samples(a, p0, spb, tab);
    __syncthreads();
    for (int x = threadIdx.x; x < live * m; x += 256) {
        const int g = x / m, e = x - g * m;
        const RowSample& r = tab[g];
        double v = r.P[r.i * m + e];
        v += r.P[(a.S - 1) * m + e] * r.shift;
        lds[g * sm + e] = v;
    }
    __syncthreads();
    const int g = threadIdx.x / d, rr = threadIdx.x - g * d;
    if (g >= live) return;
    const double* Ai = tab[g].A + rr * d;
    double ar[kRowMaxD], acc[kRowMaxD];
#pragma unroll
    for (int l = 0; l < kRowMaxD; ++l) {
        ar[l] = l < d ? Ai[l] : 0.0;
        acc[l] = 0.0;
    }
    const double* bs = lds + g * sm;
#pragma unroll
    for (int l = 0; l < kRowMaxD; ++l) {
        if (l >= d) break;
#pragma unroll
        for (int c = 0; c < kRowMaxD; ++c)
            if (c < d) acc[c] += ar[l] * bs[l * d + c];
    }
    double* y = a.Y + (p0 + g) * m + rr * d;
#pragma unroll
    for (int c = 0; c < kRowMaxD; ++c)
        if (c < d) y[c] = acc[c];
}

// dA_job_i += g b^T (row r) and temp_i = A_i^T g (row r); LDS per sample: b, g, A.
// (A per-sample table as in lm_fwd_prod_rows measured slower here: the staging
// loads then wait on LDS reads of the table.)
__global__ void __launch_bounds__(256) lm_vjp_prod_rows(LmArgs a) {
    extern __shared__ double lds[];
    const int d = a.d, m = a.m, spb = vjp_rows_spb(d), sm = rows_stride(d);
    const long long per = a.S * m;
    const long long total = a.jobs * a.S;
    const long long p0 = (long long)blockIdx.x * spb;
    for (int x = threadIdx.x; x < spb * m; x += 256) {
        const int g = x / m, e = x - g * m;
        const long long p = p0 + g;
        if (p >= total) continue;
        const long long job = p / a.S, i = p - job * a.S;
        const long long b = job / a.Q;
        const int q = (int)(job - b * a.Q);
        const double shift = a.x[q] - 1.0;
        const double* P = a.carry + ((b * a.Q + q) * (long long)(a.n - 1) + (a.level - 2)) * per;
        double v = P[i * m + e];
        v += P[(a.S - 1) * m + e] * shift;
        const int y = g * sm + e;
        lds[y] = v;                                                                  // b
        lds[spb * sm + y] = a.gbuf[(job * 2 + a.ga) * per + i * m + e];              // g
        lds[2 * spb * sm + y] = a.in[(b * a.S + i) * (long long)a.m_in + e];         // A
    }
    __syncthreads();
    const int g = threadIdx.x / d, rr = threadIdx.x - g * d;
    const long long p = p0 + g;
    if (g >= spb || p >= total) return;
    const long long job = p / a.S, i = p - job * a.S;
    const double* bs = lds + g * sm;
    const double* gs = lds + spb * sm + g * sm;
    const double* as = lds + 2 * spb * sm + g * sm;
    double gr[kRowMaxD], ac[kRowMaxD];
#pragma unroll
    for (int j = 0; j < kRowMaxD; ++j) {
        gr[j] = j < d ? gs[rr * d + j] : 0.0;  // g[r][:]
        ac[j] = j < d ? as[j * d + rr] : 0.0;  // a[:][r]
    }
    double* dA = a.dAjob + job * per + i * m + rr * d;
    double* Gb = a.gbuf + (job * 2 + (1 - a.ga)) * per + i * m + rr * d;
    // the row of dA is read before any store: the stores into Gb might alias it
    // as far as the compiler knows, so reads left in the loop would each wait
    // for a full memory round trip
    double dav[kRowMaxD];
#pragma unroll
    for (int cc = 0; cc < kRowMaxD; ++cc) dav[cc] = cc < d ? dA[cc] : 0.0;
#pragma unroll
    for (int cc = 0; cc < kRowMaxD; ++cc) {
        if (cc >= d) break;
        double v = 0.0;  // sum_j g[r][j] b[cc][j]
#pragma unroll
        for (int j = 0; j < kRowMaxD; ++j)
            if (j < d) v += gr[j] * bs[cc * d + j];
        double t2 = 0.0;  // sum_i a[i][r] g[i][cc]
#pragma unroll
        for (int ii = 0; ii < kRowMaxD; ++ii)
            if (ii < d) t2 += ac[ii] * gs[ii * d + cc];
        dA[cc] = dav[cc] + v;
        Gb[cc] = t2;
    }
}

// ---------------------------------------------------------------------------
// MFMA products for d = D in {16, 32, 64}.  A 256-thread workgroup stages the
// operands of PPB samples in LDS (k-major, so a wavefront's operand reads are
// contiguous) and computes every 16 x 16 output tile with
// v_mfma_f64_16x16x4f64 (lane l, register r of a tile holds row 4r + l/16,
// column l%16).  The MFMA sums k in its own order, so these products differ
// from the reference's l-ordered sums by rounding only.
// ---------------------------------------------------------------------------
typedef double dbl4 __attribute__((ext_vector_type(4)));

// WPP wavefronts share one product; a 4-wavefront workgroup holds 4 / WPP
// products.  The k range is staged in KS slices of DK = D / KS rows: LDS then
// holds DK of the D rows of both operands, so more workgroups fit a CU and one
// workgroup's loads overlap another's MFMAs.  Measured: at d = 32, 17 KB per
// workgroup (KS = products per workgroup) beat 34 KB and 67 KB; at d = 64, two
// slices (33 KB) beat one (67 KB) and four, which slowed the VJP.
template <int D, int W>
struct MfmaShape {
    static constexpr int TILES = (D / 16) * (D / 16);  // 16x16 tiles per product
    static constexpr int WPP = W;                      // wavefronts per product
    static constexpr int TPW = TILES / WPP;            // tiles per wavefront
    static constexpr int PPB = 4 / WPP;                // products per workgroup
    static constexpr int MM = D * D;
    static constexpr int LD = D + 1;                                            // padded LDS row
    static constexpr int KS = D == 64 ? 2 : (D == 32 ? PPB : 1);                // k slices
    static constexpr int DK = D / KS;                                           // k rows per slice
    static_assert(TILES % WPP == 0 && 4 % WPP == 0 && DK % 4 == 0, "tile split");
};

// default split: one product per wavefront up to d = 32, four wavefronts at d = 64
template <int D>
constexpr int default_wpp() { return D == 64 ? 4 : 1; }

// Operands sit in LDS k-major with the leading dimension padded to D + 1
// doubles: transposed stores (consecutive lanes step one row) and the tile
// reads (consecutive lanes step one column, four rows apart) then fall on
// different banks.
// acc += Aop * Bop over the DK rows of a slice, Aop[i][k] = LA[k*LD + i0 + i], Bop[k][j] = LB[k*LD + j0 + j]
template <int DK, int LD>
__device__ __forceinline__ void mfma_tile(const double* LA, const double* LB, int i0, int j0, int lane, dbl4& acc) {
    const int li = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int kk = 0; kk < DK / 4; ++kk) {
        const double av = LA[(4 * kk + lk) * LD + i0 + li];
        const double bv = LB[(4 * kk + lk) * LD + j0 + li];
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
    }
}

// rows k0..k0+DK-1 of a row-major D x D source: L[k - k0][j] = X[k][j] (+ T[k][j] * s)
template <int D, int DK>
__device__ __forceinline__ void stage_rows(double* L, const double* X, const double* T, double s, int k0) {
    constexpr int LD = D + 1;
    const double* Xs = X + k0 * D;
    for (int e = threadIdx.x; e < DK * D; e += 256) {
        const int r = e / D, c = e - r * D;
        double v = Xs[e];
        if (T) v += T[k0 * D + e] * s;
        L[r * LD + c] = v;
    }
}

// columns k0..k0+DK-1 of a row-major D x D source: L[k - k0][i] = X[i][k] (+ T[i][k] * s)
template <int D, int DK>
__device__ __forceinline__ void stage_cols(double* L, const double* X, const double* T, double s, int k0) {
    constexpr int LD = D + 1;
    for (int e = threadIdx.x; e < DK * D; e += 256) {
        const int r = e / DK, c = e - r * DK;
        double v = X[r * D + k0 + c];
        if (T) v += T[r * D + k0 + c] * s;
        L[c * LD + r] = v;
    }
}

template <int D>
__device__ __forceinline__ void store_tile(double* out, const dbl4& acc, int i0, int j0, int lane, bool add) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        double* o = out + (i0 + 4 * r + (lane >> 4)) * D + j0 + (lane & 15);
        if (add) *o += acc[r];
        else *o = acc[r];
    }
}

// Y_i = A_i (P_i + s T) for products p = job*S + i; per product the index
// arithmetic once (uniform over the workgroup), then its slices.
template <int D, int W>
__global__ void __launch_bounds__(256) lm_fwd_prod_mfma(LmArgs a) {
    using M = MfmaShape<D, W>;
    constexpr int LD = M::LD, LM = M::DK * LD;
    __shared__ double LA[M::PPB][LM];  // slice of A_i transposed: LA[k][i] = a[i][k0 + k]
    __shared__ double LB[M::PPB][LM];  // slice of P_i + s T: LB[k][j]
    const long long total = a.jobs * a.S;
    const long long p0 = (long long)blockIdx.x * M::PPB;
    const double* srcA[M::PPB];
    const double* srcP[M::PPB];
    const double* srcT[M::PPB];
    double shift[M::PPB];
#pragma unroll
    for (int g = 0; g < M::PPB; ++g) {
        const long long p = p0 + g;
        srcA[g] = srcP[g] = srcT[g] = nullptr;
        shift[g] = 0.0;
        if (p >= total) continue;
        const long long job = p / a.S, i = p - job * a.S;
        const long long b = job / a.Qj;
        const int q = (int)(job - b * a.Qj);
        shift[g] = a.x[q] - 1;
        const double* P = lm_slot(a, job, b, q, a.level - 1);
        srcP[g] = P + i * M::MM;
        srcT[g] = P + (a.S - 1) * M::MM;
        srcA[g] = a.in + (b * a.S + i) * (long long)M::MM;
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int g = wave / M::WPP, wp = wave - g * M::WPP;
    dbl4 acc[M::TPW];
#pragma unroll
    for (int tt = 0; tt < M::TPW; ++tt) acc[tt] = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int ks = 0; ks < M::KS; ++ks) {
        if (ks) __syncthreads();  // the previous slice's tile reads are done
#pragma unroll
        for (int h = 0; h < M::PPB; ++h) {
            if (!srcA[h]) continue;
            stage_cols<D, M::DK>(LA[h], srcA[h], nullptr, 0.0, ks * M::DK);
            stage_rows<D, M::DK>(LB[h], srcP[h], srcT[h], shift[h], ks * M::DK);
        }
        __syncthreads();
#pragma unroll
        for (int tt = 0; tt < M::TPW; ++tt) {
            const int tile = wp * M::TPW + tt;
            mfma_tile<M::DK, LD>(LA[g], LB[g], 16 * (tile / (D / 16)), 16 * (tile % (D / 16)), lane, acc[tt]);
        }
    }
    const long long p = p0 + g;
    if (p >= total) return;
    double* y = a.Y + p * M::MM;
#pragma unroll
    for (int tt = 0; tt < M::TPW; ++tt) {
        const int tile = wp * M::TPW + tt;
        store_tile<D>(y, acc[tt], 16 * (tile / (D / 16)), 16 * (tile % (D / 16)), lane, false);
    }
}

// VJP products in two phases over the same LDS: dav = g b^T (b = P_i + s T),
// added to dA_job, then temp = A_i^T g
template <int D, int W>
__global__ void __launch_bounds__(256) lm_vjp_prod_mfma(LmArgs a) {
    using M = MfmaShape<D, W>;
    constexpr int LD = M::LD, LM = M::DK * LD;
    __shared__ double LA[M::PPB][LM];
    __shared__ double LB[M::PPB][LM];
    const long long per = a.S * M::MM;
    const long long total = a.jobs * a.S;
    const long long p0 = (long long)blockIdx.x * M::PPB;
    const double* srcA[M::PPB];
    const double* srcP[M::PPB];
    const double* srcT[M::PPB];
    const double* srcG[M::PPB];
    double shift[M::PPB];
#pragma unroll
    for (int g = 0; g < M::PPB; ++g) {
        const long long p = p0 + g;
        srcA[g] = srcP[g] = srcT[g] = srcG[g] = nullptr;
        shift[g] = 0.0;
        if (p >= total) continue;
        const long long job = p / a.S, i = p - job * a.S;
        const long long b = job / a.Q;
        const int q = (int)(job - b * a.Q);
        shift[g] = a.x[q] - 1.0;
        const double* P = a.carry + ((b * a.Q + q) * (long long)(a.n - 1) + (a.level - 2)) * per;
        srcP[g] = P + i * M::MM;
        srcT[g] = P + (a.S - 1) * M::MM;
        srcG[g] = a.gbuf + (job * 2 + a.ga) * per + i * M::MM;
        srcA[g] = a.in + (b * a.S + i) * (long long)M::MM;
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int gw = wave / M::WPP, wp = wave - gw * M::WPP;
    const long long pw = p0 + gw;
    dbl4 acc[M::TPW];
    // phase 1: LA[k][i] = g[i][k], LB[k][j] = b[j][k]
#pragma unroll
    for (int tt = 0; tt < M::TPW; ++tt) acc[tt] = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int ks = 0; ks < M::KS; ++ks) {
        if (ks) __syncthreads();
#pragma unroll
        for (int h = 0; h < M::PPB; ++h) {
            if (!srcA[h]) continue;
            stage_cols<D, M::DK>(LA[h], srcG[h], nullptr, 0.0, ks * M::DK);
            stage_cols<D, M::DK>(LB[h], srcP[h], srcT[h], shift[h], ks * M::DK);
        }
        __syncthreads();
#pragma unroll
        for (int tt = 0; tt < M::TPW; ++tt) {
            const int tile = wp * M::TPW + tt;
            mfma_tile<M::DK, LD>(LA[gw], LB[gw], 16 * (tile / (D / 16)), 16 * (tile % (D / 16)), lane, acc[tt]);
        }
    }
    if (pw < total) {
        const long long job = pw / a.S, i = pw - job * a.S;
        double* dA = a.dAjob + job * per + i * M::MM;
#pragma unroll
        for (int tt = 0; tt < M::TPW; ++tt) {
            const int tile = wp * M::TPW + tt;
            store_tile<D>(dA, acc[tt], 16 * (tile / (D / 16)), 16 * (tile % (D / 16)), lane, true);
        }
    }
    // phase 2: LA[k][i] = a[k][i], LB[k][j] = g[k][j]
#pragma unroll
    for (int tt = 0; tt < M::TPW; ++tt) acc[tt] = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int ks = 0; ks < M::KS; ++ks) {
        __syncthreads();  // the previous phase's or slice's tile reads are done
#pragma unroll
        for (int h = 0; h < M::PPB; ++h) {
            if (!srcA[h]) continue;
            stage_rows<D, M::DK>(LA[h], srcA[h], nullptr, 0.0, ks * M::DK);
            stage_rows<D, M::DK>(LB[h], srcG[h], nullptr, 0.0, ks * M::DK);
        }
        __syncthreads();
#pragma unroll
        for (int tt = 0; tt < M::TPW; ++tt) {
            const int tile = wp * M::TPW + tt;
            mfma_tile<M::DK, LD>(LA[gw], LB[gw], 16 * (tile / (D / 16)), 16 * (tile % (D / 16)), lane, acc[tt]);
        }
    }
    if (pw < total) {
        const long long job = pw / a.S, i = pw - job * a.S;
        double* Gb = a.gbuf + (job * 2 + (1 - a.ga)) * per + i * M::MM;
#pragma unroll
        for (int tt = 0; tt < M::TPW; ++tt) {
            const int tile = wp * M::TPW + tt;
            store_tile<D>(Gb, acc[tt], 16 * (tile / (D / 16)), 16 * (tile % (D / 16)), lane, false);
        }
    }
}

// P^{(j)} = Prefix(Y) (j < n) and the level's GL-weighted sum, one entry of one job
template <int I>
__global__ void __launch_bounds__(256) lm_fwd_scan(LmArgs a) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= a.jobs * a.m) return;
    const long long job = idx / a.m;
    const int e = (int)(idx - job * a.m);
    const long long b = job / a.Qj;
    const int q = (int)(job - b * a.Qj);
    const long long S = a.S;
    const int m = a.m, n = a.n, j = a.level;
    const double dt = a.dt;
    const double fscale = dt * a.w[q];  // magnus.hpp:66
    const bool om1 = a.outputs && j == 1 && q == 0 && (a.op != OP_ONE || n == 1);
    const bool lvl = a.outputs && j >= 2 && (a.op != OP_ONE || j == n);
    const double c = (j == 1) ? dt : fscale;
    double* Pout = (j < n) ? lm_slot(a, job, b, q, j) : nullptr;
    const double* A = a.in + b * S * a.m_in;
    const double* Yj = a.Y + job * S * m;
    PrefixState<I> st;
    double acc = 0.0;
    constexpr int U = 16;  // samples whose inputs are loaded ahead of the recurrence
    for (long long i0 = 0; i0 < S; i0 += U) {
        double f[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long long i = i0 + u;
            f[u] = 0.0;
            if (i < S) f[u] = (j == 1) ? raw_entry(A + i * a.m_in, e, a.sc) : Yj[i * m + e];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long long i = i0 + u;
            if (i >= S) break;
            if (om1 || lvl) sum_acc<I>(acc, i, S, f[u], c);
            if (Pout) Pout[i * m + e] = st.step(i, f[u], dt);
        }
    }
    if (a.totals && Pout) a.totals[(job * n + (j - 1)) * m + e] = Pout[(S - 1) * m + e];
    if (om1) a.omega1[b * m + e] = acc;
    if (lvl) {
        if (a.op == OP_MANY) {
            a.partial[(job * a.n_out + (j - 1)) * m + e] = acc;
        } else {
            double* p = a.partial + job * a.n_out * m + e;
            const bool first = (a.op == OP_SUM) ? (j == 2) : true;  // sum: levels 2..n accumulate
            if (first) *p = acc;
            else *p += acc;
        }
    }
}

// ---------------------------------------------------------------------------
// Sample-parallel forward scan (one job per instance, op "many": the W-basis
// sweep).  Chunks c of clen samples, [s_c, e_c] with e_c = min(s_c + clen, S-1),
// share their end points.  Pass 1 runs each chunk's recurrence from a zero
// base (chunk 0 from the true start) and keeps its end value and its share of
// the level sum; pass 2 turns the end values into chunk bases (exclusive scan
// in chunk order) and adds the shares in chunk order; pass 3 reruns each
// chunk's recurrence from its base and writes the prefix (and, in the last
// chunk, the level total).  Inside a chunk the operations are the sequential
// kernel's; only the chunk bases are summed in another order.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double lm_scan_input(const LmArgs& a, int j, long long b, long long job, long long i,
                                                int e) {
    if (j == 1) return raw_entry(a.in + (b * a.S + i) * a.m_in, e, a.sc);
    return a.Y[(job * a.S + i) * a.m + e];
}

__host__ __device__ __forceinline__ int lm_nchunks(const LmArgs& a) {
    return (int)((a.S - 1 + a.clen - 1) / a.clen);
}

// walk chunk c from base `at` (chunk 0: the true start), calling out(i, P, f_i)
template <int I, typename F>
__device__ __forceinline__ double lm_chunk_walk(const LmArgs& a, int j, long long b, long long job, int e, int c,
                                                double at, F&& out) {
    const long long S = a.S, s0 = (long long)c * a.clen;
    const long long e0 = (s0 + a.clen < S - 1) ? s0 + a.clen : S - 1;
    PrefixState<I> st;
    double P;
    const double fs = lm_scan_input(a, j, b, job, s0, e);
    if (c == 0) {
        P = st.step(0, fs, a.dt);
    } else {
        st.start(fs, at);
        P = at;
    }
    out(s0, P, fs);
    constexpr int U = 8;
    for (long long i0 = s0 + 1; i0 <= e0; i0 += U) {
        double f[U];
#pragma unroll
        for (int u = 0; u < U; ++u) f[u] = (i0 + u <= e0) ? lm_scan_input(a, j, b, job, i0 + u, e) : 0.0;
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (i0 + u <= e0) {
                P = st.step(i0 + u, f[u], a.dt);
                out(i0 + u, P, f[u]);
            }
    }
    return P;
}

template <int I>
__global__ void __launch_bounds__(256) lm_fwd_scan_c1(LmArgs a) {
    const long long JM = a.jobs * a.m;
    const int C = lm_nchunks(a);
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= JM * C) return;
    const int c = (int)(idx / JM);
    const long long je = idx - (long long)c * JM;
    const long long job = je / a.m;
    const int e = (int)(je - job * a.m);
    const long long b = job;  // one job per instance
    const long long S = a.S;
    const bool last = (c == C - 1);
    double acc = 0.0;
    const double cw = a.dt * (a.level == 1 ? 1.0 : a.w[0]);
    const double end = lm_chunk_walk<I>(a, a.level, b, job, e, c, 0.0, [&](long long i, double, double f) {
        if (i < (long long)(c + 1) * a.clen || last) sum_acc<I>(acc, i, S, f, cw);
    });
    a.cscan[(long long)c * JM + je] = end;
    a.cscan[((long long)C + c) * JM + je] = acc;
}

template <int I>
__global__ void __launch_bounds__(64) lm_fwd_scan_c2(LmArgs a) {
    const long long JM = a.jobs * a.m;
    const int C = lm_nchunks(a);
    const long long je = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (je >= JM) return;
    const long long job = je / a.m;
    const int e = (int)(je - job * a.m);
    const int n = a.n, j = a.level;
    double off = 0.0, acc = 0.0;
    // the chunks in batches: a batch's loads go out together (the stores into
    // the same array would otherwise order every load behind the previous store)
    constexpr int U = 16;
    for (int c0 = 0; c0 < C; c0 += U) {
        double inc[U], part[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int c = c0 + u;
            inc[u] = c < C ? a.cscan[(long long)c * JM + je] : 0.0;
            part[u] = c < C ? a.cscan[((long long)C + c) * JM + je] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int c = c0 + u;
            if (c < C) {
                a.cscan[(long long)c * JM + je] = off;
                off += inc[u];
                acc += part[u];
            }
        }
    }
    if (!a.outputs) return;
    if (j == 1) {
        a.omega1[job * a.m + e] = acc;
    } else {
        a.partial[(job * a.n_out + (j - 1)) * (long long)a.m + e] = acc;
    }
    (void)n;
}

template <int I>
__global__ void __launch_bounds__(256) lm_fwd_scan_c3(LmArgs a) {
    const long long JM = a.jobs * a.m;
    const int C = lm_nchunks(a);
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= JM * C) return;
    const int c = (int)(idx / JM);
    const long long je = idx - (long long)c * JM;
    const long long job = je / a.m;
    const int e = (int)(je - job * a.m);
    const int j = a.level, m = a.m;
    const bool last = (c == C - 1);
    double* Pout = lm_slot(a, job, job, 0, j);
    const double at = a.cscan[(long long)c * JM + je];
    const double end = lm_chunk_walk<I>(a, j, job, job, e, c, at, [&](long long i, double P, double) {
        if (i < (long long)(c + 1) * a.clen || last) Pout[i * m + e] = P;
    });
    if (last && a.totals) a.totals[(job * a.n + (j - 1)) * (long long)m + e] = end;
}

// ---------------------------------------------------------------------------
// Chunk passes of levels >= 2 with the level's products formed on the fly
// (generic d = 8, one job per instance).  lm_fwd_prod_rows writes Y_i =
// A_i (P_i + s T) for the whole level and both chunk passes read it back; here
// each chunk pass forms its chunk's products from the input and the previous
// level directly, so the level reads A and P^(j-1) twice and writes P^(j) once
// instead of reading them once, writing Y and reading Y twice.  A wavefront
// walks one chunk (lane = entry); each batch of kFuseU samples is staged in
// the wavefront's own LDS slice (A rows padded to d + 1 doubles) and every lane
// sums its entry's product in the rows kernel's order, so Y and the walk
// (lm_chunk_walk's) are the same operations as the unfused passes.
// ---------------------------------------------------------------------------
constexpr int kFuseU = 8;                            // samples per staged batch
constexpr int kFuseD = 8;                            // the fused dimension
constexpr int kFuseM = kFuseD * kFuseD;
constexpr int kFuseSlot = kFuseD * (kFuseD + 1) + kFuseM;  // doubles per staged sample

__host__ __device__ __forceinline__ bool fused_level_ok(const LmArgs& a) {
    return a.level >= 2 && !a.sc && a.d == kFuseD && a.m_in == a.m && a.Qj == 1;
}
__host__ __device__ __forceinline__ bool fused_raw_ok(const LmArgs& a) {
    return a.level == 1 && !a.sc && a.d == kFuseD && a.m_in == a.m && a.Qj == 1;
}
__host__ __device__ __forceinline__ long long fused_items(const LmArgs& a) {
    return (long long)lm_nchunks(a) * a.jobs;
}

// RAW: level 1, whose input is the sample itself (no product, no previous level)
template <int I, bool WRITE, bool RAW>
__global__ void __launch_bounds__(256, 4) lm_fwd_chunk_fused(LmArgs a) {
    constexpr int D = kFuseD, M = kFuseM;
    __shared__ double lds[4][kFuseU * kFuseSlot];
    const int wid = threadIdx.x >> 6, e = threadIdx.x & 63, r = e / D, cc = e % D;
    const long long item = (long long)blockIdx.x * 4 + wid;
    const long long jobs = a.jobs, S = a.S, JM = jobs * M;
    const int C = lm_nchunks(a);
    if (item >= (long long)C * jobs) return;  // whole wavefronts leave together
    const int c = (int)(item / jobs);
    const long long job = item - (long long)c * jobs;
    const long long je = job * M + e;
    const int j = a.level;
    const long long s0 = (long long)c * a.clen;
    const long long e0 = (s0 + a.clen < S - 1) ? s0 + a.clen : S - 1;
    const int len = (int)(e0 - s0);  // samples s0 + 1 .. e0 follow s0
    const bool last = (c == C - 1);
    const double* Ab = a.in + (job * S + s0) * M + e;
    const double* P = RAW ? nullptr : lm_slot(a, job, job, 0, j - 1);
    const double* Pb = RAW ? nullptr : P + s0 * M + e;
    const double shift = a.x[0] - 1;
    const double Te = RAW ? 0.0 : P[(S - 1) * M + e];
    double* slot = lds[wid];
    // pass 1: end value and level-sum share of the chunk from a zero base;
    // pass 3: the prefix from the chunk's base, written out
    double* Pout = WRITE ? lm_slot(a, job, job, 0, j) + s0 * M + e : nullptr;
    const double cw = a.dt * (RAW ? 1.0 : a.w[0]);
    const double at = WRITE ? a.cscan[(long long)c * JM + je] : 0.0;
    PrefixState<I> st;
    double Pv = 0.0, acc = 0.0;
    // a batch's loads are issued one batch ahead, so they land while the
    // previous batch is multiplied and walked
    double av[kFuseU], bv[kFuseU];
#pragma unroll
    for (int u = 0; u < kFuseU; ++u) {
        av[u] = 0.0;
        bv[u] = 0.0;
        if (u <= len) {
            av[u] = Ab[u * M];
            if (!RAW) {
                double v = Pb[u * M];
                v += Te * shift;
                bv[u] = v;
            }
        }
    }
    for (int o0 = 0; o0 <= len; o0 += kFuseU) {
        double raw[kFuseU];
        if (RAW) {
#pragma unroll
            for (int u = 0; u < kFuseU; ++u) raw[u] = av[u];
        } else {
            __builtin_amdgcn_wave_barrier();  // the previous batch's reads stay before these writes
#pragma unroll
            for (int u = 0; u < kFuseU; ++u) {
                slot[u * kFuseSlot + r * (D + 1) + cc] = av[u];
                slot[u * kFuseSlot + D * (D + 1) + e] = bv[u];
            }
            __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0): the writes have landed
            __builtin_amdgcn_wave_barrier();
        }
        const int o1 = o0 + kFuseU;
        if (o1 <= len) {
            const double* an = Ab + (long long)o1 * M;
            const double* pn = RAW ? nullptr : Pb + (long long)o1 * M;
#pragma unroll
            for (int u = 0; u < kFuseU; ++u) {
                av[u] = 0.0;
                bv[u] = 0.0;
                if (o1 + u <= len) {
                    av[u] = an[u * M];
                    if (!RAW) {
                        double v = pn[u * M];
                        v += Te * shift;
                        bv[u] = v;
                    }
                }
            }
        }
        // one sample at a time: its product, then its step (unrolling this loop
        // lets the compiler hoist all of the batch's LDS reads and spill)
#pragma unroll 1
        for (int u = 0; u < kFuseU; ++u) {
            const int o = o0 + u;
            if (o > len) break;
            double y = 0.0;
            if (RAW) {
                // a runtime-indexed register array would live in scratch: pick by unrolled compare
#pragma unroll
                for (int t = 0; t < kFuseU; ++t)
                    if (t == u) y = raw[t];
            } else {
                const double* Ar = slot + u * kFuseSlot + r * (D + 1);
                const double* Bc = slot + u * kFuseSlot + D * (D + 1) + cc;
#pragma unroll
                for (int l = 0; l < D; ++l) y += Ar[l] * Bc[l * D];
            }
            const long long i = s0 + o;
            if (o == 0) {
                if (c == 0) {
                    Pv = st.step(0, y, a.dt);
                } else {
                    st.start(y, at);
                    Pv = at;
                }
            } else {
                Pv = st.step(i, y, a.dt);
            }
            if (i < s0 + a.clen || last) {
                if (WRITE) Pout[(long long)o * M] = Pv;
                else sum_acc<I>(acc, i, S, y, cw);
            }
        }
    }
    if (WRITE) {
        if (last && a.totals) a.totals[(job * a.n + (j - 1)) * (long long)M + e] = Pv;
    } else {
        a.cscan[(long long)c * JM + je] = Pv;
        a.cscan[((long long)C + c) * JM + je] = acc;
    }
}

// VJP products for d = 8 streamed by wavefronts: a wavefront takes kVwSamples
// consecutive samples of one job (lane = entry (r, c)), stages each sample's
// b = P + s T, adjoint g and input A (rows padded to d + 1) in its own LDS
// slice, and writes dA[r][c] += sum_j g[r][j] b[c][j] and Gb[r][c] =
// sum_i A[i][r] g[i][c] with coalesced 512-byte rows; the next batch's loads
// are issued before the current batch is multiplied.  The sums run in
// lm_vjp_prod_rows' order, so the results are the same bits.
constexpr int kVwSamples = 32;                        // samples per wavefront
constexpr int kVwU = 4;                               // samples per staged batch
constexpr int kVwSlot = 3 * kFuseD * (kFuseD + 1);    // doubles per staged sample

__host__ __device__ __forceinline__ bool vjp_wave_ok(const LmArgs& a) {
    return !a.sc && a.d == kFuseD && a.m_in == a.m;
}

__global__ void __launch_bounds__(256, 4) lm_vjp_prod_wave(LmArgs a) {
    constexpr int D = kFuseD, M = kFuseM, LD = D + 1;
    __shared__ double lds[4][kVwU * kVwSlot];
    const int wid = threadIdx.x >> 6, e = threadIdx.x & 63, r = e / D, cc = e % D;
    const long long S = a.S, per = S * M;
    const long long nblk = (S + kVwSamples - 1) / kVwSamples;
    const long long item = (long long)blockIdx.x * 4 + wid;
    if (item >= a.jobs * nblk) return;  // whole wavefronts leave together
    const long long job = item / nblk;
    const long long i0 = (item - job * nblk) * kVwSamples;
    const int len = (int)(S - i0 < kVwSamples ? S - i0 : kVwSamples);
    const long long b = job / a.Q;
    const int q = (int)(job - b * a.Q);
    const double shift = a.x[q] - 1.0;
    const double* P = a.carry + ((b * a.Q + q) * (long long)(a.n - 1) + (a.level - 2)) * per;
    const double Te = P[(S - 1) * M + e];
    const double* Pb = P + i0 * M + e;
    const double* Gi = a.gbuf + (job * 2 + a.ga) * per + i0 * M + e;
    const double* Ai = a.in + (b * S + i0) * M + e;
    double* dA = a.dAjob + job * per + i0 * M + e;
    double* Gb = a.gbuf + (job * 2 + (1 - a.ga)) * per + i0 * M + e;
    double* slot = lds[wid];
    double vb[kVwU], vg[kVwU], va[kVwU];
    auto load = [&](int o0) {
#pragma unroll
        for (int u = 0; u < kVwU; ++u) {
            vb[u] = 0.0;
            vg[u] = 0.0;
            va[u] = 0.0;
            if (o0 + u < len) {
                double v = Pb[(long long)(o0 + u) * M];
                v += Te * shift;
                vb[u] = v;
                vg[u] = Gi[(long long)(o0 + u) * M];
                va[u] = Ai[(long long)(o0 + u) * M];
            }
        }
    };
    load(0);
    for (int o0 = 0; o0 < len; o0 += kVwU) {
        __builtin_amdgcn_wave_barrier();  // the previous batch's reads stay before these writes
#pragma unroll
        for (int u = 0; u < kVwU; ++u) {
            double* sl = slot + u * kVwSlot + r * LD + cc;
            sl[0] = vb[u];
            sl[D * LD] = vg[u];
            sl[2 * D * LD] = va[u];
        }
        __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0): the writes have landed
        __builtin_amdgcn_wave_barrier();
        double dav[kVwU];
#pragma unroll
        for (int u = 0; u < kVwU; ++u) dav[u] = (o0 + u < len) ? dA[(long long)(o0 + u) * M] : 0.0;
        if (o0 + kVwU < len) load(o0 + kVwU);
#pragma unroll
        for (int u = 0; u < kVwU; ++u) {
            if (o0 + u >= len) break;
            const double* bs = slot + u * kVwSlot;
            const double* gs = bs + D * LD;
            const double* as = gs + D * LD;
            double v = 0.0;  // sum_j g[r][j] b[c][j]
#pragma unroll
            for (int j = 0; j < D; ++j) v += gs[r * LD + j] * bs[cc * LD + j];
            double t2 = 0.0;  // sum_i a[i][r] g[i][c]
#pragma unroll
            for (int ii = 0; ii < D; ++ii) t2 += as[ii * LD + r] * gs[ii * LD + cc];
            dA[(long long)(o0 + u) * M] = dav[u] + v;
            Gb[(long long)(o0 + u) * M] = t2;
        }
    }
}

// ---------------------------------------------------------------------------
// VJP
// ---------------------------------------------------------------------------

__device__ __forceinline__ double lm_cot(const LmArgs& a, long long b, int k, int e) {
    const double* row = a.cot + (b * a.n_out + k) * (long long)a.m_cot;
    if (a.sc == 1) return e == 0 ? 0.0 : row[e - 1];
    return row[e];
}

// barY = sum_vjp(dOut_last, dscale) (grad.hpp:127, 216, 317); dA_job = 0
template <int I>
__global__ void __launch_bounds__(256) lm_vjp_init(LmArgs a) {
    const long long per = a.S * a.m;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= a.jobs * per) return;
    const long long job = idx / per;
    const long long rem = idx - job * per;
    const long long i = rem / a.m;
    const int e = (int)(rem - i * a.m);
    const long long b = job / a.Q;
    const int q = (int)(job - b * a.Q);
    const double dscale = a.dt * a.w[q];  // grad.hpp:315
    const double o = lm_cot(a, b, (a.op == OP_MANY) ? a.n - 1 : 0, e);
    a.gbuf[(job * 2 + a.ga) * per + rem] = o * sum_vjp_coef<I>(i, a.S, dscale);
    a.dAjob[idx] = 0.0;
}

// dA_job_i += barY_i b_i^T (one entry) and temp_i = A_i^T barY_i for level k
__global__ void __launch_bounds__(256) lm_vjp_prod(LmArgs a) {
    const long long per = a.S * a.m;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= a.jobs * per) return;
    const long long job = idx / per;
    const long long rem = idx - job * per;
    const long long i = rem / a.m;
    const int e = (int)(rem - i * a.m);
    const long long b = job / a.Q;
    const int q = (int)(job - b * a.Q);
    const double shift = a.x[q] - 1.0;  // grad.hpp:314
    const double* P = a.carry + ((b * a.Q + q) * (long long)(a.n - 1) + (a.level - 2)) * per;
    const double* T = P + (a.S - 1) * a.m;
    const double* g = a.gbuf + (job * 2 + a.ga) * per + i * a.m;
    const double* Ai = a.in + (b * a.S + i) * a.m_in + (a.sc == 2 ? 1 : 0);
    double dav, temp;
    prod_vjp(Ai, P + i * a.m, T, g, e, a.d, a.sc != 0, shift, dav, temp);
    if (!(a.sc && e == 0)) a.dAjob[idx] += dav;
    a.gbuf[(job * 2 + (1 - a.ga)) * per + rem] = temp;
}

// bar_total, prefix_vjp and sum_vjp_add of level k for one entry of one job
// (grad.hpp:225-235, 305-335)
template <int I>
__global__ void __launch_bounds__(256) lm_vjp_scan(LmArgs a) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= a.jobs * a.m) return;
    const long long job = idx / a.m;
    const int e = (int)(idx - job * a.m);
    const long long b = job / a.Q;
    const int q = (int)(job - b * a.Q);
    const long long S = a.S;
    const int m = a.m, k = a.level;
    const double shift = a.x[q] - 1.0;
    const double dscale = a.dt * a.w[q];
    double* G = a.gbuf + (job * 2 + (1 - a.ga)) * S * m + e;
    // bar_total accumulated from the last sample down, as the reference's reverse loop does
    const double tl = G[(S - 1) * m];
    double bt;
    if (a.bt_fused) {
        bt = tl * (1.0 + shift);
    } else {
        bt = tl;
        bt += tl * shift;
    }
    constexpr int U = 16;
    // shift 0 (the W-basis sweep): every term is a signed zero, bt is unchanged
    if (shift != 0.0) {
        for (long long i0 = S - 2; i0 >= 0; i0 -= U) {
            double t[U];
#pragma unroll
            for (int u = 0; u < U; ++u) t[u] = (i0 - u >= 0) ? G[(i0 - u) * m] : 0.0;
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (i0 - u >= 0) bt += t[u] * shift;
        }
    }
    if (a.wbar) bt += a.wbar[(job * a.n + (k - 2)) * (long long)m + e];  // cotangent of the level total
    G[(S - 1) * m] = bt;
    prefix_vjp<I>(G, m, S, a.dt);
    if (k > 2 && a.op != OP_ONE) {  // sum_vjp_add (grad.hpp:232, 332)
        const double o = lm_cot(a, b, a.op == OP_MANY ? k - 2 : 0, e);
        constexpr int V = 8;  // loads of a batch issued before its stores
        for (long long i0 = 0; i0 < S; i0 += V) {
            double g[V];
#pragma unroll
            for (int u = 0; u < V; ++u) g[u] = (i0 + u < S) ? G[(i0 + u) * m] : 0.0;
#pragma unroll
            for (int u = 0; u < V; ++u)
                if (i0 + u < S) G[(i0 + u) * m] = sum_vjp_add<I>(g[u], i0 + u, S, o, dscale);
        }
    }
}

// ---------------------------------------------------------------------------
// Sample-parallel reverse scan, Boole, W-basis sweep (shift 0).  The reverse
// block loop of prefix_vjp_boole carries g4 (the adjoint reaching a block's
// end, additive: a block passes r0 + g1 + g2 + g3 + g4 down) and `pending`
// (linear in the lowest block's g4).  Pass 1 runs each chunk with zero inputs
// and keeps its chain increment, its outgoing pending and its raw lowest-block
// start (which the chunk below overwrites); pass 2 walks the chunks top-down
// per entry to get each chunk's true inputs; pass 3 reruns each chunk from
// them and applies sum_vjp_add to the samples it writes.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void boole_rev_block(double r0, double g1, double g2, double g3, double& g4,
                                                double& pending, double dt, double& a1, double& a2, double& a3,
                                                double& a4) {
    double a0 = r0;
    a0 += g1;
    a0 += g2;
    a0 += g3;
    a0 += g4;
    a1 = g1 * (dt / 2);
    a1 += g2 * (4 * dt / 3);
    a1 += g3 * (9 * dt / 8);
    a1 += g4 * (64 * dt / 45);
    a2 = g2 * (dt / 3);
    a2 += g3 * (9 * dt / 8);
    a2 += g4 * (24 * dt / 45);
    a3 = g3 * (3 * dt / 8);
    a3 += g4 * (64 * dt / 45);
    a4 = g4 * (14 * dt / 45);
    a4 += pending;
    double p = g1 * (dt / 2);
    p += g2 * (dt / 3);
    p += g3 * (3 * dt / 8);
    p += g4 * (14 * dt / 45);
    pending = p;
    g4 = a0;
}

__global__ void __launch_bounds__(256) lm_vjp_scan_b1(LmArgs a) {
    const long long JM = a.jobs * a.m;
    const int C = lm_nchunks(a);
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= JM * C) return;
    const int c = (int)(idx / JM);
    const long long je = idx - (long long)c * JM;
    const long long job = je / a.m;
    const int e = (int)(je - job * a.m), m = a.m;
    const long long S = a.S, s0 = (long long)c * a.clen;
    const long long e0 = (s0 + a.clen < S - 1) ? s0 + a.clen : S - 1;
    const double* G = a.gbuf + (job * 2 + (1 - a.ga)) * S * m + e;
    double g4 = 0.0, pending = 0.0, a1, a2, a3, a4;
    for (long long s = e0 - 4; s >= s0; s -= 4) {
        boole_rev_block(G[s * m], G[(s + 1) * m], G[(s + 2) * m], G[(s + 3) * m], g4, pending, a.dt, a1, a2, a3,
                        a4);
    }
    a.cscan[(long long)c * JM + je] = g4;                  // chain increment
    a.cscan[((long long)C + c) * JM + je] = pending;       // outgoing pending with zero inputs
    a.cscan[((long long)2 * C + c) * JM + je] = G[s0 * m]; // raw lowest-block start
}

__global__ void __launch_bounds__(64) lm_vjp_scan_b2(LmArgs a) {
    const long long JM = a.jobs * a.m;
    const int C = lm_nchunks(a);
    const long long je = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (je >= JM) return;
    const long long job = je / a.m;
    const int e = (int)(je - job * a.m), m = a.m, k = a.level;
    const long long S = a.S;
    const double* G = a.gbuf + (job * 2 + (1 - a.ga)) * S * m + e;
    double g4 = G[(S - 1) * m];  // bar_total with shift 0: the last sample's value
    if (a.wbar) g4 += a.wbar[(job * a.n + (k - 2)) * (long long)m + e];
    double pending = 0.0;
    const double c14 = 14 * a.dt / 45;
    constexpr int U = 16;  // chunks per batch of loads, as in lm_fwd_scan_c2
    for (int c1 = C - 1; c1 >= 0; c1 -= U) {
        double inc[U], pl[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int c = c1 - u;
            inc[u] = c >= 0 ? a.cscan[(long long)c * JM + je] : 0.0;
            pl[u] = c >= 0 ? a.cscan[((long long)C + c) * JM + je] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int c = c1 - u;
            if (c >= 0) {
                a.cscan[(long long)c * JM + je] = g4;
                a.cscan[((long long)C + c) * JM + je] = pending;
                pending = pl[u] + g4 * c14;
                g4 = inc[u] + g4;
            }
        }
    }
}

__global__ void __launch_bounds__(256) lm_vjp_scan_b3(LmArgs a) {
    const long long JM = a.jobs * a.m;
    const int C = lm_nchunks(a);
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= JM * C) return;
    const int c = (int)(idx / JM);
    const long long je = idx - (long long)c * JM;
    const long long job = je / a.m;
    const int e = (int)(je - job * a.m), m = a.m, k = a.level;
    const long long S = a.S, s0 = (long long)c * a.clen;
    const long long e0 = (s0 + a.clen < S - 1) ? s0 + a.clen : S - 1;
    double* G = a.gbuf + (job * 2 + (1 - a.ga)) * S * m + e;
    const bool add = (k > 2 && a.op != OP_ONE);
    const double o = add ? lm_cot(a, job, a.op == OP_MANY ? k - 2 : 0, e) : 0.0;
    const double dscale = a.dt * a.w[0];
    double g4 = a.cscan[(long long)c * JM + je];
    double pending = a.cscan[((long long)C + c) * JM + je];
    const double raw0 = a.cscan[((long long)2 * C + c) * JM + je];
    double a1, a2, a3, a4;
    for (long long s = e0 - 4; s >= s0; s -= 4) {
        const double r0 = (s == s0) ? raw0 : G[s * m];
        boole_rev_block(r0, G[(s + 1) * m], G[(s + 2) * m], G[(s + 3) * m], g4, pending, a.dt, a1, a2, a3, a4);
        if (add) {
            a1 = sum_vjp_add<BOOLE>(a1, s + 1, S, o, dscale);
            a2 = sum_vjp_add<BOOLE>(a2, s + 2, S, o, dscale);
            a3 = sum_vjp_add<BOOLE>(a3, s + 3, S, o, dscale);
            a4 = sum_vjp_add<BOOLE>(a4, s + 4, S, o, dscale);
        }
        G[(s + 1) * m] = a1;
        G[(s + 2) * m] = a2;
        G[(s + 3) * m] = a3;
        G[(s + 4) * m] = a4;
    }
    if (c == 0) G[0] = add ? sum_vjp_add<BOOLE>(pending, 0, S, o, dscale) : pending;
}

// dA_job += barY  (grad.hpp:144, 247, 347)
__global__ void __launch_bounds__(256) lm_vjp_final(LmArgs a) {
    const long long per = a.S * a.m;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= a.jobs * per) return;
    const long long job = idx / per;
    a.dAjob[idx] += a.gbuf[(job * 2 + a.ga) * per + (idx - job * per)];
}

// ---------------------------------------------------------------------------
// GL-lane reductions in fixed q order
// ---------------------------------------------------------------------------
// out[b] = Omega_1 + sum_q partial[b][q]   (lane reduction, magnus.hpp:86-90)
__global__ void fwd_reduce_kernel(FwdReduceArgs a) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long per = (long long)a.n_out * a.m_out;
    if (idx >= a.B * per) return;
    const long long b = idx / per;
    const int rem = (int)(idx - b * per);
    const int k = rem / a.m_out;
    const int eo = rem - k * a.m_out;
    const int e = a.sc ? eo + 1 : eo;  // SpaceCurve output drops component 0
    double v = 0.0;
    if (a.n == 1) {
        v = (k == 0) ? a.omega1[b * a.m + e] : 0.0;
    } else {
        if (a.op != OP_ONE && k == 0) v = a.omega1[b * a.m + e];
        if (!(a.op == OP_MANY && k == 0))  // Omega_1 has no GL-node contributions
            for (int q = 0; q < a.Qj; ++q) v += a.partial[((b * a.Qj + q) * a.n_out + k) * a.m + e];
    }
    a.out[idx] = v;
}

// dA[b] = sum_vjp(dOut_first, dt) + sum_q dA_q[b]
template <int I>
__global__ void vjp_reduce_kernel(VjpReduceArgs a) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long per = a.S * a.m_out;
    if (idx >= a.B * per) return;
    const long long b = idx / per;
    const long long rem = idx - b * per;
    const long long i = rem / a.m_out;
    const int eo = (int)(rem - i * a.m_out);
    const int e = a.sc ? eo + 1 : eo;
    double v = 0.0;
    if (!(a.op == OP_ONE && a.n > 1)) {
        const double* row = a.cot + b * (long long)a.n_out * a.m_cot;  // dOut[0] / dOut
        const double o = a.sc ? row[e - 1] : row[e];
        v = o * sum_vjp_coef<I>(i, a.S, a.dt);
    }
    for (int q = 0; q < a.Qj; ++q) v += a.dAjob[((b * a.Qj + q) * a.S + i) * a.m + e];
    a.dA[idx] = v;
}

inline long long blocks_for(long long n) { return (n + 255) / 256; }

// Product path of a level: D (MFMA, default split), D + W (MFMA, W wavefronts
// per product), 1 (VALU row products, d <= 16) or 0 (VALU, one thread per
// output entry).  a.tile selects:
// 0 default (MFMA for d in {16, 32, 64}), 1 VALU, 2 / 3 d = 32 split over 2 / 4
// wavefronts; a choice that does not apply to d falls back to the default.
inline int prod_path(const LmArgs& a) {
    const bool mf = !a.sc && (a.d == 16 || a.d == 32 || a.d == 64);
    const bool rows = !a.sc && a.d <= kRowMaxD;  // row products: VALU with LDS-staged operands
    if (!mf || a.tile == 1) return rows ? 1 : 0;
    if (a.d == 32 && a.tile == 2) return 32 + 2;
    if (a.d == 32 && a.tile == 3) return 32 + 4;
    return a.d;
}

// the sample-parallel scan applies to the W-basis sweep (one job per
// instance, all level sums kept) with a chunk scratch, block rules that resume
// from a base value, and at least two chunks
template <int I>
inline bool chunked_scan(const LmArgs& a) {
    return I != TRAPEZOID && a.cscan && a.clen > 0 && a.clen % 4 == 0 && a.Qj == 1 && a.op == OP_MANY &&
           a.S - 1 >= 2 * (long long)a.clen;
}

template <int D, int W>
inline void launch_mfma(void (*k)(LmArgs), const LmArgs& a, hipStream_t st) {
    constexpr int ppb = MfmaShape<D, W>::PPB;
    const long long prods = a.jobs * a.S;
    hipLaunchKernelGGL(k, dim3((prods + ppb - 1) / ppb), dim3(256), 0, st, a);
}

// Algorithmic work of one level's launches, as the timers record it (8 bytes
// per value, FMA = 2 flops): a product launch reads the samples once and the
// previous level's values and writes the products (the VJP also reads the
// carry and the adjoint and writes dA_job and the next adjoint: 5 values per
// entry), P = 2 d^3 flops per product (20 for the su(2) product); a scan
// launch reads the products and writes the prefix level, 8 flops per entry.
struct LevelCounts {
    double prod_bytes, prod_flops, scan_bytes, scan_flops;
};
inline LevelCounts level_counts(const LmArgs& a, bool vjp) {
    const double jobs = (double)a.jobs, S = (double)a.S, m = a.m;
    const double B = jobs / (vjp ? a.Q : a.Qj);
    const double P = a.sc ? 20.0 : 2.0 * a.d * a.d * (double)a.d;
    LevelCounts c;
    c.prod_flops = (vjp ? 2.0 : 1.0) * jobs * S * P;
    c.prod_bytes = 8.0 * (B * S * a.m_in + (vjp ? 5.0 : 2.0) * jobs * S * m);
    c.scan_flops = 8.0 * jobs * S * m;
    c.scan_bytes = 16.0 * jobs * S * m;
    return c;
}

template <int I>
hipError_t lm_forward_level_i(const LmArgs& a, hipStream_t st) {
    const LevelCounts cnt = level_counts(a, false);
    bool fused = false, raw = false;
    if constexpr (I != TRAPEZOID) {
        const bool chunked = chunked_scan<I>(a);
        fused = chunked && fused_level_ok(a);
        raw = chunked && fused_raw_ok(a);
    }
    if (fused || raw) {
        // levels >= 2: the products are formed inside both chunk passes (the
        // level reads the samples and the previous level and writes this one);
        // level 1: the same wavefront-per-chunk passes over the samples
        void* h = fused ? timer_begin("lm_fwd_fused", st, cnt.prod_bytes, cnt.prod_flops + cnt.scan_flops)
                        : timer_begin("lm_fwd_scan", st, cnt.scan_bytes, cnt.scan_flops);
        const long long JM = a.jobs * a.m, waves = fused_items(a);  // one chunk per wavefront
        const dim3 grid((unsigned)((waves + 3) / 4));
        if constexpr (I != TRAPEZOID) {
            if (fused) {
                hipLaunchKernelGGL((lm_fwd_chunk_fused<I, false, false>), grid, dim3(256), 0, st, a);
                hipLaunchKernelGGL((lm_fwd_scan_c2<I>), dim3((JM + 63) / 64), dim3(64), 0, st, a);
                if (a.level < a.n)
                    hipLaunchKernelGGL((lm_fwd_chunk_fused<I, true, false>), grid, dim3(256), 0, st, a);
            } else {
                hipLaunchKernelGGL((lm_fwd_chunk_fused<I, false, true>), grid, dim3(256), 0, st, a);
                hipLaunchKernelGGL((lm_fwd_scan_c2<I>), dim3((JM + 63) / 64), dim3(64), 0, st, a);
                if (a.level < a.n)
                    hipLaunchKernelGGL((lm_fwd_chunk_fused<I, true, true>), grid, dim3(256), 0, st, a);
            }
        }
        timer_end(h);
        return hipGetLastError();
    }
    if (a.level >= 2) {
        void* h = timer_begin("lm_fwd_prod", st, cnt.prod_bytes, cnt.prod_flops);
        switch (prod_path(a)) {
            case 16: launch_mfma<16, 1>(lm_fwd_prod_mfma<16, 1>, a, st); break;
            case 32: launch_mfma<32, 1>(lm_fwd_prod_mfma<32, 1>, a, st); break;
            case 32 + 2: launch_mfma<32, 2>(lm_fwd_prod_mfma<32, 2>, a, st); break;
            case 32 + 4: launch_mfma<32, 4>(lm_fwd_prod_mfma<32, 4>, a, st); break;
            case 64: launch_mfma<64, 4>(lm_fwd_prod_mfma<64, 4>, a, st); break;
            case 1: {
                const int spb = rows_spb(a.d);
                hipLaunchKernelGGL(lm_fwd_prod_rows, dim3((a.jobs * a.S + spb - 1) / spb), dim3(256),
                                   rows_lds_bytes(a.d, spb), st, a);
                break;
            }
            default: hipLaunchKernelGGL(lm_fwd_prod, dim3(blocks_for(a.jobs * a.S * a.m)), dim3(256), 0, st, a);
        }
        timer_end(h);
    }
    void* h = timer_begin("lm_fwd_scan", st, cnt.scan_bytes, cnt.scan_flops);
    bool chunked = false;
    if constexpr (I != TRAPEZOID) {
        chunked = chunked_scan<I>(a);
        if (chunked) {
            const long long JM = a.jobs * a.m, work = JM * lm_nchunks(a);
            hipLaunchKernelGGL((lm_fwd_scan_c1<I>), dim3((work + 255) / 256), dim3(256), 0, st, a);
            hipLaunchKernelGGL((lm_fwd_scan_c2<I>), dim3((JM + 63) / 64), dim3(64), 0, st, a);
            if (a.level < a.n)
                hipLaunchKernelGGL((lm_fwd_scan_c3<I>), dim3((work + 255) / 256), dim3(256), 0, st, a);
        }
    }
    if (!chunked) {
        // one wavefront per workgroup: the scans are latency-bound chains, so
        // they are spread over as many CUs (memory queues) as there are wavefronts
        hipLaunchKernelGGL((lm_fwd_scan<I>), dim3((a.jobs * a.m + 63) / 64), dim3(64), 0, st, a);
    }
    timer_end(h);
    return hipGetLastError();
}

template <int I>
hipError_t lm_vjp_level_i(const LmArgs& a, hipStream_t st) {
    const LevelCounts cnt = level_counts(a, true);
    void* h = timer_begin("lm_vjp_prod", st, cnt.prod_bytes, cnt.prod_flops);
    switch (prod_path(a)) {
        case 16: launch_mfma<16, 1>(lm_vjp_prod_mfma<16, 1>, a, st); break;
        case 32: launch_mfma<32, 1>(lm_vjp_prod_mfma<32, 1>, a, st); break;
        case 32 + 2: launch_mfma<32, 2>(lm_vjp_prod_mfma<32, 2>, a, st); break;
        case 32 + 4: launch_mfma<32, 4>(lm_vjp_prod_mfma<32, 4>, a, st); break;
        case 64: launch_mfma<64, 4>(lm_vjp_prod_mfma<64, 4>, a, st); break;
        case 1: {
            if (vjp_wave_ok(a)) {
                const long long waves = a.jobs * ((a.S + kVwSamples - 1) / kVwSamples);
                hipLaunchKernelGGL(lm_vjp_prod_wave, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, st, a);
                break;
            }
            const int spb = vjp_rows_spb(a.d);
            hipLaunchKernelGGL(lm_vjp_prod_rows, dim3((a.jobs * a.S + spb - 1) / spb), dim3(256),
                               (size_t)3 * spb * rows_stride(a.d) * sizeof(double), st, a);
            break;
        }
        default: hipLaunchKernelGGL(lm_vjp_prod, dim3(blocks_for(a.jobs * a.S * a.m)), dim3(256), 0, st, a);
    }
    timer_end(h);
    h = timer_begin("lm_vjp_scan", st, cnt.scan_bytes, cnt.scan_flops);
    if (I == BOOLE && chunked_scan<I>(a) && a.Q == 1 && a.x[0] == 1.0) {
        const long long JM = a.jobs * a.m, work = JM * lm_nchunks(a);
        hipLaunchKernelGGL(lm_vjp_scan_b1, dim3((work + 255) / 256), dim3(256), 0, st, a);
        hipLaunchKernelGGL(lm_vjp_scan_b2, dim3((JM + 63) / 64), dim3(64), 0, st, a);
        hipLaunchKernelGGL(lm_vjp_scan_b3, dim3((work + 255) / 256), dim3(256), 0, st, a);
    } else {
        hipLaunchKernelGGL((lm_vjp_scan<I>), dim3((a.jobs * a.m + 63) / 64), dim3(64), 0, st, a);
    }
    timer_end(h);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_fwd_reduce(const FwdReduceArgs& a, hipStream_t s) {
    const long long total = a.B * (long long)a.n_out * a.m_out;
    if (total == 0) return hipSuccess;
    const int block = 256;
    const long long grid = (total + block - 1) / block;
    hipLaunchKernelGGL(fwd_reduce_kernel, dim3(grid), dim3(block), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_vjp_reduce(int integ, const VjpReduceArgs& a, hipStream_t s) {
    const long long total = a.B * a.S * a.m_out;
    if (total == 0) return hipSuccess;
    const int block = 256;
    const long long grid = (total + block - 1) / block;
    switch (integ) {
        case BOOLE: hipLaunchKernelGGL((vjp_reduce_kernel<BOOLE>), dim3(grid), dim3(block), 0, s, a); break;
        case SIMPSON: hipLaunchKernelGGL((vjp_reduce_kernel<SIMPSON>), dim3(grid), dim3(block), 0, s, a); break;
        case TRAPEZOID: hipLaunchKernelGGL((vjp_reduce_kernel<TRAPEZOID>), dim3(grid), dim3(block), 0, s, a); break;
        default: hipLaunchKernelGGL((vjp_reduce_kernel<RIEMANN>), dim3(grid), dim3(block), 0, s, a); break;
    }
    return hipGetLastError();
}

hipError_t launch_lm_forward_level(int integ, const LmArgs& a, hipStream_t s) {
    switch (integ) {
        case BOOLE: return lm_forward_level_i<BOOLE>(a, s);
        case SIMPSON: return lm_forward_level_i<SIMPSON>(a, s);
        case TRAPEZOID: return lm_forward_level_i<TRAPEZOID>(a, s);
        default: return lm_forward_level_i<RIEMANN>(a, s);
    }
}

hipError_t launch_lm_vjp_init(int integ, const LmArgs& a, hipStream_t s) {
    const dim3 grid(blocks_for(a.jobs * a.S * a.m));
    void* h = timer_begin("lm_vjp_init", s);
    switch (integ) {
        case BOOLE: hipLaunchKernelGGL((lm_vjp_init<BOOLE>), grid, dim3(256), 0, s, a); break;
        case SIMPSON: hipLaunchKernelGGL((lm_vjp_init<SIMPSON>), grid, dim3(256), 0, s, a); break;
        case TRAPEZOID: hipLaunchKernelGGL((lm_vjp_init<TRAPEZOID>), grid, dim3(256), 0, s, a); break;
        default: hipLaunchKernelGGL((lm_vjp_init<RIEMANN>), grid, dim3(256), 0, s, a); break;
    }
    timer_end(h);
    return hipGetLastError();
}

hipError_t launch_lm_vjp_level(int integ, const LmArgs& a, hipStream_t s) {
    switch (integ) {
        case BOOLE: return lm_vjp_level_i<BOOLE>(a, s);
        case SIMPSON: return lm_vjp_level_i<SIMPSON>(a, s);
        case TRAPEZOID: return lm_vjp_level_i<TRAPEZOID>(a, s);
        default: return lm_vjp_level_i<RIEMANN>(a, s);
    }
}

hipError_t launch_lm_vjp_final(const LmArgs& a, hipStream_t s) {
    void* h = timer_begin("lm_vjp_final", s);
    hipLaunchKernelGGL(lm_vjp_final, dim3(blocks_for(a.jobs * a.S * a.m)), dim3(256), 0, s, a);
    timer_end(h);
    return hipGetLastError();
}

}  // namespace magnus
