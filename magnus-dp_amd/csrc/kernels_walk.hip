// kernels_walk.hip -- level walkers of the per-node recursion for d in {16,
// 32} under the Boole rule: the forward levels of a record_vjp call (or of
// the VJP::none recompute) and the reverse levels of a VJP from the carry.
//
// The level kernels of kernels_lm.hip run a level of the reference recursion
// (src/magnus.hpp:230-254, src/grad.hpp:251-349) as a product launch over
// every (job, sample), which writes the products Y to HBM, and a scan launch
// per (job, entry) that reads them back.  Here one workgroup per job
// (instance b, GL node q) walks the samples of one level once.  Wavefront w
// owns the 16 x 16 output tile w of every sample's product; it forms the tile
// with v_mfma_f64_16x16x4f64 from operands loaded straight into registers one
// sample ahead and runs the per-entry recurrence on the result in the same
// registers, so a level reads its inputs once and writes its outputs once and
// no product array exists.
//
// Operand layout.  An MFMA step takes A[i][k] from lane i + 16 kk and B[k][j]
// from lane j + 16 kk (kk = lane / 16); step s of a tile uses k = kk * D/4 + s.
// A row operand (X[i][k]) is then D/4 contiguous doubles per lane (16-byte
// loads, a wavefront reads 16 whole rows), a column operand (X[k][j]) one
// double per step with 16 lanes on 128 contiguous bytes of a row.  Lane l,
// accumulator r of a tile holds row 4 r + l / 16, column l % 16.
//
// Job order.  Workgroup g runs on XCD g % 8; job xcd_job(g) gives each XCD a
// contiguous range of jobs, so the Q jobs of an instance walk the same samples
// on one XCD at about the same time and read A through its L2 once.
//
// Forward (walk_fwd): level 1 is the raw sample, levels j >= 2 the product
// Y_i = A_i (P^(j-1)_i + s_q T), T = P^(j-1)_{S-1}; each lane runs
// PrefixState<BOOLE> and sum_acc on its four entries, writes P^(j) into the
// carry slot (j < n) and the level's sum into partial / omega1 exactly as
// lm_fwd_scan does.  The products sum k in the MFMA's order; everything else
// is the sequential scan's arithmetic.
//
// VJP (walk_rev), level k from n down to 2, from the cotangent barY of Y^(k):
// dA_job,i += barY_i (P_i + s T)^T and temp_i = A_i^T barY_i (grad.hpp:311-
// 313), then bar_total, prefix_vjp and sum_vjp_add give the cotangent of
// Y^(k-1) (grad.hpp:225-235, 305-335).  bar_total needs every temp_i before
// the reverse walk starts at the last sample; prefix_vjp is linear and maps a
// value X at the last sample to X times the sum weights (the Boole block
// recurrence passes a block's end value down whole), so the walker runs the
// block recurrence on temp alone, accumulates Delta = s_q sum_i temp_i in the
// reference's order (bar_total = temp_{S-1} + Delta) and leaves
// barY^(k-1)_i - c_i Delta in gbuf and Delta in dbar; the next level (or the
// final reduce) adds c_i Delta as it reads.  Level n forms its cotangent
// inline (lm_vjp_init's values) and writes dA_job instead of adding to it.
#include "kernels.hpp"
#include "rules.hpp"
#include "walk.hpp"

namespace magnus {

namespace {

typedef double dbl4 __attribute__((ext_vector_type(4)));

template <int D>
struct WalkShape {
    static constexpr int TT = D / 16;          // tiles per side
    static constexpr int WAVES = TT * TT;      // one output tile per wavefront
    static constexpr int THREADS = 64 * WAVES;
    static constexpr int KS = D / 4;           // MFMA steps per tile
    static constexpr int M = D * D;
};

__device__ __forceinline__ long long xcd_job(long long g, long long G) {
    const long long per = G / 8, rem = G % 8, x = g % 8, s = g / 8;
    return x * per + (x < rem ? x : rem) + s;
}

// P^(j) of job (b, q): carry[b][q][j-1] or a ping-pong buffer (the level kernels' layout)
__device__ __forceinline__ double* walk_slot(const LmArgs& a, long long job, long long b, int q, int j) {
    if (a.carry) return a.carry + ((b * a.Q + q) * (long long)(a.n - 1) + (j - 1)) * a.S * a.m;
    return a.pingpong + (job * 2 + (j & 1)) * a.S * a.m;
}

__device__ __forceinline__ double walk_cot(const LmArgs& a, long long b, int k, int e) {
    return a.cot[(b * a.n_out + k) * (long long)a.m_cot + e];
}

// KS contiguous doubles (16-byte aligned) into v[0..KS-1]
template <int KS>
__device__ __forceinline__ void load_row(double* v, const double* p) {
    const double2* p2 = reinterpret_cast<const double2*>(p);
#pragma unroll
    for (int h = 0; h < KS / 2; ++h) {
        const double2 t = p2[h];
        v[2 * h] = t.x;
        v[2 * h + 1] = t.y;
    }
}

// ---------------------------------------------------------------------------
// forward level
// ---------------------------------------------------------------------------
template <int D>
struct FwdOps {
    double a[WalkShape<D>::KS];  // A_i row operand (RAW: a[0..3] are the lane's four entries)
    double p[WalkShape<D>::KS];  // P^(j-1)_i column operand
};

// Operand buffers in flight: the loads of samples i + 1 .. i + NB - 1 are issued
// while sample i is formed.  Register-bound at d = 32 with products (a
// tighter register budget for more resident workgroups spilled and ran
// slower); level 1 holds four values per sample.
template <int D, bool RAW>
constexpr int fwd_buffers() { return RAW ? 4 : (D == 16 ? 4 : 2); }
template <int D>
constexpr int rev_buffers() { return D == 16 ? 3 : 2; }

template <int D, bool RAW>
__global__ void __launch_bounds__(WalkShape<D>::THREADS) walk_fwd(LmArgs a) {
    using W = WalkShape<D>;
    constexpr int KS = W::KS, M = W::M;
    const long long job = xcd_job(blockIdx.x, gridDim.x);
    const long long b = job / a.Qj;
    const int q = (int)(job - b * a.Qj);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int I = wave / W::TT, J = wave - (wave / W::TT) * W::TT, li = lane & 15, lk = lane >> 4;
    const long long S = a.S;
    const int n = a.n, j = a.level;
    const double dt = a.dt;
    const bool om1 = a.outputs && j == 1 && q == 0 && (a.op != OP_ONE || n == 1);
    const bool lvl = a.outputs && j >= 2 && (a.op != OP_ONE || j == n);
    const double c = (j == 1) ? dt : dt * a.w[q];
    double* Pout = (j < n) ? walk_slot(a, job, b, q, j) : nullptr;
    const double* A = a.in + b * S * M;
    int e[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) e[r] = (16 * I + 4 * r + lk) * D + 16 * J + li;
    // operand bases (sample 0) and the shifted totals s_q T of the column operand
    const double* arow = A + (16 * I + li) * D + lk * KS;
    const double* Pin = RAW ? nullptr : walk_slot(a, job, b, q, j - 1);
    const double* pcol = RAW ? nullptr : Pin + (lk * KS) * D + 16 * J + li;
    double sT[KS];
    if constexpr (!RAW) {
        const double shift = a.x[q] - 1.0;  // magnus.hpp:65
        const double* T = Pin + (S - 1) * M;
#pragma unroll
        for (int s = 0; s < KS; ++s) sT[s] = T[(lk * KS + s) * D + 16 * J + li] * shift;
    }
    auto load = [&](long long i, FwdOps<D>& o) {
        if constexpr (RAW) {
#pragma unroll
            for (int r = 0; r < 4; ++r) o.a[r] = A[i * M + e[r]];
        } else {
            load_row<KS>(o.a, arow + i * M);
#pragma unroll
            for (int s = 0; s < KS; ++s) o.p[s] = pcol[i * M + s * D];
        }
    };
    PrefixState<BOOLE> st[4];
    double sum[4] = {0.0, 0.0, 0.0, 0.0}, last[4] = {0.0, 0.0, 0.0, 0.0};
    auto step = [&](long long i, const FwdOps<D>& o, auto) {
        double f[4];
        if constexpr (RAW) {
#pragma unroll
            for (int r = 0; r < 4; ++r) f[r] = o.a[r];
        } else {
            dbl4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int s = 0; s < KS; ++s) {
                double bv = o.p[s];
                bv += sT[s];
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(o.a[s], bv, acc, 0, 0, 0);
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) f[r] = acc[r];
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            if (om1 || lvl) sum_acc<BOOLE>(sum[r], i, S, f[r], c);
            if (Pout) {
                const double P = st[r].step(i, f[r], dt);
                Pout[i * M + e[r]] = P;
                last[r] = P;
            }
        }
    };
    constexpr int NB = fwd_buffers<D, RAW>();
    FwdOps<D> o[NB];
    walk_range<NB, NB, true>(0, S - 1, o, load, step);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        if (a.totals && Pout) a.totals[(job * n + (j - 1)) * M + e[r]] = last[r];
        if (om1) a.omega1[b * M + e[r]] = sum[r];
        if (lvl) {
            if (a.op == OP_MANY) {
                a.partial[(job * a.n_out + (j - 1)) * M + e[r]] = sum[r];
            } else {
                double* p = a.partial + job * a.n_out * M + e[r];
                const bool first = (a.op == OP_SUM) ? (j == 2) : true;  // sum: levels 2..n accumulate
                if (first) *p = sum[r];
                else *p += sum[r];
            }
        }
    }
}

// ---------------------------------------------------------------------------
// VJP level
// ---------------------------------------------------------------------------
template <int D>
struct RevOps {
    double urow[WalkShape<D>::KS];  // barY_i row operand (stored part; TOP: unused)
    double prow[WalkShape<D>::KS];  // P^(k-1)_i row operand (B^T of the dA product)
    double acol[WalkShape<D>::KS];  // A_i column operand (A^T of the temp product)
    double ucol[WalkShape<D>::KS];  // barY_i column operand (stored part; TOP: unused)
    double da[4];                   // dA_job,i at the lane's entries (TOP: unused)
};

template <int D, bool TOP>
__global__ void __launch_bounds__(WalkShape<D>::THREADS) walk_rev(LmArgs a) {
    using W = WalkShape<D>;
    constexpr int KS = W::KS, M = W::M;
    const long long job = xcd_job(blockIdx.x, gridDim.x);
    const long long b = job / a.Q;
    const int q = (int)(job - b * a.Q);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int I = wave / W::TT, J = wave - (wave / W::TT) * W::TT, li = lane & 15, lk = lane >> 4;
    const long long S = a.S;
    const int n = a.n, k = a.level;
    const double dt = a.dt;
    const double shift = a.x[q] - 1.0;   // grad.hpp:314
    const double dscale = dt * a.w[q];   // grad.hpp:315
    int e[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) e[r] = (16 * I + 4 * r + lk) * D + 16 * J + li;
    const int erow = (16 * I + li) * D + lk * KS;  // row operand of barY: entries erow + s
    const int ecol = (lk * KS) * D + 16 * J + li;  // column operand of barY: entries ecol + s D
    const double* A = a.in + b * S * M;
    const double* P = a.carry + ((b * a.Q + q) * (long long)(n - 1) + (k - 2)) * S * M;  // P^(k-1)
    const double* U = a.gbuf + (job * 2 + a.ga) * S * M;            // barY^(k) less c_i Delta^(k)
    double* Uo = a.gbuf + (job * 2 + (1 - a.ga)) * S * M;           // barY^(k-1) less c_i Delta^(k-1)
    const double* Dup = a.dbar + (job * 2 + a.ga) * M;
    double* Dout = a.dbar + (job * 2 + (1 - a.ga)) * M;
    double* dAj = a.dAjob + job * S * M;
    // per-lane constants: the cotangent (TOP) or the deferred Delta at the
    // barY operand positions, and s T^T at the dA product's B operand
    double yrow[KS], ycol[KS], sTrow[KS];
    {
        const double* T = P + (S - 1) * M;
        const int prow0 = (16 * J + li) * D + lk * KS;
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            sTrow[s] = T[prow0 + s] * shift;
            if constexpr (TOP) {
                const int kr = (a.op == OP_MANY) ? n - 1 : 0;
                yrow[s] = walk_cot(a, b, kr, erow + s);
                ycol[s] = walk_cot(a, b, kr, ecol + s * D);
            } else {
                yrow[s] = Dup[erow + s];
                ycol[s] = Dup[ecol + s * D];
            }
        }
    }
    const bool addo = k > 2 && a.op != OP_ONE;  // sum_vjp_add of the level below (grad.hpp:232, 332)
    double osum[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) osum[r] = addo ? walk_cot(a, b, a.op == OP_MANY ? k - 2 : 0, e[r]) : 0.0;
    const int prow0 = (16 * J + li) * D + lk * KS;
    auto load = [&](long long i, RevOps<D>& o) {
        const long long base = i * M;
        load_row<KS>(o.prow, P + base + prow0);
#pragma unroll
        for (int s = 0; s < KS; ++s) o.acol[s] = A[base + (lk * KS + s) * D + 16 * I + li];
        if constexpr (!TOP) {
            load_row<KS>(o.urow, U + base + erow);
#pragma unroll
            for (int s = 0; s < KS; ++s) o.ucol[s] = U[base + ecol + s * D];
#pragma unroll
            for (int r = 0; r < 4; ++r) o.da[r] = dAj[base + e[r]];
        }
    };
    // reverse state per entry: the raw values of the open block, the chained
    // g4 and pending (prefix_vjp_boole), and Delta
    double g1[4], g2[4], g3[4], g4[4], pend[4], dl[4];
    auto fin = [&](double v, long long i, int r) { return addo ? sum_vjp_add<BOOLE>(v, i, S, osum[r], dscale) : v; };
    // POS: the sample's position in its Boole block (i & 3), 4 for the last
    // sample S - 1 (bar_total with shift 0, lm_vjp_scan_b2)
    auto step_at = [&](long long i, const RevOps<D>& o, auto posc) {
        constexpr int POS = decltype(posc)::value;
        // barY_i at the operand positions
        const double cy = TOP ? sum_vjp_coef<BOOLE>(i, S, dscale) : sum_vjp_coef<BOOLE>(i, S, dt);
        dbl4 dav = {0.0, 0.0, 0.0, 0.0}, tmp = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            const double ya = TOP ? yrow[s] * cy : o.urow[s] + cy * yrow[s];
            double bv = o.prow[s];
            bv += sTrow[s];
            dav = __builtin_amdgcn_mfma_f64_16x16x4f64(ya, bv, dav, 0, 0, 0);
        }
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            const double yb = TOP ? ycol[s] * cy : o.ucol[s] + cy * ycol[s];
            tmp = __builtin_amdgcn_mfma_f64_16x16x4f64(o.acol[s], yb, tmp, 0, 0, 0);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            dAj[i * M + e[r]] = TOP ? dav[r] : o.da[r] + dav[r];
            const double t = tmp[r];
            if constexpr (POS == 4) {
                g4[r] = t;
                pend[r] = 0.0;
                dl[r] = t * shift;
            } else {
                dl[r] += t * shift;
                if constexpr (POS == 3) {
                    g3[r] = t;
                } else if constexpr (POS == 2) {
                    g2[r] = t;
                } else if constexpr (POS == 1) {
                    g1[r] = t;
                } else {  // block [i, i + 4] (boole.hpp:108-151)
                    double a0 = t;
                    a0 += g1[r];
                    a0 += g2[r];
                    a0 += g3[r];
                    a0 += g4[r];
                    double a1 = g1[r] * (dt / 2);
                    a1 += g2[r] * (4 * dt / 3);
                    a1 += g3[r] * (9 * dt / 8);
                    a1 += g4[r] * (64 * dt / 45);
                    double a2 = g2[r] * (dt / 3);
                    a2 += g3[r] * (9 * dt / 8);
                    a2 += g4[r] * (24 * dt / 45);
                    double a3 = g3[r] * (3 * dt / 8);
                    a3 += g4[r] * (64 * dt / 45);
                    double a4 = g4[r] * (14 * dt / 45);
                    a4 += pend[r];
                    double p = g1[r] * (dt / 2);
                    p += g2[r] * (dt / 3);
                    p += g3[r] * (3 * dt / 8);
                    p += g4[r] * (14 * dt / 45);
                    pend[r] = p;
                    g4[r] = a0;
                    Uo[(i + 1) * M + e[r]] = fin(a1, i + 1, r);
                    Uo[(i + 2) * M + e[r]] = fin(a2, i + 2, r);
                    Uo[(i + 3) * M + e[r]] = fin(a3, i + 3, r);
                    Uo[(i + 4) * M + e[r]] = fin(a4, i + 4, r);
                }
            }
        }
    };
    // The last sample on its own, then S - 2 .. 0: whole Boole blocks from a
    // block's last inner sample, so with an unroll of a multiple of 4 the
    // unrolled position u is at block position 3 - u.
    constexpr int NB = rev_buffers<D>(), UNROLL = NB % 4 == 0 ? NB : (NB % 2 == 0 ? 4 : 4 * NB);
    {
        RevOps<D> ol;
        load(S - 1, ol);
        step_at(S - 1, ol, std::integral_constant<int, 4>{});
    }
    RevOps<D> o[NB];
    walk_range<NB, UNROLL, false>(0, S - 2, o, load, [&](long long i, const RevOps<D>& op, auto uc) {
        step_at(i, op, std::integral_constant<int, (3 - decltype(uc)::value) & 3>{});
    });
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        Uo[e[r]] = fin(pend[r], 0, r);
        Dout[e[r]] = dl[r];
    }
}

// dA[b] = sum_vjp(dOut_first, dt) + sum_q (dA_job + barY^(1)), barY^(1) =
// gbuf + c_i Delta (vjp_reduce_kernel after lm_vjp_final)
__global__ void walk_vjp_reduce_kernel(LmArgs la, VjpReduceArgs a) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long per = a.S * a.m;
    if (idx >= a.B * per) return;
    const long long b = idx / per;
    const long long rem = idx - b * per;
    const long long i = rem / a.m;
    const int e = (int)(rem - i * a.m);
    const double c = sum_vjp_coef<BOOLE>(i, a.S, a.dt);
    double v = 0.0;
    if (!(a.op == OP_ONE && a.n > 1)) v = a.cot[b * (long long)a.n_out * a.m_cot + e] * c;
    for (int q = 0; q < a.Qj; ++q) {
        const long long job = b * a.Qj + q;
        double g = la.gbuf[(job * 2 + la.ga) * per + rem];
        g += c * la.dbar[(job * 2 + la.ga) * a.m + e];
        double d = a.dAjob[job * per + rem];
        d += g;
        v += d;
    }
    a.dA[idx] = v;
}

}  // namespace

bool walk_ok(int integ, const LmArgs& a) {
    return integ == BOOLE && a.sc == 0 && (a.d == 16 || a.d == 32) && a.tile == 0 && a.Qj == a.Q && a.S >= 5;
}

hipError_t launch_walk_forward_level(const LmArgs& a, hipStream_t s) {
    if (a.jobs == 0) return hipSuccess;
    const double per = (double)a.S * a.m;
    const int j = a.level;
    const double bytes = 8.0 * per * a.jobs * ((j >= 2 ? 2 : 1) + (j < a.n ? 1 : 0));
    const double flops = (double)a.jobs * a.S * ((j >= 2 ? 2.0 * a.d * a.d * a.d : 0.0) + 8.0 * a.m);
    void* h = timer_begin("walk_fwd", s, bytes, flops);
    const dim3 grid((unsigned)a.jobs);
    if (a.d == 16) {
        if (j == 1) hipLaunchKernelGGL((walk_fwd<16, true>), grid, dim3(WalkShape<16>::THREADS), 0, s, a);
        else hipLaunchKernelGGL((walk_fwd<16, false>), grid, dim3(WalkShape<16>::THREADS), 0, s, a);
    } else {
        if (j == 1) hipLaunchKernelGGL((walk_fwd<32, true>), grid, dim3(WalkShape<32>::THREADS), 0, s, a);
        else hipLaunchKernelGGL((walk_fwd<32, false>), grid, dim3(WalkShape<32>::THREADS), 0, s, a);
    }
    timer_end(h);
    return hipGetLastError();
}

hipError_t launch_walk_vjp_level(const LmArgs& a, hipStream_t s) {
    if (a.jobs == 0) return hipSuccess;
    const double per = (double)a.S * a.m;
    const bool top = a.level == a.n;
    // reads A, P^(k-1) (and barY, dA_job below the top level); writes dA_job and barY^(k-1)
    const double bytes = 8.0 * per * a.jobs * (top ? 4 : 6);
    const double flops = (double)a.jobs * a.S * (4.0 * a.d * a.d * a.d + 16.0 * a.m);
    void* h = timer_begin("walk_rev", s, bytes, flops);
    const dim3 grid((unsigned)a.jobs);
    if (a.d == 16) {
        if (top) hipLaunchKernelGGL((walk_rev<16, true>), grid, dim3(WalkShape<16>::THREADS), 0, s, a);
        else hipLaunchKernelGGL((walk_rev<16, false>), grid, dim3(WalkShape<16>::THREADS), 0, s, a);
    } else {
        if (top) hipLaunchKernelGGL((walk_rev<32, true>), grid, dim3(WalkShape<32>::THREADS), 0, s, a);
        else hipLaunchKernelGGL((walk_rev<32, false>), grid, dim3(WalkShape<32>::THREADS), 0, s, a);
    }
    timer_end(h);
    return hipGetLastError();
}

hipError_t launch_walk_vjp_reduce(const LmArgs& la, const VjpReduceArgs& r, hipStream_t s) {
    const long long total = r.B * r.S * r.m;
    if (total == 0) return hipSuccess;
    void* h = timer_begin("walk_reduce", s, 8.0 * (double)total * (1 + 2 * r.Qj), 0.0);
    hipLaunchKernelGGL(walk_vjp_reduce_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, la, r);
    timer_end(h);
    return hipGetLastError();
}

}  // namespace magnus
