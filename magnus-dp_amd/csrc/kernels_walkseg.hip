// kernels_walkseg.hip -- the dense W-basis VJP with time-segment checkpoints,
// d in {16, 32}, Boole.
//
// The W-basis VJP (one basis sweep W_j = Prefix(A W_{j-1}) for all GL nodes,
// the fold's adjoint, then a reverse level loop) keeps the K = n - 1 basis
// levels over all samples: K S d^2 doubles per instance.  Here the samples are
// cut into G segments [s_g, s_{g+1}] (block-aligned, consecutive segments
// share their end sample) and only per-segment state is kept:
//   * a first forward pass walks the segments in order, one launch per
//     (segment, level), a workgroup per instance; each level's walk keeps the
//     Boole prefix recurrence's state (P_s, f_s) at every segment start and
//     carries the running level sum from segment to segment (the sums and
//     totals the fold needs);
//   * the reverse walks the segments from the last down; per segment it
//     rebuilds the K levels from their start states into a segment buffer and
//     runs the reverse levels k = n..2 over it, the reverse block recurrence's
//     state (g4, pending) handed from segment to segment; a segment finishes
//     the adjoint at its own first sample (g4 14 dt / 45 + pending, the value
//     the block below would write) and writes dA for its samples.
// Each entry goes through the operations of one walk over all samples in the
// same order, so the result does not depend on G (bit for bit).  Products run
// on v_mfma_f64_16x16x4f64 with the operand layout of kernels_walk.hip (wave w
// owns output tile w; MFMA step s contracts k = (lane / 16) d/4 + s).
// Reference: the W-basis form of src/magnus.hpp:230-254 and src/grad.hpp:251-
// 349 (the level recursion and its reverse), boole.hpp:24-151 (prefix and its
// transpose).
// No floating-point contraction in this file, the shared rules of rules.hpp
// included: a value computed in two places (a step inlined in two unrolled
// copies of the walk, a segment's own first output) must round alike in
// both, and the compiler would otherwise fuse products into FMAs in one copy
// and not in the other.  The pragma precedes the includes because contraction
// is decided where an operation is parsed, not where it is inlined.
#pragma clang fp contract(off)

#include "kernels.hpp"
#include "rules.hpp"
#include "walk.hpp"

namespace magnus {

namespace {

typedef double dbl4 __attribute__((ext_vector_type(4)));

template <int D>
struct SegShape {
    static constexpr int TT = D / 16;
    static constexpr int WAVES = TT * TT;
    static constexpr int THREADS = 64 * WAVES;
    static constexpr int KS = D / 4;
    static constexpr int M = D * D;
};

template <int KS>
__device__ __forceinline__ void seg_load_row(double* v, const double* p) {
    const double2* p2 = reinterpret_cast<const double2*>(p);
#pragma unroll
    for (int h = 0; h < KS / 2; ++h) {
        const double2 t = p2[h];
        v[2 * h] = t.x;
        v[2 * h + 1] = t.y;
    }
}

// prefetch depths: forward d = 16 16 VGPRs per sample, d = 32 32; reverse
// d = 16 40, d = 32 72 (one workgroup per CU at the batches this path
// serves).  The reverse's depth is a multiple of the Boole block, so that a
// sample's position in its block is known where the step is unrolled.
template <int D>
constexpr int seg_fwd_buffers() { return D == 16 ? 8 : 6; }
template <int D>
constexpr int seg_rev_buffers() { return 4; }

template <int D>
struct SegFwdOps {
    double a[SegShape<D>::KS];  // A_i row operand (RAW: a[0..3] are the lane's four entries)
    double w[SegShape<D>::KS];  // W_{j-1} at sample i, column operand
};

// Level j of segment g: W_j = Prefix(A W_{j-1}) (level 1: Prefix(A)) on the
// segment's samples from its start state, and the level sum (first pass).
template <int D, bool RAW>
__global__ void __launch_bounds__(SegShape<D>::THREADS) seg_fwd(WalkSegArgs a) {
    using W = SegShape<D>;
    constexpr int KS = W::KS, M = W::M;
    const long long b = blockIdx.x;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int I = wave / W::TT, J = wave - (wave / W::TT) * W::TT, li = lane & 15, lk = lane >> 4;
    const long long S = a.S, s0 = a.s0, s1 = a.s1, L1 = a.L + 1;
    const int n = a.n, K = n - 1, j = a.level, g = a.g, G = a.G;
    const double dt = a.dt;
    const bool pref = j < n;         // levels 1..K keep a prefix; level n is a sum only
    const bool first = a.mode == 0;  // first pass: sums, checkpoints, totals
    const double* A = a.in + b * S * M;
    double* Wout = pref ? a.wseg + ((b * K + (j - 1)) * L1) * M : nullptr;
    const double* Win = RAW ? nullptr : a.wseg + ((b * K + (j - 2)) * L1) * M;
    int e[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) e[r] = (16 * I + 4 * r + lk) * D + 16 * J + li;
    PrefixState<BOOLE> st[4];
    double sum[4] = {0.0, 0.0, 0.0, 0.0};
    long long i0 = 0;
    if (g > 0) {  // resume after the segment start, which the previous segment processed
        i0 = s0 + 1;
        if (pref) {
            const double* ck = a.ckp + (((b * G + g) * K + (j - 1)) * 2) * M;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const double P = ck[e[r]], f = ck[M + e[r]];
                st[r].start(f, P);
                Wout[e[r]] = P;
            }
        }
        if (first) {
#pragma unroll
            for (int r = 0; r < 4; ++r) sum[r] = a.sums[(b * n + (j - 1)) * M + e[r]];
        }
    }
    const double* arow = A + (16 * I + li) * D + lk * KS;
    const double* wcol = RAW ? nullptr : Win + (lk * KS) * D + 16 * J + li;
    auto load = [&](long long i, SegFwdOps<D>& o) {
        if constexpr (RAW) {
#pragma unroll
            for (int r = 0; r < 4; ++r) o.a[r] = A[i * M + e[r]];
        } else {
            seg_load_row<KS>(o.a, arow + i * M);
#pragma unroll
            for (int s = 0; s < KS; ++s) o.w[s] = wcol[(i - s0) * M + s * D];
        }
    };
    auto step = [&](long long i, const SegFwdOps<D>& o, auto) {
        double f[4];
        if constexpr (RAW) {
#pragma unroll
            for (int r = 0; r < 4; ++r) f[r] = o.a[r];
        } else {
            dbl4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int s = 0; s < KS; ++s) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(o.a[s], o.w[s], acc, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 4; ++r) f[r] = acc[r];
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            if (first) sum_acc<BOOLE>(sum[r], i, S, f[r], dt);
            if (pref) Wout[(i - s0) * M + e[r]] = st[r].step(i, f[r], dt);
        }
    };
    constexpr int NB = seg_fwd_buffers<D>();
    SegFwdOps<D> o[NB];
    walk_range<NB, NB, true>(i0, s1, o, load, step);
    if (!first) return;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        a.sums[(b * n + (j - 1)) * M + e[r]] = sum[r];
        if (j == 1 && s1 == S - 1) a.z0[b * M + e[r]] = sum[r];
        if (pref && g + 1 < G) {  // the next segment's start state: (P, f) at s1, a block end
            double* ck = a.ckp + (((b * G + g + 1) * K + (j - 1)) * 2) * M;
            ck[e[r]] = st[r].base;
            ck[M + e[r]] = st[r].f0;
        }
        if (pref && s1 == S - 1) a.totals[(b * n + (j - 1)) * M + e[r]] = st[r].base;
    }
}

// The top output of a reverse Boole block, g4 14 dt / 45 + pending: the block
// writes it, and a segment computes it for its own first sample from the
// state the block below would receive.
__device__ __forceinline__ double top_out(double g4, double pend, double dt) {
    double a4 = g4 * (14 * dt / 45);
    a4 += pend;
    return a4;
}

template <int D>
struct SegRevOps {
    double urow[SegShape<D>::KS];  // barY_i row operand (below the top level)
    double wrow[SegShape<D>::KS];  // W_{k-1} at sample i, row operand (B^T of the dA product)
    double acol[SegShape<D>::KS];  // A_i column operand (A^T of the temp product)
    double ucol[SegShape<D>::KS];  // barY_i column operand (below the top level)
    double da[4];                  // dA at the lane's entries (below the top level)
};

// Reverse level k of segment g (grad.hpp:311-335 in the W-basis form: shift 0,
// weight 1, the totals' cotangents What_bar injected at the last sample).
template <int D, bool TOP>
__global__ void __launch_bounds__(SegShape<D>::THREADS) seg_rev(WalkSegArgs a) {
    using W = SegShape<D>;
    constexpr int KS = W::KS, M = W::M;
    const long long b = blockIdx.x;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int I = wave / W::TT, J = wave - (wave / W::TT) * W::TT, li = lane & 15, lk = lane >> 4;
    const long long S = a.S, s0 = a.s0, s1 = a.s1, L1 = a.L + 1;
    const int n = a.n, K = n - 1, k = a.level, g = a.g, G = a.G;
    const double dt = a.dt;
    const bool last_seg = g == G - 1;
    const long long itop = last_seg ? S - 1 : s1 - 1;  // the segment's own samples: s0 .. itop
    int e[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) e[r] = (16 * I + 4 * r + lk) * D + 16 * J + li;
    const int erow = (16 * I + li) * D + lk * KS;
    const int ecol = (lk * KS) * D + 16 * J + li;
    const int wrow0 = (16 * J + li) * D + lk * KS;
    const double* A = a.in + b * S * M;
    const double* Wl = a.wseg + ((b * K + (k - 2)) * L1) * M;  // W_{k-1} on the segment
    const double* U = a.useg + ((b * 2 + a.ga) * L1) * M;
    double* Uo = a.useg + ((b * 2 + (1 - a.ga)) * L1) * M;
    double* dAs = a.dAseg + b * L1 * M;
    const double* zb = a.zbar + b * (long long)n * M;
    double yrow[KS], ycol[KS];
    if constexpr (TOP) {  // barY^(n) = Zbar_{n-1} c_i (lm_vjp_init with weight 1)
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            yrow[s] = zb[(n - 1) * M + erow + s];
            ycol[s] = zb[(n - 1) * M + ecol + s * D];
        }
    }
    const bool addo = k > 2;  // sum_vjp_add of Zbar_{k-2} into the level below
    double osum[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) osum[r] = addo ? zb[(k - 2) * M + e[r]] : 0.0;
    auto fin = [&](double v, long long i, int r) { return addo ? sum_vjp_add<BOOLE>(v, i, S, osum[r], dt) : v; };
    double g1[4], g2[4], g3[4], g4[4], pend[4];
    if (!last_seg) {
        const double* rs = a.rstate + ((b * K + (k - 2)) * 2) * M;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            g4[r] = rs[e[r]];
            pend[r] = rs[M + e[r]];
        }
    }
    double wb[4] = {0.0, 0.0, 0.0, 0.0};
    if (last_seg) {
#pragma unroll
        for (int r = 0; r < 4; ++r) wb[r] = a.wbar[(b * n + (k - 2)) * M + e[r]];
    }
    auto load = [&](long long i, SegRevOps<D>& o) {
        const long long t = i - s0;
        seg_load_row<KS>(o.wrow, Wl + t * M + wrow0);
#pragma unroll
        for (int s = 0; s < KS; ++s) o.acol[s] = A[i * M + (lk * KS + s) * D + 16 * I + li];
        if constexpr (!TOP) {
            seg_load_row<KS>(o.urow, U + t * M + erow);
#pragma unroll
            for (int s = 0; s < KS; ++s) o.ucol[s] = U[t * M + ecol + s * D];
#pragma unroll
            for (int r = 0; r < 4; ++r) o.da[r] = dAs[t * M + e[r]];
        }
    };
    // POS: the sample's position in its Boole block (i & 3), 4 for the last
    // sample S - 1, where the level's total cotangent enters
    auto step_at = [&](long long i, const SegRevOps<D>& o, auto posc) {
        constexpr int POS = decltype(posc)::value;
        const long long t = i - s0;
        const double cy = TOP ? sum_vjp_coef<BOOLE>(i, S, dt) : 0.0;
        dbl4 dav = {0.0, 0.0, 0.0, 0.0}, tmp = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s = 0; s < KS; ++s)
            dav = __builtin_amdgcn_mfma_f64_16x16x4f64(TOP ? yrow[s] * cy : o.urow[s], o.wrow[s], dav, 0, 0, 0);
#pragma unroll
        for (int s = 0; s < KS; ++s)
            tmp = __builtin_amdgcn_mfma_f64_16x16x4f64(o.acol[s], TOP ? ycol[s] * cy : o.ucol[s], tmp, 0, 0, 0);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            dAs[t * M + e[r]] = TOP ? dav[r] : o.da[r] + dav[r];
            const double tv = tmp[r];
            if constexpr (POS == 4) {  // bar_total with shift 0 plus the total's cotangent (lm_vjp_scan_b2)
                g4[r] = tv + wb[r];
                pend[r] = 0.0;
            } else if constexpr (POS == 3) {
                g3[r] = tv;
            } else if constexpr (POS == 2) {
                g2[r] = tv;
            } else if constexpr (POS == 1) {
                g1[r] = tv;
            } else {  // block [i, i + 4] (boole_rev_block)
                double a0 = tv;
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
                const double a4 = top_out(g4[r], pend[r], dt);
                double p = g1[r] * (dt / 2);
                p += g2[r] * (dt / 3);
                p += g3[r] * (3 * dt / 8);
                p += g4[r] * (14 * dt / 45);
                pend[r] = p;
                g4[r] = a0;
                Uo[(t + 1) * M + e[r]] = fin(a1, i + 1, r);
                Uo[(t + 2) * M + e[r]] = fin(a2, i + 2, r);
                Uo[(t + 3) * M + e[r]] = fin(a3, i + 3, r);
                Uo[(t + 4) * M + e[r]] = fin(a4, i + 4, r);
            }
        }
    };
    // every walk below starts at a block's last inner sample (i % 4 == 3) and
    // covers whole blocks, so unrolled position u is at block position 3 - u
    constexpr int NB = seg_rev_buffers<D>();
    static_assert(NB % 4 == 0, "reverse prefetch depth must be a multiple of the Boole block");
    auto step = [&](long long i, const SegRevOps<D>& o, auto uc) {
        step_at(i, o, std::integral_constant<int, (3 - decltype(uc)::value) & 3>{});
    };
    long long top = itop;
    if (last_seg) {  // the last sample first, on its own
        SegRevOps<D> ol;
        load(S - 1, ol);
        step_at(S - 1, ol, std::integral_constant<int, 4>{});
        top = S - 2;
    }
    SegRevOps<D> o[NB];
    walk_range<NB, NB, false>(s0, top, o, load, step);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        if (g == 0) {  // prefix_vjp_boole's last value: G[0] = pending
            Uo[e[r]] = fin(pend[r], 0, r);
        } else {  // the top output of block [s0 - 4, s0], from the state it would receive
            Uo[e[r]] = fin(top_out(g4[r], pend[r], dt), s0, r);
            double* rs = a.rstate + ((b * K + (k - 2)) * 2) * M;
            rs[e[r]] = g4[r];
            rs[M + e[r]] = pend[r];
        }
    }
}

// dA on the segment's own samples: sum_vjp(Zbar_0, dt) + (dA_job + barY^(1))
// (lm_vjp_final, then vjp_reduce_kernel with one node)
__global__ void seg_reduce(WalkSegArgs a) {
    const long long own = (a.g == a.G - 1 ? a.S - 1 : a.s1 - 1) - a.s0 + 1;
    const long long per = own * a.m;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= a.B * per) return;
    const long long b = idx / per;
    const long long rem = idx - b * per;
    const long long t = rem / a.m;
    const int e = (int)(rem - t * a.m);
    const long long i = a.s0 + t, L1 = a.L + 1;
    double v = a.zbar[b * (long long)a.n * a.m + e] * sum_vjp_coef<BOOLE>(i, a.S, a.dt);
    double d = a.dAseg[(b * L1 + t) * a.m + e];
    d += a.useg[((b * 2 + a.ga) * L1 + t) * a.m + e];
    v += d;
    a.dA[(b * a.S + i) * a.m + e] = v;
}

}  // namespace

bool walkseg_ok(int d, int integ, int n) { return integ == BOOLE && (d == 16 || d == 32) && n >= 2; }

hipError_t launch_walkseg_forward(const WalkSegArgs& a, hipStream_t s) {
    if (a.B == 0) return hipSuccess;
    const double len = (double)(a.s1 - (a.g > 0 ? a.s0 : -1));  // samples walked
    const int j = a.level;
    const double bytes = 8.0 * a.B * len * a.m * ((j >= 2 ? 2 : 1) + (j < a.n ? 1 : 0));
    const double flops = (double)a.B * len * ((j >= 2 ? 2.0 * a.d * a.d * a.d : 0.0) + 8.0 * a.m);
    void* h = timer_begin("walkseg_fwd", s, bytes, flops);
    const dim3 grid((unsigned)a.B);
    if (a.d == 16) {
        if (j == 1) hipLaunchKernelGGL((seg_fwd<16, true>), grid, dim3(SegShape<16>::THREADS), 0, s, a);
        else hipLaunchKernelGGL((seg_fwd<16, false>), grid, dim3(SegShape<16>::THREADS), 0, s, a);
    } else {
        if (j == 1) hipLaunchKernelGGL((seg_fwd<32, true>), grid, dim3(SegShape<32>::THREADS), 0, s, a);
        else hipLaunchKernelGGL((seg_fwd<32, false>), grid, dim3(SegShape<32>::THREADS), 0, s, a);
    }
    timer_end(h);
    return hipGetLastError();
}

hipError_t launch_walkseg_reverse(const WalkSegArgs& a, hipStream_t s) {
    if (a.B == 0) return hipSuccess;
    const bool top = a.level == a.n;
    const double len = (double)((a.g == a.G - 1 ? a.S - 1 : a.s1 - 1) - a.s0 + 1);
    const double bytes = 8.0 * a.B * len * a.m * (top ? 4 : 6);
    const double flops = (double)a.B * len * (4.0 * a.d * a.d * a.d + 16.0 * a.m);
    void* h = timer_begin("walkseg_rev", s, bytes, flops);
    const dim3 grid((unsigned)a.B);
    if (a.d == 16) {
        if (top) hipLaunchKernelGGL((seg_rev<16, true>), grid, dim3(SegShape<16>::THREADS), 0, s, a);
        else hipLaunchKernelGGL((seg_rev<16, false>), grid, dim3(SegShape<16>::THREADS), 0, s, a);
    } else {
        if (top) hipLaunchKernelGGL((seg_rev<32, true>), grid, dim3(SegShape<32>::THREADS), 0, s, a);
        else hipLaunchKernelGGL((seg_rev<32, false>), grid, dim3(SegShape<32>::THREADS), 0, s, a);
    }
    timer_end(h);
    return hipGetLastError();
}

hipError_t launch_walkseg_reduce(const WalkSegArgs& a, hipStream_t s) {
    const long long own = (a.g == a.G - 1 ? a.S - 1 : a.s1 - 1) - a.s0 + 1;
    const long long total = a.B * own * a.m;
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(seg_reduce, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace magnus
