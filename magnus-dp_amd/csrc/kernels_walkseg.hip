// kernels_walkseg.hip -- the dense W-basis VJP with time-segment checkpoints
// and the grouped walker pipeline of uneven batches, d in {16, 32, 64}, Boole.
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

// What one forward walk reads and writes: level j of one instance over the
// samples s0..s1 of one segment.  Rows of W are addressed relative to wbase
// (a segment buffer starts at its segment's first sample, a full level at 0).
struct FwdWalk {
    const double* A;     // the instance's samples [S][M]
    double* Wout;        // W_j (nullptr: the sum-only level n)
    const double* Win;   // W_{j-1} (level >= 2)
    long long wbase;
    const double* ckin;  // (P, f) at s0 to resume from (segments after the first, levels < n)
    double* ckout;       // (P, f) at s1 for the next segment, or nullptr
    double* sum;         // the running level sum (read when resuming), or nullptr
    double* z0;          // copy of the level-1 sum at the end, or nullptr
    double* total;       // W_j[S - 1], or nullptr
    long long S, s0, s1;
    bool resume;         // a segment after the first: s0 was walked by the segment before
    bool restart_row;    // resuming, write W_j[s0] from the start state (a fresh segment buffer)
    double dt;
};

// W_j = Prefix(A W_{j-1}) (level 1: Prefix(A)) on the segment's samples from
// its start state, and the level sum.
template <int D, bool RAW>
__device__ __forceinline__ void fwd_walk(const FwdWalk& w) {
    using W = SegShape<D>;
    constexpr int KS = W::KS, M = W::M;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int I = wave / W::TT, J = wave - (wave / W::TT) * W::TT, li = lane & 15, lk = lane >> 4;
    const long long S = w.S, s0 = w.s0, wb = w.wbase;
    const double dt = w.dt;
    const double* A = w.A;
    double* Wout = w.Wout;
    const bool keep_sum = w.sum != nullptr;
    int e[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) e[r] = (16 * I + 4 * r + lk) * D + 16 * J + li;
    PrefixState<BOOLE> st[4];
    double sum[4] = {0.0, 0.0, 0.0, 0.0};
    long long i0 = s0;
    if (w.resume) {  // resume after the segment start, which the previous segment processed
        i0 = s0 + 1;
        if (Wout) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const double P = w.ckin[e[r]], f = w.ckin[M + e[r]];
                st[r].start(f, P);
                if (w.restart_row) Wout[(s0 - wb) * M + e[r]] = P;
            }
        }
        if (keep_sum) {
#pragma unroll
            for (int r = 0; r < 4; ++r) sum[r] = w.sum[e[r]];
        }
    }
    const double* arow = A + (16 * I + li) * D + lk * KS;
    const double* wcol = RAW ? nullptr : w.Win + (lk * KS) * D + 16 * J + li;
    auto load = [&](long long i, SegFwdOps<D>& o) {
        if constexpr (RAW) {
#pragma unroll
            for (int r = 0; r < 4; ++r) o.a[r] = A[i * M + e[r]];
        } else {
            seg_load_row<KS>(o.a, arow + i * M);
#pragma unroll
            for (int s = 0; s < KS; ++s) o.w[s] = wcol[(i - wb) * M + s * D];
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
            if (keep_sum) sum_acc<BOOLE>(sum[r], i, S, f[r], dt);
            if (Wout) Wout[(i - wb) * M + e[r]] = st[r].step(i, f[r], dt);
        }
    };
    constexpr int NB = seg_fwd_buffers<D>();
    SegFwdOps<D> o[NB];
    walk_range<NB, NB, true>(i0, w.s1, o, load, step);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        if (keep_sum) w.sum[e[r]] = sum[r];
        if (w.z0) w.z0[e[r]] = sum[r];
        if (w.ckout) {  // the next segment's start state: (P, f) at s1, a block end
            w.ckout[e[r]] = st[r].base;
            w.ckout[M + e[r]] = st[r].f0;
        }
        if (w.total) w.total[e[r]] = st[r].base;
    }
}

// Level j of segment g of instance b: mode 0 is the first pass (sums,
// checkpoints, totals), mode 1 rebuilds the segment's levels.
__device__ __forceinline__ FwdWalk seg_fwd_walk(const WalkSegArgs& a, long long b) {
    const long long S = a.S, L1 = a.L + 1, M = a.m;
    const int n = a.n, K = n - 1, j = a.level, g = a.g, G = a.G;
    const bool pref = j < n, first = a.mode == 0, last = a.s1 == S - 1;
    FwdWalk w;
    w.A = a.in + b * S * M;
    w.Wout = pref ? a.wseg + ((b * K + (j - 1)) * L1) * M : nullptr;
    w.Win = j >= 2 ? a.wseg + ((b * K + (j - 2)) * L1) * M : nullptr;
    w.wbase = a.s0;
    w.ckin = (g > 0 && pref) ? a.ckp + (((b * G + g) * K + (j - 1)) * 2) * M : nullptr;
    w.ckout = (first && pref && g + 1 < G) ? a.ckp + (((b * G + g + 1) * K + (j - 1)) * 2) * M : nullptr;
    w.sum = first ? a.sums + (b * n + (j - 1)) * M : nullptr;
    w.z0 = (first && j == 1 && last) ? a.z0 + b * M : nullptr;
    w.total = (first && pref && last) ? a.totals + (b * n + (j - 1)) * M : nullptr;
    w.S = S;
    w.s0 = a.s0;
    w.s1 = a.s1;
    w.resume = g > 0;
    w.restart_row = true;
    w.dt = a.dt;
    return w;
}

// one workgroup per instance
template <int D, bool RAW>
__global__ void __launch_bounds__(SegShape<D>::THREADS) seg_fwd(WalkSegArgs a) {
    fwd_walk<D, RAW>(seg_fwd_walk(a, blockIdx.x));
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

// What one reverse walk reads and writes: level k of one instance over the
// own samples s0..itop of one segment (itop = S - 1 on the last segment, the
// sample before the next segment's start otherwise).  Rows of W, of the
// adjoints and of the dA accumulation are addressed relative to their bases.
struct RevWalk {
    const double* A;     // the instance's samples [S][M]
    const double* Wl;    // W_{k-1}, row i at (i - wbase)
    long long wbase;
    const double* U;     // barY^(k) below the top level, row i at (i - ubase)
    double* Uo;          // barY^(k-1) out, same rows
    long long ubase;
    double* dAs;         // dA accumulation, row i at (i - dbase)
    long long dbase;
    const double* zb;    // Zbar rows 0..n-1 of the instance
    const double* wbar;  // What_bar_{k-1} (read on the last segment)
    double* rs;          // (g4, pending) of this level: read from the segment above, written for the one below
    long long S, s0, itop;
    int n, k;
    bool last_seg, first_seg;
    double dt;
};

// Reverse level k over one segment (grad.hpp:311-335 in the W-basis form:
// shift 0, weight 1, the totals' cotangents What_bar injected at the last sample).
template <int D, bool TOP>
__device__ __forceinline__ void rev_walk(const RevWalk& w) {
    using W = SegShape<D>;
    constexpr int KS = W::KS, M = W::M;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int I = wave / W::TT, J = wave - (wave / W::TT) * W::TT, li = lane & 15, lk = lane >> 4;
    const long long S = w.S, s0 = w.s0, itop = w.itop;
    const int n = w.n, k = w.k;
    const double dt = w.dt;
    const bool last_seg = w.last_seg;
    int e[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) e[r] = (16 * I + 4 * r + lk) * D + 16 * J + li;
    const int erow = (16 * I + li) * D + lk * KS;
    const int ecol = (lk * KS) * D + 16 * J + li;
    const int wrow0 = (16 * J + li) * D + lk * KS;
    const double* A = w.A;
    const double* Wl = w.Wl - w.wbase * M;
    const double* U = w.U - w.ubase * M;
    double* Uo = w.Uo - w.ubase * M;
    double* dAs = w.dAs - w.dbase * M;
    const double* zb = w.zb;
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
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            g4[r] = w.rs[e[r]];
            pend[r] = w.rs[M + e[r]];
        }
    }
    double wb[4] = {0.0, 0.0, 0.0, 0.0};
    if (last_seg) {
#pragma unroll
        for (int r = 0; r < 4; ++r) wb[r] = w.wbar[e[r]];
    }
    auto load = [&](long long i, SegRevOps<D>& o) {
        seg_load_row<KS>(o.wrow, Wl + i * M + wrow0);
#pragma unroll
        for (int s = 0; s < KS; ++s) o.acol[s] = A[i * M + (lk * KS + s) * D + 16 * I + li];
        if constexpr (!TOP) {
            seg_load_row<KS>(o.urow, U + i * M + erow);
#pragma unroll
            for (int s = 0; s < KS; ++s) o.ucol[s] = U[i * M + ecol + s * D];
#pragma unroll
            for (int r = 0; r < 4; ++r) o.da[r] = dAs[i * M + e[r]];
        }
    };
    // POS: the sample's position in its Boole block (i & 3), 4 for the last
    // sample S - 1, where the level's total cotangent enters
    auto step_at = [&](long long i, const SegRevOps<D>& o, auto posc) {
        constexpr int POS = decltype(posc)::value;
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
            dAs[i * M + e[r]] = TOP ? dav[r] : o.da[r] + dav[r];
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
                Uo[(i + 1) * M + e[r]] = fin(a1, i + 1, r);
                Uo[(i + 2) * M + e[r]] = fin(a2, i + 2, r);
                Uo[(i + 3) * M + e[r]] = fin(a3, i + 3, r);
                Uo[(i + 4) * M + e[r]] = fin(a4, i + 4, r);
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
        if (w.first_seg) {  // prefix_vjp_boole's last value: G[0] = pending
            Uo[e[r]] = fin(pend[r], 0, r);
        } else {  // the top output of block [s0 - 4, s0], from the state it would receive
            Uo[s0 * M + e[r]] = fin(top_out(g4[r], pend[r], dt), s0, r);
            w.rs[e[r]] = g4[r];
            w.rs[M + e[r]] = pend[r];
        }
    }
}

// Reverse level k of segment g of instance b
__device__ __forceinline__ RevWalk seg_rev_walk(const WalkSegArgs& a, long long b) {
    const long long S = a.S, s0 = a.s0, L1 = a.L + 1, M = a.m;
    const int n = a.n, K = n - 1, k = a.level, g = a.g, G = a.G;
    RevWalk w;
    w.A = a.in + b * S * M;
    w.Wl = a.wseg + ((b * K + (k - 2)) * L1) * M;  // W_{k-1} on the segment
    w.wbase = s0;
    w.U = a.useg + ((b * 2 + a.ga) * L1) * M;
    w.Uo = a.useg + ((b * 2 + (1 - a.ga)) * L1) * M;
    w.ubase = s0;
    w.dAs = a.dAseg + b * L1 * M;
    w.dbase = s0;
    w.zb = a.zbar + b * (long long)n * M;
    w.wbar = a.wbar + (b * n + (k - 2)) * M;
    w.rs = a.rstate + ((b * K + (k - 2)) * 2) * M;
    w.S = S;
    w.s0 = s0;
    w.last_seg = g == G - 1;
    w.itop = w.last_seg ? S - 1 : a.s1 - 1;
    w.first_seg = g == 0;
    w.n = n;
    w.k = k;
    w.dt = a.dt;
    return w;
}

template <int D, bool TOP>
__global__ void __launch_bounds__(SegShape<D>::THREADS) seg_rev(WalkSegArgs a) {
    rev_walk<D, TOP>(seg_rev_walk(a, blockIdx.x));
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

// ---------------------------------------------------------------------------
// d = 64: sixteen wavefronts per workgroup, wave w owning the 16 x 16 output
// tile (w / 4, w % 4) as at d = 16, 32; a lane owns 4 entries.  A sample's
// matrices are staged in LDS (coalesced 16-byte global loads, 16-byte LDS
// stores) and every MFMA operand is read from there, four wavefronts sharing
// each operand tile: operands read along rows use a row stride of 68 doubles,
// operands read down columns 80, so that a read's 64 lanes spread evenly over
// the banks.  The forward double-buffers its two matrices (A_i, W_{j-1}[i]:
// 148 KB); the reverse stages three (A_i, W_{k-1}[i], barY_i: 108 KB) in one
// buffer with the next sample's share in registers.
// ---------------------------------------------------------------------------
constexpr int kT64 = 1024;  // threads
constexpr int kP_ROW = 68, kP_COL = 80;
constexpr int kS_ROW = 64 * kP_ROW, kS_COL = 64 * kP_COL;
constexpr int kLds64Fwd = 2 * (kS_ROW + kS_COL), kLds64Rev = kS_COL + 2 * kS_ROW;  // doubles

// s_barrier after this wave's LDS traffic only: __syncthreads() would also wait
// for every global store in flight
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

typedef double dbl2 __attribute__((ext_vector_type(2)));

struct Stage64 {
    dbl2 v[2];  // a thread's share of one 64 x 64 matrix: 16-byte chunks tid + 1024 h
};
__device__ __forceinline__ void stage_load(Stage64& r, const double* src) {
    const dbl2* p = reinterpret_cast<const dbl2*>(src);
#pragma unroll
    for (int h = 0; h < 2; ++h) r.v[h] = p[threadIdx.x + kT64 * h];
}
__device__ __forceinline__ void stage_store(const Stage64& r, double* lds, int stride) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int c = threadIdx.x + kT64 * h;
        *reinterpret_cast<dbl2*>(lds + (c >> 5) * stride + ((c & 31) << 1)) = r.v[h];
    }
}

struct Tile64 {
    int li, lk, R0, C0, ebase;
    __device__ Tile64() {
        const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
        li = lane & 15;
        lk = lane >> 4;
        R0 = 16 * (wave >> 2);
        C0 = 16 * (wave & 3);
        ebase = (R0 + lk) * 64 + C0 + li;
    }
};
// entry r of a lane: row R0 + 4 r + lk, column C0 + li
__host__ __device__ constexpr int off64(int r) { return r * 256; }

template <bool RAW>
__device__ __forceinline__ void fwd_walk64(const FwdWalk& w, double* lds) {
    constexpr int M = 4096;
    const Tile64 hd;
    const long long S = w.S, s0 = w.s0, s1 = w.s1, wb = w.wbase;
    const double dt = w.dt;
    const int eb = hd.ebase;
    const double* A = w.A;
    double* Wout = w.Wout;
    const bool keep_sum = w.sum != nullptr;
    PrefixState<BOOLE> st[4];
    double sum[4] = {0.0, 0.0, 0.0, 0.0};
    long long i0 = s0;
    if (w.resume) {
        i0 = s0 + 1;
        if (Wout) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const double P = w.ckin[hd.ebase + off64(q)], f = w.ckin[M + hd.ebase + off64(q)];
                st[q].start(f, P);
                if (w.restart_row) Wout[(s0 - wb) * M + (eb + off64(q))] = P;
            }
        }
        if (keep_sum) {
#pragma unroll
            for (int q = 0; q < 4; ++q) sum[q] = w.sum[hd.ebase + off64(q)];
        }
    }
    auto consume = [&](long long i, const double (&f)[4]) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (keep_sum) sum_acc<BOOLE>(sum[q], i, S, f[q], dt);
            if (Wout) Wout[(i - wb) * M + (eb + off64(q))] = st[q].step(i, f[q], dt);
        }
    };
    if constexpr (RAW) {
        struct Raw {
            double a[4];
        };
        auto load = [&](long long i, Raw& o) {
#pragma unroll
            for (int q = 0; q < 4; ++q) o.a[q] = A[i * M + (eb + off64(q))];
        };
        auto step = [&](long long i, const Raw& o, auto) { consume(i, o.a); };
        Raw o[4];
        walk_range<4, 4, true>(i0, s1, o, load, step);
    } else if (i0 <= s1) {
        const double* Ar = w.A;
        const double* Win = w.Win;
        Stage64 ra, rw;
        stage_load(ra, Ar + i0 * M);
        stage_load(rw, Win + (i0 - wb) * M);
        stage_store(ra, lds, kP_ROW);
        stage_store(rw, lds + kS_ROW, kP_COL);
        lds_barrier();
        int cur = 0;
        for (long long i = i0; i <= s1; ++i) {
            const long long nx = i < s1 ? i + 1 : i;  // past the end: a reload nobody reads
            stage_load(ra, Ar + nx * M);
            stage_load(rw, Win + (nx - wb) * M);
            const double* la = lds + cur * (kS_ROW + kS_COL);
            const double* lw = la + kS_ROW;
            dbl4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
            for (int s = 0; s < 16; ++s) {
                const int kk = 4 * s + hd.lk;
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(la[(hd.R0 + hd.li) * kP_ROW + kk],
                                                           lw[kk * kP_COL + hd.C0 + hd.li], acc, 0, 0, 0);
            }
            const double f[4] = {acc[0], acc[1], acc[2], acc[3]};
            consume(i, f);
            double* ln = lds + (cur ^ 1) * (kS_ROW + kS_COL);
            stage_store(ra, ln, kP_ROW);
            stage_store(rw, ln + kS_ROW, kP_COL);
            lds_barrier();
            cur ^= 1;
        }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int e = hd.ebase + off64(q);
        if (keep_sum) w.sum[e] = sum[q];
        if (w.z0) w.z0[e] = sum[q];
        if (w.ckout) {
            w.ckout[e] = st[q].base;
            w.ckout[M + e] = st[q].f0;
        }
        if (w.total) w.total[e] = st[q].base;
    }
}

template <bool TOP>
__device__ __forceinline__ void rev_walk64(const RevWalk& w, double* lds) {
    constexpr int M = 4096;
    constexpr int LA = 0, LW = kS_COL, LY = kS_COL + kS_ROW;  // A (read down columns), W, barY
    const Tile64 hd;
    const long long S = w.S, s0 = w.s0, itop = w.itop;
    const int n = w.n, k = w.k;
    const double dt = w.dt;
    const bool last_seg = w.last_seg;
    const int eb = hd.ebase;
    const double* A = w.A;
    const double* Wl = w.Wl - w.wbase * M;
    const double* U = w.U - w.ubase * M;
    // uniform bases, the lane's entry as a 32-bit offset (scalar base + vector offset addressing)
    double* Uo = w.Uo - w.ubase * M;
    double* dAs = w.dAs - w.dbase * M;
    // Zbar_{k-2} at the lane's entries, read where a block is finished (kept in
    // registers, its products with the rule's weights would be hoisted out of
    // the walk and outgrow the register budget)
    const bool addo = k > 2;
    const double* zrow = w.zb + (long long)(k - 2) * M + eb;
    auto fin = [&](double v, long long i, double os) { return addo ? sum_vjp_add<BOOLE>(v, i, S, os, dt) : v; };
    double g1[4], g2[4], g3[4], g4[4], pend[4];
    if (!last_seg) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            g4[q] = w.rs[eb + off64(q)];
            pend[q] = w.rs[M + eb + off64(q)];
        }
    }
    // the next sample's share in registers (its loads in flight during a step)
    Stage64 ra, rw, ru;
    auto gload = [&](long long i) {
        stage_load(ra, A + i * M);
        stage_load(rw, Wl + i * M);
        if constexpr (!TOP) stage_load(ru, U + i * M);
    };
    auto lstore = [&]() {
        stage_store(ra, lds + LA, kP_COL);
        stage_store(rw, lds + LW, kP_ROW);
        if constexpr (!TOP) stage_store(ru, lds + LY, kP_ROW);
    };
    if constexpr (TOP) {  // barY^(n) = Zbar_{n-1} c_i: the matrix stays in LDS, c_i scales the operands
        stage_load(ru, w.zb + (long long)(n - 1) * M);
        stage_store(ru, lds + LY, kP_ROW);
    }
    // POS: the sample's position in its Boole block, 4 for the last sample S - 1
    auto step_at = [&](long long i, auto posc) {
        constexpr int POS = decltype(posc)::value;
        const double cy = TOP ? sum_vjp_coef<BOOLE>(i, S, dt) : 1.0;
        double da[4];  // dA so far at the lane's entries: loaded here, needed after the products
        if constexpr (!TOP) {
#pragma unroll
            for (int q = 0; q < 4; ++q) da[q] = dAs[i * M + (eb + off64(q))];
        }
        const double* la = lds + LA;
        const double* lw = lds + LW;
        const double* ly = lds + LY;
        dbl4 dav = {0.0, 0.0, 0.0, 0.0}, tmp = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 1
        for (int s = 0; s < 16; ++s) {
            const int kk = 4 * s + hd.lk;
            // dA += barY W^T: barY row R0 + li, W row C0 + li
            double y = ly[(hd.R0 + hd.li) * kP_ROW + kk];
            // temp = A^T barY: A column R0 + li, barY column C0 + li
            double u = ly[kk * kP_ROW + hd.C0 + hd.li];
            if constexpr (TOP) {
                y *= cy;
                u *= cy;
            }
            dav = __builtin_amdgcn_mfma_f64_16x16x4f64(y, lw[(hd.C0 + hd.li) * kP_ROW + kk], dav, 0, 0, 0);
            tmp = __builtin_amdgcn_mfma_f64_16x16x4f64(la[kk * kP_COL + hd.R0 + hd.li], u, tmp, 0, 0, 0);
        }
        const double dvv[4] = {dav[0], dav[1], dav[2], dav[3]};
        const double tvv[4] = {tmp[0], tmp[1], tmp[2], tmp[3]};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const double tv = tvv[q];
            dAs[i * M + (eb + off64(q))] = TOP ? dvv[q] : da[q] + dvv[q];
            if constexpr (POS == 4) {
                g4[q] = tv + w.wbar[eb + off64(q)];
                pend[q] = 0.0;
            } else if constexpr (POS == 3) {
                g3[q] = tv;
            } else if constexpr (POS == 2) {
                g2[q] = tv;
            } else if constexpr (POS == 1) {
                g1[q] = tv;
            } else {
                double a0 = tv;
                a0 += g1[q];
                a0 += g2[q];
                a0 += g3[q];
                a0 += g4[q];
                double a1 = g1[q] * (dt / 2);
                a1 += g2[q] * (4 * dt / 3);
                a1 += g3[q] * (9 * dt / 8);
                a1 += g4[q] * (64 * dt / 45);
                double a2 = g2[q] * (dt / 3);
                a2 += g3[q] * (9 * dt / 8);
                a2 += g4[q] * (24 * dt / 45);
                double a3 = g3[q] * (3 * dt / 8);
                a3 += g4[q] * (64 * dt / 45);
                const double a4 = top_out(g4[q], pend[q], dt);
                double pp = g1[q] * (dt / 2);
                pp += g2[q] * (dt / 3);
                pp += g3[q] * (3 * dt / 8);
                pp += g4[q] * (14 * dt / 45);
                pend[q] = pp;
                g4[q] = a0;
                const double os = addo ? zrow[off64(q)] : 0.0;
                Uo[(i + 1) * M + (eb + off64(q))] = fin(a1, i + 1, os);
                Uo[(i + 2) * M + (eb + off64(q))] = fin(a2, i + 2, os);
                Uo[(i + 3) * M + (eb + off64(q))] = fin(a3, i + 3, os);
                Uo[(i + 4) * M + (eb + off64(q))] = fin(a4, i + 4, os);
            }
        }
    };
    // samples itop .. s0; the last segment's sample S - 1 first, on its own, the
    // rest in whole Boole blocks from a block's last inner sample
    long long top = itop;
    gload(itop);
    lstore();
    lds_barrier();
    if (last_seg) {
        gload(itop - 1 >= s0 ? itop - 1 : itop);
        step_at(itop, std::integral_constant<int, 4>{});
        lds_barrier();
        lstore();
        lds_barrier();
        top = itop - 1;
    }
    const long long span = top - s0 + 1;  // a multiple of 4
    for (long long b0 = 0; b0 < span; b0 += 4) {
        static_for<0, 4>([&](auto uc) {
            constexpr int u = decltype(uc)::value;
            const long long ii = top - b0 - u;
            const long long nx = ii - 1 >= s0 ? ii - 1 : ii;  // past the end: a reload nobody reads
            gload(nx);
            step_at(ii, std::integral_constant<int, 3 - u>{});
            lds_barrier();
            lstore();
            lds_barrier();
        });
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const double os = addo ? zrow[off64(q)] : 0.0;
        if (w.first_seg) {
            Uo[eb + off64(q)] = fin(pend[q], 0, os);
        } else {
            Uo[s0 * M + (eb + off64(q))] = fin(top_out(g4[q], pend[q], dt), s0, os);
            w.rs[eb + off64(q)] = g4[q];
            w.rs[M + eb + off64(q)] = pend[q];
        }
    }
}

template <bool RAW>
__global__ void __launch_bounds__(kT64) seg_fwd64(WalkSegArgs a) {
    __shared__ alignas(16) double lds[kLds64Fwd];
    fwd_walk64<RAW>(seg_fwd_walk(a, blockIdx.x), lds);
}

template <bool TOP>
__global__ void __launch_bounds__(kT64) seg_rev64(WalkSegArgs a) {
    __shared__ alignas(16) double lds[kLds64Rev];
    rev_walk64<TOP>(seg_rev_walk(a, blockIdx.x), lds);
}

// ---------------------------------------------------------------------------
// Grouped walkers: one workgroup per task (instance, segment g, level) of a
// task table, each instance with its own samples, order and full basis levels.
// ---------------------------------------------------------------------------
__device__ __forceinline__ long long gw_seg_end(long long s0, long long L, long long S) {
    return s0 + L < S - 1 ? s0 + L : S - 1;
}

template <int D>
__global__ void __launch_bounds__(SegShape<D>::THREADS) gw_fwd(GwArgs a) {
    constexpr int M = SegShape<D>::M;
    const GwTask t = a.tasks[blockIdx.x];
    const GwInst& in = a.inst[t.inst];
    const long long S = in.S, L = a.L;
    const int n = in.n, j = t.level;
    const long long s0 = (long long)t.seg * L, s1 = gw_seg_end(s0, L, S);
    const bool pref = j < n, last = s1 == S - 1;
    FwdWalk w;
    w.A = a.tab[in.id];
    w.Wout = pref ? in.W + (long long)(j - 1) * S * M : nullptr;
    w.Win = j >= 2 ? in.W + (long long)(j - 2) * S * M : nullptr;
    w.wbase = 0;
    w.ckin = pref ? in.fst + (long long)(j - 1) * 2 * M : nullptr;
    w.ckout = (pref && !last) ? in.fst + (long long)(j - 1) * 2 * M : nullptr;  // a thread rewrites its own entries
    w.sum = in.sums + (long long)(j - 1) * M;
    w.z0 = (j == 1 && last) ? in.z0 : nullptr;
    w.total = (pref && last) ? in.totals + (long long)(j - 1) * M : nullptr;
    w.S = S;
    w.s0 = s0;
    w.s1 = s1;
    w.resume = t.seg > 0;
    w.restart_row = false;  // the level's full array already holds W_j[s0]
    w.dt = in.dt;
    if (j == 1) fwd_walk<D, true>(w);
    else fwd_walk<D, false>(w);
}

template <int D>
__global__ void __launch_bounds__(SegShape<D>::THREADS) gw_rev(GwArgs a) {
    constexpr int M = SegShape<D>::M;
    const GwTask t = a.tasks[blockIdx.x];
    const GwInst& in = a.inst[t.inst];
    const long long S = in.S, L = a.L, L1 = L + 1;
    const int n = in.n, g = t.seg;
    const long long s0 = (long long)g * L, s1 = gw_seg_end(s0, L, S);
    const bool last_seg = s1 == S - 1;
    const long long itop = last_seg ? S - 1 : s1 - 1;
    double* dA = const_cast<double*>(a.tab[2 * a.count + in.id]);
    const int slot = g & 1;
    if (t.kind == 2) {  // dA on the segment's own samples: sum_vjp(Zbar_0, dt) + (dA + barY^(1))
        const double* U1 = in.U + ((long long)slot * L1) * M;
        const long long per = (itop - s0 + 1) * M;
        for (long long x = threadIdx.x; x < per; x += SegShape<D>::THREADS) {
            const long long r = x / M;
            const int e = (int)(x - r * M);
            const long long i = s0 + r;
            double v = in.zbar[e] * sum_vjp_coef<BOOLE>(i, S, in.dt);
            double d = dA[i * M + e];
            d += U1[r * M + e];
            v += d;
            dA[i * M + e] = v;
        }
        return;
    }
    const int k = t.level;
    RevWalk w;
    w.A = a.tab[in.id];
    w.Wl = in.W + (long long)(k - 2) * S * M;
    w.wbase = 0;
    w.Uo = in.U + (((long long)(k - 2) * 2 + slot) * L1) * M;  // barY^(k-1)
    w.U = k < n ? in.U + (((long long)(k - 1) * 2 + slot) * L1) * M : w.Uo;  // barY^(k) (unread at the top)
    w.ubase = s0;
    w.dAs = dA;
    w.dbase = 0;
    w.zb = in.zbar;
    w.wbar = in.wbar + (long long)(k - 2) * M;
    w.rs = in.rst + (long long)(k - 2) * 2 * M;
    w.S = S;
    w.s0 = s0;
    w.itop = itop;
    w.n = n;
    w.k = k;
    w.last_seg = last_seg;
    w.first_seg = g == 0;
    w.dt = in.dt;
    if (k == n) rev_walk<D, true>(w);
    else rev_walk<D, false>(w);
}

__global__ void __launch_bounds__(kT64) gw_fwd64(GwArgs a) {
    constexpr int M = 4096;
    __shared__ alignas(16) double lds[kLds64Fwd];
    const GwTask t = a.tasks[blockIdx.x];
    const GwInst& in = a.inst[t.inst];
    const long long S = in.S, L = a.L;
    const int n = in.n, j = t.level;
    const long long s0 = (long long)t.seg * L, s1 = gw_seg_end(s0, L, S);
    const bool pref = j < n, last = s1 == S - 1;
    FwdWalk w;
    w.A = a.tab[in.id];
    w.Wout = pref ? in.W + (long long)(j - 1) * S * M : nullptr;
    w.Win = j >= 2 ? in.W + (long long)(j - 2) * S * M : nullptr;
    w.wbase = 0;
    w.ckin = pref ? in.fst + (long long)(j - 1) * 2 * M : nullptr;
    w.ckout = (pref && !last) ? in.fst + (long long)(j - 1) * 2 * M : nullptr;
    w.sum = in.sums + (long long)(j - 1) * M;
    w.z0 = (j == 1 && last) ? in.z0 : nullptr;
    w.total = (pref && last) ? in.totals + (long long)(j - 1) * M : nullptr;
    w.S = S;
    w.s0 = s0;
    w.s1 = s1;
    w.resume = t.seg > 0;
    w.restart_row = false;
    w.dt = in.dt;
    if (j == 1) fwd_walk64<true>(w, lds);
    else fwd_walk64<false>(w, lds);
}

__global__ void __launch_bounds__(kT64) gw_rev64(GwArgs a) {
    constexpr int M = 4096;
    __shared__ alignas(16) double lds[kLds64Rev];
    const GwTask t = a.tasks[blockIdx.x];
    const GwInst& in = a.inst[t.inst];
    const long long S = in.S, L = a.L, L1 = L + 1;
    const int n = in.n, g = t.seg;
    const long long s0 = (long long)g * L, s1 = gw_seg_end(s0, L, S);
    const bool last_seg = s1 == S - 1;
    const long long itop = last_seg ? S - 1 : s1 - 1;
    double* dA = const_cast<double*>(a.tab[2 * a.count + in.id]);
    const int slot = g & 1;
    if (t.kind == 2) {
        const double* U1 = in.U + ((long long)slot * L1) * M;
        const long long per = (itop - s0 + 1) * M;
        for (long long x = threadIdx.x; x < per; x += kT64) {
            const long long r = x / M;
            const int e = (int)(x - r * M);
            const long long i = s0 + r;
            double v = in.zbar[e] * sum_vjp_coef<BOOLE>(i, S, in.dt);
            double d = dA[i * M + e];
            d += U1[r * M + e];
            v += d;
            dA[i * M + e] = v;
        }
        return;
    }
    const int k = t.level;
    RevWalk w;
    w.A = a.tab[in.id];
    w.Wl = in.W + (long long)(k - 2) * S * M;
    w.wbase = 0;
    w.Uo = in.U + (((long long)(k - 2) * 2 + slot) * L1) * M;
    w.U = k < n ? in.U + (((long long)(k - 1) * 2 + slot) * L1) * M : w.Uo;
    w.ubase = s0;
    w.dAs = dA;
    w.dbase = 0;
    w.zb = in.zbar;
    w.wbar = in.wbar + (long long)(k - 2) * M;
    w.rs = in.rst + (long long)(k - 2) * 2 * M;
    w.S = S;
    w.s0 = s0;
    w.itop = itop;
    w.n = n;
    w.k = k;
    w.last_seg = last_seg;
    w.first_seg = g == 0;
    w.dt = in.dt;
    if (k == n) rev_walk64<true>(w, lds);
    else rev_walk64<false>(w, lds);
}

}  // namespace

bool gw_ok(int d, int integ, int n) { return integ == BOOLE && (d == 16 || d == 32 || d == 64) && n >= 2; }

hipError_t launch_gw_forward(int d, const GwArgs& a, int ntasks, double bytes, double flops, hipStream_t s) {
    if (ntasks <= 0) return hipSuccess;
    void* h = timer_begin("gw_fwd", s, bytes, flops);
    if (d == 16) hipLaunchKernelGGL((gw_fwd<16>), dim3(ntasks), dim3(SegShape<16>::THREADS), 0, s, a);
    else if (d == 32) hipLaunchKernelGGL((gw_fwd<32>), dim3(ntasks), dim3(SegShape<32>::THREADS), 0, s, a);
    else hipLaunchKernelGGL(gw_fwd64, dim3(ntasks), dim3(kT64), 0, s, a);
    timer_end(h);
    return hipGetLastError();
}

hipError_t launch_gw_reverse(int d, const GwArgs& a, int ntasks, double bytes, double flops, hipStream_t s) {
    if (ntasks <= 0) return hipSuccess;
    void* h = timer_begin("gw_rev", s, bytes, flops);
    if (d == 16) hipLaunchKernelGGL((gw_rev<16>), dim3(ntasks), dim3(SegShape<16>::THREADS), 0, s, a);
    else if (d == 32) hipLaunchKernelGGL((gw_rev<32>), dim3(ntasks), dim3(SegShape<32>::THREADS), 0, s, a);
    else hipLaunchKernelGGL(gw_rev64, dim3(ntasks), dim3(kT64), 0, s, a);
    timer_end(h);
    return hipGetLastError();
}

bool walkseg_ok(int d, int integ, int n) { return integ == BOOLE && (d == 16 || d == 32 || d == 64) && n >= 2; }

hipError_t launch_walkseg_forward(const WalkSegArgs& a, hipStream_t s) {
    if (a.B == 0) return hipSuccess;
    const double len = (double)(a.s1 - (a.g > 0 ? a.s0 : -1));  // samples walked
    const int j = a.level;
    const double bytes = 8.0 * a.B * len * a.m * ((j >= 2 ? 2 : 1) + (j < a.n ? 1 : 0));
    const double flops = (double)a.B * len * ((j >= 2 ? 2.0 * a.d * a.d * a.d : 0.0) + 8.0 * a.m);
    void* h = timer_begin("walkseg_fwd", s, bytes, flops);
    const dim3 grid((unsigned)a.B);
    if (a.d == 64) {
        if (j == 1) hipLaunchKernelGGL(seg_fwd64<true>, grid, dim3(kT64), 0, s, a);
        else hipLaunchKernelGGL(seg_fwd64<false>, grid, dim3(kT64), 0, s, a);
    } else if (a.d == 16) {
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
    if (a.d == 64) {
        if (top) hipLaunchKernelGGL(seg_rev64<true>, grid, dim3(kT64), 0, s, a);
        else hipLaunchKernelGGL(seg_rev64<false>, grid, dim3(kT64), 0, s, a);
    } else if (a.d == 16) {
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
