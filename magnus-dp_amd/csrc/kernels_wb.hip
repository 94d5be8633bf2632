// kernels_wb.hip -- single-pass ("W-basis") Magnus kernels for the su(2)
// SpaceCurve path (4-vector algebra), forward and VJP.
//
// The reference evaluates, for every Gauss-Legendre node q, the level
// recursion  Y <- A,  P = Prefix(Y),  T = P[last],  Y_i <- A_i (P_i + s_q T)
// (src/magnus.hpp:230-254), which needs one full pass over the samples per
// level because T is a global total.  Prefix is linear and the product is
// bilinear and associative, so with the basis streams
//      W_1 = Prefix(A),   W_j = Prefix(A * W_{j-1})   (independent of q)
// every level's prefix is  P^(j) = W_j + s_q sum_{l<j} W_l * T_{j-l}  and
//      Omega_j = sum_q w_q ( Z_{j-1} + s_q sum_{l<j-1} Z_l * T_{j-1-l} ),
//      T_j     = W_j[last] + s_q sum_{l<j} W_l[last] * T_{j-l},
// with Z_l = Sum(A * W_l) (Z_0 = Sum(A)).
//
// Chunks.  The samples of an instance are cut into chunks (s, e].  A chunk's
// local streams V_j (zero at s) give its transfer series and sum series
//      V_c(x) = I + sum_{j=1..K} V_j[e] x^j,     Z_c(x) = sum_{l=0..K} ZV_l x^l,
// ZV_l = sum_{i in (s,e]} c_i A_i V_l[i].  In the ring of quaternion-valued
// polynomials truncated at degree K, the streams entering chunk c are
// W~_c = V_{c-1} ... V_0, and the instance totals are one associative
// reduction of the pairs (V_c, Z_c) under
//      (V_b, Z_b) o (V_a, Z_a) = (V_b V_a,  Z_b V_a + Z_a):
// the V-part of the total holds What_j = W_j[last], its Z-part holds Z_l.
// The chunk pass computes ZV_l from one extra prefix level: the prefix
// weights at e equal the sum weights c_i inside the chunk, so
// ZV_l = V_{l+1}[e] + (c_e - alpha_end) A_e V_l[e]  (minus the start-role
// weight of A_s for l = 0 in chunks after the first).
//
// VJP.  The gradient is the adjoint of  chunk summaries -> pair reduction ->
// GL algebra.  The combine kernel runs the GL adjoint, an exclusive prefix of
// the V series and an exclusive suffix of the pairs, and gives every chunk the
// cotangents of its summaries.  The reverse chunk pass is then the local
// adjoint of one chunk: it rebuilds the chunk's local streams backwards from
// V_j[e] (no per-sample storage), and folds the sum cotangents c_i ZV_bar
// into the suffix sums of the prefix transposes (the prefix weights at e are
// the c_i).  A sample on a chunk boundary belongs to two chunks; the fixup
// kernel adds the two halves.
//
// Products use the su(2) algebra of linalg/spacecurve.hpp:8-18 extended
// associatively (a0 b0 and a0 b terms), which coincides with the reference
// product whenever the left factor is a padded input sample (a0 = 0); the
// identity is (-1, 0, 0, 0).
#include "kernels.hpp"
#include "kernels_wb.hpp"
#include "rules.hpp"

namespace magnus {

namespace {

struct V4 {
    double x[4];
};

__device__ __forceinline__ V4 vzero() { return V4{{0.0, 0.0, 0.0, 0.0}}; }

__device__ __forceinline__ V4 vadd(const V4& a, const V4& b) {
    return V4{{a.x[0] + b.x[0], a.x[1] + b.x[1], a.x[2] + b.x[2], a.x[3] + b.x[3]}};
}

__device__ __forceinline__ V4 vneg(const V4& a) { return V4{{-a.x[0], -a.x[1], -a.x[2], -a.x[3]}}; }

__device__ __forceinline__ V4 vscale(double s, const V4& a) {
    return V4{{a.x[0] * s, a.x[1] * s, a.x[2] * s, a.x[3] * s}};
}

__device__ __forceinline__ V4 vaxpy(const V4& acc, double s, const V4& b) {  // acc + s*b
    return V4{{fma(s, b.x[0], acc.x[0]), fma(s, b.x[1], acc.x[1]), fma(s, b.x[2], acc.x[2]),
               fma(s, b.x[3], acc.x[3])}};
}

// full associative product (reference formula when a0 == 0)
__device__ __forceinline__ V4 qmul(const V4& a, const V4& b) {
    V4 c;
    c.x[0] = a.x[1] * b.x[1] + a.x[2] * b.x[2] + a.x[3] * b.x[3] - a.x[0] * b.x[0];
    c.x[1] = -b.x[0] * a.x[1] + a.x[2] * b.x[3] - a.x[3] * b.x[2] - a.x[0] * b.x[1];
    c.x[2] = -b.x[0] * a.x[2] + a.x[3] * b.x[1] - a.x[1] * b.x[3] - a.x[0] * b.x[2];
    c.x[3] = -b.x[0] * a.x[3] + a.x[1] * b.x[2] - a.x[2] * b.x[1] - a.x[0] * b.x[3];
    return c;
}

// acc + a * b with every term fused
__device__ __forceinline__ V4 qmul_acc(const V4& acc, const V4& a, const V4& b) {
    V4 c;
    c.x[0] = fma(-a.x[0], b.x[0], fma(a.x[3], b.x[3], fma(a.x[2], b.x[2], fma(a.x[1], b.x[1], acc.x[0]))));
    c.x[1] = fma(-a.x[0], b.x[1], fma(-a.x[3], b.x[2], fma(a.x[2], b.x[3], fma(-b.x[0], a.x[1], acc.x[1]))));
    c.x[2] = fma(-a.x[0], b.x[2], fma(-a.x[1], b.x[3], fma(a.x[3], b.x[1], fma(-b.x[0], a.x[2], acc.x[2]))));
    c.x[3] = fma(-a.x[0], b.x[3], fma(-a.x[2], b.x[1], fma(a.x[1], b.x[2], fma(-b.x[0], a.x[3], acc.x[3]))));
    return c;
}

// a * b for a pure vector a = (0, a1, a2, a3): SpaceCurve::sample_update's product
__device__ __forceinline__ V4 amul(const double* a, const V4& b) {
    V4 c;
    c.x[0] = a[0] * b.x[1] + a[1] * b.x[2] + a[2] * b.x[3];
    c.x[1] = -b.x[0] * a[0] + a[1] * b.x[3] - a[2] * b.x[2];
    c.x[2] = -b.x[0] * a[1] + a[2] * b.x[1] - a[0] * b.x[3];
    c.x[3] = -b.x[0] * a[2] + a[0] * b.x[2] - a[1] * b.x[1];
    return c;
}

__device__ __forceinline__ V4 qconj(const V4& a) { return V4{{a.x[0], -a.x[1], -a.x[2], -a.x[3]}}; }

__device__ __forceinline__ V4 pure(const double* a) { return V4{{0.0, a[0], a[1], a[2]}}; }

__device__ __forceinline__ V4 ld4(const double* p) { return V4{{p[0], p[1], p[2], p[3]}}; }
__device__ __forceinline__ void st4(double* p, const V4& v) {
    p[0] = v.x[0];
    p[1] = v.x[1];
    p[2] = v.x[2];
    p[3] = v.x[3];
}

__device__ __forceinline__ void lda(const double* A3, long long i, double* a) {
    a[0] = A3[i * 3 + 0];
    a[1] = A3[i * 3 + 1];
    a[2] = A3[i * 3 + 2];
}

// ---------------------------------------------------------------------------
// block rules.  Block [s0, s0+B]: P_{s0+r} = P_{s0} + sum_t alpha_{r,t} f_{s0+t}
// (boole.hpp:37-72, simpson.hpp:24-58, default.hpp:23-26).  fs is the
// integrand at s0, fn[0..B-1] the integrand at s0+1..s0+B.
//   step:     v (= P_{s0}) becomes P_{s0+B}; Pi[0..B-2] = P at s0+1..s0+B-1
//   back:     v (= P_{s0+B}) becomes P_{s0}
//   interior: Pi[0..B-2] = P at s0+1..s0+B-1 from v = P_{s0}
//   rev:      transpose: f-adjoints xb[0..B-1] at s0..s0+B-1 from the
//             P-adjoints g[0..B-2] at s0+1..s0+B-1 and U, the sum of the
//             P-adjoints from s0+B on (start role at s0 included, end role not)
//   end_role / start_role: alpha_{B,B} / alpha_{B,0}
// ---------------------------------------------------------------------------
template <int I>
struct BlockRule;

template <>
struct BlockRule<BOOLE> {
    static constexpr int B = 4;
    __device__ static double tail(const V4& fs, const V4* fn, int c) {  // 7 f0 + 32 f1 + 12 f2 + 32 f3 + 7 f4
        return fma(7.0, fs.x[c] + fn[3].x[c], fma(32.0, fn[0].x[c] + fn[2].x[c], 12.0 * fn[1].x[c]));
    }
    __device__ static void interior(const V4& fs, const V4* fn, const V4& v, double dt, V4* Pi) {
        for (int c = 0; c < 4; ++c) {
            const double b = v.x[c];
            Pi[0].x[c] = fma(dt / 2, fs.x[c] + fn[0].x[c], b);
            Pi[1].x[c] = fma(dt / 3, fma(4.0, fn[0].x[c], fs.x[c] + fn[1].x[c]), b);
            Pi[2].x[c] = fma(3 * dt / 8, fma(3.0, fn[0].x[c] + fn[1].x[c], fs.x[c] + fn[2].x[c]), b);
        }
    }
    __device__ static void step(const V4& fs, const V4* fn, V4& v, double dt, V4* Pi) {
        interior(fs, fn, v, dt, Pi);
        for (int c = 0; c < 4; ++c) v.x[c] = fma(2 * dt / 45, tail(fs, fn, c), v.x[c]);
    }
    __device__ static void back(const V4& fs, const V4* fn, V4& v, double dt) {
        for (int c = 0; c < 4; ++c) v.x[c] = fma(-2 * dt / 45, tail(fs, fn, c), v.x[c]);
    }
    __device__ static void rev(const V4* g, const V4& U, double dt, V4* xb) {
        for (int c = 0; c < 4; ++c) {
            const double g1 = g[0].x[c], g2 = g[1].x[c], g3 = g[2].x[c], u = U.x[c];
            const double h1 = g1 * (dt / 2);
            xb[0].x[c] = fma(14 * dt / 45, u, fma(3 * dt / 8, g3, fma(dt / 3, g2, h1)));
            xb[1].x[c] = fma(64 * dt / 45, u, fma(9 * dt / 8, g3, fma(4 * dt / 3, g2, h1)));
            xb[2].x[c] = fma(24 * dt / 45, u, fma(9 * dt / 8, g3, g2 * (dt / 3)));
            xb[3].x[c] = fma(64 * dt / 45, u, g3 * (3 * dt / 8));
        }
    }
    __device__ static double end_role(double dt) { return 14 * dt / 45; }
    __device__ static double start_role(double dt) { return 14 * dt / 45; }
    // composite sum weight at offset r of a block that does not start at 0 (boole.hpp:76-106)
    __device__ static double weight(int r, double dt) {
        return r == 0 ? 28 * dt / 45 : (r == 2 ? 24 * dt / 45 : 64 * dt / 45);
    }
};

template <>
struct BlockRule<SIMPSON> {
    static constexpr int B = 2;
    __device__ static void interior(const V4& fs, const V4* fn, const V4& v, double dt, V4* Pi) {
        for (int c = 0; c < 4; ++c) Pi[0].x[c] = fma(dt / 2, fs.x[c] + fn[0].x[c], v.x[c]);
    }
    __device__ static void step(const V4& fs, const V4* fn, V4& v, double dt, V4* Pi) {
        interior(fs, fn, v, dt, Pi);
        for (int c = 0; c < 4; ++c) v.x[c] = fma(dt / 3, fma(4.0, fn[0].x[c], fs.x[c] + fn[1].x[c]), v.x[c]);
    }
    __device__ static void back(const V4& fs, const V4* fn, V4& v, double dt) {
        for (int c = 0; c < 4; ++c) v.x[c] = fma(-dt / 3, fma(4.0, fn[0].x[c], fs.x[c] + fn[1].x[c]), v.x[c]);
    }
    __device__ static void rev(const V4* g, const V4& U, double dt, V4* xb) {
        for (int c = 0; c < 4; ++c) {
            const double h1 = g[0].x[c] * (dt / 2);
            xb[0].x[c] = fma(dt / 3, U.x[c], h1);
            xb[1].x[c] = fma(4 * dt / 3, U.x[c], h1);
        }
    }
    __device__ static double end_role(double dt) { return dt / 3; }
    __device__ static double start_role(double dt) { return dt / 3; }
    __device__ static double weight(int r, double dt) { return r == 0 ? 2 * dt / 3 : 4 * dt / 3; }  // simpson.hpp:60-75
};

template <>
struct BlockRule<RIEMANN> {
    static constexpr int B = 1;
    __device__ static void interior(const V4&, const V4*, const V4&, double, V4*) {}
    __device__ static void step(const V4&, const V4* fn, V4& v, double dt, V4*) { v = vaxpy(v, dt, fn[0]); }
    __device__ static void back(const V4&, const V4* fn, V4& v, double dt) { v = vaxpy(v, -dt, fn[0]); }
    __device__ static void rev(const V4*, const V4&, double, V4* xb) { xb[0] = vzero(); }
    __device__ static double end_role(double dt) { return dt; }
    __device__ static double start_role(double) { return 0.0; }
    __device__ static double weight(int, double dt) { return dt; }  // default.hpp:28-30
};

// ---------------------------------------------------------------------------
// Sample feed of the chunk passes (Simpson and Boole).  The lanes of a
// wavefront walk 64 different chunks, so a per-lane read of one sample touches
// 64 cache lines and the texture path, not the FP64 pipe, sets the pace.  The
// feed loads groups of kFeedG samples of every chunk with direct global->LDS
// loads (global_load_lds_dwordx4, 16 bytes per lane, the LDS address being the
// slab base + 16 * lane): six lanes read one chunk's group (96 contiguous
// bytes) and ten chunks share one instruction, whose 960 bytes fill one slab;
// each lane then reads its own chunk's samples from its slab.  No registers
// hold data in flight.  Under these rules chunk lengths are even ((S-1) % B ==
// 0, chunk length a multiple of 4), so a 16-byte piece never straddles the
// first or last sample of a chunk and every piece read lies in the chunk.
// Group g covers the chunk's samples counted up from its first feed sample
// (UP) or down from its end (!UP).
// ---------------------------------------------------------------------------
constexpr int kFeedG = 4;                                  // samples per group
constexpr int kFeedPieces = kFeedG * 3 / 2;                // 16-byte pieces per chunk and group
constexpr int kFeedPer = 64 / kFeedPieces;                 // chunks per load instruction (lanes 60..63 idle)
constexpr int kFeedOps = (64 + kFeedPer - 1) / kFeedPer;   // load instructions per group
constexpr int kFeedSlab = kFeedPer * kFeedG * 3 + 2;       // doubles per slab: 960 bytes + 16 (bank spread)

// Per wavefront the feed holds its slabs (the load targets) and a table of
// the wavefront's chunks in two separate LDS arrays, so that reads of the
// table need not wait for loads in flight into the slabs.  A table entry packs
// the element offset of the chunk's group 0 (UP: its first sample; !UP: its
// lowest sample) above 16 bits of length (the samples the feed may read,
// 0: none; chunks hold at most 512).
using FeedSrc = long long;
using FeedSlabs = double[kFeedOps][kFeedSlab];

// each lane publishes where its chunk's groups start
__device__ __forceinline__ void feed_init(FeedSrc* src, int lane, long long off, int len) {
    src[lane] = (long long)(((unsigned long long)off << 16) | (unsigned)len);
    __builtin_amdgcn_wave_barrier();
}

// issue the loads of group g for all 64 chunks of the wavefront (every lane calls)
template <bool UP>
__device__ __forceinline__ void feed_issue(FeedSlabs& slab, const FeedSrc* src, const double* A, int lane, int g) {
    const int sub = lane / kFeedPieces, k = lane - sub * kFeedPieces;
    const int first = (2 * k) / 3, last = (2 * k + 1) / 3;  // samples of the group that piece k touches
    // lgkmcnt(0): the wavefront's reads of the previous group from the slabs have
    // completed before any load can land there
    __builtin_amdgcn_s_waitcnt(0xC07F);
    FeedSrc sr[kFeedOps];  // all table reads first, then the loads
#pragma unroll
    for (int m = 0; m < kFeedOps; ++m) {
        const int o = m * kFeedPer + sub;
        sr[m] = src[o < 64 ? o : 63];
    }
    const long long step = (long long)(3 * kFeedG) * g;
#pragma unroll
    for (int m = 0; m < kFeedOps; ++m) {
        const int o = m * kFeedPer + sub;
        const int len = (int)(sr[m] & 0xffff);
        const long long off = sr[m] >> 16;
        const bool ok = sub < kFeedPer && o < 64 && (UP ? (kFeedG * g + last < len) : (first >= kFeedG * (g + 1) - len));
        if (ok) {
            const double* p = A + (UP ? off + step : off - step) + 2 * k;
            __builtin_amdgcn_global_load_lds(p, (__attribute__((address_space(3))) void*)&slab[m][0], 16, 0, 0);
        }
    }
}

// the group's loads have landed (vmcnt(0)); every lane of the wavefront may read the slabs
__device__ __forceinline__ void feed_wait() {
    __builtin_amdgcn_s_waitcnt(0x0F70);
    __builtin_amdgcn_wave_barrier();
}

// this lane's chunk in the slabs: sample i, component c at [3 i + c]
__device__ __forceinline__ const double* feed_row(const FeedSlabs& slab, int lane) {
    const int m = lane / kFeedPer;
    return &slab[m][(lane - m * kFeedPer) * kFeedG * 3];
}

__device__ __forceinline__ long long chunk_begin(const WbArgs& a, int c) { return (long long)c * a.LC; }
__device__ __forceinline__ long long chunk_end(const WbArgs& a, int c) {
    long long e = (long long)(c + 1) * a.LC;
    return e < a.S - 1 ? e : a.S - 1;
}

// ---------------------------------------------------------------------------
// chunk pass: local streams V_1..V_{K+1} over (s, e] from zero bases.
// summary per (b, c): [Vhat_1..Vhat_K, ZV_0..ZV_K] (4 doubles each)
// ---------------------------------------------------------------------------

// one block [s0, s0+B] of the chunk pass; ab = samples s0+1..s0+B.  On entry
// base[l] = V_{l+1}[s0] and f0[l] = its integrand at s0; on exit both at s0+B.
template <int I, int K>
__device__ __forceinline__ void fwd_block(const double (*ab)[3], V4* base, V4* f0, double dt) {
    using R = BlockRule<I>;
    constexpr int B = R::B;
    constexpr int L = K + 1;
    V4 fn[B];  // integrand of the current level at s0+1..s0+B
#pragma unroll
    for (int r = 0; r < B; ++r) fn[r] = pure(ab[r]);
#pragma unroll
    for (int l = 0; l < L; ++l) {
        V4 Pi[B > 1 ? B - 1 : 1];
        R::step(f0[l], fn, base[l], dt, Pi);
        if (l == 0) {  // V_1 = Prefix(A) is a pure vector: drop its zero component
            base[0].x[0] = 0.0;
#pragma unroll
            for (int r = 0; r + 1 < B; ++r) Pi[r].x[0] = 0.0;
        }
        f0[l] = fn[B - 1];
        if (l + 1 < L) {
#pragma unroll
            for (int r = 0; r + 1 < B; ++r) fn[r] = amul(ab[r], Pi[r]);
            fn[B - 1] = amul(ab[B - 1], base[l]);
        }
    }
}

template <int I, int K>
__global__ void __launch_bounds__(256, 2) wb_fwd_pass1(WbArgs a) {
    using R = BlockRule<I>;
    constexpr int B = R::B;
    constexpr int L = K + 1;
    constexpr bool kFed = B % 2 == 0;  // Simpson, Boole: samples through the feed
    __shared__ FeedSlabs feed_slab[kFed ? 4 : 1];
    __shared__ FeedSrc feed_src[kFed ? 4 : 1][64];
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = t < a.B * a.C;  // with the feed every lane stays: its loads are cooperative
    if (!kFed && !valid) return;
    const long long b = valid ? t / a.C : 0;
    const int c = valid ? (int)(t - b * a.C) : 0;
    const long long S = a.S;
    const double dt = a.dt;
    const double* A3 = a.in + b * S * 3;
    const long long s = chunk_begin(a, c), e = chunk_end(a, c);

    V4 base[L], f0[L];  // level l: stream V_{l+1}; f0[l] = its integrand at the block start
#pragma unroll
    for (int l = 0; l < L; ++l) {
        base[l] = vzero();
        f0[l] = vzero();
    }
    double as[3];
    lda(A3, s, as);
    f0[0] = pure(as);
    if (I == RIEMANN && c == 0) {  // P_0 = dt f_0 is a step of its own
        V4 X = f0[0];
#pragma unroll
        for (int l = 0; l < L; ++l) {
            base[l] = vscale(dt, X);
            X = amul(as, base[l]);
        }
    }
    if constexpr (kFed) {
        // group g = samples s+1+4g .. s+4+4g; the wavefront runs the largest group count
        const int lane = (int)(threadIdx.x & 63);
        FeedSlabs& slab = feed_slab[threadIdx.x >> 6];
        FeedSrc* src = feed_src[threadIdx.x >> 6];
        const int len = valid ? (int)(e - s) : 0;
        feed_init(src, lane, (b * S + s + 1) * 3, len);
        int ng = (len + kFeedG - 1) / kFeedG;
        for (int off = 32; off > 0; off >>= 1) ng = max(ng, __shfl_xor(ng, off, 64));
        const double* row = feed_row(slab, lane);
        if (ng > 0) {
            feed_issue<true>(slab, src, a.in, lane, 0);
            feed_wait();
        }
        for (int g = 0; g < ng; ++g) {
            double gs[kFeedG][3];
#pragma unroll
            for (int i = 0; i < kFeedG; ++i) {
                gs[i][0] = row[3 * i];
                gs[i][1] = row[3 * i + 1];
                gs[i][2] = row[3 * i + 2];
            }
            __builtin_amdgcn_wave_barrier();
            if (g + 1 < ng) feed_issue<true>(slab, src, a.in, lane, g + 1);  // lands during this group
#pragma unroll
            for (int jb = 0; jb < kFeedG / B; ++jb)
                if (kFeedG * g + jb * B < len) fwd_block<I, K>(gs + jb * B, base, f0, dt);
            feed_wait();
        }
        if (!valid) return;
    } else {
        double nx[B][3];
#pragma unroll
        for (int r = 0; r < B; ++r) lda(A3, s + 1 + r, nx[r]);
#pragma unroll 2
        for (long long s0 = s; s0 < e; s0 += B) {
            double ab[B][3];
#pragma unroll
            for (int r = 0; r < B; ++r) {
                ab[r][0] = nx[r][0];
                ab[r][1] = nx[r][1];
                ab[r][2] = nx[r][2];
            }
            if (s0 + B < e) {  // next block's samples, in flight during this block
#pragma unroll
                for (int r = 0; r < B; ++r) lda(A3, s0 + B + 1 + r, nx[r]);
            }
            fwd_block<I, K>(ab, base, f0, dt);
        }
    }
    // f0[l] = A_e V_l[e] (V_0 = identity), base[l] = V_{l+1}[e]
    const double cfix = (I == RIEMANN || e == S - 1) ? 0.0 : R::end_role(dt);
    double* out = a.fsum + t * (long long)(2 * K + 1) * 4;
#pragma unroll
    for (int k = 0; k < K; ++k) st4(out + k * 4, base[k]);
    V4 z0 = vaxpy(base[0], cfix, f0[0]);
    if (c > 0) z0 = vaxpy(z0, -R::start_role(dt), pure(as));
    st4(out + K * 4, z0);
#pragma unroll
    for (int k = 1; k <= K; ++k) st4(out + (K + k) * 4, vaxpy(base[k], cfix, f0[k]));
}

// ---------------------------------------------------------------------------
// truncated series.  USer: I + sum_{j=1..K} v[j-1] x^j;  FSer: sum_{j=0..K} v[j] x^j
// ---------------------------------------------------------------------------
template <int K>
struct USer {
    V4 v[K];
};
template <int K>
struct FSer {
    V4 v[K + 1];
};
template <int K>
struct Pair {
    USer<K> V;
    FSer<K> Z;
};

template <int K>
__device__ __forceinline__ USer<K> umul(const USer<K>& a, const USer<K>& b) {
    USer<K> h;
#pragma unroll
    for (int m = 1; m <= K; ++m) {
        V4 acc = vadd(a.v[m - 1], b.v[m - 1]);
#pragma unroll
        for (int j = 1; j < m; ++j) acc = vadd(acc, qmul(a.v[j - 1], b.v[m - j - 1]));
        h.v[m - 1] = acc;
    }
    return h;
}

template <int K>
__device__ __forceinline__ Pair<K> compose(const Pair<K>& later, const Pair<K>& earlier) {
    Pair<K> r;
    r.V = umul(later.V, earlier.V);
#pragma unroll
    for (int m = 0; m <= K; ++m) {
        V4 acc = vadd(later.Z.v[m], earlier.Z.v[m]);
#pragma unroll
        for (int j = 0; j < m; ++j) acc = vadd(acc, qmul(later.Z.v[j], earlier.V.v[m - j - 1]));
        r.Z.v[m] = acc;
    }
    return r;
}

template <int K>
__device__ __forceinline__ Pair<K> pair_identity() {
    Pair<K> p;
#pragma unroll
    for (int j = 0; j < K; ++j) p.V.v[j] = vzero();
#pragma unroll
    for (int j = 0; j <= K; ++j) p.Z.v[j] = vzero();
    return p;
}

template <int K>
__device__ __forceinline__ Pair<K> load_pair(const double* p) {
    Pair<K> r;
#pragma unroll
    for (int j = 0; j < K; ++j) r.V.v[j] = ld4(p + j * 4);
#pragma unroll
    for (int j = 0; j <= K; ++j) r.Z.v[j] = ld4(p + (K + j) * 4);
    return r;
}

// lane shuffles within segments of w lanes (w a power of two dividing 64)
__device__ __forceinline__ V4 shfl4(const V4& v, int src, int w) {
    return V4{{__shfl(v.x[0], src, w), __shfl(v.x[1], src, w), __shfl(v.x[2], src, w), __shfl(v.x[3], src, w)}};
}
__device__ __forceinline__ V4 shfl_down4(const V4& v, int d, int w) {
    return V4{{__shfl_down(v.x[0], d, w), __shfl_down(v.x[1], d, w), __shfl_down(v.x[2], d, w),
               __shfl_down(v.x[3], d, w)}};
}
__device__ __forceinline__ V4 shfl_up4(const V4& v, int d, int w) {
    return V4{{__shfl_up(v.x[0], d, w), __shfl_up(v.x[1], d, w), __shfl_up(v.x[2], d, w),
               __shfl_up(v.x[3], d, w)}};
}

template <int K>
__device__ __forceinline__ Pair<K> shfl_down_pair(const Pair<K>& p, int d, int w) {
    Pair<K> r;
#pragma unroll
    for (int j = 0; j < K; ++j) r.V.v[j] = shfl_down4(p.V.v[j], d, w);
#pragma unroll
    for (int j = 0; j <= K; ++j) r.Z.v[j] = shfl_down4(p.Z.v[j], d, w);
    return r;
}

// ---------------------------------------------------------------------------
// combine: one wavefront per instance, each lane owns m consecutive chunks.
// Suffix scan of the pairs over the lanes -> totals (lane 0) -> GL algebra
// (lane q evaluates node q) -> outputs.  VJP: GL adjoint, prefix scan of the
// V series, then per chunk the cotangents of its summary:
//   Wbar_{c+1} = adj_right(Y_{c+1}; Zbar) + adj_right(S_{c+1}; What_bar)
//   Vhat_bar_c = adj_left(W~_c; Wbar_{c+1}),   ZV_bar_c = adj_left(W~_c; Zbar)
// with (S_{c+1}, Y_{c+1}) the composition of the chunks after c and
// W~_c the product of the chunks before c.
// ---------------------------------------------------------------------------
template <int K, bool VJP>
__global__ void __launch_bounds__(256) wb_combine(WbArgs a) {
    constexpr int n = K + 1;
    constexpr int stride = (2 * K + 1) * 4;
    // segment of SW lanes per instance (SW = a.cw, a power of two); shuffles stay in the segment
    const int SW = a.cw;
    const int lane = (int)(threadIdx.x & (SW - 1));
    const long long b = ((long long)blockIdx.x * 4 + (threadIdx.x >> 6)) * (64 / SW) + (threadIdx.x & 63) / SW;
    if (b >= a.B) return;  // uniform over the segment
    const int C = a.C;
    const int m = (C + SW - 1) / SW;
    const int nl = (C + m - 1) / m;
    const int c0 = lane * m;
    const int c1 = c0 + m < C ? c0 + m : C;
    const bool act = lane < nl;
    const double* fs = a.fsum + b * (long long)C * stride;

    Pair<K> pr = pair_identity<K>();
    if (act) {
        pr = load_pair<K>(fs + (long long)c0 * stride);
        for (int c = c0 + 1; c < c1; ++c) pr = compose<K>(load_pair<K>(fs + (long long)c * stride), pr);
    }
    Pair<K> suf = pr;
    for (int d = 1; d < nl; d <<= 1) {
        const Pair<K> o = shfl_down_pair<K>(suf, d, SW);
        if (lane + d < nl) suf = compose<K>(o, suf);
    }
    // instance totals: Z_0..K and What_1..K
    V4 Z[K + 1], Wh[K];
#pragma unroll
    for (int j = 0; j <= K; ++j) Z[j] = shfl4(suf.Z.v[j], 0, SW);
#pragma unroll
    for (int j = 0; j < K; ++j) Wh[j] = shfl4(suf.V.v[j], 0, SW);

    const int Q = a.Q;
    double sq = 0.0, wq = 0.0;
    for (int q = 0; q < Q; ++q)
        if (lane == q) {
            sq = a.x[q] - 1.0;
            wq = a.w[q];
        }
    // GL algebra for node `lane`: T_j = What_j + s sum_{l<j} What_l T_{j-l};
    // Omega_j += w (Z_{j-1} + s sum_{l<=j-2} Z_l T_{j-1-l})
    V4 T[K + 1];
    T[0] = vzero();
#pragma unroll
    for (int j = 1; j <= K; ++j) {
        V4 acc = Wh[j - 1];
#pragma unroll
        for (int l = 1; l < j; ++l) acc = vaxpy(acc, sq, qmul(Wh[l - 1], T[j - l]));
        T[j] = acc;
    }
    V4 part[K];  // Omega_2..Omega_n of this node
#pragma unroll
    for (int j = 2; j <= n; ++j) {
        V4 acc = Z[j - 1];
#pragma unroll
        for (int l = 0; l <= j - 2; ++l) acc = vaxpy(acc, sq, qmul(Z[l], T[j - 1 - l]));
        part[j - 2] = (lane < Q) ? vscale(wq, acc) : vzero();
    }
    V4 Om[K];
#pragma unroll
    for (int j = 0; j < K; ++j) Om[j] = vzero();
    for (int q = 0; q < Q; ++q) {
#pragma unroll
        for (int j = 0; j < K; ++j) Om[j] = vadd(Om[j], shfl4(part[j], q, SW));
    }
    if (a.out && lane == 0) {  // outputs: components 1..3
        double* o = a.out + b * (long long)a.n_out * 3;
        if (a.op == OP_MANY) {
            for (int cc = 0; cc < 3; ++cc) o[cc] = Z[0].x[cc + 1];
#pragma unroll
            for (int j = 1; j < n; ++j)
                for (int cc = 0; cc < 3; ++cc) o[j * 3 + cc] = Om[j - 1].x[cc + 1];
        } else if (a.op == OP_SUM) {
            V4 acc = Z[0];
#pragma unroll
            for (int j = 0; j < K; ++j) acc = vadd(acc, Om[j]);
            for (int cc = 0; cc < 3; ++cc) o[cc] = acc.x[cc + 1];
        } else {
            for (int cc = 0; cc < 3; ++cc) o[cc] = Om[K - 1].x[cc + 1];
        }
    }
    if constexpr (VJP) {
        // ---- adjoint of the GL algebra, node `lane` ----
        const double* cot = a.cot + b * (long long)a.n_out * 3;
        V4 Ob[n];
#pragma unroll
        for (int j = 0; j < n; ++j) {
            Ob[j] = vzero();
            const double* row = nullptr;
            if (a.op == OP_MANY) row = cot + j * 3;
            else if (a.op == OP_SUM || j == n - 1) row = cot;
            if (row) Ob[j] = pure(row);
        }
        V4 Zp[K + 1], Wp[K], Tb[K + 1];
#pragma unroll
        for (int j = 0; j <= K; ++j) {
            Zp[j] = vzero();
            Tb[j] = vzero();
        }
#pragma unroll
        for (int j = 0; j < K; ++j) Wp[j] = vzero();
        if (lane < Q) {
#pragma unroll
            for (int j = 2; j <= n; ++j) {
                const V4 g = vscale(wq, Ob[j - 1]);
                Zp[j - 1] = vadd(Zp[j - 1], g);
#pragma unroll
                for (int l = 0; l <= j - 2; ++l) {
                    Zp[l] = vaxpy(Zp[l], sq, qmul(g, qconj(T[j - 1 - l])));
                    Tb[j - 1 - l] = vaxpy(Tb[j - 1 - l], sq, qmul(qconj(Z[l]), g));
                }
            }
#pragma unroll
            for (int j = K; j >= 1; --j) {
                Wp[j - 1] = vadd(Wp[j - 1], Tb[j]);
#pragma unroll
                for (int l = 1; l < j; ++l) {
                    Wp[l - 1] = vaxpy(Wp[l - 1], sq, qmul(Tb[j], qconj(T[j - l])));
                    Tb[j - l] = vaxpy(Tb[j - l], sq, qmul(qconj(Wh[l - 1]), Tb[j]));
                }
            }
        }
        V4 Zbar[K + 1], Wbar[K];
        Zbar[0] = Ob[0];
#pragma unroll
        for (int j = 1; j <= K; ++j) Zbar[j] = vzero();
#pragma unroll
        for (int j = 0; j < K; ++j) Wbar[j] = vzero();
        for (int q = 0; q < Q; ++q) {
#pragma unroll
            for (int j = 0; j <= K; ++j) Zbar[j] = vadd(Zbar[j], shfl4(Zp[j], q, SW));
#pragma unroll
            for (int j = 0; j < K; ++j) Wbar[j] = vadd(Wbar[j], shfl4(Wp[j], q, SW));
        }
        // exclusive suffix (chunks after this lane) and exclusive prefix of V
        Pair<K> sf = shfl_down_pair<K>(suf, 1, SW);
        if (lane + 1 >= nl) sf = pair_identity<K>();
        USer<K> pre = pr.V;
        for (int d = 1; d < nl; d <<= 1) {
            USer<K> o;
#pragma unroll
            for (int j = 0; j < K; ++j) o.v[j] = shfl_up4(pre.v[j], d, SW);
            if (lane >= d) pre = umul<K>(pre, o);
        }
        USer<K> W;
#pragma unroll
        for (int j = 0; j < K; ++j) W.v[j] = shfl_up4(pre.v[j], 1, SW);
        if (lane == 0) {
#pragma unroll
            for (int j = 0; j < K; ++j) W.v[j] = vzero();
        }
        if (!act) return;
        double* bs = a.bases + b * (long long)C * K * 4;
        for (int c = c0; c < c1; ++c) {  // W~_c for the lane's chunks
#pragma unroll
            for (int j = 0; j < K; ++j) st4(bs + ((long long)c * K + j) * 4, W.v[j]);
            if (c + 1 < c1) {
                USer<K> v;
#pragma unroll
                for (int j = 0; j < K; ++j) v.v[j] = ld4(fs + (long long)c * stride + j * 4);
                W = umul<K>(v, W);
            }
        }
        double* so = a.src + b * (long long)C * stride;
        for (int c = c1 - 1; c >= c0; --c) {
            // Wbar_{c+1}, coefficients 1..K
            V4 wb[K];
#pragma unroll
            for (int l = 1; l <= K; ++l) {
                V4 acc = Wbar[l - 1];
#pragma unroll
                for (int mm = l + 1; mm <= K; ++mm) acc = vadd(acc, qmul(qconj(sf.V.v[mm - l - 1]), Wbar[mm - 1]));
#pragma unroll
                for (int mm = l; mm <= K; ++mm) acc = vadd(acc, qmul(qconj(sf.Z.v[mm - l]), Zbar[mm]));
                wb[l - 1] = acc;
            }
#pragma unroll
            for (int j = 0; j < K; ++j) W.v[j] = ld4(bs + ((long long)c * K + j) * 4);
            double* o = so + (long long)c * stride;
#pragma unroll
            for (int j = 1; j <= K; ++j) {
                V4 acc = wb[j - 1];
#pragma unroll
                for (int mm = j + 1; mm <= K; ++mm) acc = vadd(acc, qmul(wb[mm - 1], qconj(W.v[mm - j - 1])));
                st4(o + (j - 1) * 4, acc);
            }
#pragma unroll
            for (int j = 0; j <= K; ++j) {
                V4 acc = Zbar[j];
#pragma unroll
                for (int mm = j + 1; mm <= K; ++mm) acc = vadd(acc, qmul(Zbar[mm], qconj(W.v[mm - j - 1])));
                st4(o + (K + j) * 4, acc);
            }
            if (c > c0) sf = compose<K>(sf, load_pair<K>(fs + (long long)c * stride));
        }
    }
}

// ---------------------------------------------------------------------------
// reverse chunk pass: local adjoint of chunk (s, e] given its summary
// cotangents; local streams rebuilt backwards from Vhat.  The values at e and
// at s go to `edge` for the fixup; dA for s < i < e is staged per wavefront in
// LDS, 8 samples (192 bytes) per lane, and written by the whole wavefront with
// stores that cover full 64-byte segments (one lane's 48-byte block writes
// would straddle segments and be written twice).
// ---------------------------------------------------------------------------
constexpr int kRevGroup = 8;                          // samples per lane and flush
constexpr int kRevRow = kRevGroup * 3;                // doubles per lane row
constexpr int kRevWaves = 4;                          // wavefronts per workgroup

// per lane: the element offset in dA of slot 0 of its current group, packed
// above 16 bits of lim (signed): slots q <= lim lie at or below the chunk start s
using RevDest = long long;

struct RevStage {
    double row[kRevWaves][64 * kRevRow];
    RevDest dst[kRevWaves][64];
};

// One block [s0, s0+B] of the reverse chunk pass.  On entry Vend[k] holds
// V_{k+1} at s0+B and U[k] the suffix sum of level k+1 from s0+B on; on exit
// they refer to s0.  FIRST: the block starting at the chunk start s, whose
// value at s (r = 0) is the chunk's start half and goes to `pend`; there the
// local streams vanish at s and s is not the end of a block of this chunk
// (except under Riemann, where s = 0 is a step of its own).
// dA at s0+r goes to slot (s0 + r - e) mod 8 of the lane's staging row.
template <int I, int K, bool FIRST>
__device__ __forceinline__ void rev_block(long long s0, long long e, const double (*ab)[3], double* aend, V4* Vend,
                                          V4* U, const V4& zK, double dt, double* row, V4& pend) {
    using R = BlockRule<I>;
    constexpr int B = R::B;
    constexpr int BI = B > 1 ? B - 1 : 1;
    constexpr bool zero_start = FIRST && I != RIEMANN;  // local streams vanish at s (no P_0 step)
    constexpr bool end_role0 = !FIRST || I == RIEMANN;  // s0 is also the end of the block before
    // rebuild V_1..V_K at s0..s0+B-1, levels ascending: Vend[k] becomes V_{k+1}[s0]
    // in place, Vi[k][r-1] holds V_{k+1}[s0+r]
    V4 Vi[K][BI];
    V4 fB = pure(aend);  // integrand of the level at s0+B
#pragma unroll
    for (int k = 0; k < K; ++k) {
        V4 fs, fn[B];
        if (k == 0) {
            fs = pure(ab[0]);
#pragma unroll
            for (int r = 1; r < B; ++r) fn[r - 1] = pure(ab[r]);
        } else {
            fs = amul(ab[0], Vend[k - 1]);
#pragma unroll
            for (int r = 1; r < B; ++r) fn[r - 1] = amul(ab[r], Vi[k - 1][r - 1]);
        }
        fn[B - 1] = fB;
        if (k + 1 < K) fB = amul(aend, Vend[k]);
        if (zero_start) Vend[k] = vzero();
        else R::back(fs, fn, Vend[k], dt);
        R::interior(fs, fn, Vend[k], dt, Vi[k]);
        if (k == 0) {  // V_1 = Prefix(A) is a pure vector: drop its zero component
            Vend[0].x[0] = 0.0;
#pragma unroll
            for (int r = 0; r + 1 < B; ++r) Vi[0][r].x[0] = 0.0;
        }
    }
    // reverse sweep of the block, top level down.  (The sum weight at s in a
    // FIRST block never reaches the output: there V = 0 and the end role is off.)
    V4 Xr[B], dAb[B];
#pragma unroll
    for (int r = 0; r < B; ++r) {
        Xr[r] = vscale(R::weight(r, dt), zK);
        dAb[r] = vzero();
    }
#pragma unroll
    for (int k = K - 1; k >= 0; --k) {
        V4 G[B];
#pragma unroll
        for (int r = 0; r < B; ++r) {
            if (!(zero_start && r == 0)) dAb[r] = qmul_acc(dAb[r], Xr[r], qconj(r == 0 ? Vend[k] : Vi[k][r - 1]));
            G[r] = vneg(amul(ab[r], Xr[r]));
        }
        V4 xb[B];
        R::rev(G + 1, U[k], dt, xb);
        V4 us = U[k];
#pragma unroll
        for (int r = B - 1; r >= 0; --r) us = vadd(us, G[r]);
        U[k] = us;
        if (end_role0) xb[0] = vaxpy(xb[0], R::end_role(dt), us);
#pragma unroll
        for (int r = 0; r < B; ++r) Xr[r] = xb[r];
    }
    const int q0 = (int)((s0 - e) & (kRevGroup - 1));
#pragma unroll
    for (int r = 0; r < B; ++r) {
        const V4 v = vadd(dAb[r], Xr[r]);  // Xr holds Xbar_1
        if (FIRST && r == 0) {
            pend = v;
        } else {
            double* w = row + (q0 + r) * 3;
            w[0] = v.x[1];
            w[1] = v.x[2];
            w[2] = v.x[3];
        }
    }
    aend[0] = ab[0][0];
    aend[1] = ab[0][1];
    aend[2] = ab[0][2];
}

// Write the staged group of every lane o of the wavefront: slot q of lane o
// goes to dA[off_o + 3 q + comp] when q > lim_o (the sample lies above the
// chunk start).  Piece idx = 64 r + lane is double k = idx mod 24 of owner lane
// o = idx / 24, so the 64 pieces of one store cover whole 64-byte segments;
// pieces r and r + 3 have the same k and owners 8 lanes apart.
__device__ __forceinline__ void rev_flush(const double* rows, const RevDest* dst, double* dA, int lane) {
    asm volatile("" : "+v"(lane));  // keep the per-piece index math inside the flush
    int ob[3], kk[3], qq[3];
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        const int idx = m * 64 + lane;
        ob[m] = idx / kRevRow;
        kk[m] = idx - ob[m] * kRevRow;
        qq[m] = kk[m] / 3;
    }
#pragma unroll
    for (int r0 = 0; r0 < kRevRow; r0 += 8) {  // 8 pieces: all LDS reads first, then the stores
        RevDest d[8];
        double v[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int r = r0 + i;
            d[i] = dst[ob[r % 3] + 8 * (r / 3)];
            v[i] = rows[r * 64 + lane];
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int m = (r0 + i) % 3;
            if (qq[m] > (int)(short)(d[i] & 0xffff)) dA[(d[i] >> 16) + kk[m]] = v[i];
        }
    }
}

// publish this lane's current group before a flush
__device__ __forceinline__ void rev_publish(RevStage& st, int wv, int lane, long long pe, int len, int grp) {
    const int top = kRevGroup * (grp + 1);
    st.dst[wv][lane] = (long long)(((unsigned long long)(pe - 3 * top) << 16) | ((unsigned)(top - len) & 0xffffu));
}

template <int I, int K>
__global__ void __launch_bounds__(256, (I == BOOLE || (I == SIMPSON && K >= 6)) ? 1 : 2) wb_rev(WbArgs a) {
    using R = BlockRule<I>;
    constexpr int B = R::B;
    constexpr bool kFed = B % 2 == 0;  // Simpson, Boole: samples through the feed
    static_assert(kRevGroup % B == 0, "a block must not straddle two staging groups");
    static_assert(kRevGroup % kFeedG == 0, "a feed group must not straddle two staging groups");
    __shared__ RevStage st;
    __shared__ FeedSlabs feed_slab[kFed ? kRevWaves : 1];
    __shared__ FeedSrc feed_src[kFed ? kRevWaves : 1][64];
    const int lane = (int)(threadIdx.x & 63), wv = (int)(threadIdx.x >> 6);
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = t < a.B * a.C;  // every lane stays: the flushes are cooperative
    const long long b = valid ? t / a.C : 0;
    const int c = valid ? (int)(t - b * a.C) : 0;
    const long long S = a.S;
    const double dt = a.dt;
    const double er = R::end_role(dt);
    const double* A3 = a.in + b * S * 3;
    const long long s = chunk_begin(a, c), e = chunk_end(a, c);
    const long long tt = valid ? t : 0;
    const double* sp = a.src + tt * (long long)(2 * K + 1) * 4;
    const double* fp = a.fsum + tt * (long long)(2 * K + 1) * 4;
    double* row = st.row[wv] + lane * kRevRow;
    const long long pe = (b * S + e) * 3;  // dA offset of the chunk end
    const int len = valid ? (int)(e - s) : 0;

    V4 Vend[K], U[K];
#pragma unroll
    for (int k = 0; k < K; ++k) Vend[k] = ld4(fp + k * 4);
    double aend[3];
    lda(A3, e, aend);
    const V4 zK = ld4(sp + 2 * K * 4);  // ZV_bar_K
    {
        // the chunk end e: Xbar_{k+1} known, U_k = Gbar_k + Vhat_bar_k + ZV_bar_{k-1}
        // (the last term is the folded sum cotangent), Xbar_k = er U_k + (c_e - er) ZV_bar_{k-1}
        const double ce = sum_vjp_coef<I>(e, S, dt);
        V4 Xu = vscale(ce, zK);
        V4 d = vzero();
#pragma unroll
        for (int k = K; k >= 1; --k) {
            const V4 zk = ld4(sp + (K + k - 1) * 4);
            d = qmul_acc(d, Xu, qconj(Vend[k - 1]));
            const V4 uv = vadd(vadd(vneg(amul(aend, Xu)), ld4(sp + (k - 1) * 4)), zk);
            U[k - 1] = uv;
            Xu = vaxpy(vscale(er, uv), ce - er, zk);
        }
        d = vadd(d, Xu);
        if (valid) {
            double* ed = a.edge + t * 6;
            ed[0] = d.x[1];
            ed[1] = d.x[2];
            ed[2] = d.x[3];
        }
    }
    // blocks s0 = e-B, e-2B, ... above s in a wavefront-uniform loop, then the
    // block starting at s, unless the chunk's start sample has no role
    // (Riemann after chunk 0)
    const bool has_first = valid && (I != RIEMANN || c == 0);
    const int nb = valid ? (int)((e - s) / B) - (has_first ? 0 : 1) : 0;  // blocks in all
    const int nbn = has_first ? nb - 1 : nb;                              // blocks above s
    V4 pend = vzero();
    if constexpr (kFed) {
        // feed group g = samples e-4g-4 .. e-4g-1 (blocks j with jB / 4 == g), two
        // feed groups per staging group; the block at s comes last, from its own loads
        constexpr int kPer = kRevGroup / kFeedG;
        FeedSlabs& slab = feed_slab[wv];
        FeedSrc* src = feed_src[wv];
        feed_init(src, lane, (b * S + e - kFeedG) * 3, len);
        int ng = (nbn * B + kFeedG - 1) / kFeedG;
        for (int off = 32; off > 0; off >>= 1) ng = max(ng, __shfl_xor(ng, off, 64));
        const double* frow = feed_row(slab, lane);
        if (ng > 0) {
            feed_issue<false>(slab, src, a.in, lane, 0);
            feed_wait();
        }
        for (int g = 0; g < ng; ++g) {
            double gs[kFeedG][3];  // samples e-4g-4+i
#pragma unroll
            for (int i = 0; i < kFeedG; ++i) {
                gs[i][0] = frow[3 * i];
                gs[i][1] = frow[3 * i + 1];
                gs[i][2] = frow[3 * i + 2];
            }
            __builtin_amdgcn_wave_barrier();
            if (g + 1 < ng) feed_issue<false>(slab, src, a.in, lane, g + 1);  // lands during this group
#pragma unroll
            for (int jb = 0; jb < kFeedG / B; ++jb) {
                const int j = g * (kFeedG / B) + jb;
                if (j < nbn)
                    rev_block<I, K, false>(e - (long long)(j + 1) * B, e, gs + kFeedG - (jb + 1) * B, aend, Vend, U,
                                           zK, dt, row, pend);
            }
            feed_wait();  // before the stores, so that it waits for the feed only
            if ((g + 1) % kPer == 0 || g + 1 == ng) {
                const int grp = g / kPer;
                rev_publish(st, wv, lane, pe, kRevGroup * grp < nbn * B ? len : 0, grp);
                rev_flush(st.row[wv], st.dst[wv], a.dA, lane);
            }
        }
        if (has_first) {
            double ab[B][3];
#pragma unroll
            for (int r = 0; r < B; ++r) lda(A3, s + r, ab[r]);
            rev_block<I, K, true>(s, e, ab, aend, Vend, U, zK, dt, row, pend);
        }
        rev_publish(st, wv, lane, pe, len, (nbn * B) / kRevGroup);
        rev_flush(st.row[wv], st.dst[wv], a.dA, lane);
    } else {
        int nbw = nbn;
        for (int off = 32; off > 0; off >>= 1) nbw = max(nbw, __shfl_xor(nbw, off, 64));
        double nx[B][3];
        if (nb > 0) {
#pragma unroll
            for (int r = 0; r < B; ++r) lda(A3, e - B + r, nx[r]);
        }
        int grp = 0;
        for (int j = 0; j < nbw; ++j) {
            if (j < nbn) {
                const long long s0 = e - (long long)(j + 1) * B;
                double ab[B][3];
#pragma unroll
                for (int r = 0; r < B; ++r) {
                    ab[r][0] = nx[r][0];
                    ab[r][1] = nx[r][1];
                    ab[r][2] = nx[r][2];
                }
                if (j + 1 < nb) {  // next block's samples, in flight during this block
#pragma unroll
                    for (int r = 0; r < B; ++r) lda(A3, s0 - B + r, nx[r]);
                }
                rev_block<I, K, false>(s0, e, ab, aend, Vend, U, zK, dt, row, pend);
                grp = (j * B) / kRevGroup;
            }
            if (((j + 1) * B) % kRevGroup == 0 || j + 1 == nbw) {
                rev_publish(st, wv, lane, pe, len, grp);
                rev_flush(st.row[wv], st.dst[wv], a.dA, lane);
            }
        }
        if (has_first) {
            rev_block<I, K, true>(s, e, nx, aend, Vend, U, zK, dt, row, pend);
            grp = (nbn * B) / kRevGroup;
        }
        rev_publish(st, wv, lane, pe, len, grp);
        rev_flush(st.row[wv], st.dst[wv], a.dA, lane);
    }
    // the start role of A_s saw the folded ZV_bar_0; in chunks after the first
    // that sample's sum weight belongs to the previous chunk
    if (c > 0 && I != RIEMANN) pend = vaxpy(pend, -R::start_role(dt), ld4(sp + K * 4));
    if (valid) {
        double* ed = a.edge + t * 6;
        ed[3] = pend.x[1];
        ed[4] = pend.x[2];
        ed[5] = pend.x[3];
    }
}

// dA at the chunk boundaries: value at e from chunk c plus the start value of chunk c+1
__global__ void __launch_bounds__(256) wb_fixup(WbArgs a) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.B * a.C) return;
    const long long b = t / a.C;
    const int c = (int)(t - b * a.C);
    const long long e = chunk_end(a, c);
    const double* ed = a.edge + t * 6;
    double v0 = ed[0], v1 = ed[1], v2 = ed[2];
    if (c + 1 < a.C) {
        v0 += ed[9];
        v1 += ed[10];
        v2 += ed[11];
    }
    double* o = a.dA + (b * a.S + e) * 3;
    o[0] = v0;
    o[1] = v1;
    o[2] = v2;
    if (c == 0) {
        double* z = a.dA + b * a.S * 3;
        z[0] = ed[3];
        z[1] = ed[4];
        z[2] = ed[5];
    }
}

template <int I, int K>
hipError_t launch_wb_k(const WbArgs& a, bool vjp, bool summaries_given, hipStream_t st) {
    const long long threads = a.B * a.C;
    const int blk = 256;
    const long long grid = (threads + blk - 1) / blk;
    // combine: one segment of cw lanes per instance, cw the smallest power of two
    // holding the chunks and the GL nodes (at most a wavefront)
    WbArgs ca = a;
    ca.cw = 1;
    while (ca.cw < 64 && (ca.cw < a.C || ca.cw < a.Q)) ca.cw <<= 1;
    const long long per_block = 4 * (64 / ca.cw);
    const long long cgrid = (a.B + per_block - 1) / per_block;
    void* h;
    if (!summaries_given) {
        h = timer_begin(vjp ? "wb_fwd_pass1_vjp" : "wb_fwd_pass1", st);
        hipLaunchKernelGGL((wb_fwd_pass1<I, K>), dim3(grid), dim3(blk), 0, st, a);
        timer_end(h);
    }
    if (!vjp) {
        h = timer_begin("wb_combine", st);
        hipLaunchKernelGGL((wb_combine<K, false>), dim3(cgrid), dim3(256), 0, st, ca);
        timer_end(h);
        return hipGetLastError();
    }
    h = timer_begin("wb_combine_vjp", st);
    hipLaunchKernelGGL((wb_combine<K, true>), dim3(cgrid), dim3(256), 0, st, ca);
    timer_end(h);
    h = timer_begin("wb_rev", st);
    hipLaunchKernelGGL((wb_rev<I, K>), dim3(grid), dim3(blk), 0, st, a);
    timer_end(h);
    h = timer_begin("wb_fixup", st);
    hipLaunchKernelGGL(wb_fixup, dim3(grid), dim3(blk), 0, st, a);
    timer_end(h);
    return hipGetLastError();
}

template <int I>
hipError_t launch_wb_i(const WbArgs& a, bool vjp, bool sg, hipStream_t st) {
    switch (a.K) {
        case 1: return launch_wb_k<I, 1>(a, vjp, sg, st);
        case 2: return launch_wb_k<I, 2>(a, vjp, sg, st);
        case 3: return launch_wb_k<I, 3>(a, vjp, sg, st);
        case 4: return launch_wb_k<I, 4>(a, vjp, sg, st);
        case 5: return launch_wb_k<I, 5>(a, vjp, sg, st);
        case 6: return launch_wb_k<I, 6>(a, vjp, sg, st);
        case 7: return launch_wb_k<I, 7>(a, vjp, sg, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

hipError_t launch_wb(int integ, const WbArgs& a, bool vjp, bool summaries_given, hipStream_t st) {
    switch (integ) {
        case BOOLE: return launch_wb_i<BOOLE>(a, vjp, summaries_given, st);
        case SIMPSON: return launch_wb_i<SIMPSON>(a, vjp, summaries_given, st);
        case RIEMANN: return launch_wb_i<RIEMANN>(a, vjp, summaries_given, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace magnus
