// kernels_wbd8.hip -- single-pass basis sweep of the dense W-basis path at
// d = 8 under the Boole rule (the resolution controller's problems).
//
// The basis streams W_1 = Prefix(A), W_j = Prefix(A W_{j-1}) (j = 1..K,
// K = n - 1), their totals What_j = W_j[S-1] and the sums Z_0 = Sum(A),
// Z_l = Sum(A W_l) (l = 1..K) are what the GL fold (kernels_wbd.hip) needs;
// the level kernels build them one level at a time, each level a pass over
// HBM.  Here one wavefront walks the samples of two instances once and keeps
// every level's state in registers: per sample it forms the K products
// A_i W_j[i] in sequence (each level's prefix at sample i needs that level's
// product at i, src/magnus.hpp:241-251) and steps every level's prefix and
// sum.  A sample is read once; the levels are written only when a VJP needs
// them (LEVELS).
//
// Products on the matrix cores.  v_mfma_f64_16x16x4f64 multiplies a 16 x 4
// slice by a 4 x 16 slice (lane l holds A[l%16][l/16] and B[l/16][l%16];
// result register r of lane l is row 4r + l/16, column l%16).  With the two
// instances stacked, [A_a; A_b] (16 x 8) times [W_a | W_b] (8 x 16) takes two
// instructions, and the diagonal 8 x 8 blocks are A_a W_a and A_b W_b: each
// output entry pairs one instance's row with the same instance's column, so
// the instances never mix (the off-diagonal blocks are discarded).  Lane l
// owns column c = l%8 of instance l/8%2, rows g = l/16 and g + 4: exactly the
// two W entries it feeds as the B operand (k = g and k = g + 4) and the two
// product entries it gets back, so a level's result feeds the next level's
// product with no data movement between lanes.  Half of each instruction's
// work is the discarded blocks, so the sweep tops out at half the FP64 matrix
// peak; it reads 1 KB per sample pair where the level kernels move about five
// level-sized arrays per level.
//
// Rounding: the MFMA sums k in its own order and the Boole rule is evaluated
// as running partial sums (the same weights as boole.hpp:24-74, regrouped),
// so results differ from the reference's sequential order by rounding only.
// Each instance's results depend on its own data alone.
#include "kernels.hpp"
#include "rules.hpp"

namespace magnus {

namespace {

typedef double dbl4 __attribute__((ext_vector_type(4)));

constexpr int kM = 64;  // entries of a d = 8 sample

// Boole weights times dt: trapezoid, Simpson, 3/8 and Boole rules of a block
struct BooleW {
    double h2, t13, t43, e38, e98, b14, b64, b24, b28;
    __device__ explicit BooleW(double dt)
        : h2(dt / 2), t13(dt / 3), t43(4 * dt / 3), e38(3 * dt / 8), e98(9 * dt / 8), b14(14 * dt / 45),
          b64(64 * dt / 45), b24(24 * dt / 45), b28(28 * dt / 45) {}
};

// One scalar stream of a prefix level.  At a block start (u, v) = (base, f0);
// after position p of the block, the partial sums base + sum_{q<=p} w_{P,q} f_q
// of the positions P > p still to come.
struct Acc {
    double u, v, w;
};

// Position P (1..4) of a Boole block: returns the prefix at that position.
template <int P>
__device__ __forceinline__ double brule(Acc& s, double f, const BooleW& c) {
    if constexpr (P == 1) {
        const double b = s.u, f0 = s.v;
        const double r = fma(f, c.h2, fma(f0, c.h2, b));
        s.u = fma(f, c.t43, fma(f0, c.t13, b));
        s.v = fma(f, c.e98, fma(f0, c.e38, b));
        s.w = fma(f, c.b64, fma(f0, c.b14, b));
        return r;
    } else if constexpr (P == 2) {
        const double r = fma(f, c.t13, s.u);
        s.v = fma(f, c.e98, s.v);
        s.w = fma(f, c.b24, s.w);
        return r;
    } else if constexpr (P == 3) {
        const double r = fma(f, c.e38, s.v);
        s.w = fma(f, c.b64, s.w);
        return r;
    } else {
        const double r = fma(f, c.b14, s.w);
        s.u = r;
        s.v = f;
        return r;
    }
}

// the four values a lane reads of one sample: its A operand entries A[c][g],
// A[c][g+4] and its own entries A[g][c], A[g+4][c]
struct Samp {
    double o0, o1, y0, y1;
};

__device__ __forceinline__ Samp load_samp(const double* __restrict__ p, int c, int g) {
    Samp s;
    s.o0 = p[c * 8 + g];
    s.o1 = p[c * 8 + 4 + g];
    s.y0 = p[g * 8 + c];
    s.y1 = p[(g + 4) * 8 + c];
    return s;
}

template <int K, bool LEVELS>
struct Sweep {
    Acc st[K][2];      // level j+1's prefix state, entries (g, c) and (g + 4, c)
    double z[K + 1][2];  // Z_0..Z_K
    double* W;         // level output base of the instance (LEVELS), W_j at W + (j-1) S 64
    long long S;
    int c, g, inst;
    bool live;

    // sample i at block position P; cz its Boole sum weight
    template <int P>
    __device__ __forceinline__ void step(const Samp& s, long long i, double cz, const BooleW& bw) {
        double y0 = s.y0, y1 = s.y1;
        z[0][0] = fma(cz, y0, z[0][0]);
        z[0][1] = fma(cz, y1, z[0][1]);
        double w0 = brule<P>(st[0][0], y0, bw);
        double w1 = brule<P>(st[0][1], y1, bw);
        if constexpr (LEVELS) store(0, i, w0, w1);
#pragma unroll
        for (int j = 1; j <= K; ++j) {
            dbl4 d = __builtin_amdgcn_mfma_f64_16x16x4f64(s.o0, w0, dbl4{0.0, 0.0, 0.0, 0.0}, 0, 0, 0);
            d = __builtin_amdgcn_mfma_f64_16x16x4f64(s.o1, w1, d, 0, 0, 0);
            y0 = inst ? d[2] : d[0];
            y1 = inst ? d[3] : d[1];
            z[j][0] = fma(cz, y0, z[j][0]);
            z[j][1] = fma(cz, y1, z[j][1]);
            if (j < K) {
                w0 = brule<P>(st[j][0], y0, bw);
                w1 = brule<P>(st[j][1], y1, bw);
                if constexpr (LEVELS) store(j, i, w0, w1);
            }
        }
    }

    __device__ __forceinline__ void store(int j, long long i, double w0, double w1) {
        if (!live) return;
        double* p = W + ((long long)j * S + i) * kM;
        p[g * 8 + c] = w0;
        p[(g + 4) * 8 + c] = w1;
    }
};

// one wavefront per pair of instances (2w, 2w + 1); a missing partner's lanes
// read the live instance and store nothing
template <int K, bool LEVELS>
__global__ void __launch_bounds__(64) wbd8_sweep(LmArgs a) {
    const int lane = threadIdx.x;
    const int c = lane & 7, inst = (lane >> 3) & 1, g = lane >> 4;
    const long long b = 2LL * blockIdx.x + inst;
    const bool live = b < a.jobs;
    const long long bl = live ? b : b - 1;
    const long long S = a.S;
    const double* __restrict__ A = a.in + bl * S * kM;
    const BooleW bw(a.dt);

    Sweep<K, LEVELS> sw;
    sw.S = S;
    sw.c = c;
    sw.g = g;
    sw.inst = inst;
    sw.live = live;
    sw.W = LEVELS ? a.carry + bl * (long long)K * S * kM : nullptr;

    // sample 0: every prefix is 0 there, so the products of levels 2..K+1 are 0
    {
        const Samp s0 = load_samp(A, c, g);
        sw.z[0][0] = bw.b14 * s0.y0;
        sw.z[0][1] = bw.b14 * s0.y1;
        sw.st[0][0] = Acc{0.0, s0.y0, 0.0};
        sw.st[0][1] = Acc{0.0, s0.y1, 0.0};
#pragma unroll
        for (int j = 1; j <= K; ++j) sw.z[j][0] = sw.z[j][1] = 0.0;
#pragma unroll
        for (int j = 1; j < K; ++j) sw.st[j][0] = sw.st[j][1] = Acc{0.0, 0.0, 0.0};
        if constexpr (LEVELS) {
#pragma unroll
            for (int j = 0; j < K; ++j) sw.store(j, 0, 0.0, 0.0);
        }
    }
    const long long nb = (S - 1) / 4;
    Samp cur[4], nxt[4] = {};
#pragma unroll
    for (int p = 0; p < 4; ++p) cur[p] = load_samp(A + (1 + p) * (long long)kM, c, g);
    for (long long blk = 0; blk < nb; ++blk) {
        const long long i0 = 4 * blk;
        if (blk + 1 < nb) {  // the next block's samples are in flight during this block
#pragma unroll
            for (int p = 0; p < 4; ++p) nxt[p] = load_samp(A + (i0 + 5 + p) * (long long)kM, c, g);
        }
        sw.template step<1>(cur[0], i0 + 1, bw.b64, bw);
        sw.template step<2>(cur[1], i0 + 2, bw.b24, bw);
        sw.template step<3>(cur[2], i0 + 3, bw.b64, bw);
        sw.template step<4>(cur[3], i0 + 4, blk + 1 < nb ? bw.b28 : bw.b14, bw);
#pragma unroll
        for (int p = 0; p < 4; ++p) cur[p] = nxt[p];
    }
    if (!live) return;
    const long long M = kM, n = a.n;
    const int e0 = g * 8 + c, e1 = (g + 4) * 8 + c;
    a.omega1[b * M + e0] = sw.z[0][0];
    a.omega1[b * M + e1] = sw.z[0][1];
#pragma unroll
    for (int l = 1; l <= K; ++l) {
        a.partial[(b * n + l) * M + e0] = sw.z[l][0];
        a.partial[(b * n + l) * M + e1] = sw.z[l][1];
    }
#pragma unroll
    for (int j = 1; j <= K; ++j) {  // What_j = W_j[S-1]: the last block's position 4
        a.totals[(b * n + (j - 1)) * M + e0] = sw.st[j - 1][0].u;
        a.totals[(b * n + (j - 1)) * M + e1] = sw.st[j - 1][1].u;
    }
}

template <bool LEVELS>
hipError_t launch_k(const LmArgs& a, hipStream_t s) {
    const dim3 grid((unsigned)((a.jobs + 1) / 2)), blk(64);
    switch (a.n - 1) {
        case 1: hipLaunchKernelGGL((wbd8_sweep<1, LEVELS>), grid, blk, 0, s, a); break;
        case 2: hipLaunchKernelGGL((wbd8_sweep<2, LEVELS>), grid, blk, 0, s, a); break;
        case 3: hipLaunchKernelGGL((wbd8_sweep<3, LEVELS>), grid, blk, 0, s, a); break;
        case 4: hipLaunchKernelGGL((wbd8_sweep<4, LEVELS>), grid, blk, 0, s, a); break;
        case 5: hipLaunchKernelGGL((wbd8_sweep<5, LEVELS>), grid, blk, 0, s, a); break;
        case 6: hipLaunchKernelGGL((wbd8_sweep<6, LEVELS>), grid, blk, 0, s, a); break;
        case 7: hipLaunchKernelGGL((wbd8_sweep<7, LEVELS>), grid, blk, 0, s, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace

bool wbd8_sweep_ok(int d, int integ, int n) { return d == 8 && integ == BOOLE && n >= 2 && n - 1 <= kWbd8MaxK; }

hipError_t launch_wbd8_sweep(const LmArgs& a, hipStream_t s) {
    if (a.jobs == 0) return hipSuccess;
    if (a.d != 8 || a.S < 5 || (a.S - 1) % 4 != 0 || a.n < 2 || a.n - 1 > kWbd8MaxK) return hipErrorInvalidValue;
    // algorithmic work: the samples read once (and the K levels written when
    // kept); per sample K products of 2 d^3 flops plus a prefix step and a sum
    // of 8 flops per entry for each of the K + 1 levels
    const double K = a.n - 1, S = (double)a.S, B = (double)a.jobs;
    const double bytes = 8.0 * B * S * kM * (1.0 + (a.carry ? K : 0.0));
    const double flops = B * S * (K * (2.0 * 512.0 + 8.0 * kM) + 8.0 * kM);
    void* h = timer_begin("wbd8_sweep", s, bytes, flops);
    hipError_t e = a.carry ? launch_k<true>(a, s) : launch_k<false>(a, s);
    timer_end(h);
    return e;
}

}  // namespace magnus
