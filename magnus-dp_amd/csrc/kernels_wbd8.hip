// kernels_wbd8.hip -- single-pass basis sweep of the dense W-basis path at
// d = 8 under the Boole rule (the resolution controller's problems).
//
// The basis streams W_1 = Prefix(A), W_j = Prefix(A W_{j-1}) (j = 1..K,
// K = n - 1), their totals What_j = W_j[S-1] and the sums Z_0 = Sum(A),
// Z_l = Sum(A W_l) (l = 1..K) are what the GL fold (kernels_wbd.hip) needs;
// the level kernels build them one level at a time, each level a pass over
// HBM.  Here one workgroup of K wavefronts walks the samples of two instances
// once, one wavefront per level: level j's prefix at sample i needs that
// level's product A_i W_{j-1}[i] (src/magnus.hpp:241-251), so the waves run as
// a pipeline of Boole blocks -- at step t wave w takes block t - w, its input
// block formed by wave w - 1 in the step before and handed over in LDS -- and
// every level's state stays in its wave's registers.  A sample comes from HBM
// once (the later waves find it in the caches); the levels are written only
// when a VJP needs them (LEVELS).
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
// product entries it gets back, so a level block handed from one wave to the
// next goes from each lane to the same lane.  Half of each instruction's
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

// ---------------------------------------------------------------------------
// Sample ring.  The samples of a workgroup's instance pair reach LDS through
// direct global->LDS loads (global_load_lds_dwordx4: lane l's 16 bytes land
// at the instruction's LDS base + 16 l) issued by one extra wavefront, the
// loader, kRingAhead Boole blocks ahead of the first level's wave; a block
// (4 samples of both instances, 4 KB) is four instructions, one per sample.
// Within a sample's 1 KB the 16-byte chunk (A[r][2h], A[r][2h+1]) of instance
// inst sits at chunk position chunk_pos(inst, r, h), a GF(2)-linear bijection
// of the bits (inst, r, h) chosen so that the waves' reads -- A[c][g] and
// A[c][g + 4] (the A operand), A[g][c] and A[g + 4][c] (own entries) -- hit
// distinct bank pairs in every half-wave (64 banks) and in every 16-lane
// quarter (32 banks) except for the A operand's read, which is then two-way
// (its 16 lanes read one half of 8 chunks); the loader lane that fills
// position l reads chunk chunk_src(l).  The loader is the
// only wave with vector memory loads in flight, so it can wait for a block
// with a vmcnt that counts only those loads (the blocks issued after it), and
// the workgroup barrier that follows publishes the block to every wave.  The
// waves of the pipeline read their blocks from the ring up to K - 1 steps
// later, so the ring holds K + kRingAhead blocks.
// ---------------------------------------------------------------------------
constexpr int kRingAhead = 6;

// the bits (inst, r0, r1, r2, h0, h1) of a chunk, low first, map to these
// position masks (XORed); chunk_src applies the inverse map
__device__ __forceinline__ int chunk_pos(int inst, int r, int h) {
    const int x = inst | (r << 1) | (h << 4);
    return (-(x & 1) & 15) ^ (-((x >> 1) & 1) & 42) ^ (-((x >> 2) & 1) & 54) ^ (-((x >> 3) & 1) & 24) ^
           (-((x >> 4) & 1) & 38) ^ (-((x >> 5) & 1) & 36);
}
__device__ __forceinline__ int chunk_src(int pos) {
    return (-(pos & 1) & 35) ^ (-((pos >> 1) & 1) & 48) ^ (-((pos >> 2) & 1) & 14) ^ (-((pos >> 3) & 1) & 28) ^
           (-((pos >> 4) & 1) & 20) ^ (-((pos >> 5) & 1) & 46);
}
// the LDS word of entry (r, col) of instance inst within a ring sample
__device__ __forceinline__ int ring_word(int inst, int r, int col) {
    return 2 * chunk_pos(inst, r, col >> 1) + (col & 1);
}

template <int N>
__device__ __forceinline__ void wait_vm() {  // vmcnt <= N (expcnt, lgkmcnt unconstrained)
    static_assert(N < 64, "vmcnt has 6 bits");
    __builtin_amdgcn_s_waitcnt((N & 15) | 0x0F70 | ((N >> 4) << 14));
}

// vmcnt <= 4 * blocks: the loader's loads of the `blocks` later blocks may still be in flight
__device__ __forceinline__ void wait_blocks(int blocks) {
    switch (blocks) {
        case 0: wait_vm<0>(); break;
        case 1: wait_vm<4>(); break;
        case 2: wait_vm<8>(); break;
        case 3: wait_vm<12>(); break;
        case 4: wait_vm<16>(); break;
        case 5: wait_vm<20>(); break;
        default: wait_vm<24>(); break;
    }
    static_assert(kRingAhead - 1 <= 6, "wait_blocks covers kRingAhead - 1 later blocks");
}

// Workgroup barrier of the pipelines: this wave's LDS writes complete, then
// s_barrier.  __syncthreads() would also wait for every vector memory access
// in flight (its release fence), which would drain the loader's ring loads
// at every step; the "memory" clobber keeps the compiler from moving memory
// accesses across the barrier.
__device__ __forceinline__ void step_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// one ring slot: [4 samples][64 chunks of both instances]
using RingSlot = double[4][2 * kM];

// the loader's loads of Boole block blk (samples 4 blk + 1 .. 4 blk + 4) into
// `slot`; src: this lane's chunk in sample 0 of its instance (every lane of
// the loader calls)
__device__ __forceinline__ void ring_issue(RingSlot& slot, const double* src, long long blk) {
#pragma unroll
    for (int p = 0; p < 4; ++p)
        __builtin_amdgcn_global_load_lds(src + (4 * blk + 1 + p) * (long long)kM,
                                         (__attribute__((address_space(3))) void*)&slot[p][0], 16, 0, 0);
}

// Per-instance-pair context shared by the waves of a workgroup.
struct Pair {
    long long b, bl, S, nb;  // this lane's instance (bl: the live one it reads), samples, blocks
    int c, g, inst;          // lane = column c of instance inst, rows g and g + 4
    bool live;               // the lane's instance exists (an odd batch leaves a partner missing)
};

// The loader of a pipeline: ring index r holds block blk_of(r); it keeps
// kRingAhead - 1 blocks in flight ahead of wave 0 and, at step t, waits for
// ring index t (wave 0's block) before the step's first barrier.
template <int R, typename BlkOf>
__device__ __forceinline__ void ring_loader(RingSlot* ring, const double* A0, const double* A1, long long nb,
                                            long long steps, int lane, BlkOf blk_of) {
    const int x = chunk_src(lane);  // (inst, r, h) of the chunk this lane places
    const double* src = ((x & 1) ? A1 : A0) + ((x >> 1) & 7) * 8 + 2 * (x >> 4);
    for (long long r = 0; r < kRingAhead - 1 && r < nb; ++r) ring_issue(ring[r % R], src, blk_of(r));
    for (long long t = 0; t < steps; ++t) {
        const long long r = t + kRingAhead - 1;
        if (r < nb) ring_issue(ring[r % R], src, blk_of(r));
        if (t < nb) wait_blocks((int)((r < nb ? r : nb - 1) - t));
        step_barrier();
        step_barrier();
    }
}

// ---------------------------------------------------------------------------
// Forward sweep on v_mfma_f64_4x4x4f64.  That instruction runs four
// independent 4 x 4 x 4 blocks: block b's A[i][k] is lane 16k + 4b + i, its
// B[k][j] lane 16k + 4b + j and its D[i][j] lane 16i + 4b + j.  An 8 x 8
// product Y = A W is four output blocks (I, J) over two halves kappa of the
// contraction: block b = 2I + J of instruction kappa multiplies rows
// 2i + I of A, columns 2k + kappa, by rows 2k + kappa of W, columns 4J + j.
// Lane L = 16k + 8I + 4J + j then owns Y[2k + I][4J + j]; the B operands
// of the two instructions are rows 2k and 2k + 1 of column 4J + j, one of
// them the lane's own entry, the other owned by lane L ^ 8 (a DPP rotation
// by 8 within each row of 16 lanes fetches it), and the A operands of lane
// L (i = L % 4) are one 16-byte chunk, A[2i + I][2k .. 2k + 1].  No lane
// computes anything discarded, so one instance fills a wavefront and a
// product is two instructions.
//
// One workgroup per instance: K level wavefronts and a loader.  The ring
// holds the instance's samples (a block of 4 samples is 2 KB, two loads of
// 1 KB: lane l places chunk position l % 32 of sample l / 32); within a
// sample, chunk (r, h) = (A[r][2h], A[r][2h + 1]) sits at position
// 4r + (h ^ ((r >> 1) & 3)), so that the A operand chunks of every 8 lanes
// and the own entries of every 16 lanes fall in distinct banks.  The waves
// hand their level blocks on in LDS, indexed by lane, in two buffers used in
// alternate steps, so one barrier per step orders every hand-over.
// ---------------------------------------------------------------------------
constexpr int kAhead4 = 6;  // blocks the loader keeps in flight ahead of wave 0

__device__ __forceinline__ int chunk4_pos(int r, int h) { return 4 * r + (h ^ ((r >> 1) & 3)); }
__device__ __forceinline__ int ring4_word(int r, int col) { return 2 * chunk4_pos(r, col >> 1) + (col & 1); }

// one ring slot: [4 samples][64 entries], 16-byte aligned for the chunk reads
struct alignas(16) Ring4Slot {
    double v[4][kM];
};

// vmcnt <= 2 * blocks: the loader's loads of the `blocks` later blocks may still be in flight
__device__ __forceinline__ void wait_blocks4(int blocks) {
    switch (blocks) {
        case 0: wait_vm<0>(); break;
        case 1: wait_vm<2>(); break;
        case 2: wait_vm<4>(); break;
        case 3: wait_vm<6>(); break;
        case 4: wait_vm<8>(); break;
        default: wait_vm<10>(); break;
    }
    static_assert(kAhead4 - 1 <= 5, "wait_blocks4 covers kAhead4 - 1 later blocks");
}

// the loader wave of a one-instance pipeline: at step t it issues block
// t + kAhead4 - 1 and waits for block t + 1 (wave 0's next block) before the
// step's barrier; the prologue issues the first blocks and waits for block 0
// before the barrier that opens the pipeline.  Ring index r holds block
// blk_of(r).
template <int R, typename BlkOf>
__device__ __forceinline__ void ring4_loader(Ring4Slot* ring, const double* A, long long nb, long long steps,
                                             int lane, BlkOf blk_of) {
    const int pos = lane & 31, r = pos >> 2, h = (pos & 3) ^ ((r >> 1) & 3);
    const double* src = A + (1 + (lane >> 5)) * (long long)kM + r * 8 + 2 * h;  // sample 4 blk + 1 + lane / 32
    auto issue = [&](long long ri) {
        const long long blk = blk_of(ri);
        Ring4Slot& slot = ring[ri % R];
#pragma unroll
        for (int q = 0; q < 2; ++q)
            __builtin_amdgcn_global_load_lds(src + (4 * blk + 2 * q) * (long long)kM,
                                             (__attribute__((address_space(3))) void*)&slot.v[2 * q][0], 16, 0, 0);
    };
    const long long pre = nb < kAhead4 - 1 ? nb : kAhead4 - 1;
    for (long long ri = 0; ri < pre; ++ri) issue(ri);
    wait_blocks4((int)(pre - 1));
    step_barrier();
    for (long long t = 0; t < steps; ++t) {
        const long long ri = t + kAhead4 - 1;
        if (ri < nb) issue(ri);
        if (t + 1 < nb) wait_blocks4((int)((ri < nb ? ri : nb - 1) - (t + 1)));
        step_barrier();
    }
}

// The B operand pair (W[2k][c], W[2k + 1][c]) of lane L from its own entry
// W[2k + I][c]: a DPP row rotation by 8 brings lane L ^ 8's entry, and the
// bank mask keeps the own value in the lanes (I = 0 for the first, I = 1 for
// the second) that own it.
__device__ __forceinline__ int dpp_ror8(int old, int src, int banks) {
    return banks == 0xC ? __builtin_amdgcn_update_dpp(old, src, 0x128, 0xF, 0xC, false)
                        : __builtin_amdgcn_update_dpp(old, src, 0x128, 0xF, 0x3, false);
}
__device__ __forceinline__ double2 bpair_dpp(double own) {
    const unsigned long long v = __builtin_bit_cast(unsigned long long, own);
    const int lo = (int)(unsigned)v, hi = (int)(unsigned)(v >> 32);
    auto join = [](int l, int h) {
        return __builtin_bit_cast(double, ((unsigned long long)(unsigned)h << 32) | (unsigned)l);
    };
    return double2{join(dpp_ror8(lo, lo, 0xC), dpp_ror8(hi, hi, 0xC)),   // rows 2k: I = 1 lanes take L ^ 8
                   join(dpp_ror8(lo, lo, 0x3), dpp_ror8(hi, hi, 0x3))};  // rows 2k + 1: I = 0 lanes take L ^ 8
}

// Y[2k + I][4J + j] of Y = A W from the A chunk a (A[2i + I][2k], A[2i + I][2k + 1])
// and the B pair wp (W[2k][4J + j], W[2k + 1][4J + j])
__device__ __forceinline__ double prod4(double2 a, double2 wp) {
    const double d = __builtin_amdgcn_mfma_f64_4x4x4f64(a.x, wp.x, 0.0, 0, 0, 0);
    return __builtin_amdgcn_mfma_f64_4x4x4f64(a.y, wp.y, d, 0, 0, 0);
}

// Hand-over layout of a level block: per sample, lane L's own entry sits at
// word 2 q + I with q = L without its bit 3, so that lane L reads its B pair
// (the entries of lanes L & ~8 and L | 8) as one 16-byte chunk.
__device__ __forceinline__ int hand_word(int lane) { return 2 * ((lane & 7) | ((lane >> 4) << 3)) + ((lane >> 3) & 1); }

// One wave of the forward pipeline, level j = w + 1.  FIRST: the wave of level
// 1's prefix (W_1 = Prefix(A), Z_0 = Sum(A)), which also forms Y_2 = A W_1;
// otherwise it forms Y_{j+1} = A W_j from the block of W_j the previous wave
// handed over.  MAKES: it steps level j + 1's prefix (j < K) and hands that
// block on.  The role is a template parameter and the step indices are
// wave-uniform (scalar registers), so that a step is straight-line code: the
// four samples' reads and products are independent and overlap; only the
// prefix recurrences run sample after sample.
template <int K, bool LEVELS, bool FIRST, bool MAKES>
__device__ __forceinline__ void sweep4_wave(const LmArgs& a, int b, int w, int lane, const Ring4Slot* ring,
                                            double (*hand)[K > 1 ? K - 1 : 1][4][64]) {
    constexpr int R = K + kAhead4 - 1;
    const long long S = a.S;
    const int nb = (int)((S - 1) / 4);
    const int j = w + 1;
    const int k = lane >> 4, I = (lane >> 3) & 1, J = (lane >> 2) & 1, i = lane & 3;
    const int own = (2 * k + I) * 8 + 4 * J + i;  // own entry (2k + I, 4J + j), j = lane % 4
    const int ra = 2 * chunk4_pos(2 * i + I, k);  // the A operand chunk
    const int ro = ring4_word(2 * k + I, 4 * J + i);
    const int hw = hand_word(lane), hp = hw & ~1;  // own word and B pair chunk in a hand-over
    const double* __restrict__ A = a.in + (long long)b * S * kM;
    double* __restrict__ W = LEVELS ? a.carry + (long long)b * K * S * kM : nullptr;
    const BooleW bw(a.dt);
    Acc s1, sn = Acc{0.0, 0.0, 0.0};  // prefix states of W_1 (FIRST) and W_{j+1} (MAKES)
    double z1 = 0.0, zn = 0.0;        // Z_0 (FIRST) and Z_j

    // sample 0: every prefix is 0 there, so the products of levels 2..K+1 are 0
    if constexpr (FIRST) {
        const double y0 = A[own];
        s1 = Acc{0.0, y0, 0.0};
        z1 = bw.b14 * y0;
    }
    if constexpr (LEVELS) {
        if (FIRST) W[own] = 0.0;
        if (MAKES) W[(long long)j * S * kM + own] = 0.0;
    }
    step_barrier();  // block 0 is in the ring
    int slot = 0;    // ring slot of this wave's next block
    for (int t = 0; t < nb + K - 1; ++t) {
        const int blk = t - w;
        if (blk >= 0 && blk < nb) {
            const Ring4Slot& sl = ring[slot];
            slot = slot + 1 == R ? 0 : slot + 1;
            double2 av[4], wp[4];
            double y[4];
#pragma unroll
            for (int p = 0; p < 4; ++p) av[p] = *reinterpret_cast<const double2*>(&sl.v[p][ra]);
            const double c4 = blk + 1 < nb ? bw.b28 : bw.b14;
            const double cz[4] = {bw.b64, bw.b24, bw.b64, c4};
            if constexpr (FIRST) {  // W_1 = Prefix(A) and Z_0 over the block
                double ao[4], wv[4];
#pragma unroll
                for (int p = 0; p < 4; ++p) {
                    ao[p] = sl.v[p][ro];
                    z1 = fma(cz[p], ao[p], z1);
                }
                wv[0] = brule<1>(s1, ao[0], bw);
                wv[1] = brule<2>(s1, ao[1], bw);
                wv[2] = brule<3>(s1, ao[2], bw);
                wv[3] = brule<4>(s1, ao[3], bw);
#pragma unroll
                for (int p = 0; p < 4; ++p) wp[p] = bpair_dpp(wv[p]);
                if constexpr (LEVELS) {
                    double* Wb = W + (4LL * blk + 1) * kM + own;
#pragma unroll
                    for (int p = 0; p < 4; ++p) Wb[p * kM] = wv[p];
                }
            } else {  // the previous wave formed this block in the previous step
#pragma unroll
                for (int p = 0; p < 4; ++p) wp[p] = *reinterpret_cast<const double2*>(&hand[(t - 1) & 1][w - 1][p][hp]);
            }
#pragma unroll
            for (int p = 0; p < 4; ++p) y[p] = prod4(av[p], wp[p]);  // Y_{j+1} = A W_j
#pragma unroll
            for (int p = 0; p < 4; ++p) zn = fma(cz[p], y[p], zn);
            if constexpr (MAKES) {  // W_{j+1} = Prefix(Y_{j+1}) over the block, handed on
                double v[4];
                v[0] = brule<1>(sn, y[0], bw);
                v[1] = brule<2>(sn, y[1], bw);
                v[2] = brule<3>(sn, y[2], bw);
                v[3] = brule<4>(sn, y[3], bw);
#pragma unroll
                for (int p = 0; p < 4; ++p) hand[t & 1][w][p][hw] = v[p];
                if constexpr (LEVELS) {  // the level this wave forms, kept for the reverse sweep
                    double* Wb = W + ((long long)j * S + 4LL * blk + 1) * kM + own;
#pragma unroll
                    for (int p = 0; p < 4; ++p) Wb[p * kM] = v[p];
                }
            }
        }
        step_barrier();  // this step's hand-overs and the loader's next block are visible
    }
    const long long M = kM, n = a.n;
    if constexpr (FIRST) {
        a.omega1[b * M + own] = z1;
        a.totals[(b * n + 0) * M + own] = s1.u;  // What_1 = W_1[S-1]
    }
    a.partial[(b * n + j) * M + own] = zn;
    if constexpr (MAKES) a.totals[(b * n + j) * M + own] = sn.u;  // What_{j+1}: the last block's position 4
}

// one workgroup per instance: K level wavefronts and the loader (wave K)
template <int K, bool LEVELS>
__global__ void __launch_bounds__(64 * (K + 1)) __attribute__((amdgpu_waves_per_eu(6))) wbd8_sweep(LmArgs a) {
    constexpr int R = K + kAhead4 - 1;  // ring slots
    __shared__ Ring4Slot ring[R];
    // level blocks handed from wave w to wave w + 1: [step parity][w][4 samples][lane]
    __shared__ alignas(16) double hand[2][K > 1 ? K - 1 : 1][4][64];
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int b = blockIdx.x;
    const long long nb = (a.S - 1) / 4;
    if (w == K) {
        ring4_loader<R>(ring, a.in + (long long)b * a.S * kM, nb, nb + K - 1, lane, [](long long r) { return r; });
    } else if (w == 0) {
        sweep4_wave<K, LEVELS, true, (K > 1)>(a, b, w, lane, ring, hand);
    } else if (w < K - 1) {
        sweep4_wave<K, LEVELS, false, true>(a, b, w, lane, ring, hand);
    } else {
        sweep4_wave<K, LEVELS, false, false>(a, b, w, lane, ring, hand);
    }
}

template <bool LEVELS>
hipError_t launch_k(const LmArgs& a, hipStream_t s) {
    const dim3 grid((unsigned)a.jobs);
    switch (a.n - 1) {
        case 1: hipLaunchKernelGGL((wbd8_sweep<1, LEVELS>), grid, dim3(128), 0, s, a); break;
        case 2: hipLaunchKernelGGL((wbd8_sweep<2, LEVELS>), grid, dim3(192), 0, s, a); break;
        case 3: hipLaunchKernelGGL((wbd8_sweep<3, LEVELS>), grid, dim3(256), 0, s, a); break;
        case 4: hipLaunchKernelGGL((wbd8_sweep<4, LEVELS>), grid, dim3(320), 0, s, a); break;
        case 5: hipLaunchKernelGGL((wbd8_sweep<5, LEVELS>), grid, dim3(384), 0, s, a); break;
        case 6: hipLaunchKernelGGL((wbd8_sweep<6, LEVELS>), grid, dim3(448), 0, s, a); break;
        case 7: hipLaunchKernelGGL((wbd8_sweep<7, LEVELS>), grid, dim3(512), 0, s, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Reverse sweep.  With cotangents Zbar_l of the sums and What_bar_j of the
// totals (the GL fold's adjoint), the adjoint of the sweep is, per sample i,
//   Ybar_{K+1}[i] = c_i Zbar_K,
//   Wbar_j[i]     = A_i^T Ybar_{j+1}[i]  (+ What_bar_j at i = S-1),
//   Ybar_j        = Prefix^T(Wbar_j) + c Zbar_{j-1},            j = K..1,
//   dA_i          = Ybar_1[i] + sum_j Ybar_{j+1}[i] W_j[i]^T,
// the reverse of src/grad.hpp:311-348 with the basis in place of the nodes
// (c_i the Boole sum weights).  Prefix^T runs backwards through each block
// (boole.hpp:108-151): a block [s, s+4] needs its four Wbar values, the sum
// of the block above's values (which reaches its top sample through the
// block base) and that block's pending contribution to its bottom sample, so
// a level carries two values per entry from block to block.
//
// One wavefront per level again, level K in wave 0, plus the sample ring's
// loader: the waves walk the blocks from the last down as a pipeline (the
// ring is filled in that order), each handing the next wave its block of
// Ybar_j (every entry, so that the next wave can read it transposed for the
// product Ybar W^T, whose contraction runs over Ybar's columns) and the dA
// sum so far; the wave of level 1 adds Ybar_1 and writes dA.  The levels are
// added to dA in a fixed order (K first), so results are reproducible.
// ---------------------------------------------------------------------------

// the Boole sum weight of sample i (sum_vjp_coef, boole.hpp:153-199)
__device__ __forceinline__ double boole_c(long long i, long long S, const BooleW& c) {
    if (i == 0 || i == S - 1) return c.b14;
    const int r = (int)(i & 3);
    return r == 2 ? c.b24 : (r == 0 ? c.b28 : c.b64);
}

__device__ __forceinline__ double pick0(const dbl4& d, int inst) { return inst ? d[2] : d[0]; }

// LDS word of entry (r, col) of instance inst in a handed Ybar block: each
// row's 16 words hold both instances, the instance half chosen by the row's
// parity and the column XOR-swizzled by the row, so that the own entries
// (g, c), (g + 4, c) and the transposed (c, g), (c, g + 4) hit distinct bank
// pairs in every 16-lane quarter (32 banks) and half-wave (64 banks)
__device__ __forceinline__ int ybar_word(int inst, int r, int col) { return 16 * r + 8 * ((r ^ inst) & 1) + (col ^ r); }
__device__ __forceinline__ double pick1(const dbl4& d, int inst) { return inst ? d[3] : d[1]; }

// One wave of the reverse pipeline, level j = K - w.  TOP: level K, whose
// input Ybar_{K+1} = c Zbar_K needs no hand-over; LAST: level 1, which adds
// Ybar_1 and writes dA (the others hand Ybar_j and the dA sum on).
template <int K, bool TOP, bool LAST>
__device__ __forceinline__ void rev_wave(const Wbd8RevArgs& a, const Pair& q, int w, int lane, const RingSlot* ring,
                                         double (*ybh)[4][2 * kM], double (*dah)[4][2][64]) {
    constexpr int R = K + kRingAhead;
    const long long S = q.S, nb = q.nb, n = a.n, bl = q.bl;
    const int j = K - w;
    const int c = q.c, g = q.g, inst = q.inst;
    const int o0 = g * 8 + c, o1 = (g + 4) * 8 + c;  // own entries (g, c), (g + 4, c)
    const int t0 = c * 8 + g, t1 = c * 8 + g + 4;    // transposed entries (c, g), (c, g + 4)
    const int ro0 = ring_word(inst, g, c), ro1 = ring_word(inst, g + 4, c);
    // words of the own and the transposed entries in a handed Ybar block
    const int yo0 = ybar_word(inst, g, c), yo1 = ybar_word(inst, g + 4, c);
    const int yt0 = ybar_word(inst, c, g), yt1 = ybar_word(inst, c, g + 4);
    const double* __restrict__ Wj = a.W + (bl * (long long)K + (j - 1)) * S * kM;
    double* __restrict__ dA = a.dA + bl * S * kM;
    const BooleW bw(a.dt);

    // Zbar_{j-1} (own entries) and, for level K, Zbar_K (own and transposed)
    const double* zr = a.zbar + bl * n * kM;
    const double zb0 = zr[(j - 1) * kM + o0], zb1 = zr[(j - 1) * kM + o1];
    double zk0 = 0.0, zk1 = 0.0, zkt0 = 0.0, zkt1 = 0.0;
    if constexpr (TOP) {
        zk0 = zr[K * kM + o0];
        zk1 = zr[K * kM + o1];
        zkt0 = zr[K * kM + t0];
        zkt1 = zr[K * kM + t1];
    }
    double up0 = 0.0, up1 = 0.0, pend0 = 0.0, pend1 = 0.0;

    // W_j's transposed entries of a block, two blocks in flight ahead of the one in
    // use.  The three buffers change roles by renaming (the active steps are
    // unrolled by three), never by copying: a copy of a register whose load is in
    // flight would wait for that load.
    auto loadw = [&](double (*y)[2], long long blk) {
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            const long long o = (4 * blk + 1 + p) * (long long)kM;
            y[p][0] = Wj[o + t0];
            y[p][1] = Wj[o + t1];
        }
    };
    double wa[4][2] = {}, wbuf[4][2] = {}, wc[4][2] = {};
    loadw(wa, nb - 1);
    if (nb > 1) loadw(wbuf, nb - 2);
    double yo[4][2] = {}, yt[4][2] = {}, part[4][2] = {};

    // one pipeline step; r = ring index of this wave's block (or out of range:
    // the wave only takes part in the barriers), wt: W_j entries of the block,
    // wpf: the buffer for the block two below
    auto step = [&](long long r, double (*wt)[2], double (*wpf)[2]) {
        const long long blk = nb - 1 - r;
        const bool act = r >= 0 && blk >= 0;
        if (!TOP && act) {  // the previous wave handed this block over in the previous step
#pragma unroll
            for (int p = 0; p < 4; ++p) {
                yo[p][0] = ybh[w - 1][p][yo0];
                yo[p][1] = ybh[w - 1][p][yo1];
                yt[p][0] = ybh[w - 1][p][yt0];
                yt[p][1] = ybh[w - 1][p][yt1];
                part[p][0] = dah[w - 1][p][0][lane];
                part[p][1] = dah[w - 1][p][1][lane];
            }
        }
        step_barrier();  // the block is in the ring; every read of the hand blocks precedes their rewrite
        if (act) {
            const long long s = 4 * blk;
            if (blk > 1) loadw(wpf, blk - 2);
            double cz[4];
#pragma unroll
            for (int p = 0; p < 4; ++p) cz[p] = boole_c(s + 1 + p, S, bw);
            if constexpr (TOP) {  // Ybar_{K+1} = c Zbar_K
#pragma unroll
                for (int p = 0; p < 4; ++p) {
                    yo[p][0] = cz[p] * zk0;
                    yo[p][1] = cz[p] * zk1;
                    yt[p][0] = cz[p] * zkt0;
                    yt[p][1] = cz[p] * zkt1;
                }
            }
            const RingSlot& sl = ring[r % R];
            double at0[4], at1[4];  // the A^T operand: A's own entries
#pragma unroll
            for (int p = 0; p < 4; ++p) {
                at0[p] = sl[p][ro0];
                at1[p] = sl[p][ro1];
            }
            double wb[4][2];
            dbl4 acc[4];
#pragma unroll
            for (int p = 0; p < 4; ++p) {
                // Wbar_j = A^T Ybar_{j+1}
                dbl4 d = __builtin_amdgcn_mfma_f64_16x16x4f64(at0[p], yo[p][0], dbl4{0.0, 0.0, 0.0, 0.0}, 0, 0, 0);
                d = __builtin_amdgcn_mfma_f64_16x16x4f64(at1[p], yo[p][1], d, 0, 0, 0);
                wb[p][0] = pick0(d, inst);
                wb[p][1] = pick1(d, inst);
                // dA so far + Ybar_{j+1} W_j^T
                if constexpr (TOP) acc[p] = dbl4{0.0, 0.0, 0.0, 0.0};
                else acc[p] = inst ? dbl4{0.0, 0.0, part[p][0], part[p][1]} : dbl4{part[p][0], part[p][1], 0.0, 0.0};
                acc[p] = __builtin_amdgcn_mfma_f64_16x16x4f64(yt[p][0], wt[p][0], acc[p], 0, 0, 0);
                acc[p] = __builtin_amdgcn_mfma_f64_16x16x4f64(yt[p][1], wt[p][1], acc[p], 0, 0, 0);
            }
            if (blk == nb - 1) {  // the total's cotangent enters at the last sample
                const double* wr = a.wbar + (bl * n + (j - 1)) * kM;
                wb[3][0] += wr[o0];
                wb[3][1] += wr[o1];
            }
            // Ybar_j = Prefix^T(Wbar_j) + c Zbar_{j-1} over the block
            double yn[4][2];
#pragma unroll
            for (int qq = 0; qq < 2; ++qq) {
                double& up = qq ? up1 : up0;
                double& pend = qq ? pend1 : pend0;
                const double zq = qq ? zb1 : zb0;
                const double g1 = wb[0][qq], g2 = wb[1][qq], g3 = wb[2][qq];
                const double g4 = wb[3][qq] + up;
                const double f4 = fma(bw.b14, g4, pend);
                const double f3 = fma(bw.e38, g3, bw.b64 * g4);
                const double f2 = fma(bw.t13, g2, fma(bw.e98, g3, bw.b24 * g4));
                const double f1 = fma(bw.h2, g1, fma(bw.t43, g2, fma(bw.e98, g3, bw.b64 * g4)));
                up = g1 + g2 + g3 + g4;
                pend = fma(bw.h2, g1, fma(bw.t13, g2, fma(bw.e38, g3, bw.b14 * g4)));
                yn[0][qq] = fma(cz[0], zq, f1);
                yn[1][qq] = fma(cz[1], zq, f2);
                yn[2][qq] = fma(cz[2], zq, f3);
                yn[3][qq] = fma(cz[3], zq, f4);
            }
            if constexpr (!LAST) {
#pragma unroll
                for (int p = 0; p < 4; ++p) {
                    ybh[w][p][yo0] = yn[p][0];
                    ybh[w][p][yo1] = yn[p][1];
                    dah[w][p][0][lane] = pick0(acc[p], inst);
                    dah[w][p][1][lane] = pick1(acc[p], inst);
                }
            } else if (q.live) {  // dA = Ybar_1 + sum_j Ybar_{j+1} W_j^T
#pragma unroll
                for (int p = 0; p < 4; ++p) {
                    double* o = dA + (s + 1 + p) * kM;
                    o[o0] = pick0(acc[p], inst) + yn[p][0];
                    o[o1] = pick1(acc[p], inst) + yn[p][1];
                }
            }
        }
        step_barrier();  // the blocks handed over in this step are visible to the next
    };
    // steps before this wave's first block, its nb blocks (buffers rotating by
    // three), then the steps after its last block
    for (long long t = 0; t < w; ++t) step(-1, wa, wc);
    long long r = 0;
    for (; r + 3 <= nb; r += 3) {
        step(r, wa, wc);
        step(r + 1, wbuf, wa);
        step(r + 2, wc, wbuf);
    }
    if (r < nb) step(r, wa, wc);
    if (r + 1 < nb) step(r + 1, wbuf, wa);
    for (long long t = w + nb; t < nb + K - 1; ++t) step(-1, wa, wc);
    // sample 0: the prefixes are constant there (W_j[0] = 0), so dA_0 = Ybar_1[0]
    // = level 1's pending from the lowest block + c_0 Zbar_0
    if (LAST && q.live) {
        dA[o0] = fma(bw.b14, zb0, pend0);
        dA[o1] = fma(bw.b14, zb1, pend1);
    }
}

template <int K>
__global__ void __launch_bounds__(64 * (K + 1)) wbd8_rev(Wbd8RevArgs a) {
    constexpr int H = K > 1 ? K - 1 : 1;
    constexpr int R = K + kRingAhead;     // sample ring slots (the loader is wave K)
    __shared__ RingSlot ring[R];
    __shared__ double ybh[H][4][2 * kM];  // Ybar_j blocks handed from wave w to wave w + 1 (ybar_word)
    __shared__ double dah[H][4][2][64];   // dA sums so far, by lane
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    Pair q;
    q.c = lane & 7;
    q.inst = (lane >> 3) & 1;
    q.g = lane >> 4;
    q.b = 2LL * blockIdx.x + q.inst;
    q.live = q.b < a.B;
    q.bl = q.live ? q.b : q.b - 1;
    q.S = a.S;
    q.nb = (a.S - 1) / 4;
    if (w == K) {  // ring index r holds block nb - 1 - r (the blocks are walked top-down)
        const long long b0 = 2LL * blockIdx.x, b1 = b0 + 1 < a.B ? b0 + 1 : b0;
        const long long nb = q.nb;
        ring_loader<R>(ring, a.in + b0 * q.S * kM, a.in + b1 * q.S * kM, nb, nb + K - 1, lane,
                       [nb](long long r) { return nb - 1 - r; });
    } else if (K == 1) {
        rev_wave<K, true, true>(a, q, w, lane, ring, ybh, dah);
    } else if (w == 0) {
        rev_wave<K, true, false>(a, q, w, lane, ring, ybh, dah);
    } else if (w < K - 1) {
        rev_wave<K, false, false>(a, q, w, lane, ring, ybh, dah);
    } else {
        rev_wave<K, false, true>(a, q, w, lane, ring, ybh, dah);
    }
}

template <int K>
hipError_t launch_rev_k(const Wbd8RevArgs& a, hipStream_t s) {
    hipLaunchKernelGGL((wbd8_rev<K>), dim3((unsigned)((a.B + 1) / 2)), dim3(64 * (K + 1)), 0, s, a);
    return hipGetLastError();
}

}  // namespace

bool wbd8_sweep_ok(int d, int integ, int n) { return d == 8 && integ == BOOLE && n >= 2 && n - 1 <= kWbd8MaxK; }

hipError_t launch_wbd8_sweep(const LmArgs& a, hipStream_t s) {
    if (a.jobs == 0) return hipSuccess;
    if (a.d != 8 || a.S < 5 || (a.S - 1) % 4 != 0 || a.S > (1LL << 30) || a.jobs > (1LL << 30) || a.n < 2 ||
        a.n - 1 > kWbd8MaxK)
        return hipErrorInvalidValue;
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

namespace magnus {

hipError_t launch_wbd8_rev(const Wbd8RevArgs& a, hipStream_t s) {
    if (a.B == 0) return hipSuccess;
    if (a.S < 5 || (a.S - 1) % 4 != 0 || a.n < 2 || a.n - 1 > kWbd8MaxK) return hipErrorInvalidValue;
    // algorithmic work: the samples and the K levels read once, dA written once;
    // per sample and level two products (A^T Ybar and Ybar W^T) and the
    // transposed prefix step and sum cotangent, 10 flops per entry
    const double K = a.n - 1, S = (double)a.S, B = (double)a.B;
    const double bytes = 8.0 * B * S * kM * (K + 2.0);
    const double flops = B * S * K * (2.0 * 2.0 * 512.0 + 10.0 * kM);
    void* h = timer_begin("wbd8_rev", s, bytes, flops);
    hipError_t e;
    switch (a.n - 1) {
        case 1: e = launch_rev_k<1>(a, s); break;
        case 2: e = launch_rev_k<2>(a, s); break;
        case 3: e = launch_rev_k<3>(a, s); break;
        case 4: e = launch_rev_k<4>(a, s); break;
        case 5: e = launch_rev_k<5>(a, s); break;
        case 6: e = launch_rev_k<6>(a, s); break;
        case 7: e = launch_rev_k<7>(a, s); break;
        default: e = hipErrorInvalidValue;
    }
    timer_end(h);
    return e;
}

}  // namespace magnus
