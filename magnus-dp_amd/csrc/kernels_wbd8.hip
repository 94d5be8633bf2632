// kernels_wbd8.hip -- single-pass basis sweeps of the dense W-basis path at
// d = 8 under the Boole rule (the resolution controller's problems).
//
// The basis streams W_1 = Prefix(A), W_j = Prefix(A W_{j-1}) (j = 1..K,
// K = n - 1), their totals What_j = W_j[S-1] and the sums Z_0 = Sum(A),
// Z_l = Sum(A W_l) (l = 1..K) are what the GL fold (kernels_wbd.hip) needs;
// the level kernels build them one level at a time, each level a pass over
// HBM.  Here one workgroup walks an instance's samples once with one
// wavefront per level: level j's prefix at sample i needs that level's
// product A_i W_{j-1}[i] (src/magnus.hpp:241-251), so the waves run as a
// pipeline of Boole blocks -- at step t wave w takes block t - w, its input
// block formed by wave w - 1 in the step before and handed over in LDS --
// and every level's state stays in its wave's registers.  The levels are
// written only when a VJP needs them (LEVELS), and then the even ones only:
// the reverse sweep walks the blocks backwards the same way, one wave per
// level again, and the wave of an odd level rebuilds it backwards from its
// total What_j and the kept level below it (W_1 from the samples).
//
// Products on the matrix cores.  v_mfma_f64_4x4x4f64 runs four independent
// 4 x 4 x 4 blocks: block b's A[i][k] is lane 16k + 4b + i, its B[k][j] lane
// 16k + 4b + j and its D[i][j] lane 16i + 4b + j.  An 8 x 8 product Y = X V
// is four output blocks (I, J) over two halves kappa of the contraction:
// block b = 2I + J of instruction kappa multiplies rows 2i + I of X, columns
// 2k + kappa, by rows 2k + kappa of V, columns 4J + j.  Lane
// L = 16k + 8I + 4J + j then owns Y[2k + I][4J + j]; the B operands of the
// two instructions are rows 2k and 2k + 1 of column 4J + j (the "B pair":
// the lane's own entry and that of lane L ^ 8), and the A operands of lane L
// (i = L % 4) are X[2i + I][2k] and X[2i + I][2k + 1] (the "A chunk").  No
// lane computes anything that is discarded: an instance fills a wavefront
// and a product is two instructions.
//
// Rounding: the MFMA sums the contraction in its own order and the Boole rule
// is evaluated as running partial sums (the same weights as boole.hpp:24-74,
// regrouped), so results differ from the reference's sequential order by
// rounding only.  Each instance's results depend on its own data alone.
#include "kernels.hpp"
#include "walk.hpp"
#include "rules.hpp"

namespace magnus {

namespace {

constexpr int kM = 64;  // entries of a d = 8 sample

// levels a VJP keeps in HBM: the even ones, W_2, W_4, ..., and the index of even level j among them
__host__ __device__ constexpr int kept_levels(int K) { return K / 2; }
__host__ __device__ constexpr int kept_index(int j) { return j / 2 - 1; }

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
// Sample ring.  An instance's samples reach LDS through direct global->LDS
// loads (global_load_lds_dwordx4: lane l's 16 bytes land at the
// instruction's LDS base + 16 l) issued by one extra wavefront, the loader,
// kAhead blocks ahead of the first level's wave; a block of 4 samples is
// 2 KB, two instructions of 1 KB (lane l places chunk position l % 32 of
// sample l / 32).  Within a sample the chunk (A[r][2h], A[r][2h + 1]) sits at
// position 4r + (h ^ ((r >> 1) & 3)), so that the A chunks the forward sweep
// reads (8 lanes at a time) and the single entries both sweeps read (16
// lanes at a time) fall in distinct banks.  The loader is the only wave with
// vector memory loads in flight, so it waits for a block with a vmcnt that
// counts only those loads (the blocks issued after it), and the workgroup
// barrier that follows publishes the block to every wave.
// ---------------------------------------------------------------------------
constexpr int kAhead = 6;  // blocks the loader keeps in flight ahead of wave 0

__device__ __forceinline__ int chunk_pos(int r, int h) { return 4 * r + (h ^ ((r >> 1) & 3)); }
// the LDS word of entry (r, col) within a ring sample
__device__ __forceinline__ int ring_word(int r, int col) { return 2 * chunk_pos(r, col >> 1) + (col & 1); }

// one ring slot: [4 samples][64 entries], 16-byte aligned for the chunk reads
struct alignas(16) RingSlot {
    double v[4][kM];
};

template <int N>
__device__ __forceinline__ void wait_vm() {  // vmcnt <= N (expcnt, lgkmcnt unconstrained)
    static_assert(N < 64, "vmcnt has 6 bits");
    __builtin_amdgcn_s_waitcnt((N & 15) | 0x0F70 | ((N >> 4) << 14));
}

// vmcnt <= 2 * blocks: the loader's loads of the `blocks` later blocks may still be in flight
__device__ __forceinline__ void wait_blocks(int blocks) {
    switch (blocks) {
        case 0: wait_vm<0>(); break;
        case 1: wait_vm<2>(); break;
        case 2: wait_vm<4>(); break;
        case 3: wait_vm<6>(); break;
        case 4: wait_vm<8>(); break;
        default: wait_vm<10>(); break;
    }
    static_assert(kAhead - 1 <= 5, "wait_blocks covers kAhead - 1 later blocks");
}

// Workgroup barrier of the pipelines: this wave's LDS writes complete, then
// s_barrier.  __syncthreads() would also wait for every vector memory access
// in flight (its release fence), which would drain the loader's ring loads
// at every step; the "memory" clobber keeps the compiler from moving memory
// accesses across the barrier.
__device__ __forceinline__ void step_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// The loader wave of a pipeline: ring index r holds block blk_of(r).  At step
// t it issues ring index t + kAhead - 1 and waits for index t + LEAD before
// the step's barrier (wave 0 reads index t at step t, and with LEAD = 2 index
// t + 1 as well); the prologue issues the first indices and waits for indices
// 0..LEAD-1 before the barrier that opens the pipeline.  The waves read an
// index up to K - 1 steps after wave 0, and the loader rewrites a slot only
// after the barrier that ends the step of its last reader, so the ring holds
// K + kAhead - 1 slots.
template <int R, int LEAD, typename BlkOf>
__device__ __forceinline__ void ring_loader(RingSlot* ring, const double* A, long long nb, long long steps, int lane,
                                            BlkOf blk_of) {
    static_assert(LEAD >= 1 && LEAD < kAhead, "the loader must issue ahead of what it waits for");
    const int pos = lane & 31, r = pos >> 2, h = (pos & 3) ^ ((r >> 1) & 3);
    const double* src = A + (1 + (lane >> 5)) * (long long)kM + r * 8 + 2 * h;  // sample 4 blk + 1 + lane / 32
    auto issue = [&](long long ri) {
        const long long blk = blk_of(ri);
        RingSlot& slot = ring[ri % R];
#pragma unroll
        for (int q = 0; q < 2; ++q)
            __builtin_amdgcn_global_load_lds(src + (4 * blk + 2 * q) * (long long)kM,
                                             (__attribute__((address_space(3))) void*)&slot.v[2 * q][0], 16, 0, 0);
    };
    // index `need` (and those before it) complete; the ones issued after it may be in flight
    auto wait_for = [&](long long need, long long last_issued) {
        wait_blocks((int)(last_issued - (need < last_issued ? need : last_issued)));
    };
    const long long pre = nb < kAhead - 1 ? nb : kAhead - 1;
    for (long long ri = 0; ri < pre; ++ri) issue(ri);
    wait_for(LEAD - 1, pre - 1);
    step_barrier();
    for (long long t = 0; t < steps; ++t) {
        const long long ri = t + kAhead - 1;
        if (ri < nb) issue(ri);
        if (t + LEAD < nb) wait_for(t + LEAD, ri < nb ? ri : nb - 1);
        step_barrier();
    }
}

// The sampled matrix A(t_i) = C_0 + sum_k (sin_k C_{2k-1} + cos_k C_{2k}) at one
// entry, from that entry's coefficients c[0..2M] and the grid's sin/cos row
// tk[0..2M-1]: one fused multiply-add per term in k order, as the sampling
// kernels (kernels_sample.hip) accumulate, so the values are theirs bit for bit.
__device__ __forceinline__ double fourier_value(const double* c, const double* tk, int modes) {
    double v = c[0];
#pragma unroll
    for (int k = 1; k <= kWbd8MaxModes; ++k) {
        if (k > modes) break;
        v = __fma_rn(tk[2 * k - 2], c[2 * k - 1], v);
        v = __fma_rn(tk[2 * k - 1], c[2 * k], v);
    }
    return v;
}
// entry e of sample i of an instance, its coefficients read from global memory
__device__ __forceinline__ double fourier_entry(const double* coef, const double* tab, int modes, long long i, int e) {
    double c[2 * kWbd8MaxModes + 1];
#pragma unroll
    for (int q = 0; q < 2 * kWbd8MaxModes + 1; ++q) c[q] = q <= 2 * modes ? coef[q * kM + e] : 0.0;
    return fourier_value(c, tab + i * 2 * modes, modes);
}

// lane `src`'s value of v, in every lane
__device__ __forceinline__ double lane_bcast(double v, int src) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, src);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), src);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// acc + (lane J of this lane's 16-lane row of t) * c, one fused multiply-add:
// v_fmac_f64 with the row_newbcast DPP source modifier (FP64 DPP broadcasts a
// row lane to the whole row).
template <int J>
__device__ __forceinline__ double fma_row_bcast(double t, double c, double acc) {
    static_assert(J >= 0 && J < 16, "a lane of the row");
    asm("v_fmac_f64_dpp %0, %1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf bound_ctrl:1"
        : "+v"(acc)
        : "v"(t), "v"(c), "n"(J));
    return acc;
}

// The loader of a pipeline whose samples are formed from M Fourier modes:
// lane l keeps entry l's 2M + 1 coefficients in registers and, for each ring
// index, writes that entry of the block's four samples (4 blk + 1 .. 4 blk + 4)
// to the slot, the four as independent fused multiply-add chains.  The
// block's rows of the sin/cos table (4 x 2M <= 64 contiguous doubles, value x
// = 2M p + 2(k - 1) + {0: sin, 1: cos}) sit in ceil(M / 2) registers, value x
// in register x / 16 at lane x % 16 of every 16-lane row (each row loads the
// same 128 bytes), so each multiply-add takes its table value straight from
// the register by a row broadcast: no LDS, no scalar registers, one
// instruction per term.  The next index's rows load while this one is formed.
//
// The step protocol is ring_loader's: indices 0 .. kAhead - 2, the barrier
// that opens the pipeline, then at step t index t + kAhead - 1 and the step's
// barrier (which publishes the writes, complete before it).  Every index up to
// the last step's is formed, those past the last block from the last block's
// rows: their slots are free by the protocol and no wave reads them, and the
// loop then has no branch around a load (walk_range), so its waits count only
// the loads issued since.
template <int M>
struct TabRows {
    double r[M > 0 ? (M + 1) / 2 : 1];
};

template <int R, int M, typename BlkOf>
__device__ __forceinline__ void ring_fourier_m(RingSlot* ring, const double* coef, const double* tab, long long nb,
                                               long long steps, int lane, BlkOf blk_of) {
    constexpr int NR = (M + 1) / 2;  // table registers per block
    const int word = ring_word(lane >> 3, lane & 7);
    double c[2 * M + 1];
#pragma unroll
    for (int q = 0; q <= 2 * M; ++q) c[q] = coef[q * kM + lane];
    auto load = [&](long long ri, TabRows<M>& t) {
        if constexpr (M > 0) {
            const double* rows = tab + (4 * blk_of(ri < nb ? ri : nb - 1) + 1) * (2 * M);
#pragma unroll
            for (int q = 0; q < NR; ++q) {
                const int x = 16 * q + (lane & 15);
                t.r[q] = rows[x < 8 * M ? x : 0];
            }
        }
    };
    auto step = [&](long long ri, TabRows<M>& t, auto) {
        RingSlot& slot = ring[ri % R];
        double v[4];
#pragma unroll
        for (int p = 0; p < 4; ++p) v[p] = c[0];
        static_for<1, M + 1>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            static_for<0, 4>([&](auto pc) {
                constexpr int p = decltype(pc)::value, x = 2 * M * p + 2 * (k - 1);
                v[p] = fma_row_bcast<x % 16>(t.r[x / 16], c[2 * k - 1], v[p]);
            });
            static_for<0, 4>([&](auto pc) {
                constexpr int p = decltype(pc)::value, x = 2 * M * p + 2 * (k - 1) + 1;
                v[p] = fma_row_bcast<x % 16>(t.r[x / 16], c[2 * k], v[p]);
            });
        });
#pragma unroll
        for (int p = 0; p < 4; ++p) slot.v[p][word] = v[p];
        if (ri >= kAhead - 2) step_barrier();
    };
    TabRows<M> tr[2] = {};
    walk_range<2, 2, true>(0, steps + kAhead - 2, tr, load, step);
}

template <int R, typename BlkOf>
__device__ __forceinline__ void ring_fourier(RingSlot* ring, const double* coef, const double* tab, int modes,
                                             long long nb, long long steps, int lane, BlkOf blk_of) {
    switch (modes) {
        case 0: ring_fourier_m<R, 0>(ring, coef, tab, nb, steps, lane, blk_of); break;
        case 1: ring_fourier_m<R, 1>(ring, coef, tab, nb, steps, lane, blk_of); break;
        case 2: ring_fourier_m<R, 2>(ring, coef, tab, nb, steps, lane, blk_of); break;
        case 3: ring_fourier_m<R, 3>(ring, coef, tab, nb, steps, lane, blk_of); break;
        case 4: ring_fourier_m<R, 4>(ring, coef, tab, nb, steps, lane, blk_of); break;
        case 5: ring_fourier_m<R, 5>(ring, coef, tab, nb, steps, lane, blk_of); break;
        case 6: ring_fourier_m<R, 6>(ring, coef, tab, nb, steps, lane, blk_of); break;
        case 7: ring_fourier_m<R, 7>(ring, coef, tab, nb, steps, lane, blk_of); break;
        default: ring_fourier_m<R, kWbd8MaxModes>(ring, coef, tab, nb, steps, lane, blk_of); break;
    }
    static_assert(kWbd8MaxModes == 8, "one case per mode count");
}

// The B pair (V[2k][c], V[2k + 1][c]) of lane L from its own entry
// V[2k + I][c]: a DPP row rotation by 8 brings lane L ^ 8's entry, and the
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
    return double2{join(dpp_ror8(lo, lo, 0xC), dpp_ror8(hi, hi, 0xC)),   // row 2k: I = 1 lanes take L ^ 8
                   join(dpp_ror8(lo, lo, 0x3), dpp_ror8(hi, hi, 0x3))};  // row 2k + 1: I = 0 lanes take L ^ 8
}

// (V[2k][col], V[2k + 1][col]) of lane L from its own V[2k + J][col] (lane
// L ^ 4 holds the other: row rotations by 4 and 12 with bank masks)
__device__ __forceinline__ double2 pair4_dpp(double own) {
    const unsigned long long v = __builtin_bit_cast(unsigned long long, own);
    const int lo = (int)(unsigned)v, hi = (int)(unsigned)(v >> 32);
    auto join = [](int l, int h) {
        return __builtin_bit_cast(double, ((unsigned long long)(unsigned)h << 32) | (unsigned)l);
    };
    return double2{join(__builtin_amdgcn_update_dpp(lo, lo, 0x124, 0xF, 0xA, false),   // J = 1 lanes take L - 4
                        __builtin_amdgcn_update_dpp(hi, hi, 0x124, 0xF, 0xA, false)),
                   join(__builtin_amdgcn_update_dpp(lo, lo, 0x12C, 0xF, 0x5, false),   // J = 0 lanes take L + 4
                        __builtin_amdgcn_update_dpp(hi, hi, 0x12C, 0xF, 0x5, false))};
}

// Y[2k + I][4J + j] (+ c) of Y = X V from the A chunk x and the B pair v
__device__ __forceinline__ double prod(double2 x, double2 v, double c = 0.0) {
    const double d = __builtin_amdgcn_mfma_f64_4x4x4f64(x.x, v.x, c, 0, 0, 0);
    return __builtin_amdgcn_mfma_f64_4x4x4f64(x.y, v.y, d, 0, 0, 0);
}

// Lane fields of the product layout.
struct Lane {
    int k, I, J, i;  // lane = 16k + 8I + 4J + i (i is the column j of the own entry)
    int c, own;      // own entry (2k + I, c), c = 4J + i
    __device__ explicit Lane(int lane)
        : k(lane >> 4), I((lane >> 3) & 1), J((lane >> 2) & 1), i(lane & 3), c(4 * J + i), own((2 * k + I) * 8 + c) {}
};

// ---------------------------------------------------------------------------
// Forward sweep.  Hand-over layout of a level block: per sample, lane L's own
// entry (2k + I, c) sits at word 16k + 2c + I, so that a lane reads its B
// pair as one 16-byte chunk.  The waves hand their level blocks on in two
// buffers used in alternate steps, so one barrier per step orders every
// hand-over.
// ---------------------------------------------------------------------------

// One wave of the forward pipeline, level j = w + 1.  FIRST: the wave of level
// 1's prefix (W_1 = Prefix(A), Z_0 = Sum(A)), which also forms Y_2 = A W_1;
// otherwise it forms Y_{j+1} = A W_j from the block of W_j the previous wave
// handed over.  MAKES: it steps level j + 1's prefix (j < K) and hands that
// block on.  The role is a template parameter and the step indices are
// wave-uniform (scalar registers), so that a step is straight-line code: the
// four samples' reads and products are independent and overlap; only the
// prefix recurrences run sample after sample.
template <int K, bool LEVELS, bool FIRST, bool MAKES>
__device__ __forceinline__ void sweep_wave(const LmArgs& a, int b, int w, const Lane& L, const RingSlot* ring,
                                           double (*hand)[K > 1 ? K - 1 : 1][4][64]) {
    constexpr int R = K + kAhead - 1;
    const long long S = a.S;
    const int nb = (int)((S - 1) / 4);
    const int j = w + 1;
    const int own = L.own;
    const int ra = 2 * chunk_pos(2 * L.i + L.I, L.k);  // the A chunk A[2i + I][2k ..]
    const int ro = ring_word(2 * L.k + L.I, L.c);       // the own entry
    const int hw = 16 * L.k + 2 * L.c + L.I, hp = hw - L.I;
    const double* __restrict__ A = a.in + (long long)b * S * kM;
    double* __restrict__ W = LEVELS ? a.carry + (long long)b * kept_levels(K) * S * kM : nullptr;
    const BooleW bw(a.dt);
    Acc s1, sn = Acc{0.0, 0.0, 0.0};  // prefix states of W_1 (FIRST) and W_{j+1} (MAKES)
    double z1 = 0.0, zn = 0.0;        // Z_0 (FIRST) and Z_j

    // sample 0: every prefix is 0 there, so the products of levels 2..K+1 are 0
    if constexpr (FIRST) {
        const double y0 = a.fcoef ? fourier_entry(a.fcoef + (long long)b * (2 * a.fmodes + 1) * kM, a.ftab, a.fmodes, 0, own)
                                  : A[own];
        s1 = Acc{0.0, y0, 0.0};
        z1 = bw.b14 * y0;
    }
    const bool keeps = LEVELS && MAKES && ((j + 1) & 1) == 0;  // W_{j+1} is an even level
    if (keeps) W[(long long)kept_index(j + 1) * S * kM + own] = 0.0;
    step_barrier();  // block 0 is in the ring
    int slot = 0;    // ring slot of this wave's next block
    for (int t = 0; t < nb + K - 1; ++t) {
        const int blk = t - w;
        if (blk >= 0 && blk < nb) {
            const RingSlot& sl = ring[slot];
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
            } else {  // the previous wave formed this block in the previous step
#pragma unroll
                for (int p = 0; p < 4; ++p) wp[p] = *reinterpret_cast<const double2*>(&hand[(t - 1) & 1][w - 1][p][hp]);
            }
#pragma unroll
            for (int p = 0; p < 4; ++p) y[p] = prod(av[p], wp[p]);  // Y_{j+1} = A W_j
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
                if (keeps) {  // the level this wave forms, kept for the reverse sweep
                    double* Wb = W + ((long long)kept_index(j + 1) * S + 4LL * blk + 1) * kM + own;
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
    constexpr int R = K + kAhead - 1;  // ring slots
    __shared__ RingSlot ring[R];
    // level blocks handed from wave w to wave w + 1: [step parity][w][4 samples][64 words]
    __shared__ alignas(16) double hand[2][K > 1 ? K - 1 : 1][4][64];
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int b = blockIdx.x;
    const long long nb = (a.S - 1) / 4;
    const Lane L(lane);
    if (w == K) {
        if (a.fcoef)
            ring_fourier<R>(ring, a.fcoef + (long long)b * (2 * a.fmodes + 1) * kM, a.ftab, a.fmodes, nb, nb + K - 1,
                            lane, [](long long r) { return r; });
        else
            ring_loader<R, 1>(ring, a.in + (long long)b * a.S * kM, nb, nb + K - 1, lane, [](long long r) { return r; });
    } else if (w == 0) {
        sweep_wave<K, LEVELS, true, (K > 1)>(a, b, w, L, ring, hand);
    } else if (w < K - 1) {
        sweep_wave<K, LEVELS, false, true>(a, b, w, L, ring, hand);
    } else {
        sweep_wave<K, LEVELS, false, false>(a, b, w, L, ring, hand);
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
// Ybar_j and the dA sum so far; the wave of level 1 adds Ybar_1 and writes
// dA.  The levels are added to dA in a fixed order (K first), so results are
// reproducible.  Wbar_j = A^T Ybar_{j+1} takes A^T's chunk (two entries of a
// column of A, from the ring) and Ybar's B pair; Ybar W^T takes Ybar's A
// chunk (two entries of a row) and W^T's B pair (W[4J + j][2k ..], 16
// contiguous bytes of the level in HBM).  A handed Ybar block keeps entry
// (r, c) at word 16 (r >> 1) + 2 (c ^ (2 (r >> 1) & 7)) + (r & 1): the B pair
// is one 16-byte chunk and both reads and the writes are free of bank
// conflicts.
// ---------------------------------------------------------------------------

__device__ __forceinline__ int ybar_word(int r, int c) { return 16 * (r >> 1) + 2 * (c ^ ((2 * (r >> 1)) & 7)) + (r & 1); }

// the Boole sum weight of sample i (sum_vjp_coef, boole.hpp:153-199)
__device__ __forceinline__ double boole_c(long long i, long long S, const BooleW& c) {
    if (i == 0 || i == S - 1) return c.b14;
    const int r = (int)(i & 3);
    return r == 2 ? c.b24 : (r == 0 ? c.b28 : c.b64);
}

// One wave of the reverse pipeline, level j = K - w.  TOP: level K, whose
// input Ybar_{K+1} = c Zbar_K needs no hand-over; LAST: level 1, which adds
// Ybar_1 and writes dA (the others hand Ybar_j and the dA sum on).  MODE: how
// the wave gets W_j's entries W_j[c][2k + I] (its half of W_j^T's B pair):
// 0 (even j) reads them from the kept level; 1 (j = 1) and 2 (odd j >= 3)
// rebuild the level backwards: from a block's end value, the block's base is
// that value minus the block's Boole sum of the integrand Y_j, and its inner
// positions follow from the base with the forward's rules, starting from the
// total What_j.  Y_1 = A comes from the ring; Y_j = A W_{j-1} is formed as
// Y_j^T = W_{j-1}^T A^T, whose entries (2k + I, c) are the ones needed, from
// the kept level W_{j-1} (each lane reads one entry, lane L ^ 4 the other of
// its pair) and A's rows from the ring.  A block's first sample is the last of
// the block below, one ring slot above and in the next buffer of W_{j-1}.
template <int K, bool TOP, bool LAST, int MODE>
__device__ __forceinline__ void rev_wave(const Wbd8RevArgs& a, int b, int w, int lane, const Lane& L,
                                         const RingSlot* ring, double (*ybh)[K > 1 ? K - 1 : 1][4][64],
                                         double (*dah)[K > 1 ? K - 1 : 1][4][64]) {
    constexpr int R = K + kAhead - 1;
    const long long S = a.S, n = a.n;
    const int nb = (int)((S - 1) / 4);
    const int j = K - w;
    const int k = L.k, I = L.I, i = L.i, c = L.c, own = L.own;
    // ring words of the A^T chunk A[2k][2i + I], A[2k + 1][2i + I]
    const int rt0 = ring_word(2 * k, 2 * i + I), rt1 = ring_word(2 * k + 1, 2 * i + I);
    // Ybar hand-over words: own entry, B pair, A chunk (2i + I, 2k), (2i + I, 2k + 1)
    const int yw = ybar_word(2 * k + I, c), yp = ybar_word(2 * k, c);
    const int ya0 = ybar_word(2 * i + I, 2 * k), ya1 = ybar_word(2 * i + I, 2 * k + 1);
    static_assert(MODE != 1 || LAST, "the samples rebuild level 1 only");
    // MODE 0: this lane's half W_j[c][2k + I] of W_j^T's B pair (lane L ^ 8
    // loads the other half); MODE 2: this lane's entry W_{j-1}[2k + J][2i + I]
    const double* __restrict__ Wj =
        MODE == 1 ? nullptr
        : MODE == 0
            ? a.W + ((long long)b * kept_levels(K) + kept_index(j)) * S * kM + c * 8 + 2 * k + I
            : a.W + ((long long)b * kept_levels(K) + kept_index(j - 1)) * S * kM + (2 * k + L.J) * 8 + 2 * i + I;
    const int rq = ring_word(c, 2 * k + I);  // MODE 1: A[c][2k + I] in a ring sample
    const int rc = 2 * chunk_pos(c, k);      // MODE 2: A's row chunk A[c][2k .. 2k + 1]
    double wend = 0.0, a0t = 0.0;  // MODE 1, 2: W_j[c][2k + I] at the current block's end; MODE 1: A_0 there
    if constexpr (MODE != 0) wend = a.totals[((long long)b * n + (j - 1)) * kM + c * 8 + 2 * k + I];
    if constexpr (MODE == 1)
        a0t = a.fcoef ? fourier_entry(a.fcoef + (long long)b * (2 * a.fmodes + 1) * kM, a.ftab, a.fmodes, 0, c * 8 + 2 * k + I)
                      : a.in[(long long)b * S * kM + c * 8 + 2 * k + I];
    double* __restrict__ dA = a.dA + (long long)b * S * kM;
    const BooleW bw(a.dt);

    const double* zr = a.zbar + (long long)b * n * kM;
    const double zb = zr[(j - 1) * kM + own];  // Zbar_{j-1}
    double2 zkp = double2{0.0, 0.0}, zkc = double2{0.0, 0.0};  // TOP: Zbar_K's B pair and A chunk
    if constexpr (TOP) {
        zkp = double2{zr[K * kM + 2 * k * 8 + c], zr[K * kM + (2 * k + 1) * 8 + c]};
        zkc = double2{zr[K * kM + (2 * i + I) * 8 + 2 * k], zr[K * kM + (2 * i + I) * 8 + 2 * k + 1]};
    }
    double up = 0.0, pend = 0.0;

    // W_j's halves of a block, two blocks in flight ahead of the one in use.
    // The three buffers change roles by renaming (the active steps are
    // unrolled by three), never by copying: a copy of a register whose load is
    // in flight would wait for that load.
    auto loadw = [&](double* y, int blk) {
        if constexpr (MODE != 1) {
#pragma unroll
            for (int p = 0; p < 4; ++p) y[p] = Wj[(4LL * blk + 1 + p) * kM];
        }
    };
    double wa[4] = {}, wbuf[4] = {}, wc[4] = {};
    loadw(wa, nb - 1);
    if (nb > 1) loadw(wbuf, nb - 2);
    int slot = 0;  // ring slot of this wave's next block
    double ftop = 0.0;  // MODE 2: Y_j at the top sample of the next block (this block's bottom)

    // one pipeline step; r = this wave's reverse position (block nb - 1 - r;
    // out of range: the wave only takes part in the barrier), wt: the loaded
    // entries of the block (MODE 2: of W_{j-1}, which the step turns into W_j's
    // halves), wn: those of the block below, wpf: the buffer for the block two
    // below
    auto step = [&](int r, double* wt, const double* wn, double* wpf) {
        const int blk = nb - 1 - r;
        if (r >= 0 && blk >= 0) {
            const int t = r + w;  // the pipeline step
            const RingSlot& sl = ring[slot];
            slot = slot + 1 == R ? 0 : slot + 1;
            if (blk > 1) loadw(wpf, blk - 2);
            if constexpr (MODE != 0) {  // W_j over the block, from its end value wend
                double f[5];  // the integrand Y_j[c][2k + I] at the block's samples 4 blk .. 4 blk + 4
                if constexpr (MODE == 1) {
                    f[0] = blk > 0 ? ring[slot].v[3][rq] : a0t;
#pragma unroll
                    for (int p = 0; p < 4; ++p) f[p + 1] = sl.v[p][rq];
                } else {
                    // Y_j[0] = A_0 W_{j-1}[0] = 0; otherwise the block below's last sample,
                    // which is also that block's f[4] in the next step (the wave's first
                    // block forms its own)
#pragma unroll
                    for (int p = 0; p < 3; ++p)
                        f[p + 1] = prod(pair4_dpp(wt[p]), *reinterpret_cast<const double2*>(&sl.v[p][rc]));
                    f[4] = r == 0 ? prod(pair4_dpp(wt[3]), *reinterpret_cast<const double2*>(&sl.v[3][rc])) : ftop;
                    f[0] = blk > 0 ? prod(pair4_dpp(wn[3]), *reinterpret_cast<const double2*>(&ring[slot].v[3][rc]))
                                   : 0.0;
                    ftop = f[0];
                }
                const double base =
                    wend - fma(bw.b14, f[4], fma(bw.b64, f[3], fma(bw.b24, f[2], fma(bw.b64, f[1], bw.b14 * f[0]))));
                Acc st = Acc{base, f[0], 0.0};
                wt[0] = brule<1>(st, f[1], bw);
                wt[1] = brule<2>(st, f[2], bw);
                wt[2] = brule<3>(st, f[3], bw);
                wt[3] = wend;
                wend = base;
            }
            const long long s = 4LL * blk;
            double cz[4];
#pragma unroll
            for (int p = 0; p < 4; ++p) cz[p] = boole_c(s + 1 + p, S, bw);
            double2 yb[4], yc[4];
            double part[4];
            if constexpr (TOP) {  // Ybar_{K+1} = c Zbar_K
#pragma unroll
                for (int p = 0; p < 4; ++p) {
                    yb[p] = double2{cz[p] * zkp.x, cz[p] * zkp.y};
                    yc[p] = double2{cz[p] * zkc.x, cz[p] * zkc.y};
                    part[p] = 0.0;
                }
            } else {  // the previous wave handed Ybar_{j+1} and the dA sum over in the previous step
#pragma unroll
                for (int p = 0; p < 4; ++p) {
                    const double* yh = ybh[(t - 1) & 1][w - 1][p];
                    yb[p] = *reinterpret_cast<const double2*>(&yh[yp]);
                    yc[p] = double2{yh[ya0], yh[ya1]};
                    part[p] = dah[(t - 1) & 1][w - 1][p][lane];
                }
            }
            double wb[4], acc[4];
#pragma unroll
            for (int p = 0; p < 4; ++p) {
                const double2 at = double2{sl.v[p][rt0], sl.v[p][rt1]};
                wb[p] = prod(at, yb[p]);              // Wbar_j = A^T Ybar_{j+1}
                acc[p] = prod(yc[p], bpair_dpp(wt[p]), part[p]);  // dA so far + Ybar_{j+1} W_j^T
            }
            if (blk == nb - 1) wb[3] += a.wbar[((long long)b * n + (j - 1)) * kM + own];  // the total's cotangent
            // Ybar_j = Prefix^T(Wbar_j) + c Zbar_{j-1} over the block
            const double g1 = wb[0], g2 = wb[1], g3 = wb[2];
            const double g4 = wb[3] + up;
            const double f4 = fma(bw.b14, g4, pend);
            const double f3 = fma(bw.e38, g3, bw.b64 * g4);
            const double f2 = fma(bw.t13, g2, fma(bw.e98, g3, bw.b24 * g4));
            const double f1 = fma(bw.h2, g1, fma(bw.t43, g2, fma(bw.e98, g3, bw.b64 * g4)));
            up = g1 + g2 + g3 + g4;
            pend = fma(bw.h2, g1, fma(bw.t13, g2, fma(bw.e38, g3, bw.b14 * g4)));
            double yn[4];
            yn[0] = fma(cz[0], zb, f1);
            yn[1] = fma(cz[1], zb, f2);
            yn[2] = fma(cz[2], zb, f3);
            yn[3] = fma(cz[3], zb, f4);
            if constexpr (!LAST) {
#pragma unroll
                for (int p = 0; p < 4; ++p) {
                    ybh[t & 1][w][p][yw] = yn[p];
                    dah[t & 1][w][p][lane] = acc[p];
                }
            } else {  // dA = Ybar_1 + sum_j Ybar_{j+1} W_j^T
                double* o = dA + (s + 1) * kM + own;
#pragma unroll
                for (int p = 0; p < 4; ++p) o[p * kM] = acc[p] + yn[p];
            }
        }
        step_barrier();  // this step's hand-overs and the loader's next block are visible
    };
    step_barrier();  // ring index 0 is in the ring
    // steps before this wave's first block, its nb blocks (buffers rotating by
    // three), then the steps after its last block
    for (int t = 0; t < w; ++t) step(-1, wa, wbuf, wc);
    int r = 0;
    for (; r + 3 <= nb; r += 3) {
        step(r, wa, wbuf, wc);
        step(r + 1, wbuf, wc, wa);
        step(r + 2, wc, wa, wbuf);
    }
    if (r < nb) step(r, wa, wbuf, wc);
    if (r + 1 < nb) step(r + 1, wbuf, wc, wa);
    for (int t = w + nb; t < nb + K - 1; ++t) step(-1, wa, wbuf, wc);
    // sample 0: the prefixes are constant there (W_j[0] = 0), so dA_0 = Ybar_1[0]
    // = level 1's pending from the lowest block + c_0 Zbar_0
    if constexpr (LAST) dA[own] = fma(bw.b14, zb, pend);
}

// Which pipeline role (0..K-1: the level wave at that lag, K: the loader)
// each hardware wavefront of a reverse workgroup takes.  Wavefront h runs on
// SIMD h % 4, so h and h + 4 share one.  The waves of odd levels >= 3 rebuild
// their level on top of the adjoint and are the busiest, the loader computes
// nothing: roles are paired heaviest with lightest so that the busy ones
// spread over the SIMDs.  (The forward sweep keeps wave h at lag h: its
// level waves carry near equal work, and a remapping measured slower.)
template <int K>
struct RevRoles {
    int role[K + 1];
    static constexpr int weight(int p) { return p == K ? 0 : (((K - p) & 1) && K - p >= 3 ? 2 : 1); }
    constexpr RevRoles() : role() {
        constexpr int N = K + 1;
        int order[N] = {};  // roles by weight, heaviest first (stable)
        int cnt = 0;
        for (int wgt = 2; wgt >= 0; --wgt)
            for (int p = 0; p <= K; ++p)
                if (weight(p) == wgt) order[cnt++] = p;
        int lo = 0, hi = N - 1;
        for (int sm = 0; sm < 4 && sm < N; ++sm) {
            role[sm] = order[lo++];
            if (sm + 4 < N) role[sm + 4] = order[hi--];  // a SIMD with two wavefronts: the lightest left
        }
    }
};

// one workgroup per instance: K level wavefronts and the loader (role K)
template <int K>
__global__ void __launch_bounds__(64 * (K + 1)) __attribute__((amdgpu_waves_per_eu(4))) wbd8_rev(Wbd8RevArgs a) {
    constexpr int H = K > 1 ? K - 1 : 1;
    constexpr int R = K + kAhead - 1;  // ring slots
    __shared__ RingSlot ring[R];
    // Ybar_j blocks (ybar_word) and dA sums (by lane) handed from wave w to wave w + 1
    __shared__ alignas(16) double ybh[2][H][4][64];
    __shared__ double dah[2][H][4][64];
    constexpr RevRoles<K> roles;
    const int hw = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int w = roles.role[hw];
    const int b = blockIdx.x;
    const long long nb = (a.S - 1) / 4;
    const Lane L(lane);
    if (w == K) {  // ring index r holds block nb - 1 - r (the blocks are walked top-down)
        if (a.fcoef)
            ring_fourier<R>(ring, a.fcoef + (long long)b * (2 * a.fmodes + 1) * kM, a.ftab, a.fmodes, nb, nb + K - 1,
                            lane, [nb](long long r) { return nb - 1 - r; });
        else
            ring_loader<R, 2>(ring, a.in + (long long)b * a.S * kM, nb, nb + K - 1, lane,
                              [nb](long long r) { return nb - 1 - r; });
    } else if (K == 1) {
        rev_wave<K, true, true, 1>(a, b, w, lane, L, ring, ybh, dah);
    } else if (w == 0) {
        rev_wave<K, true, false, (K & 1) ? 2 : 0>(a, b, w, lane, L, ring, ybh, dah);
    } else if (w == K - 1) {
        rev_wave<K, false, true, 1>(a, b, w, lane, L, ring, ybh, dah);
    } else if ((K - w) & 1) {
        rev_wave<K, false, false, 2>(a, b, w, lane, L, ring, ybh, dah);
    } else {
        rev_wave<K, false, false, 0>(a, b, w, lane, L, ring, ybh, dah);
    }
}

template <int K>
hipError_t launch_rev_k(const Wbd8RevArgs& a, hipStream_t s) {
    hipLaunchKernelGGL((wbd8_rev<K>), dim3((unsigned)a.B), dim3(64 * (K + 1)), 0, s, a);
    return hipGetLastError();
}

}  // namespace

bool wbd8_sweep_ok(int d, int integ, int n) { return d == 8 && integ == BOOLE && n >= 2 && n - 1 <= kWbd8MaxK; }

hipError_t launch_wbd8_sweep(const LmArgs& a, hipStream_t s) {
    if (a.jobs == 0) return hipSuccess;
    if (a.d != 8 || a.S < 5 || (a.S - 1) % 4 != 0 || a.S > (1LL << 30) || a.jobs > (1LL << 30) || a.n < 2 ||
        a.n - 1 > kWbd8MaxK)
        return hipErrorInvalidValue;
    // algorithmic work: the samples read once (and the kept levels W_2..W_K
    // written for a VJP); per sample K products of 2 d^3 flops plus a prefix step and a sum
    // of 8 flops per entry for each of the K + 1 levels
    const double K = a.n - 1, S = (double)a.S, B = (double)a.jobs;
    const double bytes = 8.0 * B * S * kM * ((a.fcoef ? 0.0 : 1.0) + (a.carry ? kept_levels(a.n - 1) : 0.0));
    const double flops = B * S * (K * (2.0 * 512.0 + 8.0 * kM) + 8.0 * kM);
    void* h = timer_begin("wbd8_sweep", s, bytes, flops);
    hipError_t e = a.carry ? launch_k<true>(a, s) : launch_k<false>(a, s);
    timer_end(h);
    return e;
}

hipError_t launch_wbd8_rev(const Wbd8RevArgs& a, hipStream_t s) {
    if (a.B == 0) return hipSuccess;
    if (a.S < 5 || (a.S - 1) % 4 != 0 || a.S > (1LL << 30) || a.B > (1LL << 30) || a.n < 2 ||
        a.n - 1 > kWbd8MaxK)
        return hipErrorInvalidValue;
    // algorithmic work: the samples and the kept levels (W_2..W_K) read once, dA written once;
    // per sample and level two products (A^T Ybar and Ybar W^T) and the
    // transposed prefix step and sum cotangent, 10 flops per entry
    const double K = a.n - 1, S = (double)a.S, B = (double)a.B;
    const double bytes = 8.0 * B * S * kM * (kept_levels(a.n - 1) + (a.fcoef ? 1.0 : 2.0));
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
