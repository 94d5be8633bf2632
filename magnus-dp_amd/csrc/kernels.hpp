// kernels.hpp -- argument blocks and launchers of the device kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace magnus {

constexpr int kArgMaxQ = 64;

// GL-lane reductions (kernels_lm.hip)
struct FwdReduceArgs {
    const double* partial;  // [B][Qj][n_out][m]
    const double* omega1;   // [B][m]
    double* out;            // [B][n_out][m_out]
    long long B;
    int n, m, m_out, n_out, Qj, op, sc;
};

struct VjpReduceArgs {
    const double* dAjob;  // [B][Qj][S][m]
    const double* cot;    // [B][n_out][m_cot]
    double* dA;           // [B][S][m_out]
    long long B;
    long long S;
    int n, m, m_out, m_cot, n_out, Qj, op, sc;
    double dt;
};

// Level-parallel kernels (kernels_lm.hip).  One job = (instance, GL node); a
// level's per-sample products run for all (job, sample, entry) at once, its
// prefix scans and sums per (job, entry).
struct LmArgs {
    const double* in;    // [B][S][m_in]
    const double* cot;   // [B][n_out][m_cot]        (VJP)
    double* carry;       // [B][Q][n-1][S][m] or nullptr
    double* pingpong;    // [jobs][2][S][m] level buffers when carry == nullptr
    double* Y;           // [jobs][S][m] products of the current level (forward)
    double* partial;     // [jobs][n_out][m]
    double* omega1;      // [B][m]
    double* gbuf;        // [jobs][2][S][m] adjoint level buffers (VJP)
    double* dAjob;       // [jobs][S][m]               (VJP)
    double* totals;      // [jobs][n][m] last prefix value of each level, or nullptr (forward)
    const double* wbar;  // [jobs][n][m] cotangents of the level totals, or nullptr (VJP)
    double* cscan;       // [3][nchunks][jobs*m] chunk scratch of the sample-parallel scans, or nullptr
    double* dbar;        // [jobs][2][m] deferred cotangents of the level totals (level walkers' VJP)
    // d = 8 single-pass sweeps: samples formed from Fourier coefficients instead of read from `in`
    const double* fcoef; // [jobs][2 fmodes + 1][64] or nullptr
    const double* ftab;  // [S][fmodes][2]: sin, cos of 2 pi k t_i (fourier_table)
    int fmodes;
    int clen;            // samples per chunk of the sample-parallel scan (multiple of 4)
    long long S;
    long long jobs;
    int n, d, m, m_in, m_cot, Q, Qj, n_out, op;
    int sc;        // 0 generic, 1 padded su(2) (S,3) input, 2 su(2) on (S,2,2) input
    int outputs;   // forward: 0 records the carry only
    int bt_fused;  // VJP: Fixed2/SpaceCurve bar_total = temp * (1 + x); generic temp + temp * x
    int level;     // forward level j (1..n) / VJP level k (n..2)
    int ga;        // VJP: gbuf half holding the current adjoint
    int tile;      // product path: 0 default, 1 VALU, 2 / 3 d = 32 MFMA split over 2 / 4 wavefronts
    double dt;
    double w[kArgMaxQ];
    double x[kArgMaxQ];
};

// Per-kernel timing hooks (implemented by the C-ABI layer).  When timing is
// enabled, timer_begin records a start event on `s` and returns a handle that
// timer_end closes; both are no-ops (nullptr) otherwise.  A launch site may
// state the algorithmic HBM bytes and flops of what it launches (FMA = 2
// flops); the totals are reported beside the durations.
void* timer_begin(const char* name, hipStream_t s, double bytes = 0.0, double flops = 0.0);
void timer_end(void* handle);

hipError_t launch_fwd_reduce(const FwdReduceArgs& a, hipStream_t s);
hipError_t launch_vjp_reduce(int integ, const VjpReduceArgs& a, hipStream_t s);

// forward level a.level: products (level >= 2) then scan/sums
hipError_t launch_lm_forward_level(int integ, const LmArgs& a, hipStream_t s);
// VJP: init (adjoint of the last level's sum), one level a.level, final dA += adjoint
hipError_t launch_lm_vjp_init(int integ, const LmArgs& a, hipStream_t s);
hipError_t launch_lm_vjp_level(int integ, const LmArgs& a, hipStream_t s);
hipError_t launch_lm_vjp_final(const LmArgs& a, hipStream_t s);

// Level walkers of the per-node recursion (kernels_walk.hip) for d in {16,
// 32} under the Boole rule: one workgroup per job walks one level's samples
// once, products on the matrix cores fed straight into the prefix recurrence.
// walk_ok: the problem and LmArgs they apply to (generic matrices, per-node
// jobs, default product path).  The forward level replaces
// launch_lm_forward_level (same outputs).  The VJP runs launch_walk_vjp_level
// for k = n..2 (no init or final launch; ga alternates as for the level
// kernels) and then launch_walk_vjp_reduce, which replaces launch_vjp_reduce:
// a level leaves the next adjoint in gbuf without the total's cotangent
// (grad.hpp:305-313), which is deferred in dbar and added where the adjoint is
// read.
bool walk_ok(int integ, const LmArgs& a);
hipError_t launch_walk_forward_level(const LmArgs& a, hipStream_t s);
hipError_t launch_walk_vjp_level(const LmArgs& a, hipStream_t s);
hipError_t launch_walk_vjp_reduce(const LmArgs& a, const VjpReduceArgs& r, hipStream_t s);

// Segment walkers of the dense W-basis VJP with time-segment checkpoints
// (kernels_walkseg.hip), d in {16, 32}, Boole: the samples are cut into G
// segments [s_g, s_{g+1}] of block-aligned length; a workgroup per instance
// walks one level of one segment.  The forward keeps, per segment start and
// level, the prefix recurrence's state (P, f) and the running level sums, so
// the reverse rebuilds one segment's levels at a time from its checkpoint and
// walks them back with the reverse recurrence's state (g4, pending) handed
// from segment to segment.  Same operations, in the same order, as one walk
// over all samples, so the result does not depend on G.
struct WalkSegArgs {
    const double* in;      // [B][S][m]
    double* wseg;          // [B][K][L + 1][m]: W_1..W_K at samples s_g + t, t = 0..L
    double* ckp;           // [B][G][K][2][m]: (P, f) of levels 1..K at s_g (g >= 1)
    double* sums;          // [B][n][m]: running level sums Z_0..Z_{n-1}
    double* z0;            // [B][m]: Z_0 (the level-1 sum) once the last segment is done
    double* totals;        // [B][n][m]: rows 0..K-1 = W_j[S - 1]
    double* useg;          // [B][2][L + 1][m]: adjoint halves on the segment
    double* dAseg;         // [B][L + 1][m]
    double* rstate;        // [B][K][2][m]: (g4, pending) of reverse level k at row k - 2
    const double* zbar;    // [B][n][m]
    const double* wbar;    // [B][n][m]
    double* dA;            // [B][S][m]
    long long B, S;
    long long s0, s1;      // the segment's first and last sample
    int L, G, g, n, d, m, level, ga, mode;  // mode (forward): 0 first pass (sums, checkpoints, totals), 1 rebuild
    double dt;
};
bool walkseg_ok(int d, int integ, int n);
hipError_t launch_walkseg_forward(const WalkSegArgs& a, hipStream_t s);
hipError_t launch_walkseg_reverse(const WalkSegArgs& a, hipStream_t s);
hipError_t launch_walkseg_reduce(const WalkSegArgs& a, hipStream_t s);

// Grouped walkers of the dense W-basis path (kernels_walkseg.hip), d in {16,
// 32}, Boole, op many or sum: many instances of one dim, each with its own
// samples and order, run as a pipeline of launches over task tables.  The
// samples of an instance are cut into segments of L samples; a task walks one
// level of one segment of one instance (one workgroup).  Forward task
// (i, g, j) reads W_{j-1} on segment g and the state level j left at the end
// of segment g - 1, so the launch of step t runs every task with g + j - 1 = t;
// reverse task (i, g, k) reads barY^(k) on segment g and the state level k
// left at the start of segment g + 1, so step t runs (G_i - 1 - g) + (n_i - k)
// = t, and the dA of segment g (kind 2) runs one step after its level 2.  The
// adjoints barY^(k) alternate between two segment slots by segment parity.
struct GwInst {
    double* W;           // [K][S][m] basis levels W_1..W_K
    double* fst;         // [K][2][m] (P, f) of each level at the end of the last walked segment
    double* U;           // [K][2][L + 1][m] barY^(1..K), two segment slots
    double* rst;         // [K][2][m] (g4, pending) of each reverse level
    double* sums;        // [n][m] running sums Z_0..Z_{n-1}
    double* z0;          // [m] Z_0 once the last segment is walked
    double* totals;      // [n][m] rows 0..K-1: W_j[S - 1]
    const double* zbar;  // [n][m]
    const double* wbar;  // [n][m]
    long long S;
    int n;
    int id;              // the instance's row in the call's pointer table
    double dt;
};
struct GwTask {
    int inst, seg, level;
    int kind;  // 0 forward level, 1 reverse level, 2 dA of the segment's own samples
};
struct GwArgs {
    const GwTask* tasks;       // this launch's tasks
    const GwInst* inst;
    const double* const* tab;  // [3][count]: inputs, outputs or cotangents, dA
    long long count;
    int L;                     // segment length (a multiple of 4)
};
bool gw_ok(int d, int integ, int n);
hipError_t launch_gw_forward(int d, const GwArgs& a, int ntasks, double bytes, double flops, hipStream_t s);
hipError_t launch_gw_reverse(int d, const GwArgs& a, int ntasks, double bytes, double flops, hipStream_t s);

// Single-pass basis sweep of the dense W-basis path at d = 8 under the Boole
// rule (kernels_wbd8.hip): one wavefront per two instances walks the samples
// once and writes the sums Z_0 (omega1), Z_1..Z_K (partial rows 1..K, row
// stride n) and the totals What_1..What_K (totals rows 0..K-1), K = n - 1 <=
// kWbd8MaxK; with carry != nullptr it also writes the levels W_1..W_K
// ([B][K][S][64]) for the reverse sweep.  Uses in, carry, partial, omega1,
// totals, S, jobs (= instances), n, d, dt of LmArgs.
constexpr int kWbd8MaxK = 7;
constexpr int kWbd8MaxModes = 8;  // Fourier-formed samples: at most this many modes
bool wbd8_sweep_ok(int d, int integ, int n);
hipError_t launch_wbd8_sweep(const LmArgs& a, hipStream_t s);

// Reverse single-pass sweep of the same problems (kernels_wbd8.hip): from the
// kept levels W_1..W_K and the fold's cotangents of the sums (zbar rows 0..K)
// and totals (wbar rows 0..K-1), one wavefront per two instances walks the
// samples from the last to the first and writes the whole dA (no per-level
// arrays).
struct Wbd8RevArgs {
    const double* in;    // [B][S][64], or nullptr with fcoef set
    const double* fcoef; // [B][2 fmodes + 1][64]: the samples formed from Fourier coefficients, or nullptr
    const double* ftab;  // [S][fmodes][2]
    int fmodes;
    const double* W;       // [B][K / 2][S][64]: the even levels W_2, W_4, ...
    const double* totals;  // [B][n][64]: rows 0..K-1 are What_1..What_K, from which the reverse rebuilds the odd levels
    const double* zbar;  // [B][n][64]
    const double* wbar;  // [B][n][64]
    double* dA;          // [B][S][64]
    long long B, S;
    int n;
    double dt;
};
hipError_t launch_wbd8_rev(const Wbd8RevArgs& a, hipStream_t s);

// Dense W-basis GL fold (kernels_wbd.hip).  The level kernels run one job per
// instance with shift 0 and weight 1, giving the basis streams W_j = Prefix(A W_{j-1})
// (W_1 = Prefix(A)), their totals What_j and the sums Z_l = Sum(A W_l) (Z_0 = Sum(A)).
// Per GL node q: T_1 = What_1, T_j = What_j + s_q sum_{l<j} What_l T_{j-l};
// Omega_1 = Z_0, Omega_j = sum_q w_q (Z_{j-1} + s_q sum_{l<=j-2} Z_l T_{j-1-l}).
struct WbdArgs {
    const double* z;       // [B][n][m]  rows 1..K: Z_1..Z_K (row 0 unused)
    const double* z0;      // [B][m]     Z_0
    const double* totals;  // [B][n][m]  rows 0..K-1: What_1..What_K
    const double* cot;     // [B][n_out][m]              (VJP)
    double* out;           // [B][n_out][m]              (forward)
    double* zbar;          // [B][n][m]  rows 0..K: Zbar_0..Zbar_K      (VJP)
    double* wbar;          // [B][n][m]  rows 0..K-1: What_bar_1..K     (VJP)
    double* part;          // [B][Q][n][m] node partials: Omega_j (forward), Zbar_l (VJP)
    double* part2;         // [B][Q][n][m] node partials of What_bar_j  (VJP)
    double* work;          // [B][Q][2K+1][m] per-node scratch (T, T_bar)
    long long B;
    long long ostride;     // doubles between instances in out (0: n_out * m)
    int n, d, m, Q, op, n_out;
    int rows;              // row stride of z / totals / zbar / wbar (0: n); rows >= n
    double w[kArgMaxQ];
    double x[kArgMaxQ];
    // all-orders fold (launch_wbd_fold_orders): node f of the flattened GL
    // tables of orders lo..n has weight w[f], node x[f] and order korder[f];
    // part / work rows are per (instance, f); out row k-1 gets order k's sum
    int nflat, lo, kout;  // kout: the order of out's row 0
    int korder[kArgMaxQ];
};
hipError_t launch_wbd_fold(const WbdArgs& a, hipStream_t s);      // -> out
hipError_t launch_wbd_fold_orders(const WbdArgs& a, hipStream_t s);  // -> out rows lo-kout..n-kout (op sum)
hipError_t launch_wbd_fold_vjp(const WbdArgs& a, hipStream_t s);  // -> zbar, wbar

}  // namespace magnus
