/*
 * magnus_capi.h -- C ABI of the MI355X (gfx950) Magnus hot path.
 *
 * Library: magnus-dp_amd/magnus/libmagnus_hip.so.  All compute entry points
 * take DEVICE pointers (HIP device memory of the current device) and an
 * optional hipStream_t passed as `void*` (NULL = the default stream).  Work is
 * enqueued on that stream; the call returns after enqueueing (asynchronous),
 * except that argument errors are reported synchronously.  FP64 only.
 *
 * The entry points replace the reference's composer boundary:
 *   magnus_forward_f64     <- detail::run_raw_impl              (src/composer/backends.hpp:182-220)
 *   magnus_vjp_f64         <- detail::run_vjp_raw_impl          (src/composer/backends.hpp:222-251)
 *   magnus_sc_forward_f64  <- detail::run_spacecurve_raw_impl   (src/composer/backends.hpp:253-301)
 *   magnus_sc_vjp_f64      <- detail::run_spacecurve_vjp_raw_impl (src/composer/backends.hpp:303-357)
 * with one addition: `count` independent instances of the same problem,
 * packed contiguously, are solved by one call (the reference solves one).
 * count == 1 is exactly the reference call.
 *
 * Layouts (row-major, sample-major, as in the reference):
 *   matrix input     in    [count][samples][dim][dim]
 *   SpaceCurve input in    [count][samples][3]
 *   output (one/sum) out   [count][dim][dim]          SC: [count][3]
 *   output (many)    out   [count][n][dim][dim]       SC: [count][n][3]
 *   carry            carry [count][Q][n-1][samples][dim][dim]   Q = (n+3)/2
 *                          (SC: dim = 2, i.e. 4 scalars per sample)
 *                          src/graddata.hpp:27-41, src/composer/backends.hpp:104-106
 *   cotangent        same shape as the output
 *   dA               same shape as the input
 *
 * Error contract: a nonzero status is returned and magnus_last_error() (a
 * thread-local string) holds the reference's message text for the same
 * condition, e.g. "invalid dispatch option (based on params)".
 */
#ifndef MAGNUS_CAPI_H
#define MAGNUS_CAPI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define MAGNUS_OK 0
#define MAGNUS_EINVAL 1  /* std::invalid_argument in the reference -> ValueError */
#define MAGNUS_ERANGE 2  /* std::out_of_range -> IndexError                      */
#define MAGNUS_ENOMEM 3  /* device allocation failed                             */
#define MAGNUS_EDEVICE 4 /* HIP runtime / launch error                           */

/* kernel ops, src/composer/dispatch.hpp:285-293 */
#define MAGNUS_OP_ONE 0
#define MAGNUS_OP_MANY 1
#define MAGNUS_OP_SUM 2

/* "Auto" for every registry */
#define MAGNUS_AUTO (-1)

/* matrix backends, src/linalg/backends.hpp:63-68 */
#define MAGNUS_MAT_FIXED2 0
#define MAGNUS_MAT_MANUAL 1
#define MAGNUS_MAT_BLAS 2
#define MAGNUS_MAT_SPACECURVE 3

/* integrators, src/integration/backends.hpp:53-58 */
#define MAGNUS_INT_BOOLE 0
#define MAGNUS_INT_SIMPSON 1
#define MAGNUS_INT_TRAPEZOID 2
#define MAGNUS_INT_RIEMANN 3

/* GL lane executors, src/util/gausslegendre_backends.hpp:55-59 (validated only:
 * on the device the GL nodes are a grid dimension) */
#define MAGNUS_GL_THREADPOOL 0
#define MAGNUS_GL_OPENMP 1
#define MAGNUS_GL_SERIAL 2

/* Problem description; mirrors ParamsBase + the resolved dispatch indices
 * (src/composer/dispatch.hpp:30-36, src/composer/backends.cpp:65-102). */
typedef struct magnus_problem {
    uint64_t n;       /* Magnus order, >= 1                       */
    uint64_t dim;     /* matrix dimension (ignored by the SC calls) */
    uint64_t samples; /* uniform samples on [t0, tf], >= 2          */
    double t0;
    double tf;
    int32_t op;             /* MAGNUS_OP_*                          */
    int32_t matrix_backend; /* MAGNUS_MAT_* or MAGNUS_AUTO          */
    int32_t integrator;     /* MAGNUS_INT_* or MAGNUS_AUTO          */
    int32_t gl_backend;     /* MAGNUS_GL_*  or MAGNUS_AUTO          */
} magnus_problem;

/* ---- name resolution: type_list::resolve (dispatch.hpp:119-129) and
 *      op_from_str (dispatch.hpp:295-301).  "Auto" -> MAGNUS_AUTO. ---- */
int magnus_resolve_op(const char* name, int32_t* out);
int magnus_resolve_matrix_backend(const char* name, int32_t* out);
int magnus_resolve_integrator(const char* name, int32_t* out);
int magnus_resolve_gl_backend(const char* name, int32_t* out);

/* dispatch names in registry order, NULL past the end (pybind.cpp:38-55) */
const char* magnus_matrix_backend_name(int32_t idx);
const char* magnus_integrator_name(int32_t idx);
const char* magnus_gl_backend_name(int32_t idx);
const char* magnus_op_name(int32_t idx);
const char* magnus_numeric_backend_name(int32_t idx);

/* ---- Gauss-Legendre table: pybind.cpp:57-59 (max_order) and
 *      pybind.cpp:102-116 + util/gausslegendre.cpp:74-111 (replace).
 *      Host arrays, flattened by order: order o occupies o entries. ---- */
uint64_t magnus_max_order(void);
int magnus_replace_gl_table(uint64_t max_order, const double* weights, const double* nodes);
int magnus_gl_order(uint64_t order, double* weights, double* nodes);

/* ---- sizes (scalars per instance) ---- */
uint64_t magnus_carry_elems(const magnus_problem* p);    /* Q*(n-1)*S*dim*dim   */
uint64_t magnus_output_elems(const magnus_problem* p);   /* n_out*dim*dim       */
uint64_t magnus_sc_carry_elems(const magnus_problem* p); /* Q*(n-1)*S*4         */
uint64_t magnus_sc_output_elems(const magnus_problem* p);/* n_out*3             */

/* ---- validation only (no device work): the checks and messages the compute
 *      calls apply before touching memory.  spacecurve != 0 validates the
 *      SC entry points; forward != 0 also validates gl_backend. ---- */
int magnus_validate(const magnus_problem* p, int spacecurve, int forward);

/* ---- compute: device pointers; carry may be NULL ---- */

/* Forward.  carry != NULL records the VJP carry (record_vjp=True). */
int magnus_forward_f64(const magnus_problem* p, uint64_t count, const double* in, double* out,
                       double* carry, void* stream);

/* Reverse sweep.  carry == NULL recomputes the forward internally
 * (VJP::none, src/grad.hpp:14-56). */
int magnus_vjp_f64(const magnus_problem* p, uint64_t count, const double* in, const double* cot,
                   const double* carry, double* dA, void* stream);

/* SpaceCurve (su(2)) variants: (S,3) input padded to 4-vectors on the fly. */
int magnus_sc_forward_f64(const magnus_problem* p, uint64_t count, const double* in3,
                          double* out3, double* carry, void* stream);
int magnus_sc_vjp_f64(const magnus_problem* p, uint64_t count, const double* in3,
                      const double* cot3, const double* carry, double* dA3, void* stream);

/* Magnus sums of every order from one basis sweep (no counterpart in the
 * reference, which solves one order per call): out [count][n][dim][dim], row
 * k-1 = the op="sum" output of order k = 1..p->n (each with its own GL nodes),
 * the same values magnus_forward_f64 returns for order k and op "sum" up to
 * the reassociation of the dense W-basis path.  p->op is ignored; matrix
 * problems only.  For resolution control (choosing n per instance). */
int magnus_forward_orders_f64(const magnus_problem* p, uint64_t count, const double* in, double* out, void* stream);

/* The same for orders first..p->n only (1 <= first <= n): out [count][n-first+1][dim][dim],
 * row k-first = order k.  A resolution judge that compares orders n and n+1 folds two
 * orders instead of all of them. */
int magnus_forward_orders_from_f64(const magnus_problem* p, uint64_t count, uint64_t first, const double* in,
                                   double* out, void* stream);

/* ---- compact checkpoints of the single-pass SpaceCurve path (no counterpart in
 *      the reference, whose saved state is the full carry above).  The forward
 *      keeps its per-chunk summaries (2(n-1)+1 4-vectors per chunk of samples,
 *      magnus_sc_checkpoint_elems doubles for `count` instances); a VJP given them
 *      skips its forward sweep and returns the same dA as magnus_sc_vjp_f64 with
 *      carry == NULL.  For the autograd front end (torch_ops.magnus_sc).  The
 *      checkpoint is valid only with the same problem, count and input. ---- */

/* 0 when the problem does not take the single-pass path (Trapezoid, n = 1, n > 8) */
uint64_t magnus_sc_checkpoint_elems(const magnus_problem* p, uint64_t count);
int magnus_sc_forward_ckpt_f64(const magnus_problem* p, uint64_t count, const double* in3, double* out3,
                               double* ckpt, void* stream);
int magnus_sc_vjp_ckpt_f64(const magnus_problem* p, uint64_t count, const double* in3, const double* cot3,
                           const double* ckpt, double* dA3, void* stream);

/* ---- kept basis state of the dense W-basis path (no counterpart in the
 *      reference, whose saved state is the full carry above).  Matrix problems
 *      with op "many" or "sum" and n >= 2 (op "one" and n = 1 return 0 /
 *      EINVAL) solve through one basis sweep -- levels W_1..W_{n-1}, their
 *      totals and the sums -- and a GL fold.  The forward can keep that
 *      state, (n-1)*S*d*d + (2n+1)*d*d doubles per instance against the
 *      carry's Q*(n-1)*S*d*d (d = 8 under the Boole rule with n <= 8: the
 *      (n-1)/2 even levels, the odd ones being rebuilt from their totals),
 *      and a VJP given it skips its own sweep; its dA is
 *      the same as magnus_vjp_f64 with carry == NULL (same kernels on the same
 *      state).  The state is valid only with the same problem, count and input. ---- */
uint64_t magnus_basis_state_elems(const magnus_problem* p, uint64_t count);
int magnus_forward_basis_f64(const magnus_problem* p, uint64_t count, const double* in, double* out, double* state,
                             void* stream);
int magnus_vjp_basis_f64(const magnus_problem* p, uint64_t count, const double* in, const double* cot,
                         const double* state, double* dA, void* stream);

/* ---- dense VJP with time-segment checkpoints (no counterpart in the
 *      reference; north-star "checkpointed prefix carries in place of full
 *      vjp_data when memory requires it").  The VJP of magnus_vjp_f64 with
 *      carry == NULL (VJP::none, src/grad.hpp:95-110 with
 *      src/composer/backends.hpp:222-251) for dense problems of dim 16, 32 or 64
 *      under the Boole rule with op "many" or "sum" and n >= 2 (EINVAL
 *      otherwise), keeping the basis levels for one segment of samples at a
 *      time: the first pass keeps each level's prefix state at every segment
 *      start, the reverse rebuilds one segment's levels from it and walks them
 *      back.  segments = 0 picks the segment length that balances the segment
 *      buffers against the checkpoints (about 4-6 % of the n - 1 full levels at
 *      S = 4097 and 8193); any count of segments gives the same dA bit for bit.
 *      magnus_vjp_segmented_scratch_bytes: the device bytes one call allocates
 *      (0 where the path does not apply); magnus_vjp_segment_length: the
 *      segment length in samples that a segment count gives. ---- */
uint64_t magnus_vjp_segmented_scratch_bytes(const magnus_problem* p, uint64_t count, uint64_t segments);
uint64_t magnus_vjp_segment_length(const magnus_problem* p, uint64_t segments);
int magnus_vjp_segmented_f64(const magnus_problem* p, uint64_t count, const double* in, const double* cot, double* dA,
                             uint64_t segments, void* stream);

/* ---- Fourier-form input (no counterpart in the reference, whose samples come
 *      from a Python callable evaluated on the grid, src/api/magnus/__init__.py:
 *      58-85).  The basis-state forward, the VJP from that state and the
 *      all-orders forward of dim-8 problems under the Boole rule (2 <= n <= 8)
 *      for A(t) = C_0 + sum_{k=1..modes} (sin(2 pi k t) C_{2k-1} + cos(2 pi k t)
 *      C_{2k}), coeffs [count][2 modes + 1][8][8], modes <= 8 (ERANGE above):
 *      the single-pass sweeps form each sample from the coefficients instead of
 *      reading a sample array, so nothing is sampled into HBM.  Same sin/cos
 *      table and fused multiply-adds as magnus_sample_fourier_f64, so the
 *      results are those of sampling first and calling magnus_forward_basis_f64 /
 *      magnus_vjp_basis_f64 / magnus_forward_orders_from_f64, bit for bit.
 *      Other problems: EINVAL. ---- */
int magnus_fourier_forward_basis_f64(const magnus_problem* p, uint64_t count, const double* coeffs, uint64_t modes,
                                     double* out, double* state, void* stream);
int magnus_fourier_vjp_basis_f64(const magnus_problem* p, uint64_t count, const double* coeffs, uint64_t modes,
                                 const double* cot, const double* state, double* dA, void* stream);
int magnus_fourier_forward_orders_from_f64(const magnus_problem* p, uint64_t count, uint64_t first,
                                           const double* coeffs, uint64_t modes, double* out, void* stream);

/* ---- SpaceCurve VJP that allocates nothing.  No counterpart in the reference:
 *      for callers that cannot allocate device memory (code under HIP stream
 *      capture, tight serving loops).  Same arguments, kernels and results as
 *      magnus_sc_vjp_f64 with carry == NULL on the single-pass path (su(2)
 *      input, Boole/Simpson/Riemann, 2 <= n <= 8; other problems return
 *      EINVAL).  Its scratch is a fixed 128 MiB device array that each GPU the
 *      library runs on holds until the process ends (the 4096-trajectory,
 *      16385-sample, n = 6 batch needs 120 MB); a batch needing more returns
 *      ERANGE.  Calls that overlap on one GPU share that array and race. ---- */
int magnus_sc_vjp_noalloc_f64(const magnus_problem* p, uint64_t count, const double* in3, const double* cot3,
                              double* dA3, void* stream);

/* ---- sampling A(t) on the device, on the reference's grid
 *      t_i = np.linspace(t0, tf, samples) (replaces evaluating a Python callable
 *      there, src/api/magnus/__init__.py:58-85, for these two families).
 *      The grid and the sin/cos arguments follow numpy's order and rounding;
 *      the matrix form adds its terms in k order with one fused multiply-add
 *      each, so samples differ from a numpy evaluation through sin/cos and
 *      those roundings (a few ulp).  modes <= 32. ---- */

/* out[b][i][e] = C[b][0][e] + sum_{k=1..modes} ( sin(2 pi k t_i) C[b][2k-1][e]
 *                                               + cos(2 pi k t_i) C[b][2k][e] ),
 * coeffs [count][2*modes+1][entries], out [count][samples][entries]; entries = d*d. */
int magnus_sample_fourier_f64(uint64_t count, uint64_t samples, double t0, double tf, uint64_t entries,
                              uint64_t modes, const double* coeffs, double* out, void* stream);

/* out[b][i][c] = sum_{k=1..modes} amp[b][k-1][c] sin(2 pi k t_i + phase[b][k-1][c]),
 * amp/phase [count][modes][3], out [count][samples][3] (SpaceCurve input). */
int magnus_sample_sc_fourier_f64(uint64_t count, uint64_t samples, double t0, double tf, uint64_t modes,
                                 const double* amp, const double* phase, double* out, void* stream);

/* ---- propagators exp(Omega) of Magnus outputs (the step after the path: the
 *      reference's tests/bench.py:93 applies scipy.linalg.expm to the Magnus sum). ---- */

/* Dense: out[b] = expm(in[b]) for real d x d matrices, in/out [count][d][d].
 * Scaling and squaring with Pade degree 3..13 (Higham 2005). */
int magnus_expm_f64(uint64_t count, uint64_t dim, const double* in, double* out, void* stream);

/* SpaceCurve: omega (3-vector) stands for -i omega.sigma; out[b] = exp(-i omega.sigma)
 * as a complex 2x2 matrix, [count][2][2][re, im]. */
int magnus_sc_expm_f64(uint64_t count, const double* in3, double* out, void* stream);

/* ---- uneven batches (no counterpart in the reference, which solves one problem
 *      per call).  A plan holds `count` matrix problems of any dim, samples,
 *      order, op and integrator.  Dense problems of dim 16, 32 or 64 under the
 *      Boole rule with op "many" or "sum" and n >= 2 run as one grouped
 *      pipeline: each is cut into segments of samples, and one launch per
 *      pipeline step and dim runs one level of one segment of every such
 *      instance that the step reaches (grouped walkers; the GL fold once per
 *      (dim, order, op) group).  They join it in order of their chain of steps
 *      (samples + (n - 1) 128, weighted by dim^3), shortest first, while their
 *      basis levels fit in the workspace left after a reserve for the branches
 *      (an eighth, and at least one instance per branch of the largest bucket
 *      left); the rest, the longest chains, run in buckets.  The other problems
 *      are grouped into buckets of equal shape, each bucket split into
 *      sub-batches that fit its workspace and spread over `branches` HIP
 *      streams (longest first), each sub-batch a gather of the instances'
 *      buffers, the level-parallel kernels and a scatter of the results.  With
 *      use_graph != 0 the whole sequence is captured once into a hipGraph and
 *      each call is one graph launch.  Buffers are per-instance DEVICE pointers
 *      passed as HOST arrays of `count` pointers (layouts as above with
 *      count = 1).  Bucketed instances give the results of per-instance
 *      magnus_forward_f64 / magnus_vjp_f64 calls (carry == NULL) with the
 *      plan's product path per bucket, bit for bit; pipeline instances agree
 *      with them to rounding (the same recurrences, products on the matrix
 *      cores in another summation order) and do not depend on the rest of the
 *      batch.  Calls on one plan must not overlap (one stream). ---- */
typedef struct magnus_uneven_plan magnus_uneven_plan;

/* workspace_bytes: device scratch for all branches (0: half of the free device
 * memory); branches: 1..8 concurrent streams.  EINVAL names the first invalid
 * problem ("instance i: ..."); ENOMEM if one instance does not fit. */
int magnus_uneven_create(const magnus_problem* probs, uint64_t count, uint64_t workspace_bytes, int branches,
                         magnus_uneven_plan** out);
void magnus_uneven_destroy(magnus_uneven_plan* plan);
int magnus_uneven_forward_f64(magnus_uneven_plan* plan, const double* const* in, double* const* out,
                              int use_graph, void* stream);
int magnus_uneven_vjp_f64(magnus_uneven_plan* plan, const double* const* in, const double* const* cot,
                          double* const* dA, int use_graph, void* stream);

/* Autotuner: for every bucket, times the VJP (forward recompute + reverse
 * sweep) of its first sub-batch under each product path worth trying for its
 * dim (dim 16: FP64 MFMA or VALU row products; dim 32: MFMA with 1, 2 or 4
 * wavefronts per product) on the given buffers, keeps the fastest, and
 * rebalances the branches with the measured times; for the grouped pipeline,
 * times its whole VJP with segments of 64, 128 and 256 samples and keeps the
 * fastest.  Synchronizes the stream. */
int magnus_uneven_autotune(magnus_uneven_plan* plan, const double* const* in, const double* const* cot,
                           double* const* dA, void* stream);

/* Tuning table as text: "pipeline L ms" (the grouped pipeline's segment length in
 * samples and its measured VJP time) when the plan has a pipeline, then one bucket
 * per line: "dim samples n tile ms_per_instance".
 * magnus_uneven_tuning writes at most `len` bytes (NUL-terminated) and returns
 * the full length; magnus_uneven_set_tuning applies the lines whose
 * (dim, samples, n) match a bucket (tile: 0 default, 1 VALU, 2 / 3 MFMA with
 * 2 / 4 wavefronts per product at dim 32) and a "pipeline L" line (L a multiple
 * of 4 up to 256). */
uint64_t magnus_uneven_tuning(const magnus_uneven_plan* plan, char* buf, uint64_t len);
int magnus_uneven_set_tuning(magnus_uneven_plan* plan, const char* text);

/* plan facts: 0 buckets, 1 sub-batches, 2 branches in use, 3 workspace bytes,
 * 4 instances in the grouped pipeline, 5 pipeline waves */
uint64_t magnus_uneven_info(const magnus_uneven_plan* plan, int what);
/* 1 if instance i runs in the grouped pipeline, 0 if in a bucket (or out of range) */
int magnus_uneven_in_pipeline(const magnus_uneven_plan* plan, uint64_t i);

/* ---- resolution controller (no counterpart in the reference).  Features of a
 *      coarse sampling and a 3-layer MLP (in -> hidden -> hidden -> out, ReLU)
 *      trained on the device: FP32 master parameters, bfloat16 weight copies
 *      read by the layer kernels, FP32 activations and accumulation, Adam
 *      (beta 0.9 / 0.999, eps 1e-8).  x, logits, labels and loss_rows are
 *      device pointers; parameter and gradient copies are host arrays laid out
 *      W1 [hidden][in], b1, W2 [hidden][hidden], b2, W3 [out][hidden], b3.
 *      Calls on one network must not overlap. ---- */

/* features [count][magnus_ctrl_feature_count()] (float) of A sampled on a
 * grid of 2^k + 1 points (9..129) of [t0, tf], in [count][grid][dim][dim], dim <= 16 */
uint64_t magnus_ctrl_feature_count(void);
int magnus_ctrl_features_f64(uint64_t count, uint64_t grid, uint64_t dim, double t0, double tf, const double* in,
                             float* out, void* stream);

typedef struct magnus_mlp magnus_mlp;
int magnus_mlp_create(uint32_t in_dim, uint32_t hidden, uint32_t out_dim, uint64_t max_batch, uint64_t seed,
                      magnus_mlp** out);
void magnus_mlp_destroy(magnus_mlp* net);
uint64_t magnus_mlp_param_count(const magnus_mlp* net);
int magnus_mlp_get_params(const magnus_mlp* net, float* host, uint64_t len);
/* also refreshes the bfloat16 copies and resets the optimizer state */
int magnus_mlp_set_params(magnus_mlp* net, const float* host, uint64_t len);
/* gradients of the last training step (before its parameter update) */
int magnus_mlp_get_grads(const magnus_mlp* net, float* host, uint64_t len);
int magnus_mlp_forward(magnus_mlp* net, uint64_t batch, const float* x, float* logits, void* stream);
/* softmax cross-entropy (mean over the batch) on int32 class labels, backward,
 * one Adam step at learning rate lr; loss_rows [batch] gets each row's loss */
int magnus_mlp_train_step(magnus_mlp* net, uint64_t batch, const float* x, const int32_t* labels, float lr,
                          float* loss_rows, void* stream);

/* ---- kernel timing (diagnostics).  While enabled, every main kernel launch
 *      is bracketed by a HIP event pair recorded on its launch stream.
 *      magnus_timing_read waits for the recorded events and returns, for the
 *      kernel family `name` (e.g. "wb_rev", "lm_fwd_prod"), the number of
 *      launches and their total duration in milliseconds since the last reset. ---- */
int magnus_timing_enable(int on);
int magnus_timing_read(const char* name, uint64_t* launches, double* total_ms);
/* the algorithmic HBM bytes and flops of the same launches (FMA = 2 flops), as
 * their launch sites state them (0 where a site states none); with the
 * duration from magnus_timing_read this gives a kernel family's achieved rate
 * when its launches differ in size.  No counterpart in the reference. */
int magnus_timing_counts(const char* name, double* bytes, double* flops);
void magnus_timing_reset(void);

/* Library scratch comes from a per-(device, stream) cache of device blocks that
 * grows to the largest call seen and is kept for reuse (no allocation on the
 * steady-state path).  This waits for the devices and frees every cache. */
int magnus_release_scratch(void);

/* Bound on the bytes a call may take from its stream's scratch cache (0: no
 * bound, the default), for sharing the device with the caller's own memory.
 * A call that would pass it fails with ENOMEM, except the carry-free dense VJP
 * (magnus_vjp_f64 with carry == NULL), which then keeps its basis levels for
 * one time segment at a time where magnus_vjp_segmented_f64 applies.  The
 * bound counts the cache's cursor position, so memory the cache already holds
 * beyond it stays unused.  No counterpart in the reference. */
int magnus_set_scratch_limit(uint64_t bytes);
uint64_t magnus_scratch_limit(void);

/* thread-local message of the last failing call on this thread */
const char* magnus_last_error(void);

/* name of the kernel family the last successful call on this thread used
 * (diagnostics: "wb" single-pass SpaceCurve, "lm" level-parallel) */
const char* magnus_last_kernel(void);

#ifdef __cplusplus
}
#endif

#endif /* MAGNUS_CAPI_H */
