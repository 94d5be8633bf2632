/*
 * magnus_oracle.h -- CPU restatement of the reference Magnus hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it may be loaded by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and by
 * nothing else.  The product (magnus-dp_amd/) never links, loads or calls it.
 *
 * Every routine follows the reference's own operation order (file:line cited at
 * each definition in magnus_oracle.c), in FP64, so that it stands in for the
 * reference CPU path, which cannot be compiled in this image.
 *
 * Enumerations follow the reference registries:
 *   integrators   integration/backends.hpp:53-58  {Boole, Simpson, Trapezoid, Riemann}
 *   matrix        linalg/backends.hpp:63-68        {Fixed2, Manual, Blas, SpaceCurve}
 *   ops           composer/dispatch.hpp:285-293    {one, many, sum}
 */
#ifndef MAGNUS_ORACLE_H
#define MAGNUS_ORACLE_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_BOOLE = 0, OR_SIMPSON = 1, OR_TRAPEZOID = 2, OR_RIEMANN = 3 };
enum { OR_FIXED2 = 0, OR_MANUAL = 1, OR_BLAS = 2, OR_SPACECURVE = 3 };
enum { OR_ONE = 0, OR_MANY = 1, OR_SUM = 2 };

/* Gauss-Legendre table on [0,1] (own Newton iteration on P_n). */
int oracle_gl_init(int max_order);
int oracle_gl_max_order(void);
int oracle_gl_get(int order, double* weights, double* nodes);
int oracle_gl_replace(int max_order, const double* weights, const double* nodes);

/* Integrator primitives on a span of `len` samples with `m` scalars each. */
int oracle_prefix(int integ, size_t len, size_t m, double dt, double* A);
int oracle_sum(int integ, size_t len, size_t m, double dt, const double* A, double* out);
int oracle_prefix_vjp(int integ, size_t len, size_t m, double dt, double* A);
int oracle_sum_vjp(int integ, size_t len, size_t m, double dt, double* A, const double* out);
int oracle_sum_vjp_add(int integ, size_t len, size_t m, double dt, double* A, const double* out);

/* Per-sample product update and its VJP (matrix policy chosen by `mat`). */
int oracle_sample_update(int mat, size_t dim, size_t len, const double* A, double* Y,
                         const double* total, double shift);
int oracle_sample_update_vjp(int mat, size_t dim, size_t len, double* dA, const double* A,
                             double* barY, const double* prefix, double shift);

/* Drivers.  `lanes` = 1 reproduces GL_forloop; lanes > 1 reproduces GL_openmp. */
int oracle_forward(int n, int op, int mat, int integ, int lanes, size_t samples, size_t dim,
                   double t0, double tf, const double* in, double* out, double* carry);
int oracle_vjp(int n, int op, int mat, int integ, size_t samples, size_t dim, double t0,
               double tf, const double* in, const double* cot, const double* carry, double* dA);
int oracle_sc_forward(int n, int op, int integ, int lanes, size_t samples, double t0, double tf,
                      const double* in3, double* out3, double* carry);
int oracle_sc_vjp(int n, int op, int integ, size_t samples, double t0, double tf,
                  const double* in3, const double* cot3, const double* carry, double* dA3);

/* Batched drivers for CPU timing: `count` independent instances packed
 * contiguously; instances are spread over `threads` OpenMP threads. */
int oracle_forward_batch(int n, int op, int mat, int integ, int lanes, int threads, size_t count,
                         size_t samples, size_t dim, double t0, double tf, const double* in,
                         double* out, double* carry);
int oracle_vjp_batch(int n, int op, int mat, int integ, int threads, size_t count, size_t samples,
                     size_t dim, double t0, double tf, const double* in, const double* cot,
                     const double* carry, double* dA);
int oracle_sc_forward_batch(int n, int op, int integ, int lanes, int threads, size_t count,
                            size_t samples, double t0, double tf, const double* in3, double* out3,
                            double* carry);
int oracle_sc_vjp_batch(int n, int op, int integ, int threads, size_t count, size_t samples,
                        double t0, double tf, const double* in3, const double* cot3,
                        const double* carry, double* dA3);

const char* oracle_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
