/*
 * magnus_oracle.c -- CPU restatement of the reference Magnus hot path (FP64).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product and the CPU
 * baseline timed beside it.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load this library.
 *
 * Each function restates the reference routine named in its comment, keeping
 * the reference's loop order and its order of floating-point operations
 * (scale first, then the listed adds, base last), so that the only expected
 * differences from the reference binary are FMA-contraction choices of the
 * host compiler.  Citations are relative to the reference repository root.
 *
 * Pins (tests/test_oracle_*.py): GL table vs numpy leggauss; integrator
 * exactness on polynomials; adjoint identities <L x, y> = <x, L^T y> for every
 * *_vjp; Omega_2 closed form for A(t) = A0 + t A1; commuting families;
 * SpaceCurve == complex 2x2 product form; VJP vs central finite differences;
 * carry == recompute; one == many[n-1] and sum == sum(many).
 */
#include "magnus_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static __thread char g_err[256];

static int fail(const char* msg) {
    snprintf(g_err, sizeof g_err, "%s", msg);
    return 1;
}

const char* oracle_last_error(void) { return g_err; }

/* ------------------------------------------------------------------------ */
/* Gauss-Legendre table.  Reference: util/gen_nodes.py:11-47 builds nodes and  */
/* weights with numpy leggauss, mapped x -> (x+1)/2, w -> w/2, ascending.      */
/* Here: Newton iteration on P_N from Chebyshev-like starting points.          */
/* ------------------------------------------------------------------------ */
#define GL_CAP 128
/* The table is process-wide and built lazily on first use; callers may drive */
/* the oracle from several threads at once (one instance per thread), so     */
/* every read and every swap of the table holds g_gl_mu.                      */
static pthread_mutex_t g_gl_mu = PTHREAD_MUTEX_INITIALIZER;
static double* g_gl_w = NULL;
static double* g_gl_x = NULL;
static int g_gl_max = 0;

static size_t gl_offset(int order) { return (size_t)(order - 1) * (size_t)order / 2; }

static void legendre_nodes(int N, double* x_out, double* w_out) {
    /* roots of P_N in (-1,1), ascending */
    for (int i = 0; i < N; ++i) {
        double x = cos(M_PI * (i + 0.75) / (N + 0.5));
        double dp = 1.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = x;
            for (int k = 2; k <= N; ++k) {
                double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            double pn = (N == 1) ? x : p1;
            double pnm1 = (N == 1) ? 1.0 : p0;
            dp = N * (x * pn - pnm1) / (x * x - 1.0);
            double dx = pn / dp;
            x -= dx;
            if (fabs(dx) < 1e-17) break;
        }
        /* final derivative at converged x */
        {
            double p0 = 1.0, p1 = x;
            for (int k = 2; k <= N; ++k) {
                double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            double pn = (N == 1) ? x : p1;
            double pnm1 = (N == 1) ? 1.0 : p0;
            dp = N * (x * pn - pnm1) / (x * x - 1.0);
        }
        double w = 2.0 / ((1.0 - x * x) * dp * dp);
        /* i runs from the largest root down; store ascending */
        x_out[N - 1 - i] = x;
        w_out[N - 1 - i] = w;
    }
    /* symmetrise as numpy does (x -> (x - x[::-1])/2, w -> (w + w[::-1])/2) */
    for (int i = 0; i < N / 2; ++i) {
        double a = 0.5 * (x_out[i] - x_out[N - 1 - i]);
        x_out[i] = a;
        x_out[N - 1 - i] = -a;
        double b = 0.5 * (w_out[i] + w_out[N - 1 - i]);
        w_out[i] = b;
        w_out[N - 1 - i] = b;
    }
    if (N % 2 == 1) x_out[N / 2] = 0.0;
}

/* Installs a freshly built table (w, x own their memory) and frees the old one. */
static void gl_swap(int max_order, double* w, double* x) {
    pthread_mutex_lock(&g_gl_mu);
    double* old_w = g_gl_w;
    double* old_x = g_gl_x;
    g_gl_w = w;
    g_gl_x = x;
    g_gl_max = max_order;
    pthread_mutex_unlock(&g_gl_mu);
    free(old_w);
    free(old_x);
}

int oracle_gl_init(int max_order) {
    if (max_order < 1 || max_order > GL_CAP) return fail("max_order out of range");
    size_t total = gl_offset(max_order + 1);
    double* w = (double*)malloc(total * sizeof(double));
    double* x = (double*)malloc(total * sizeof(double));
    if (!w || !x) { free(w); free(x); return fail("out of memory"); }
    for (int N = 1; N <= max_order; ++N) {
        double* xo = x + gl_offset(N);
        double* wo = w + gl_offset(N);
        legendre_nodes(N, xo, wo);
        for (int i = 0; i < N; ++i) {
            xo[i] = (xo[i] + 1.0) / 2.0;
            wo[i] = wo[i] / 2.0;
        }
    }
    gl_swap(max_order, w, x);
    return 0;
}

int oracle_gl_replace(int max_order, const double* weights, const double* nodes) {
    /* reference: api/pybind.cpp:102-116 (flattened, order-contiguous) */
    if (max_order < 1) return fail("max_order must be at least 1");
    size_t total = gl_offset(max_order + 1);
    double* w = (double*)malloc(total * sizeof(double));
    double* x = (double*)malloc(total * sizeof(double));
    if (!w || !x) { free(w); free(x); return fail("out of memory"); }
    memcpy(w, weights, total * sizeof(double));
    memcpy(x, nodes, total * sizeof(double));
    gl_swap(max_order, w, x);
    return 0;
}

/* Largest order of the table, building the default (orders 1..64) if none is */
/* installed yet; 0 if that build fails. */
static int gl_ready(void) {
    pthread_mutex_lock(&g_gl_mu);
    int have = g_gl_w != NULL, mx = g_gl_max;
    pthread_mutex_unlock(&g_gl_mu);
    if (have) return mx;
    if (oracle_gl_init(64)) return 0;
    pthread_mutex_lock(&g_gl_mu);
    mx = g_gl_max;
    pthread_mutex_unlock(&g_gl_mu);
    return mx;
}

int oracle_gl_max_order(void) { return gl_ready(); }

int oracle_gl_get(int order, double* weights, double* nodes) {
    if (!gl_ready()) return 1;
    pthread_mutex_lock(&g_gl_mu);
    if (order < 1 || order > g_gl_max) {
        pthread_mutex_unlock(&g_gl_mu);
        return fail("Gauss-Legendre order is outside the available table");
    }
    memcpy(weights, g_gl_w + gl_offset(order), (size_t)order * sizeof(double));
    memcpy(nodes, g_gl_x + gl_offset(order), (size_t)order * sizeof(double));
    pthread_mutex_unlock(&g_gl_mu);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Elementwise matrix helpers (linalg/manual.hpp:44-81)                        */
/* ------------------------------------------------------------------------ */
static inline void v_add(size_t m, double* a, const double* b, double x) {
    for (size_t i = 0; i < m; ++i) a[i] += b[i] * x;
}
static inline void v_add1(size_t m, double* a, const double* b) {
    for (size_t i = 0; i < m; ++i) a[i] += b[i] * 1.0;
}
static inline void v_scale(size_t m, double* a, double x) {
    for (size_t i = 0; i < m; ++i) a[i] *= x;
}
static inline void v_copy(size_t m, const double* s, double* d) { memcpy(d, s, m * sizeof(double)); }
static inline void v_zero(size_t m, double* d) { memset(d, 0, m * sizeof(double)); }

#define ROW(A, i) ((A) + (size_t)(i) * m)

/* ------------------------------------------------------------------------ */
/* Integrators                                                                 */
/* ------------------------------------------------------------------------ */

/* integration/boole.hpp:24-74 */
static void boole_prefix(size_t len, size_t m, double dt, double* A, double* scratch) {
    double* f0 = scratch;
    double* f1 = scratch + m;
    double* f2 = scratch + 2 * m;
    double* f3 = scratch + 3 * m;
    double* f4 = scratch + 4 * m;
    v_copy(m, ROW(A, 0), f0);
    v_zero(m, ROW(A, 0));
    size_t start = 0;
    while (start + 1 < len) {
        double* base = ROW(A, start);
        double* a1 = ROW(A, start + 1);
        double* a2 = ROW(A, start + 2);
        double* a3 = ROW(A, start + 3);
        double* a4 = ROW(A, start + 4);
        v_copy(m, a1, f1);
        v_scale(m, a1, dt / 2);
        v_add(m, a1, f0, dt / 2);
        v_add1(m, a1, base);

        v_copy(m, a2, f2);
        v_scale(m, a2, dt / 3);
        v_add(m, a2, f0, dt / 3);
        v_add(m, a2, f1, 4 * dt / 3);
        v_add1(m, a2, base);

        v_copy(m, a3, f3);
        v_scale(m, a3, 3 * dt / 8);
        v_add(m, a3, f0, 3 * dt / 8);
        v_add(m, a3, f1, 9 * dt / 8);
        v_add(m, a3, f2, 9 * dt / 8);
        v_add1(m, a3, base);

        v_copy(m, a4, f4);
        v_scale(m, a4, 14 * dt / 45);
        v_add(m, a4, f0, 14 * dt / 45);
        v_add(m, a4, f1, 64 * dt / 45);
        v_add(m, a4, f2, 24 * dt / 45);
        v_add(m, a4, f3, 64 * dt / 45);
        v_add1(m, a4, base);

        v_copy(m, f4, f0);
        start += 4;
    }
}

/* integration/boole.hpp:76-106 */
static void boole_sum(size_t len, size_t m, double dt, const double* A, double* out) {
    size_t intervals = len - 1;
    size_t boole_intervals = intervals - intervals % 4;
    for (size_t i = 0; i < boole_intervals; i += 4) {
        v_add(m, out, ROW(A, i), 14 * dt / 45);
        v_add(m, out, ROW(A, i + 1), 64 * dt / 45);
        v_add(m, out, ROW(A, i + 2), 24 * dt / 45);
        v_add(m, out, ROW(A, i + 3), 64 * dt / 45);
        v_add(m, out, ROW(A, i + 4), 14 * dt / 45);
    }
    size_t tail = intervals - boole_intervals;
    size_t s = boole_intervals;
    if (tail == 1) {
        v_add(m, out, ROW(A, s), dt / 2);
        v_add(m, out, ROW(A, s + 1), dt / 2);
    } else if (tail == 2) {
        v_add(m, out, ROW(A, s), dt / 3);
        v_add(m, out, ROW(A, s + 1), 4 * dt / 3);
        v_add(m, out, ROW(A, s + 2), dt / 3);
    } else if (tail == 3) {
        v_add(m, out, ROW(A, s), 3 * dt / 8);
        v_add(m, out, ROW(A, s + 1), 9 * dt / 8);
        v_add(m, out, ROW(A, s + 2), 9 * dt / 8);
        v_add(m, out, ROW(A, s + 3), 3 * dt / 8);
    }
}

/* integration/boole.hpp:108-151 */
static void boole_prefix_vjp(size_t len, size_t m, double dt, double* A, double* scratch) {
    double* pending = scratch;
    double* g1 = scratch + m;
    double* g2 = scratch + 2 * m;
    double* g3 = scratch + 3 * m;
    double* g4 = scratch + 4 * m;
    v_zero(m, pending);
    size_t start = len - 5;
    for (;;) {
        v_copy(m, ROW(A, start + 1), g1);
        v_copy(m, ROW(A, start + 2), g2);
        v_copy(m, ROW(A, start + 3), g3);
        v_copy(m, ROW(A, start + 4), g4);

        double* a0 = ROW(A, start);
        v_add1(m, a0, g1);
        v_add1(m, a0, g2);
        v_add1(m, a0, g3);
        v_add1(m, a0, g4);

        double* a1 = ROW(A, start + 1);
        v_scale(m, a1, dt / 2);
        v_add(m, a1, g2, 4 * dt / 3);
        v_add(m, a1, g3, 9 * dt / 8);
        v_add(m, a1, g4, 64 * dt / 45);

        double* a2 = ROW(A, start + 2);
        v_scale(m, a2, dt / 3);
        v_add(m, a2, g3, 9 * dt / 8);
        v_add(m, a2, g4, 24 * dt / 45);

        double* a3 = ROW(A, start + 3);
        v_scale(m, a3, 3 * dt / 8);
        v_add(m, a3, g4, 64 * dt / 45);

        double* a4 = ROW(A, start + 4);
        v_scale(m, a4, 14 * dt / 45);
        v_add1(m, a4, pending);

        v_copy(m, g1, pending);
        v_scale(m, pending, dt / 2);
        v_add(m, pending, g2, dt / 3);
        v_add(m, pending, g3, 3 * dt / 8);
        v_add(m, pending, g4, 14 * dt / 45);

        if (start == 0) break;
        start -= 4;
    }
    v_copy(m, pending, ROW(A, 0));
}

/* integration/boole.hpp:153-199 */
static void boole_sum_vjp(size_t len, size_t m, double dt, double* A, const double* out) {
    size_t intervals = len - 1;
    size_t boole_intervals = intervals - intervals % 4;
    if (boole_intervals == 0) {
        v_zero(m, ROW(A, 0));
    } else {
        v_copy(m, out, ROW(A, 0));
        v_scale(m, ROW(A, 0), 14 * dt / 45);
        for (size_t i = 0; i < boole_intervals; i += 4) {
            v_copy(m, out, ROW(A, i + 1));
            v_scale(m, ROW(A, i + 1), 64 * dt / 45);
            v_copy(m, out, ROW(A, i + 2));
            v_scale(m, ROW(A, i + 2), 24 * dt / 45);
            v_copy(m, out, ROW(A, i + 3));
            v_scale(m, ROW(A, i + 3), 64 * dt / 45);
        }
        for (size_t i = 4; i < boole_intervals; i += 4) {
            v_copy(m, out, ROW(A, i));
            v_scale(m, ROW(A, i), 28 * dt / 45);
        }
        v_copy(m, out, ROW(A, boole_intervals));
        v_scale(m, ROW(A, boole_intervals), 14 * dt / 45);
    }
    size_t tail = intervals - boole_intervals;
    size_t s = boole_intervals;
    if (tail == 1) {
        v_add(m, ROW(A, s), out, dt / 2);
        v_copy(m, out, ROW(A, s + 1));
        v_scale(m, ROW(A, s + 1), dt / 2);
    } else if (tail == 2) {
        v_add(m, ROW(A, s), out, dt / 3);
        v_copy(m, out, ROW(A, s + 1));
        v_scale(m, ROW(A, s + 1), 4 * dt / 3);
        v_copy(m, out, ROW(A, s + 2));
        v_scale(m, ROW(A, s + 2), dt / 3);
    } else if (tail == 3) {
        v_add(m, ROW(A, s), out, 3 * dt / 8);
        v_copy(m, out, ROW(A, s + 1));
        v_scale(m, ROW(A, s + 1), 9 * dt / 8);
        v_copy(m, out, ROW(A, s + 2));
        v_scale(m, ROW(A, s + 2), 9 * dt / 8);
        v_copy(m, out, ROW(A, s + 3));
        v_scale(m, ROW(A, s + 3), 3 * dt / 8);
    }
}

/* integration/boole.hpp:201-228 */
static void boole_sum_vjp_add(size_t len, size_t m, double dt, double* A, const double* out) {
    size_t intervals = len - 1;
    size_t boole_intervals = intervals - intervals % 4;
    for (size_t i = 0; i < boole_intervals; i += 4) {
        v_add(m, ROW(A, i), out, 14 * dt / 45);
        v_add(m, ROW(A, i + 1), out, 64 * dt / 45);
        v_add(m, ROW(A, i + 2), out, 24 * dt / 45);
        v_add(m, ROW(A, i + 3), out, 64 * dt / 45);
        v_add(m, ROW(A, i + 4), out, 14 * dt / 45);
    }
    size_t tail = intervals - boole_intervals;
    size_t s = boole_intervals;
    if (tail == 1) {
        v_add(m, ROW(A, s), out, dt / 2);
        v_add(m, ROW(A, s + 1), out, dt / 2);
    } else if (tail == 2) {
        v_add(m, ROW(A, s), out, dt / 3);
        v_add(m, ROW(A, s + 1), out, 4 * dt / 3);
        v_add(m, ROW(A, s + 2), out, dt / 3);
    } else if (tail == 3) {
        v_add(m, ROW(A, s), out, 3 * dt / 8);
        v_add(m, ROW(A, s + 1), out, 9 * dt / 8);
        v_add(m, ROW(A, s + 2), out, 9 * dt / 8);
        v_add(m, ROW(A, s + 3), out, 3 * dt / 8);
    }
}

/* integration/simpson.hpp:24-58 */
static void simpson_prefix(size_t len, size_t m, double dt, double* A, double* scratch) {
    double* tmp = scratch;
    double* prev2 = scratch + m;
    double* prev1 = scratch + 2 * m;
    double half_dt = dt / 2;
    double third_dt = dt / 3;
    v_copy(m, ROW(A, 0), prev2);
    v_zero(m, ROW(A, 0));
    v_copy(m, ROW(A, 1), prev1);
    v_scale(m, ROW(A, 1), half_dt);
    v_add(m, ROW(A, 1), prev2, half_dt);
    for (size_t i = 2; i < len; ++i) {
        v_copy(m, ROW(A, i), tmp);
        if (i % 2 == 0) {
            v_scale(m, ROW(A, i), third_dt);
            v_add(m, ROW(A, i), prev2, third_dt);
            v_add(m, ROW(A, i), prev1, 4 * third_dt);
            v_add1(m, ROW(A, i), ROW(A, i - 2));
        } else {
            v_scale(m, ROW(A, i), half_dt);
            v_add(m, ROW(A, i), prev1, half_dt);
            v_add1(m, ROW(A, i), ROW(A, i - 1));
        }
        v_copy(m, prev1, prev2);
        v_copy(m, tmp, prev1);
    }
}

/* integration/simpson.hpp:60-75 */
static void simpson_sum(size_t len, size_t m, double dt, const double* A, double* out) {
    size_t simpson_last = (len % 2 == 1) ? len - 1 : len - 2;
    v_add(m, out, ROW(A, 0), dt / 3);
    for (size_t i = 1; i < simpson_last; ++i)
        v_add(m, out, ROW(A, i), (i % 2 == 1) ? 4 * dt / 3 : 2 * dt / 3);
    v_add(m, out, ROW(A, simpson_last), dt / 3);
    if (simpson_last + 1 < len) {
        v_add(m, out, ROW(A, simpson_last), dt / 2);
        v_add(m, out, ROW(A, simpson_last + 1), dt / 2);
    }
}

/* integration/simpson.hpp:77-123 (std::swap of the two views == pointer swap) */
static void simpson_prefix_vjp(size_t len, size_t m, double dt, double* A, double* scratch) {
    double* bar = scratch;
    double* pending = scratch + m;
    double* next_pending = scratch + 2 * m;
    double half_dt = dt / 2;
    double third_dt = dt / 3;
    v_zero(m, pending);
    v_zero(m, next_pending);

#define REVERSE_ODD(I)                                   \
    do {                                                 \
        size_t ii = (I);                                 \
        v_copy(m, ROW(A, ii), bar);                      \
        v_add1(m, ROW(A, ii - 1), bar);                  \
        v_scale(m, ROW(A, ii), half_dt);                 \
        v_add1(m, ROW(A, ii), pending);                  \
        v_add(m, next_pending, bar, half_dt);            \
        v_zero(m, pending);                              \
        double* sw = pending; pending = next_pending; next_pending = sw; \
    } while (0)
#define REVERSE_EVEN(I)                                  \
    do {                                                 \
        size_t ii = (I);                                 \
        v_copy(m, ROW(A, ii), bar);                      \
        v_add1(m, ROW(A, ii - 2), bar);                  \
        v_scale(m, ROW(A, ii), third_dt);                \
        v_add1(m, ROW(A, ii), pending);                  \
        v_add(m, next_pending, bar, 4 * third_dt);       \
        v_copy(m, bar, pending);                         \
        v_scale(m, pending, third_dt);                   \
        double* sw = pending; pending = next_pending; next_pending = sw; \
    } while (0)

    size_t i = len - 1;
    if (i % 2 == 1) {
        REVERSE_ODD(i);
        --i;
    }
    while (i > 2) {
        REVERSE_EVEN(i);
        REVERSE_ODD(i - 1);
        i -= 2;
    }
    REVERSE_EVEN(2);
#undef REVERSE_ODD
#undef REVERSE_EVEN

    v_copy(m, ROW(A, 1), bar);
    v_scale(m, ROW(A, 1), half_dt);
    v_add1(m, ROW(A, 1), pending);
    v_copy(m, bar, ROW(A, 0));
    v_scale(m, ROW(A, 0), half_dt);
    v_add1(m, ROW(A, 0), next_pending);
}

/* integration/simpson.hpp:125-147 */
static void simpson_sum_vjp(size_t len, size_t m, double dt, double* A, const double* out) {
    size_t simpson_last = (len % 2 == 1) ? len - 1 : len - 2;
    v_copy(m, out, ROW(A, 0));
    v_scale(m, ROW(A, 0), dt / 3);
    for (size_t i = 1; i < simpson_last; i += 2) {
        v_copy(m, out, ROW(A, i));
        v_scale(m, ROW(A, i), 4 * dt / 3);
    }
    for (size_t i = 2; i < simpson_last; i += 2) {
        v_copy(m, out, ROW(A, i));
        v_scale(m, ROW(A, i), 2 * dt / 3);
    }
    v_copy(m, out, ROW(A, simpson_last));
    v_scale(m, ROW(A, simpson_last), dt / 3);
    if (simpson_last + 1 < len) {
        v_add(m, ROW(A, simpson_last), out, dt / 2);
        v_copy(m, out, ROW(A, simpson_last + 1));
        v_scale(m, ROW(A, simpson_last + 1), dt / 2);
    }
}

/* integration/simpson.hpp:149-162 */
static void simpson_sum_vjp_add(size_t len, size_t m, double dt, double* A, const double* out) {
    size_t simpson_last = (len % 2 == 1) ? len - 1 : len - 2;
    v_add(m, ROW(A, 0), out, dt / 3);
    for (size_t i = 1; i < simpson_last; i += 2) v_add(m, ROW(A, i), out, 4 * dt / 3);
    for (size_t i = 2; i < simpson_last; i += 2) v_add(m, ROW(A, i), out, 2 * dt / 3);
    v_add(m, ROW(A, simpson_last), out, dt / 3);
    if (simpson_last + 1 < len) {
        v_add(m, ROW(A, simpson_last), out, dt / 2);
        v_add(m, ROW(A, simpson_last + 1), out, dt / 2);
    }
}

/* integration/trapezoid.hpp:23-33, including the reference's reconstruction
 * of f_i from prefix differences (line 31). */
static void trapezoid_prefix(size_t len, size_t m, double dt, double* A, double* scratch) {
    double* tmp = scratch;
    v_copy(m, ROW(A, 0), tmp);
    v_zero(m, ROW(A, 0));
    for (size_t i = 1; i < len; ++i) {
        v_scale(m, ROW(A, i), dt / 2);
        v_add(m, ROW(A, i), tmp, dt / 2);
        v_add1(m, ROW(A, i), ROW(A, i - 1));
        v_scale(m, tmp, -1);
        v_add(m, tmp, ROW(A, i), 2 / dt);
        v_add(m, tmp, ROW(A, i - 1), -2 / dt);
    }
}

/* integration/trapezoid.hpp:35-41 */
static void trapezoid_sum(size_t len, size_t m, double dt, const double* A, double* out) {
    v_add(m, out, ROW(A, 0), dt / 2);
    for (size_t i = 1; i + 1 < len; ++i) v_add(m, out, ROW(A, i), dt);
    v_add(m, out, ROW(A, len - 1), dt / 2);
}

/* integration/trapezoid.hpp:43-52 */
static void trapezoid_prefix_vjp(size_t len, size_t m, double dt, double* A, double* scratch) {
    double* tmp = scratch;
    double half_dt = dt / 2;
    v_zero(m, tmp);
    for (size_t i = len - 1; i > 0; --i) {
        v_scale(m, ROW(A, i), half_dt);
        v_add1(m, ROW(A, i), tmp);
        v_scale(m, tmp, -1);
        v_add(m, tmp, ROW(A, i), 2);
    }
    v_copy(m, tmp, ROW(A, 0));
    v_scale(m, ROW(A, 0), 0.5);
}

/* integration/trapezoid.hpp:54-64 */
static void trapezoid_sum_vjp(size_t len, size_t m, double dt, double* A, const double* out) {
    v_copy(m, out, ROW(A, 0));
    v_scale(m, ROW(A, 0), dt / 2);
    for (size_t i = 1; i + 1 < len; ++i) {
        v_copy(m, out, ROW(A, i));
        v_scale(m, ROW(A, i), dt);
    }
    v_copy(m, out, ROW(A, len - 1));
    v_scale(m, ROW(A, len - 1), dt / 2);
}

/* integration/trapezoid.hpp:66-71 */
static void trapezoid_sum_vjp_add(size_t len, size_t m, double dt, double* A, const double* out) {
    v_add(m, ROW(A, 0), out, dt / 2);
    for (size_t i = 1; i + 1 < len; ++i) v_add(m, ROW(A, i), out, dt);
    v_add(m, ROW(A, len - 1), out, dt / 2);
}

/* integration/default.hpp:23-26 ("Riemann") */
static void riemann_prefix(size_t len, size_t m, double dt, double* A) {
    v_scale(m, ROW(A, 0), dt);
    for (size_t i = 1; i < len; ++i) {
        v_scale(m, ROW(A, i), dt);
        v_add1(m, ROW(A, i), ROW(A, i - 1));
    }
}

/* integration/default.hpp:28-30 */
static void riemann_sum(size_t len, size_t m, double dt, const double* A, double* out) {
    for (size_t i = 0; i < len; ++i) v_add(m, out, ROW(A, i), dt);
}

/* integration/default.hpp:32-38 */
static void riemann_prefix_vjp(size_t len, size_t m, double dt, double* A) {
    for (size_t i = len - 1; i > 0; --i) {
        v_add1(m, ROW(A, i - 1), ROW(A, i));
        v_scale(m, ROW(A, i), dt);
    }
    v_scale(m, ROW(A, 0), dt);
}

/* integration/default.hpp:40-49 */
static void riemann_sum_vjp(size_t len, size_t m, double dt, double* A, const double* out) {
    for (size_t i = 0; i < len; ++i) {
        v_copy(m, out, ROW(A, i));
        v_scale(m, ROW(A, i), dt);
    }
}
static void riemann_sum_vjp_add(size_t len, size_t m, double dt, double* A, const double* out) {
    for (size_t i = 0; i < len; ++i) v_add(m, ROW(A, i), out, dt);
}

/* Integrator dispatch.  `scratch` holds at least 5*m doubles. */
static void I_prefix(int integ, size_t len, size_t m, double dt, double* A, double* scratch) {
    switch (integ) {
        case OR_BOOLE: boole_prefix(len, m, dt, A, scratch); break;
        case OR_SIMPSON: simpson_prefix(len, m, dt, A, scratch); break;
        case OR_TRAPEZOID: trapezoid_prefix(len, m, dt, A, scratch); break;
        default: riemann_prefix(len, m, dt, A); break;
    }
}
static void I_sum(int integ, size_t len, size_t m, double dt, const double* A, double* out) {
    switch (integ) {
        case OR_BOOLE: boole_sum(len, m, dt, A, out); break;
        case OR_SIMPSON: simpson_sum(len, m, dt, A, out); break;
        case OR_TRAPEZOID: trapezoid_sum(len, m, dt, A, out); break;
        default: riemann_sum(len, m, dt, A, out); break;
    }
}
static void I_prefix_vjp(int integ, size_t len, size_t m, double dt, double* A, double* scratch) {
    switch (integ) {
        case OR_BOOLE: boole_prefix_vjp(len, m, dt, A, scratch); break;
        case OR_SIMPSON: simpson_prefix_vjp(len, m, dt, A, scratch); break;
        case OR_TRAPEZOID: trapezoid_prefix_vjp(len, m, dt, A, scratch); break;
        default: riemann_prefix_vjp(len, m, dt, A); break;
    }
}
static void I_sum_vjp(int integ, size_t len, size_t m, double dt, double* A, const double* out) {
    switch (integ) {
        case OR_BOOLE: boole_sum_vjp(len, m, dt, A, out); break;
        case OR_SIMPSON: simpson_sum_vjp(len, m, dt, A, out); break;
        case OR_TRAPEZOID: trapezoid_sum_vjp(len, m, dt, A, out); break;
        default: riemann_sum_vjp(len, m, dt, A, out); break;
    }
}
static void I_sum_vjp_add(int integ, size_t len, size_t m, double dt, double* A, const double* out) {
    switch (integ) {
        case OR_BOOLE: boole_sum_vjp_add(len, m, dt, A, out); break;
        case OR_SIMPSON: simpson_sum_vjp_add(len, m, dt, A, out); break;
        case OR_TRAPEZOID: trapezoid_sum_vjp_add(len, m, dt, A, out); break;
        default: riemann_sum_vjp_add(len, m, dt, A, out); break;
    }
}

static int divisibility(int integ) {
    /* boole.hpp:18, simpson.hpp:18, trapezoid.hpp:18, default.hpp:15 */
    switch (integ) {
        case OR_BOOLE: return 4;
        case OR_SIMPSON: return 2;
        case OR_TRAPEZOID: return 2;
        default: return 1;
    }
}

/* ------------------------------------------------------------------------ */
/* Matrix products                                                            */
/* ------------------------------------------------------------------------ */

/* linalg/manual.hpp:8-16 (also the semantics of cblas_dgemm in blas.hpp:8-82) */
static void manual_matmul(size_t dim, const double* a, const double* b, double* out) {
    for (size_t i = 0; i < dim; ++i)
        for (size_t j = 0; j < dim; ++j) {
            out[i * dim + j] = 0;
            for (size_t k = 0; k < dim; ++k) out[i * dim + j] += a[i * dim + k] * b[k * dim + j];
        }
}

/* linalg/manual.hpp:19-42 (blas.hpp:190-216 computes the same two products) */
static void manual_matmul_vjp(size_t dim, double* da, const double* a, const double* dout,
                              const double* b, double* db) {
    for (size_t i = 0; i < dim; ++i)
        for (size_t k = 0; k < dim; ++k) {
            double value = 0;
            for (size_t j = 0; j < dim; ++j) value += dout[i * dim + j] * b[k * dim + j];
            da[i * dim + k] += value;
        }
    for (size_t k = 0; k < dim; ++k)
        for (size_t j = 0; j < dim; ++j) {
            double value = 0;
            for (size_t i = 0; i < dim; ++i) value += a[i * dim + k] * dout[i * dim + j];
            db[k * dim + j] = value;
        }
}

/* linalg/generic.hpp:23-42 */
static void generic_sample_update(size_t dim, size_t len, const double* A, double* Y,
                                  const double* total, double shift, double* temp) {
    size_t m = dim * dim;
    for (size_t i = 0; i < len; ++i) {
        const double* a = A + i * m;
        double* y = Y + i * m;
        v_add(m, y, total, shift);
        manual_matmul(dim, a, y, temp);
        v_copy(m, temp, y);
    }
}

/* linalg/generic.hpp:50-86 (b and db alias `temp`, as in the reference) */
static void generic_sample_update_vjp(size_t dim, size_t len, double* dA, const double* A,
                                      double* barY, const double* prefix, double shift,
                                      double* temp) {
    size_t m = dim * dim;
    size_t last = len - 1;
    const double* total = prefix + last * m;
#define REVERSE_SAMPLE(SMP)                                              \
    do {                                                                 \
        size_t sm = (SMP);                                               \
        v_copy(m, prefix + sm * m, temp);                                \
        v_add(m, temp, total, shift);                                    \
        manual_matmul_vjp(dim, dA + sm * m, A + sm * m, barY + sm * m, temp, temp); \
    } while (0)
    REVERSE_SAMPLE(last);
    double* bar_total = barY + last * m;
    v_copy(m, temp, bar_total);
    v_add(m, bar_total, temp, shift);
    for (size_t sample = last; sample-- > 0;) {
        REVERSE_SAMPLE(sample);
        v_add(m, bar_total, temp, shift);
        v_copy(m, temp, barY + sample * m);
    }
#undef REVERSE_SAMPLE
}

/* linalg/fixed.hpp:119-154, Dim == 2 body */
static void fixed2_sample_update(size_t len, const double* A, double* Y, const double* total,
                                 double shift) {
    double x = shift;
    for (size_t i = 0; i < len; ++i) {
        const double* a = A + i * 4;
        double* y = Y + i * 4;
        double b0 = y[0] + total[0] * x;
        double b1 = y[1] + total[1] * x;
        double b2 = y[2] + total[2] * x;
        double b3 = y[3] + total[3] * x;
        y[0] = a[0] * b0 + a[1] * b2;
        y[1] = a[0] * b1 + a[1] * b3;
        y[2] = a[2] * b0 + a[3] * b2;
        y[3] = a[2] * b1 + a[3] * b3;
    }
}

/* linalg/fixed.hpp:157-226, Dim == 2 body */
static void fixed2_sample_update_vjp(size_t len, double* dA, const double* A, double* barY,
                                     const double* prefix, double shift) {
    double x = shift;
    double one_plus_x = 1.0 + shift;
    size_t last = len - 1;
    const double* total = prefix + last * 4;
    double temp[4];
#define REV2(SMP)                                                  \
    do {                                                           \
        size_t sm = (SMP);                                         \
        double* da = dA + sm * 4;                                  \
        const double* a = A + sm * 4;                              \
        const double* g = barY + sm * 4;                           \
        const double* p = prefix + sm * 4;                         \
        double b0 = p[0] + total[0] * x;                           \
        double b1 = p[1] + total[1] * x;                           \
        double b2 = p[2] + total[2] * x;                           \
        double b3 = p[3] + total[3] * x;                           \
        double g0 = g[0], g1 = g[1], g2 = g[2], g3 = g[3];         \
        da[0] += g0 * b0 + g1 * b1;                                \
        da[1] += g0 * b2 + g1 * b3;                                \
        da[2] += g2 * b0 + g3 * b1;                                \
        da[3] += g2 * b2 + g3 * b3;                                \
        temp[0] = a[0] * g0 + a[2] * g2;                           \
        temp[1] = a[0] * g1 + a[2] * g3;                           \
        temp[2] = a[1] * g0 + a[3] * g2;                           \
        temp[3] = a[1] * g1 + a[3] * g3;                           \
    } while (0)
    REV2(last);
    double* bar_total = barY + last * 4;
    for (int c = 0; c < 4; ++c) bar_total[c] = temp[c] * one_plus_x;
    for (size_t sample = last; sample-- > 0;) {
        REV2(sample);
        double* g = barY + sample * 4;
        for (int c = 0; c < 4; ++c) bar_total[c] += temp[c] * x;
        for (int c = 0; c < 4; ++c) g[c] = temp[c];
    }
#undef REV2
}

/* linalg/spacecurve.hpp:104-129 */
static void sc_sample_update(size_t len, const double* A, double* Y, const double* total,
                             double shift) {
    double x = shift;
    for (size_t i = 0; i < len; ++i) {
        const double* a = A + i * 4;
        double* y = Y + i * 4;
        double b0 = y[0] + total[0] * x;
        double b1 = y[1] + total[1] * x;
        double b2 = y[2] + total[2] * x;
        double b3 = y[3] + total[3] * x;
        y[0] = a[1] * b1 + a[2] * b2 + a[3] * b3;
        y[1] = -b0 * a[1] + a[2] * b3 - a[3] * b2;
        y[2] = -b0 * a[2] + a[3] * b1 - a[1] * b3;
        y[3] = -b0 * a[3] + a[1] * b2 - a[2] * b1;
    }
}

/* linalg/spacecurve.hpp:132-191 */
static void sc_sample_update_vjp(size_t len, double* dA, const double* A, double* barY,
                                 const double* prefix, double shift) {
    double x = shift;
    double one_plus_x = 1.0 + shift;
    size_t last = len - 1;
    const double* total = prefix + last * 4;
    double temp[4];
#define REVSC(SMP)                                                 \
    do {                                                           \
        size_t sm = (SMP);                                         \
        double* da = dA + sm * 4;                                  \
        const double* a = A + sm * 4;                              \
        const double* g = barY + sm * 4;                           \
        const double* p = prefix + sm * 4;                         \
        double b0 = p[0] + total[0] * x;                           \
        double b1 = p[1] + total[1] * x;                           \
        double b2 = p[2] + total[2] * x;                           \
        double b3 = p[3] + total[3] * x;                           \
        double g0 = g[0], g1 = g[1], g2 = g[2], g3 = g[3];         \
        da[1] += g0 * b1 - g1 * b0 - g2 * b3 + g3 * b2;            \
        da[2] += g0 * b2 + g1 * b3 - g2 * b0 - g3 * b1;            \
        da[3] += g0 * b3 - g1 * b2 + g2 * b1 - g3 * b0;            \
        temp[0] = -g1 * a[1] - g2 * a[2] - g3 * a[3];              \
        temp[1] = g0 * a[1] + g2 * a[3] - g3 * a[2];               \
        temp[2] = g0 * a[2] - g1 * a[3] + g3 * a[1];               \
        temp[3] = g0 * a[3] + g1 * a[2] - g2 * a[1];               \
    } while (0)
    REVSC(last);
    double* bar_total = barY + last * 4;
    for (int c = 0; c < 4; ++c) bar_total[c] = temp[c] * one_plus_x;
    for (size_t sample = last; sample-- > 0;) {
        REVSC(sample);
        double* g = barY + sample * 4;
        for (int c = 0; c < 4; ++c) bar_total[c] += temp[c] * x;
        for (int c = 0; c < 4; ++c) g[c] = temp[c];
    }
#undef REVSC
}

/* MatrixSpan::sample_update(_vjp) forwarding (linalg/matrix.hpp:367-401) */
static void M_update(int mat, size_t dim, size_t len, const double* A, double* Y,
                     const double* total, double shift, double* temp) {
    if (mat == OR_SPACECURVE) sc_sample_update(len, A, Y, total, shift);
    else if (mat == OR_FIXED2) fixed2_sample_update(len, A, Y, total, shift);
    else generic_sample_update(dim, len, A, Y, total, shift, temp);
}
static void M_update_vjp(int mat, size_t dim, size_t len, double* dA, const double* A,
                         double* barY, const double* prefix, double shift, double* temp) {
    if (mat == OR_SPACECURVE) sc_sample_update_vjp(len, dA, A, barY, prefix, shift);
    else if (mat == OR_FIXED2) fixed2_sample_update_vjp(len, dA, A, barY, prefix, shift);
    else generic_sample_update_vjp(dim, len, dA, A, barY, prefix, shift, temp);
}

/* ------------------------------------------------------------------------ */
/* Exposed primitives                                                          */
/* ------------------------------------------------------------------------ */
int oracle_prefix(int integ, size_t len, size_t m, double dt, double* A) {
    double* scratch = (double*)malloc(5 * m * sizeof(double));
    if (!scratch) return fail("out of memory");
    I_prefix(integ, len, m, dt, A, scratch);
    free(scratch);
    return 0;
}
int oracle_sum(int integ, size_t len, size_t m, double dt, const double* A, double* out) {
    I_sum(integ, len, m, dt, A, out);
    return 0;
}
int oracle_prefix_vjp(int integ, size_t len, size_t m, double dt, double* A) {
    double* scratch = (double*)malloc(5 * m * sizeof(double));
    if (!scratch) return fail("out of memory");
    I_prefix_vjp(integ, len, m, dt, A, scratch);
    free(scratch);
    return 0;
}
int oracle_sum_vjp(int integ, size_t len, size_t m, double dt, double* A, const double* out) {
    I_sum_vjp(integ, len, m, dt, A, out);
    return 0;
}
int oracle_sum_vjp_add(int integ, size_t len, size_t m, double dt, double* A, const double* out) {
    I_sum_vjp_add(integ, len, m, dt, A, out);
    return 0;
}
int oracle_sample_update(int mat, size_t dim, size_t len, const double* A, double* Y,
                         const double* total, double shift) {
    size_t m = (mat == OR_SPACECURVE) ? 4 : dim * dim;
    double* temp = (double*)malloc(m * sizeof(double));
    if (!temp) return fail("out of memory");
    M_update(mat, dim, len, A, Y, total, shift, temp);
    free(temp);
    return 0;
}
int oracle_sample_update_vjp(int mat, size_t dim, size_t len, double* dA, const double* A,
                             double* barY, const double* prefix, double shift) {
    size_t m = (mat == OR_SPACECURVE) ? 4 : dim * dim;
    double* temp = (double*)malloc(m * sizeof(double));
    if (!temp) return fail("out of memory");
    M_update_vjp(mat, dim, len, dA, A, barY, prefix, shift, temp);
    free(temp);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Validation, mirroring composer/backends.hpp:108-122 and backends.cpp:26-61  */
/* ------------------------------------------------------------------------ */
static int validate(int n, int op, int mat, int integ, size_t samples, size_t dim) {
    if (n < 1) return fail("n must be at least 1");
    if (samples < 2) return fail("data must contain at least two samples");
    if (dim == 0) return fail("data must contain square matrices");
    if (op < 0 || op > 2) return fail("could not deduce kernel op");
    if (integ < 0 || integ > 3) return fail("Could not resolve dispatch name option");
    if (mat < 0 || mat > 3) return fail("Could not resolve dispatch name option");
    if (mat == OR_FIXED2 && dim != 2) return fail("invalid dispatch option (based on params)");
    if (mat == OR_SPACECURVE && dim != 2) return fail("invalid dispatch option (based on params)");
    int gl_max = gl_ready();
    if (!gl_max) return 1;
    if (n > 1 && (n + 3) / 2 > gl_max)
        return fail("requested magnus order exceeds GL table. Consider providing a larger table.");
    if ((samples - 1) % (size_t)divisibility(integ) != 0)
        return fail("sample interval count violates integrator divisibility");
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Forward drivers: Magnus::one / many / sum (magnus.hpp:12-261)               */
/* ------------------------------------------------------------------------ */
static int forward_impl(int n, int op, int mat, int integ, int lanes_req, size_t S, size_t dim,
                        double t0, double tf, const double* A, double* out, double* carry) {
    size_t m = (mat == OR_SPACECURVE) ? 4 : dim * dim;
    double dt = (tf - t0) / (double)(S - 1); /* magnus.hpp:32 */
    size_t n_out = (op == OR_MANY) ? (size_t)n : 1;
    v_zero(n_out * m, out);

    if (n == 1) { /* magnus.hpp:36-40, 118-123, 203-207 */
        I_sum(integ, S, m, dt, A, out);
        return 0;
    }
    size_t Q = (size_t)(n + 3) / 2; /* util/extra.hpp:11-13 */
    size_t K = (size_t)n - 1;       /* util/extra.hpp:15-17 */
    double* w = (double*)malloc(Q * sizeof(double));
    double* x = (double*)malloc(Q * sizeof(double));
    oracle_gl_get((int)Q, w, x);

    size_t L = 1;
    if (lanes_req > 1) L = (size_t)lanes_req < Q ? (size_t)lanes_req : Q;

    /* per-lane workspaces (util/extra.hpp:51-86): Y, total copy, integrator scratch */
    double* ws = (double*)calloc(L * (S * m + m + 5 * m + n_out * m), sizeof(double));
    if (!ws) { free(w); free(x); return fail("out of memory"); }

    /* lane 0 writes straight into `out` (magnus.hpp:57, 140, 224) */
    if (op != OR_ONE) I_sum(integ, S, m, dt, A, out); /* Omega_1: magnus.hpp:143-144, 227-228 */

#ifdef _OPENMP
#pragma omp parallel for num_threads((int)L) schedule(static, 1) if (L > 1)
#endif
    for (size_t lane = 0; lane < L; ++lane) {
        double* Y = ws + lane * (S * m + m + 5 * m + n_out * m);
        double* total_copy = Y + S * m;
        double* scratch = total_copy + m;
        double* partial = (lane == 0) ? out : scratch + 5 * m;
        /* OpenMP schedule(static) split of [0, Q) into contiguous lane ranges */
        size_t base = Q / L, rem = Q % L;
        size_t q0 = lane * base + (lane < rem ? lane : rem);
        size_t q1 = q0 + base + (lane < rem ? 1 : 0);
        for (size_t q = q0; q < q1; ++q) {
            double w_q = w[q], x_q = x[q];
            double shift = x_q - 1;           /* magnus.hpp:65 */
            double final_scale = dt * w_q;    /* magnus.hpp:66 */
            v_copy(S * m, A, Y);              /* magnus.hpp:68 */
            double* temp = scratch;           /* integrator.borrow_scratch() == scratch[0] */
            for (size_t k = 2; k <= (size_t)n; ++k) {
                I_prefix(integ, S, m, dt, Y, scratch);
                if (carry) v_copy(S * m, Y, carry + (q * K + (k - 2)) * S * m); /* graddata.hpp:101-104 */
                v_copy(m, ROW(Y, S - 1), total_copy);
                M_update(mat, dim, S, A, Y, total_copy, shift, temp);
                if (op == OR_MANY) I_sum(integ, S, m, final_scale, Y, partial + (k - 1) * m);
                else if (op == OR_SUM) I_sum(integ, S, m, final_scale, Y, partial);
            }
            if (op == OR_ONE) I_sum(integ, S, m, final_scale, Y, partial); /* magnus.hpp:81-82 */
        }
    }
    /* lane tree reduction, magnus.hpp:86-90 */
    for (size_t stride = 1; stride < L; stride *= 2)
        for (size_t i = 0; i + stride < L; i += 2 * stride) {
            double* dst = (i == 0) ? out : ws + i * (S * m + m + 5 * m + n_out * m) + S * m + m + 5 * m;
            double* src = ws + (i + stride) * (S * m + m + 5 * m + n_out * m) + S * m + m + 5 * m;
            for (size_t e = 0; e < n_out * m; ++e) dst[e] += src[e];
        }
    free(ws);
    free(w);
    free(x);
    return 0;
}

int oracle_forward(int n, int op, int mat, int integ, int lanes, size_t samples, size_t dim,
                   double t0, double tf, const double* in, double* out, double* carry) {
    if (validate(n, op, mat, integ, samples, dim)) return 1;
    size_t m = (mat == OR_SPACECURVE) ? 4 : dim * dim;
    /* run_raw copies the input first (composer/backends.hpp:380-382) */
    double* A = (double*)malloc(samples * m * sizeof(double));
    if (!A) return fail("out of memory");
    memcpy(A, in, samples * m * sizeof(double));
    int rc = forward_impl(n, op, mat, integ, lanes, samples, dim, t0, tf, A, out, carry);
    free(A);
    return rc;
}

/* VJP::none (grad.hpp:14-56): forward recompute writing only the carry */
static void vjp_none(int n, int mat, int integ, size_t S, size_t dim, double t0, double tf,
                     const double* A, double* carry, double* Y, double* scratch) {
    size_t m = (mat == OR_SPACECURVE) ? 4 : dim * dim;
    double dt = (tf - t0) / (double)(S - 1);
    size_t Q = (size_t)(n + 3) / 2, K = (size_t)n - 1;
    double w[GL_CAP], x[GL_CAP];
    oracle_gl_get((int)Q, w, x);
    double* total_copy = scratch + 5 * m;
    double* temp = scratch + 6 * m; /* FwdPathWorkspace::temp (borrowed scratch) */
    for (size_t q = 0; q < Q; ++q) {
        double shift = x[q] - 1;
        v_copy(S * m, A, Y);
        for (size_t k = 2; k <= (size_t)n; ++k) {
            I_prefix(integ, S, m, dt, Y, scratch);
            v_copy(S * m, Y, carry + (q * K + (k - 2)) * S * m);
            v_copy(m, ROW(Y, S - 1), total_copy);
            M_update(mat, dim, S, A, Y, total_copy, shift, temp);
        }
    }
}

/* VJP::one / many / sum (grad.hpp:58-349); serial GL loop as dispatched */
static int vjp_impl(int n, int op, int mat, int integ, size_t S, size_t dim, double t0, double tf,
                    const double* A, const double* dOut, const double* carry_in, double* dA) {
    size_t m = (mat == OR_SPACECURVE) ? 4 : dim * dim;
    double dt = (tf - t0) / (double)(S - 1);

    if (op == OR_ONE && n == 1) { /* grad.hpp:85-88 */
        I_sum_vjp(integ, S, m, dt, dA, dOut);
        return 0;
    }
    if (op == OR_ONE) v_zero(S * m, dA);             /* grad.hpp:90 */
    else I_sum_vjp(integ, S, m, dt, dA, dOut);       /* grad.hpp:175-176, 278 */
    if (n == 1) return 0;

    size_t Q = (size_t)(n + 3) / 2, K = (size_t)n - 1;
    double* scratch = (double*)calloc(8 * m, sizeof(double));
    double* barY = (double*)calloc(S * m, sizeof(double));
    double* local = NULL;
    const double* carry = carry_in;
    if (!scratch || !barY) { free(scratch); free(barY); return fail("out of memory"); }
    if (!carry) {
        local = (double*)calloc(Q * K * S * m, sizeof(double));
        double* Y = (double*)calloc(S * m, sizeof(double));
        if (!local || !Y) { free(local); free(Y); free(scratch); free(barY); return fail("out of memory"); }
        vjp_none(n, mat, integ, S, dim, t0, tf, A, local, Y, scratch);
        free(Y);
        carry = local;
    }
    double w[GL_CAP], x[GL_CAP];
    oracle_gl_get((int)Q, w, x);
    double* tmp = scratch; /* integrator.borrow_scratch() == scratch[0] */

    for (size_t q = 0; q < Q; ++q) {
        double shift = x[q] - 1.0;
        double dscale = dt * w[q];
        if (op == OR_ONE) {
            I_sum_vjp(integ, S, m, dscale, barY, dOut);
            for (size_t k = (size_t)n; k > 1; --k) {
                const double* P = carry + (q * K + (k - 2)) * S * m;
                M_update_vjp(mat, dim, S, dA, A, barY, P, shift, tmp + 5 * m);
                I_prefix_vjp(integ, S, m, dt, barY, scratch);
            }
        } else {
            const double* last_dOut = (op == OR_MANY) ? dOut + ((size_t)n - 1) * m : dOut;
            I_sum_vjp(integ, S, m, dscale, barY, last_dOut);
            for (size_t k = (size_t)n; k > 2; --k) {
                const double* P = carry + (q * K + (k - 2)) * S * m;
                M_update_vjp(mat, dim, S, dA, A, barY, P, shift, tmp + 5 * m);
                I_prefix_vjp(integ, S, m, dt, barY, scratch);
                const double* direct = (op == OR_MANY) ? dOut + (k - 2) * m : dOut;
                I_sum_vjp_add(integ, S, m, dscale, barY, direct);
            }
            const double* P = carry + (q * K) * S * m;
            M_update_vjp(mat, dim, S, dA, A, barY, P, shift, tmp + 5 * m);
            I_prefix_vjp(integ, S, m, dt, barY, scratch);
        }
        for (size_t e = 0; e < S * m; ++e) dA[e] += barY[e]; /* grad.hpp:144, 247, 347 */
    }
    free(local);
    free(scratch);
    free(barY);
    return 0;
}

int oracle_vjp(int n, int op, int mat, int integ, size_t samples, size_t dim, double t0,
               double tf, const double* in, const double* cot, const double* carry, double* dA) {
    if (validate(n, op, mat, integ, samples, dim)) return 1;
    return vjp_impl(n, op, mat, integ, samples, dim, t0, tf, in, cot, carry, dA);
}

/* ------------------------------------------------------------------------ */
/* SpaceCurve wrappers: pad (S,3) -> (S,4), unpad outputs                      */
/* composer/backends.hpp:253-357                                               */
/* ------------------------------------------------------------------------ */
int oracle_sc_forward(int n, int op, int integ, int lanes, size_t samples, double t0, double tf,
                      const double* in3, double* out3, double* carry) {
    if (validate(n, op, OR_SPACECURVE, integ, samples, 2)) return 1;
    size_t n_out = (op == OR_MANY) ? (size_t)n : 1;
    double* padded = (double*)malloc(samples * 4 * sizeof(double));
    double* padded_out = (double*)calloc(n_out * 4, sizeof(double));
    if (!padded || !padded_out) { free(padded); free(padded_out); return fail("out of memory"); }
    for (size_t i = 0; i < samples; ++i) {
        padded[4 * i] = 0.0;
        padded[4 * i + 1] = in3[3 * i];
        padded[4 * i + 2] = in3[3 * i + 1];
        padded[4 * i + 3] = in3[3 * i + 2];
    }
    int rc = forward_impl(n, op, OR_SPACECURVE, integ, lanes, samples, 2, t0, tf, padded,
                          padded_out, carry);
    for (size_t i = 0; i < n_out; ++i) {
        out3[3 * i] = padded_out[4 * i + 1];
        out3[3 * i + 1] = padded_out[4 * i + 2];
        out3[3 * i + 2] = padded_out[4 * i + 3];
    }
    free(padded);
    free(padded_out);
    return rc;
}

int oracle_sc_vjp(int n, int op, int integ, size_t samples, double t0, double tf,
                  const double* in3, const double* cot3, const double* carry, double* dA3) {
    if (validate(n, op, OR_SPACECURVE, integ, samples, 2)) return 1;
    size_t n_out = (op == OR_MANY) ? (size_t)n : 1;
    double* pin = (double*)malloc(samples * 4 * sizeof(double));
    double* pcot = (double*)malloc(n_out * 4 * sizeof(double));
    double* pout = (double*)calloc(samples * 4, sizeof(double));
    if (!pin || !pcot || !pout) { free(pin); free(pcot); free(pout); return fail("out of memory"); }
    for (size_t i = 0; i < samples; ++i) {
        pin[4 * i] = 0.0;
        pin[4 * i + 1] = in3[3 * i];
        pin[4 * i + 2] = in3[3 * i + 1];
        pin[4 * i + 3] = in3[3 * i + 2];
    }
    for (size_t i = 0; i < n_out; ++i) {
        pcot[4 * i] = 0.0;
        pcot[4 * i + 1] = cot3[3 * i];
        pcot[4 * i + 2] = cot3[3 * i + 1];
        pcot[4 * i + 3] = cot3[3 * i + 2];
    }
    int rc = vjp_impl(n, op, OR_SPACECURVE, integ, samples, 2, t0, tf, pin, pcot, carry, pout);
    for (size_t i = 0; i < samples; ++i) {
        dA3[3 * i] = pout[4 * i + 1];
        dA3[3 * i + 1] = pout[4 * i + 2];
        dA3[3 * i + 2] = pout[4 * i + 3];
    }
    free(pin);
    free(pcot);
    free(pout);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Batched timing drivers (instances spread over OpenMP threads)              */
/* ------------------------------------------------------------------------ */
static size_t carry_elems(int n, size_t S, size_t m) {
    if (n < 2) return 0;
    return (size_t)((n + 3) / 2) * (size_t)(n - 1) * S * m;
}

int oracle_forward_batch(int n, int op, int mat, int integ, int lanes, int threads, size_t count,
                         size_t samples, size_t dim, double t0, double tf, const double* in,
                         double* out, double* carry) {
    if (validate(n, op, mat, integ, samples, dim)) return 1;
    size_t m = (mat == OR_SPACECURVE) ? 4 : dim * dim;
    size_t n_out = (op == OR_MANY) ? (size_t)n : 1;
    size_t ce = carry_elems(n, samples, m);
    int rc = 0;
#ifdef _OPENMP
#pragma omp parallel for num_threads(threads > 0 ? threads : 1) schedule(dynamic, 1) reduction(| : rc)
#endif
    for (size_t b = 0; b < count; ++b)
        rc |= oracle_forward(n, op, mat, integ, lanes, samples, dim, t0, tf, in + b * samples * m,
                             out + b * n_out * m, carry ? carry + b * ce : NULL);
    return rc;
}

int oracle_vjp_batch(int n, int op, int mat, int integ, int threads, size_t count, size_t samples,
                     size_t dim, double t0, double tf, const double* in, const double* cot,
                     const double* carry, double* dA) {
    if (validate(n, op, mat, integ, samples, dim)) return 1;
    size_t m = (mat == OR_SPACECURVE) ? 4 : dim * dim;
    size_t n_out = (op == OR_MANY) ? (size_t)n : 1;
    size_t ce = carry_elems(n, samples, m);
    int rc = 0;
#ifdef _OPENMP
#pragma omp parallel for num_threads(threads > 0 ? threads : 1) schedule(dynamic, 1) reduction(| : rc)
#endif
    for (size_t b = 0; b < count; ++b)
        rc |= oracle_vjp(n, op, mat, integ, samples, dim, t0, tf, in + b * samples * m,
                         cot + b * n_out * m, carry ? carry + b * ce : NULL, dA + b * samples * m);
    return rc;
}

int oracle_sc_forward_batch(int n, int op, int integ, int lanes, int threads, size_t count,
                            size_t samples, double t0, double tf, const double* in3, double* out3,
                            double* carry) {
    if (validate(n, op, OR_SPACECURVE, integ, samples, 2)) return 1;
    size_t n_out = (op == OR_MANY) ? (size_t)n : 1;
    size_t ce = carry_elems(n, samples, 4);
    int rc = 0;
#ifdef _OPENMP
#pragma omp parallel for num_threads(threads > 0 ? threads : 1) schedule(dynamic, 1) reduction(| : rc)
#endif
    for (size_t b = 0; b < count; ++b)
        rc |= oracle_sc_forward(n, op, integ, lanes, samples, t0, tf, in3 + b * samples * 3,
                                out3 + b * n_out * 3, carry ? carry + b * ce : NULL);
    return rc;
}

int oracle_sc_vjp_batch(int n, int op, int integ, int threads, size_t count, size_t samples,
                        double t0, double tf, const double* in3, const double* cot3,
                        const double* carry, double* dA3) {
    if (validate(n, op, OR_SPACECURVE, integ, samples, 2)) return 1;
    size_t n_out = (op == OR_MANY) ? (size_t)n : 1;
    size_t ce = carry_elems(n, samples, 4);
    int rc = 0;
#ifdef _OPENMP
#pragma omp parallel for num_threads(threads > 0 ? threads : 1) schedule(dynamic, 1) reduction(| : rc)
#endif
    for (size_t b = 0; b < count; ++b)
        rc |= oracle_sc_vjp(n, op, integ, samples, t0, tf, in3 + b * samples * 3,
                            cot3 + b * n_out * 3, carry ? carry + b * ce : NULL,
                            dA3 + b * samples * 3);
    return rc;
}
