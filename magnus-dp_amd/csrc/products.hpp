// products.hpp -- per-sample products of the level recursion and their
// adjoints, one output scalar at a time, in the reference's operation order.
#pragma once

#include <hip/hip_runtime.h>

namespace magnus {

// generic_sample_update (linalg/generic.hpp:23-42, manual.hpp:8-16):
// y_rc = sum_l a_rl * (p_lc + t_lc * shift), accumulated in l order from 0.
__device__ __forceinline__ double prod_generic(const double* __restrict__ a, const double* __restrict__ p,
                                               const double* __restrict__ t, int d, int r, int c,
                                               double shift) {
    double acc = 0.0;
    for (int l = 0; l < d; ++l) {
        double b = p[l * d + c];
        b += t[l * d + c] * shift;
        acc += a[r * d + l] * b;
    }
    return acc;
}

// SpaceCurve::sample_update (linalg/spacecurve.hpp:104-129); a = (0, a1, a2, a3)
__device__ __forceinline__ double prod_sc(const double* __restrict__ a3, const double* __restrict__ p,
                                          const double* __restrict__ t, int e, double shift) {
    double b0 = p[0];
    b0 += t[0] * shift;
    double b1 = p[1];
    b1 += t[1] * shift;
    double b2 = p[2];
    b2 += t[2] * shift;
    double b3 = p[3];
    b3 += t[3] * shift;
    const double a1 = a3[0], a2 = a3[1], a3_ = a3[2];
    if (e == 0) return a1 * b1 + a2 * b2 + a3_ * b3;
    if (e == 1) return -b0 * a1 + a2 * b3 - a3_ * b2;
    if (e == 2) return -b0 * a2 + a3_ * b1 - a1 * b3;
    return -b0 * a3_ + a1 * b2 - a2 * b1;
}

// raw input scalar e of sample i; sc == 1: (S,3) SpaceCurve input padded with a
// leading 0 (composer/backends.hpp:269-277); otherwise the stored value.
__device__ __forceinline__ double raw_entry(const double* __restrict__ Ai, int e, int sc) {
    if (sc == 1) return e == 0 ? 0.0 : Ai[e - 1];
    return Ai[e];
}

// sample_update_vjp for one entry e of a sample: dav = dA_e contribution,
// temp = (A^T g)_e, with b = p + t * shift (generic.hpp:50-86, manual.hpp:19-42,
// spacecurve.hpp:132-191; SpaceCurve leaves dA's component 0 untouched).
__device__ __forceinline__ void prod_vjp(const double* __restrict__ Ai, const double* __restrict__ Pi,
                                         const double* __restrict__ T, const double* __restrict__ g, int e, int d,
                                         bool sc, double shift, double& dav, double& temp) {
    if (sc) {
        double b0 = Pi[0]; b0 += T[0] * shift;
        double b1 = Pi[1]; b1 += T[1] * shift;
        double b2 = Pi[2]; b2 += T[2] * shift;
        double b3 = Pi[3]; b3 += T[3] * shift;
        const double g0 = g[0], g1 = g[1], g2 = g[2], g3 = g[3];
        const double a1 = Ai[0], a2 = Ai[1], a3 = Ai[2];
        if (e == 0) {
            dav = 0.0;
            temp = -g1 * a1 - g2 * a2 - g3 * a3;
        } else if (e == 1) {
            dav = g0 * b1 - g1 * b0 - g2 * b3 + g3 * b2;
            temp = g0 * a1 + g2 * a3 - g3 * a2;
        } else if (e == 2) {
            dav = g0 * b2 + g1 * b3 - g2 * b0 - g3 * b1;
            temp = g0 * a2 - g1 * a3 + g3 * a1;
        } else {
            dav = g0 * b3 - g1 * b2 + g2 * b1 - g3 * b0;
            temp = g0 * a3 + g1 * a2 - g2 * a1;
        }
        return;
    }
    const int r = e / d, cc = e - r * d;
    // da[r][cc] += sum_j g[r][j] * b[cc][j]
    double v = 0.0;
    for (int jj = 0; jj < d; ++jj) {
        double bb = Pi[cc * d + jj];
        bb += T[cc * d + jj] * shift;
        v += g[r * d + jj] * bb;
    }
    dav = v;
    // temp[r][cc] = sum_i' a[i'][r] * g[i'][cc]
    double t2 = 0.0;
    for (int ii = 0; ii < d; ++ii) t2 += Ai[ii * d + r] * g[ii * d + cc];
    temp = t2;
}

}  // namespace magnus
