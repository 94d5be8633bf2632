// rules.hpp -- per-scalar integrator rules shared by the device kernels.
//
// Every integrator in the reference acts elementwise on the d*d scalars of a
// sample, so each rule here is the reference loop for ONE scalar stream,
// keeping its order of floating-point operations (scale first, the listed
// adds next, the block base last).  The kernels run these recurrences for many
// scalar streams (matrix entries x GL nodes x instances) in parallel.
#pragma once

#include <hip/hip_runtime.h>

namespace magnus {

enum Integ : int { BOOLE = 0, SIMPSON = 1, TRAPEZOID = 2, RIEMANN = 3 };
enum Op : int { OP_ONE = 0, OP_MANY = 1, OP_SUM = 2 };

// ---------------------------------------------------------------------------
// integrator.sum weights: out += f_i * c_i(scale)
//   boole.hpp:76-106 (block of 4; interior boundaries are added twice, once
//   from each neighbouring block), simpson.hpp:60-75, trapezoid.hpp:35-41,
//   default.hpp:28-30.  S satisfies the divisibility rule, so no tails occur.
// ---------------------------------------------------------------------------
template <int I>
__device__ __forceinline__ void sum_acc(double& acc, long long i, long long S, double f, double c) {
    if constexpr (I == BOOLE) {
        if (i == 0) {
            acc += f * (14 * c / 45);
        } else {
            const int r = (int)(i & 3);
            if (r == 1) acc += f * (64 * c / 45);
            else if (r == 2) acc += f * (24 * c / 45);
            else if (r == 3) acc += f * (64 * c / 45);
            else {
                acc += f * (14 * c / 45);
                if (i != S - 1) acc += f * (14 * c / 45);
            }
        }
    } else if constexpr (I == SIMPSON) {
        if (i == 0 || i == S - 1) acc += f * (c / 3);
        else if (i & 1) acc += f * (4 * c / 3);
        else acc += f * (2 * c / 3);
    } else if constexpr (I == TRAPEZOID) {
        if (i == 0 || i == S - 1) acc += f * (c / 2);
        else acc += f * c;
    } else {
        acc += f * c;
    }
}

// integrator.sum_vjp coefficient: A_i = out * w_i  (boole.hpp:153-199,
// simpson.hpp:125-147, trapezoid.hpp:54-64, default.hpp:40-44)
template <int I>
__device__ __forceinline__ double sum_vjp_coef(long long i, long long S, double c) {
    if constexpr (I == BOOLE) {
        if (i == 0 || i == S - 1) return 14 * c / 45;
        const int r = (int)(i & 3);
        if (r == 1 || r == 3) return 64 * c / 45;
        if (r == 2) return 24 * c / 45;
        return 28 * c / 45;
    } else if constexpr (I == SIMPSON) {
        if (i == 0 || i == S - 1) return c / 3;
        return (i & 1) ? 4 * c / 3 : 2 * c / 3;
    } else if constexpr (I == TRAPEZOID) {
        return (i == 0 || i == S - 1) ? c / 2 : c;
    } else {
        return c;
    }
}

// integrator.sum_vjp_add: A_i += out * w_i, with Boole's interior boundary
// receiving its two contributions as two adds (boole.hpp:201-228).
template <int I>
__device__ __forceinline__ double sum_vjp_add(double g, long long i, long long S, double out, double c) {
    if constexpr (I == BOOLE) {
        if (i == 0) return g + out * (14 * c / 45);
        const int r = (int)(i & 3);
        if (r == 1 || r == 3) return g + out * (64 * c / 45);
        if (r == 2) return g + out * (24 * c / 45);
        g = g + out * (14 * c / 45);
        if (i != S - 1) g = g + out * (14 * c / 45);
        return g;
    } else if constexpr (I == SIMPSON) {
        if (i == 0 || i == S - 1) return g + out * (c / 3);
        return g + out * ((i & 1) ? 4 * c / 3 : 2 * c / 3);
    } else if constexpr (I == TRAPEZOID) {
        return g + out * ((i == 0 || i == S - 1) ? c / 2 : c);
    } else {
        return g + out * c;
    }
}

// ---------------------------------------------------------------------------
// Sequential prefix recurrences (one scalar stream, sample by sample).
// step(i, f) returns P_i given the raw value f_i = Y_i.
// ---------------------------------------------------------------------------
template <int I>
struct PrefixState;

// boole.hpp:24-74: block [s, s+4] from base P[s]; trapezoid at +1, Simpson at
// +2, 3/8 at +3, Boole at +4; each value is scale(f) + f0.. + base.
template <>
struct PrefixState<BOOLE> {
    double f0 = 0, f1 = 0, f2 = 0, f3 = 0, base = 0;
    __device__ __forceinline__ double step(long long i, double f, double dt) {
        if (i == 0) {
            f0 = f;
            base = 0.0;
            return 0.0;
        }
        const int r = (int)(i & 3);
        double P;
        if (r == 1) {
            f1 = f;
            P = f * (dt / 2);
            P += f0 * (dt / 2);
            P += base;
        } else if (r == 2) {
            f2 = f;
            P = f * (dt / 3);
            P += f0 * (dt / 3);
            P += f1 * (4 * dt / 3);
            P += base;
        } else if (r == 3) {
            f3 = f;
            P = f * (3 * dt / 8);
            P += f0 * (3 * dt / 8);
            P += f1 * (9 * dt / 8);
            P += f2 * (9 * dt / 8);
            P += base;
        } else {
            P = f * (14 * dt / 45);
            P += f0 * (14 * dt / 45);
            P += f1 * (64 * dt / 45);
            P += f2 * (24 * dt / 45);
            P += f3 * (64 * dt / 45);
            P += base;
            base = P;
            f0 = f;
        }
        return P;
    }
    // resume at a block start s (s % 4 == 0) whose prefix value is `at`
    __device__ __forceinline__ void start(double fs, double at) {
        f0 = fs;
        base = at;
    }
};

// simpson.hpp:24-58: odd i = trapezoid from i-1, even i = Simpson panel from i-2.
template <>
struct PrefixState<SIMPSON> {
    double prev1 = 0, prev2 = 0, Pm1 = 0, Pm2 = 0;
    __device__ __forceinline__ double step(long long i, double f, double dt) {
        const double half_dt = dt / 2;
        const double third_dt = dt / 3;
        if (i == 0) {
            prev2 = f;
            Pm1 = 0.0;
            return 0.0;
        }
        double P;
        if (i == 1) {
            P = f * half_dt;
            P += prev2 * half_dt;
            prev1 = f;
            Pm2 = 0.0;
            Pm1 = P;
            return P;
        }
        if ((i & 1) == 0) {
            P = f * third_dt;
            P += prev2 * third_dt;
            P += prev1 * (4 * third_dt);
            P += Pm2;
        } else {
            P = f * half_dt;
            P += prev1 * half_dt;
            P += Pm1;
        }
        prev2 = prev1;
        prev1 = f;
        Pm2 = Pm1;
        Pm1 = P;
        return P;
    }
    // resume at an even sample s >= 2 whose prefix value is `at`
    __device__ __forceinline__ void start(double fs, double at) {
        prev1 = fs;
        Pm1 = at;
    }
};

// trapezoid.hpp:23-33, including the reconstruction of f_{i} from prefix
// differences (line 31) that the reference performs.
template <>
struct PrefixState<TRAPEZOID> {
    double tmp = 0, Pm1 = 0;
    __device__ __forceinline__ double step(long long i, double f, double dt) {
        if (i == 0) {
            tmp = f;
            Pm1 = 0.0;
            return 0.0;
        }
        double P = f * (dt / 2);
        P += tmp * (dt / 2);
        P += Pm1;
        tmp *= -1;
        tmp += P * (2 / dt);
        tmp += Pm1 * (-2 / dt);
        Pm1 = P;
        return P;
    }
};

// default.hpp:23-26
template <>
struct PrefixState<RIEMANN> {
    double Pm1 = 0;
    __device__ __forceinline__ double step(long long i, double f, double dt) {
        double P = f * dt;
        if (i > 0) P += Pm1;
        Pm1 = P;
        return P;
    }
    __device__ __forceinline__ void start(double, double at) { Pm1 = at; }
};

// ---------------------------------------------------------------------------
// Reverse sweeps (prefix_vjp) on one scalar stream G[i * stride], in place.
// ---------------------------------------------------------------------------

// boole.hpp:108-151.  Blocks are taken in groups of kVjpGroup, highest
// first; a group's raw values are loaded before any of its stores (a block
// writes G[s+1..s+4], which the lower blocks of the group do not read: their
// G[s+4] is the chained g4), so the loads of a group are in flight together.
constexpr int kVjpGroup = 8;
__device__ __forceinline__ void prefix_vjp_boole(double* G, long long stride, long long S, double dt) {
    double pending = 0.0;
    const long long nb = (S - 1) / 4;  // blocks [4b, 4b+4], b = nb-1 .. 0
    double g4 = G[(S - 1) * stride];
    for (long long top = nb - 1; top >= 0; top -= kVjpGroup) {
        double r0[kVjpGroup], r1[kVjpGroup], r2[kVjpGroup], r3[kVjpGroup];
#pragma unroll
        for (int u = 0; u < kVjpGroup; ++u) {
            const long long s = 4 * (top - u);
            r0[u] = r1[u] = r2[u] = r3[u] = 0.0;
            if (top - u >= 0) {
                r0[u] = G[s * stride];
                r1[u] = G[(s + 1) * stride];
                r2[u] = G[(s + 2) * stride];
                r3[u] = G[(s + 3) * stride];
            }
        }
#pragma unroll
        for (int u = 0; u < kVjpGroup; ++u) {
            if (top - u < 0) break;
            const long long s = 4 * (top - u);
            const double g1 = r1[u], g2 = r2[u], g3 = r3[u];
            double a0 = r0[u];
            a0 += g1;
            a0 += g2;
            a0 += g3;
            a0 += g4;
            double a1 = g1 * (dt / 2);
            a1 += g2 * (4 * dt / 3);
            a1 += g3 * (9 * dt / 8);
            a1 += g4 * (64 * dt / 45);
            double a2 = g2 * (dt / 3);
            a2 += g3 * (9 * dt / 8);
            a2 += g4 * (24 * dt / 45);
            double a3 = g3 * (3 * dt / 8);
            a3 += g4 * (64 * dt / 45);
            double a4 = g4 * (14 * dt / 45);
            a4 += pending;
            G[(s + 1) * stride] = a1;
            G[(s + 2) * stride] = a2;
            G[(s + 3) * stride] = a3;
            G[(s + 4) * stride] = a4;
            double p = g1 * (dt / 2);
            p += g2 * (dt / 3);
            p += g3 * (3 * dt / 8);
            p += g4 * (14 * dt / 45);
            pending = p;
            g4 = a0;  // G[s] after its update is g4 of the block below
        }
    }
    G[0] = pending;
}

// simpson.hpp:77-123 (the view swap of pending/next_pending is a value swap)
__device__ __forceinline__ void prefix_vjp_simpson(double* G, long long stride, long long S, double dt) {
    const double half_dt = dt / 2;
    const double third_dt = dt / 3;
    double pending = 0.0, next_pending = 0.0;
    auto reverse_odd = [&](long long i) {
        const double bar = G[i * stride];
        G[(i - 1) * stride] += bar;
        double v = bar * half_dt;
        v += pending;
        G[i * stride] = v;
        next_pending += bar * half_dt;
        pending = next_pending;
        next_pending = 0.0;
    };
    auto reverse_even = [&](long long i) {
        const double bar = G[i * stride];
        G[(i - 2) * stride] += bar;
        double v = bar * third_dt;
        v += pending;
        G[i * stride] = v;
        next_pending += bar * (4 * third_dt);
        const double np = bar * third_dt;
        pending = next_pending;
        next_pending = np;
    };
    long long i = S - 1;
    if (i & 1) {
        reverse_odd(i);
        --i;
    }
    while (i > 2) {
        reverse_even(i);
        reverse_odd(i - 1);
        i -= 2;
    }
    reverse_even(2);
    const double bar = G[stride];
    double v1 = bar * half_dt;
    v1 += pending;
    G[stride] = v1;
    double v0 = bar * half_dt;
    v0 += next_pending;
    G[0] = v0;
}

// trapezoid.hpp:43-52 (reconstruction-style recurrence of the reference)
__device__ __forceinline__ void prefix_vjp_trapezoid(double* G, long long stride, long long S, double dt) {
    const double half_dt = dt / 2;
    double tmp = 0.0;
    for (long long i = S - 1; i > 0; --i) {
        double g = G[i * stride] * half_dt;
        g += tmp;
        G[i * stride] = g;
        tmp *= -1;
        tmp += g * 2;
    }
    G[0] = tmp * 0.5;
}

// default.hpp:32-38
__device__ __forceinline__ void prefix_vjp_riemann(double* G, long long stride, long long S, double dt) {
    double carry = G[(S - 1) * stride];
    for (long long i = S - 1; i > 0; --i) {
        double prev = G[(i - 1) * stride];
        prev += carry;
        G[i * stride] = carry * dt;
        carry = prev;
    }
    G[0] = carry * dt;
}

template <int I>
__device__ __forceinline__ void prefix_vjp(double* G, long long stride, long long S, double dt) {
    if constexpr (I == BOOLE) prefix_vjp_boole(G, stride, S, dt);
    else if constexpr (I == SIMPSON) prefix_vjp_simpson(G, stride, S, dt);
    else if constexpr (I == TRAPEZOID) prefix_vjp_trapezoid(G, stride, S, dt);
    else prefix_vjp_riemann(G, stride, S, dt);
}

}  // namespace magnus
