// kernels_ep.hip -- entry-parallel Magnus kernels (forward, VJP, reductions).
//
// Work decomposition: one "job" = (instance b, Gauss-Legendre node q).  A job
// owns S x m scalars per level (m = d*d, or 4 for the su(2) 4-vectors); its
// threads own disjoint matrix entries and run the reference's per-entry
// recurrences sample by sample, so prefix scans, sums and reverse sweeps keep
// the reference's order of operations.  Per-sample products read the other
// entries of the same sample from the level buffer written at the previous
// level (workgroup barrier between levels).  Several jobs share a workgroup
// when m is small.  The q-partials are reduced by a second kernel in fixed q
// order (deterministic, no atomics).
//
// Reference: Magnus::{one,many,sum} (src/magnus.hpp:12-261),
//            VJP::{none,one,many,sum} (src/grad.hpp:14-349).
#include "kernels.hpp"
#include "rules.hpp"

namespace magnus {

// ---------------------------------------------------------------------------
// per-sample products
// ---------------------------------------------------------------------------

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

// ---------------------------------------------------------------------------
// forward
// ---------------------------------------------------------------------------
template <int I, int EPT>
__global__ void __launch_bounds__(1024) ep_forward_kernel(EpFwdArgs a) {
    const int tpj = a.tpj;
    const int jl = threadIdx.x / tpj;
    const int lt = threadIdx.x - jl * tpj;
    const long long job = (long long)blockIdx.x * a.jpw + jl;
    const bool active = (jl < a.jpw) && (job < a.jobs);
    const long long S = a.S;
    const int m = a.m, d = a.d, n = a.n;
    const int scm = a.sc;                 // 0 generic, 1 padded su(2), 2 su(2) on (S,2,2) input
    const bool sc = scm != 0;
    const int aoff = (scm == 2) ? 1 : 0;  // a1 of the 4-vector
    const long long b = active ? job / a.Qj : 0;
    const int q = active ? (int)(job % a.Qj) : 0;
    const double dt = a.dt;
    const double shift = a.x[q] - 1;           // magnus.hpp:65
    const double fscale = dt * a.w[q];         // magnus.hpp:66
    const double* __restrict__ A = a.in + b * S * a.m_in;

    auto slot = [&](int j) -> double* {        // P^{(j)} = carry[q][j-1]
        if (a.carry) return a.carry + ((b * a.Q + q) * (long long)(n - 1) + (j - 1)) * S * m;
        return a.pingpong + (job * 2 + (j & 1)) * S * m;
    };

    double acc_sum[EPT];
#pragma unroll
    for (int u = 0; u < EPT; ++u) acc_sum[u] = 0.0;

    for (int j = 1; j <= n; ++j) {
        if (active) {
            const double* __restrict__ Pin = (j >= 2) ? slot(j - 1) : nullptr;
            const double* __restrict__ T = (j >= 2) ? Pin + (S - 1) * m : nullptr;
            double* __restrict__ Pout = (j < n) ? slot(j) : nullptr;
            // which output this level feeds
            const bool om1 = a.outputs && j == 1 && q == 0 && (a.op != OP_ONE || n == 1);
            const bool lvl = a.outputs && j >= 2 && (a.op != OP_ONE || j == n);
            const double c = (j == 1) ? dt : fscale;
            PrefixState<I> st[EPT];
            double acc[EPT];
#pragma unroll
            for (int u = 0; u < EPT; ++u) acc[u] = 0.0;
            for (long long i = 0; i < S; ++i) {
                const double* __restrict__ Ai = A + i * a.m_in;
#pragma unroll
                for (int u = 0; u < EPT; ++u) {
                    const int e = lt + u * tpj;
                    if (e < m) {
                        double f;
                        if (j == 1) {
                            f = raw_entry(Ai, e, scm);
                        } else if (sc) {
                            f = prod_sc(Ai + aoff, Pin + i * m, T, e, shift);
                        } else {
                            const int r = e / d, cc = e - r * d;
                            f = prod_generic(Ai, Pin + i * m, T, d, r, cc, shift);
                        }
                        if (om1 || lvl) sum_acc<I>(acc[u], i, S, f, c);
                        if (Pout) Pout[i * m + e] = st[u].step(i, f, dt);
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < EPT; ++u) {
                const int e = lt + u * tpj;
                if (e >= m) continue;
                if (om1) a.omega1[b * m + e] = acc[u];
                if (lvl) {
                    if (a.op == OP_MANY) a.partial[(job * a.n_out + (j - 1)) * m + e] = acc[u];
                    else acc_sum[u] += acc[u];
                }
            }
        }
        __syncthreads();
    }
    if (active && a.outputs && n > 1 && a.op != OP_MANY) {
#pragma unroll
        for (int u = 0; u < EPT; ++u) {
            const int e = lt + u * tpj;
            if (e < m) a.partial[job * a.n_out * m + e] = acc_sum[u];
        }
    }
}

// out[b] = Omega_1 + sum_q partial[b][q]   (lane reduction, magnus.hpp:86-90)
__global__ void fwd_reduce_kernel(FwdReduceArgs a) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long per = (long long)a.n_out * a.m_out;
    if (idx >= a.B * per) return;
    const long long b = idx / per;
    const int rem = (int)(idx - b * per);
    const int k = rem / a.m_out;
    const int eo = rem - k * a.m_out;
    const int e = a.sc ? eo + 1 : eo;  // SpaceCurve output drops component 0
    double v = 0.0;
    if (a.n == 1) {
        v = (k == 0) ? a.omega1[b * a.m + e] : 0.0;
    } else {
        if (a.op != OP_ONE && k == 0) v = a.omega1[b * a.m + e];
        if (!(a.op == OP_MANY && k == 0))  // Omega_1 has no GL-node contributions
            for (int q = 0; q < a.Qj; ++q) v += a.partial[((b * a.Qj + q) * a.n_out + k) * a.m + e];
    }
    a.out[idx] = v;
}

// ---------------------------------------------------------------------------
// VJP
// ---------------------------------------------------------------------------

// da += g b^T and temp = a^T g for one owned entry (manual.hpp:19-42)
template <int I, int EPT>
__global__ void __launch_bounds__(1024) ep_vjp_kernel(EpVjpArgs a) {
    const int tpj = a.tpj;
    const int jl = threadIdx.x / tpj;
    const int lt = threadIdx.x - jl * tpj;
    const long long job = (long long)blockIdx.x * a.jpw + jl;
    const bool active = (jl < a.jpw) && (job < a.jobs);
    const long long S = a.S;
    const int m = a.m, d = a.d, n = a.n;
    const int scm = a.sc;
    const bool sc = scm != 0;
    const int aoff = (scm == 2) ? 1 : 0;
    const long long b = active ? job / a.Q : 0;
    const int q = active ? (int)(job % a.Q) : 0;
    const double dt = a.dt;
    const double shift = a.x[q] - 1.0;         // grad.hpp:314
    const double dscale = dt * a.w[q];         // grad.hpp:315
    const double* __restrict__ A = a.in + b * S * a.m_in;
    const double* __restrict__ cot = a.cot + b * (long long)a.n_out * a.m_cot;
    double* Ga = a.gbuf + (job * 2 + 0) * S * m;
    double* Gb = a.gbuf + (job * 2 + 1) * S * m;
    double* __restrict__ dA = a.dAjob + job * S * m;

    auto cot_entry = [&](int k, int e) -> double {  // padded cotangent row k
        const double* row = cot + (long long)k * a.m_cot;
        if (scm == 1) return e == 0 ? 0.0 : row[e - 1];
        return row[e];
    };

    // barY = sum_vjp(dOut_last, dscale)  (grad.hpp:127, 216, 317)
    if (active) {
        const int klast = (a.op == OP_MANY) ? n - 1 : 0;
        for (int u = 0; u < EPT; ++u) {
            const int e = lt + u * tpj;
            if (e >= m) break;
            const double o = cot_entry(klast, e);
            for (long long i = 0; i < S; ++i) {
                Ga[i * m + e] = o * sum_vjp_coef<I>(i, S, dscale);
                dA[i * m + e] = 0.0;
            }
        }
    }
    __syncthreads();

    for (int k = n; k >= 2; --k) {
        if (active) {
            const double* __restrict__ P = a.carry + ((b * a.Q + q) * (long long)(n - 1) + (k - 2)) * S * m;
            const double* __restrict__ T = P + (S - 1) * m;
            double bt[EPT];
            // sample_update_vjp (generic.hpp:50-86, fixed.hpp:167-217, spacecurve.hpp:132-191)
            for (long long i = S - 1; i >= 0; --i) {
                const double* __restrict__ Ai = A + i * a.m_in;
                const double* __restrict__ Pi = P + i * m;
                const double* __restrict__ g = Ga + i * m;
#pragma unroll
                for (int u = 0; u < EPT; ++u) {
                    const int e = lt + u * tpj;
                    if (e >= m) continue;
                    double dav, temp;
                    if (sc) {
                        double b0 = Pi[0]; b0 += T[0] * shift;
                        double b1 = Pi[1]; b1 += T[1] * shift;
                        double b2 = Pi[2]; b2 += T[2] * shift;
                        double b3 = Pi[3]; b3 += T[3] * shift;
                        const double g0 = g[0], g1 = g[1], g2 = g[2], g3 = g[3];
                        const double a1 = Ai[aoff], a2 = Ai[aoff + 1], a3 = Ai[aoff + 2];
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
                        if (e != 0) dA[i * m + e] += dav;
                    } else {
                        const int r = e / d, cc = e - r * d;
                        // da[r][cc] += sum_j g[r][j] * b[cc][j]
                        double v = 0.0;
                        for (int jj = 0; jj < d; ++jj) {
                            double bb = Pi[cc * d + jj];
                            bb += T[cc * d + jj] * shift;
                            v += g[r * d + jj] * bb;
                        }
                        dA[i * m + e] += v;
                        // temp[r][cc] = sum_i' a[i'][r] * g[i'][cc]
                        double t2 = 0.0;
                        for (int ii = 0; ii < d; ++ii) t2 += Ai[ii * d + r] * g[ii * d + cc];
                        temp = t2;
                    }
                    if (i == S - 1) {
                        if (a.bt_fused) {
                            bt[u] = temp * (1.0 + shift);
                        } else {
                            bt[u] = temp;
                            bt[u] += temp * shift;
                        }
                    } else {
                        bt[u] += temp * shift;
                        Gb[i * m + e] = temp;
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < EPT; ++u) {
                const int e = lt + u * tpj;
                if (e >= m) continue;
                Gb[(S - 1) * m + e] = bt[u];
                prefix_vjp<I>(Gb + e, m, S, dt);
                if (k > 2 && a.op != OP_ONE) {  // sum_vjp_add (grad.hpp:232, 332)
                    const double o = cot_entry(a.op == OP_MANY ? k - 2 : 0, e);
                    for (long long i = 0; i < S; ++i)
                        Gb[i * m + e] = sum_vjp_add<I>(Gb[i * m + e], i, S, o, dscale);
                }
            }
        }
        double* t = Ga;
        Ga = Gb;
        Gb = t;
        __syncthreads();
    }
    // dA += barY (grad.hpp:144, 247, 347)
    if (active) {
        for (int u = 0; u < EPT; ++u) {
            const int e = lt + u * tpj;
            if (e >= m) break;
            for (long long i = 0; i < S; ++i) dA[i * m + e] += Ga[i * m + e];
        }
    }
}

// dA[b] = sum_vjp(dOut_first, dt) + sum_q dA_q[b]
template <int I>
__global__ void vjp_reduce_kernel(VjpReduceArgs a) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long per = a.S * a.m_out;
    if (idx >= a.B * per) return;
    const long long b = idx / per;
    const long long rem = idx - b * per;
    const long long i = rem / a.m_out;
    const int eo = (int)(rem - i * a.m_out);
    const int e = a.sc ? eo + 1 : eo;
    double v = 0.0;
    if (!(a.op == OP_ONE && a.n > 1)) {
        const double* row = a.cot + b * (long long)a.n_out * a.m_cot;  // dOut[0] / dOut
        const double o = a.sc ? row[e - 1] : row[e];
        v = o * sum_vjp_coef<I>(i, a.S, a.dt);
    }
    for (int q = 0; q < a.Qj; ++q) v += a.dAjob[((b * a.Qj + q) * a.S + i) * a.m + e];
    a.dA[idx] = v;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <int I>
static hipError_t launch_ep_forward_i(const EpFwdArgs& a, int block, long long grid, int ept, hipStream_t s) {
    switch (ept) {
        case 1: hipLaunchKernelGGL((ep_forward_kernel<I, 1>), dim3(grid), dim3(block), 0, s, a); break;
        case 2: hipLaunchKernelGGL((ep_forward_kernel<I, 2>), dim3(grid), dim3(block), 0, s, a); break;
        case 4: hipLaunchKernelGGL((ep_forward_kernel<I, 4>), dim3(grid), dim3(block), 0, s, a); break;
        case 8: hipLaunchKernelGGL((ep_forward_kernel<I, 8>), dim3(grid), dim3(block), 0, s, a); break;
        default: hipLaunchKernelGGL((ep_forward_kernel<I, 16>), dim3(grid), dim3(block), 0, s, a); break;
    }
    return hipGetLastError();
}

hipError_t launch_ep_forward(int integ, const EpFwdArgs& a, int block, long long grid, int ept, hipStream_t s) {
    switch (integ) {
        case BOOLE: return launch_ep_forward_i<BOOLE>(a, block, grid, ept, s);
        case SIMPSON: return launch_ep_forward_i<SIMPSON>(a, block, grid, ept, s);
        case TRAPEZOID: return launch_ep_forward_i<TRAPEZOID>(a, block, grid, ept, s);
        default: return launch_ep_forward_i<RIEMANN>(a, block, grid, ept, s);
    }
}

template <int I>
static hipError_t launch_ep_vjp_i(const EpVjpArgs& a, int block, long long grid, int ept, hipStream_t s) {
    switch (ept) {
        case 1: hipLaunchKernelGGL((ep_vjp_kernel<I, 1>), dim3(grid), dim3(block), 0, s, a); break;
        case 2: hipLaunchKernelGGL((ep_vjp_kernel<I, 2>), dim3(grid), dim3(block), 0, s, a); break;
        case 4: hipLaunchKernelGGL((ep_vjp_kernel<I, 4>), dim3(grid), dim3(block), 0, s, a); break;
        case 8: hipLaunchKernelGGL((ep_vjp_kernel<I, 8>), dim3(grid), dim3(block), 0, s, a); break;
        default: hipLaunchKernelGGL((ep_vjp_kernel<I, 16>), dim3(grid), dim3(block), 0, s, a); break;
    }
    return hipGetLastError();
}

hipError_t launch_ep_vjp(int integ, const EpVjpArgs& a, int block, long long grid, int ept, hipStream_t s) {
    switch (integ) {
        case BOOLE: return launch_ep_vjp_i<BOOLE>(a, block, grid, ept, s);
        case SIMPSON: return launch_ep_vjp_i<SIMPSON>(a, block, grid, ept, s);
        case TRAPEZOID: return launch_ep_vjp_i<TRAPEZOID>(a, block, grid, ept, s);
        default: return launch_ep_vjp_i<RIEMANN>(a, block, grid, ept, s);
    }
}

hipError_t launch_fwd_reduce(const FwdReduceArgs& a, hipStream_t s) {
    const long long total = a.B * (long long)a.n_out * a.m_out;
    if (total == 0) return hipSuccess;
    const int block = 256;
    const long long grid = (total + block - 1) / block;
    hipLaunchKernelGGL(fwd_reduce_kernel, dim3(grid), dim3(block), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_vjp_reduce(int integ, const VjpReduceArgs& a, hipStream_t s) {
    const long long total = a.B * a.S * a.m_out;
    if (total == 0) return hipSuccess;
    const int block = 256;
    const long long grid = (total + block - 1) / block;
    switch (integ) {
        case BOOLE: hipLaunchKernelGGL((vjp_reduce_kernel<BOOLE>), dim3(grid), dim3(block), 0, s, a); break;
        case SIMPSON: hipLaunchKernelGGL((vjp_reduce_kernel<SIMPSON>), dim3(grid), dim3(block), 0, s, a); break;
        case TRAPEZOID: hipLaunchKernelGGL((vjp_reduce_kernel<TRAPEZOID>), dim3(grid), dim3(block), 0, s, a); break;
        default: hipLaunchKernelGGL((vjp_reduce_kernel<RIEMANN>), dim3(grid), dim3(block), 0, s, a); break;
    }
    return hipGetLastError();
}

}  // namespace magnus
