// kernels_expm.hip -- batched FP64 propagators exp(Omega), the step after the
// Magnus path (the reference's tests/bench.py:93 applies scipy.linalg.expm to
// the Magnus sum).
//
// Dense d x d: scaling and squaring with diagonal Pade approximants (N. J.
// Higham, "The scaling and squaring method for the matrix exponential
// revisited", SIAM J. Matrix Anal. Appl. 26(4), 2005, Algorithm 2.3): the
// degree m in {3, 5, 7, 9, 13} is the smallest whose theta_m bounds the
// 1-norm, otherwise m = 13 after scaling by 2^-s; r_m = (V - U)^{-1} (V + U)
// by LU with partial pivoting; then s squarings.  One workgroup per matrix,
// work matrices in a per-matrix scratch block of global memory.
//
// SpaceCurve: Omega = omega (3-vector) stands for -i omega.sigma, and
// exp(-i omega.sigma) = cos|omega| I - i sin|omega| (omega/|omega|).sigma.
#include "kernels_expm.hpp"

namespace magnus {

namespace {

constexpr int kExpmThreads = 256;

// Pade coefficients b_0..b_m (Higham 2005, (2.11)) and theta_m (Table 2.3)
__constant__ double kPade3[4] = {120.0, 60.0, 12.0, 1.0};
__constant__ double kPade5[6] = {30240.0, 15120.0, 3360.0, 420.0, 30.0, 1.0};
__constant__ double kPade7[8] = {17297280.0, 8648640.0, 1995840.0, 277200.0, 25200.0, 1512.0, 56.0, 1.0};
__constant__ double kPade9[10] = {17643225600.0, 8821612800.0, 2075673600.0, 302702400.0, 30270240.0,
                                  2162160.0, 110880.0, 3960.0, 90.0, 1.0};
__constant__ double kPade13[14] = {64764752532480000.0, 32382376266240000.0, 7771770303897600.0,
                                   1187353796428800.0, 129060195264000.0, 10559470521600.0, 670442572800.0,
                                   33522128640.0, 1323241920.0, 40840800.0, 960960.0, 16380.0, 182.0, 1.0};
constexpr double kTheta3 = 1.495585217958292e-2;
constexpr double kTheta5 = 2.539398330063230e-1;
constexpr double kTheta7 = 9.504178996162932e-1;
constexpr double kTheta9 = 2.097847961257068e0;
constexpr double kTheta13 = 5.371920351148152e0;

// C = A * B (d x d, row-major), all threads of the workgroup
__device__ void wg_matmul(const double* A, const double* B, double* C, int d) {
    for (int e = threadIdx.x; e < d * d; e += blockDim.x) {
        const int r = e / d, c = e - r * d;
        double acc = 0.0;
        for (int l = 0; l < d; ++l) acc = fma(A[r * d + l], B[l * d + c], acc);
        C[e] = acc;
    }
    __syncthreads();
}

__global__ void __launch_bounds__(kExpmThreads) expm_kernel(ExpmArgs a) {
    const int d = a.d;
    const long long mm = (long long)d * d;
    const long long b = blockIdx.x;
    const double* Ain = a.in + b * mm;
    double* W = a.work + b * mm * 8;  // A, A2, A4, A6, U, V, T, spare
    double *A = W, *A2 = W + mm, *A4 = W + 2 * mm, *A6 = W + 3 * mm, *U = W + 4 * mm, *V = W + 5 * mm,
           *T = W + 6 * mm;
    __shared__ double red[kExpmThreads];
    __shared__ int ired[kExpmThreads];
    __shared__ double s_norm;
    __shared__ int s_piv;

    // 1-norm (max column sum)
    double cmax = 0.0;
    for (int c = threadIdx.x; c < d; c += blockDim.x) {
        double sum = 0.0;
        for (int r = 0; r < d; ++r) sum += fabs(Ain[r * d + c]);
        cmax = fmax(cmax, sum);
    }
    red[threadIdx.x] = cmax;
    __syncthreads();
    for (int w = kExpmThreads / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) s_norm = red[0];
    __syncthreads();
    const double nrm = s_norm;
    int m = 13, s = 0;
    if (nrm <= kTheta3) m = 3;
    else if (nrm <= kTheta5) m = 5;
    else if (nrm <= kTheta7) m = 7;
    else if (nrm <= kTheta9) m = 9;
    else if (isfinite(nrm)) {  // non-finite input: no scaling (the result is non-finite anyway)
        const double ratio = nrm / kTheta13;
        if (ratio > 1.0) s = (int)ceil(log2(ratio));
        if (s > 1100) s = 1100;  // 2^-1100 underflows every double: more squarings change nothing
    }
    const double scale = ldexp(1.0, -s);
    for (long long e = threadIdx.x; e < mm; e += blockDim.x) A[e] = Ain[e] * scale;
    __syncthreads();
    wg_matmul(A, A, A2, d);
    const double* bc = m == 3 ? kPade3 : m == 5 ? kPade5 : m == 7 ? kPade7 : m == 9 ? kPade9 : kPade13;
    if (m >= 5) wg_matmul(A2, A2, A4, d);
    if (m >= 7) wg_matmul(A4, A2, A6, d);
    if (m == 13) {
        // U = A [A6 (b13 A6 + b11 A4 + b9 A2) + b7 A6 + b5 A4 + b3 A2 + b1 I]
        // V =    A6 (b12 A6 + b10 A4 + b8 A2) + b6 A6 + b4 A4 + b2 A2 + b0 I
        for (long long e = threadIdx.x; e < mm; e += blockDim.x) {
            T[e] = bc[13] * A6[e] + bc[11] * A4[e] + bc[9] * A2[e];
            V[e] = bc[12] * A6[e] + bc[10] * A4[e] + bc[8] * A2[e];
        }
        __syncthreads();
        wg_matmul(A6, T, U, d);  // U holds A6 * T
        for (long long e = threadIdx.x; e < mm; e += blockDim.x) {
            const int r = (int)(e / d), c = (int)(e - (long long)r * d);
            T[e] = U[e] + bc[7] * A6[e] + bc[5] * A4[e] + bc[3] * A2[e] + (r == c ? bc[1] : 0.0);
        }
        __syncthreads();
        wg_matmul(A6, V, U, d);  // U holds A6 * V
        for (long long e = threadIdx.x; e < mm; e += blockDim.x) {
            const int r = (int)(e / d), c = (int)(e - (long long)r * d);
            V[e] = U[e] + bc[6] * A6[e] + bc[4] * A4[e] + bc[2] * A2[e] + (r == c ? bc[0] : 0.0);
        }
        __syncthreads();
        wg_matmul(A, T, U, d);
    } else {
        // U = A sum_{k odd} b_k A^{k-1},  V = sum_{k even} b_k A^k
        for (long long e = threadIdx.x; e < mm; e += blockDim.x) {
            const int r = (int)(e / d), c = (int)(e - (long long)r * d);
            const double id = r == c ? 1.0 : 0.0;
            double u = bc[1] * id + bc[3] * A2[e], v = bc[0] * id + bc[2] * A2[e];
            if (m >= 5) {
                u += bc[5] * A4[e];
                v += bc[4] * A4[e];
            }
            if (m >= 7) {
                u += bc[7] * A6[e];
                v += bc[6] * A6[e];
            }
            T[e] = u;
            V[e] = v;
        }
        __syncthreads();
        if (m == 9) {  // the A^8 terms: A8 = A4 A4 into A2's slot (A2 no longer needed)
            wg_matmul(A4, A4, A2, d);
            for (long long e = threadIdx.x; e < mm; e += blockDim.x) {
                T[e] += bc[9] * A2[e];
                V[e] += bc[8] * A2[e];
            }
            __syncthreads();
        }
        wg_matmul(A, T, U, d);
    }
    // P = V + U (into T), Q = V - U (into V); solve Q R = P
    double* P = T;
    double* Q = V;
    for (long long e = threadIdx.x; e < mm; e += blockDim.x) {
        const double u = U[e], v = V[e];
        P[e] = v + u;
        Q[e] = v - u;
    }
    __syncthreads();
    for (int k = 0; k < d; ++k) {
        // pivot: largest |Q[i][k]|, i >= k (first on ties)
        double best = -1.0;
        int bi = k;
        for (int i = k + threadIdx.x; i < d; i += blockDim.x) {
            const double v = fabs(Q[i * d + k]);
            if (v > best) {
                best = v;
                bi = i;
            }
        }
        red[threadIdx.x] = best;
        ired[threadIdx.x] = bi;
        __syncthreads();
        for (int w = kExpmThreads / 2; w > 0; w >>= 1) {
            if (threadIdx.x < w) {
                const double o = red[threadIdx.x + w];
                const int oi = ired[threadIdx.x + w];
                if (o > red[threadIdx.x] || (o == red[threadIdx.x] && oi < ired[threadIdx.x])) {
                    red[threadIdx.x] = o;
                    ired[threadIdx.x] = oi;
                }
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) s_piv = ired[0];
        __syncthreads();
        const int p = s_piv;
        if (p != k) {
            for (int j = threadIdx.x; j < d; j += blockDim.x) {
                double t = Q[k * d + j];
                Q[k * d + j] = Q[p * d + j];
                Q[p * d + j] = t;
                t = P[k * d + j];
                P[k * d + j] = P[p * d + j];
                P[p * d + j] = t;
            }
            __syncthreads();
        }
        const double piv = Q[k * d + k];
        // eliminate below: rows i > k, columns j > k of Q and all columns of P
        const int rows = d - k - 1;
        for (int x = threadIdx.x; x < rows * (2 * d); x += blockDim.x) {
            const int i = k + 1 + x / (2 * d);
            const int j = x % (2 * d);
            const double f = Q[i * d + k] / piv;
            if (j < d) {
                if (j > k) Q[i * d + j] = fma(-f, Q[k * d + j], Q[i * d + j]);
            } else {
                P[i * d + (j - d)] = fma(-f, P[k * d + (j - d)], P[i * d + (j - d)]);
            }
        }
        __syncthreads();
        // the multipliers Q[i][k] are read above by every thread of row i before any
        // write to column k; clear them now that the row updates are done
        for (int i = k + 1 + threadIdx.x; i < d; i += blockDim.x) Q[i * d + k] = 0.0;
        __syncthreads();
    }
    // back substitution, one right-hand side (column of P) per thread
    for (int c = threadIdx.x; c < d; c += blockDim.x) {
        for (int k = d - 1; k >= 0; --k) {
            double v = P[k * d + c];
            for (int j = k + 1; j < d; ++j) v = fma(-Q[k * d + j], P[j * d + c], v);
            P[k * d + c] = v / Q[k * d + k];
        }
    }
    __syncthreads();
    // s squarings, ping-pong between P and U
    double* R = P;
    double* S2 = U;
    for (int q = 0; q < s; ++q) {
        wg_matmul(R, R, S2, d);
        double* t = R;
        R = S2;
        S2 = t;
    }
    double* out = a.out + b * mm;
    for (long long e = threadIdx.x; e < mm; e += blockDim.x) out[e] = R[e];
}

// exp(-i omega.sigma) as a complex 2x2 (re, im interleaved), one thread per instance
__global__ void __launch_bounds__(256) sc_expm_kernel(ExpmArgs a) {
    const long long b = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= a.B) return;
    const double w1 = a.in[b * 3 + 0], w2 = a.in[b * 3 + 1], w3 = a.in[b * 3 + 2];
    const double th = sqrt(w1 * w1 + w2 * w2 + w3 * w3);
    const double c = cos(th);
    double sn;  // sin(th) / th
    if (th < 1e-4) {
        const double t2 = th * th;
        sn = 1.0 - t2 / 6.0 + t2 * t2 / 120.0;
    } else {
        sn = sin(th) / th;
    }
    // -i omega.sigma = [[-i w3, -w2 - i w1], [w2 - i w1, i w3]]
    double* o = a.out + b * 8;
    o[0] = c;        o[1] = -sn * w3;  // U00
    o[2] = -sn * w2; o[3] = -sn * w1;  // U01
    o[4] = sn * w2;  o[5] = -sn * w1;  // U10
    o[6] = c;        o[7] = sn * w3;   // U11
}

}  // namespace

hipError_t launch_expm(const ExpmArgs& a, hipStream_t s) {
    void* h = timer_begin("expm", s);
    hipLaunchKernelGGL(expm_kernel, dim3(a.B), dim3(kExpmThreads), 0, s, a);
    timer_end(h);
    return hipGetLastError();
}

hipError_t launch_sc_expm(const ExpmArgs& a, hipStream_t s) {
    void* h = timer_begin("sc_expm", s);
    hipLaunchKernelGGL(sc_expm_kernel, dim3((a.B + 255) / 256), dim3(256), 0, s, a);
    timer_end(h);
    return hipGetLastError();
}

}  // namespace magnus
