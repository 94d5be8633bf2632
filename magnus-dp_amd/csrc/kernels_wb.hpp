// kernels_wb.hpp -- arguments of the single-pass (W-basis) SpaceCurve kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace magnus {

constexpr int kWbMaxK = 7;  // n <= 8

struct WbArgs {
    const double* in;   // [B][S][3]
    const double* cot;  // [B][n_out][3]            (VJP)
    double* out;        // [B][n_out][3] or nullptr
    double* fsum;       // [B][C][2K+1][4]          chunk summaries (forward)
    double* bases;      // [B][C][K][4] or nullptr  stream values at chunk starts
    double* what;       // [B][K][4] or nullptr     stream totals W_j[S-1]
    double* src;        // [B][2K+1][4] or nullptr  adjoint sources (Zbar_0..K, What_bar_1..K)
    double* wbuf;       // [B][K][LC][C][4]         streams, interleaved by chunk
    double* w0;         // [B][K][4]                streams at sample 0
    double* rsum;       // [B][C][2K-1][4]          chunk summaries (reverse)
    double* uin;        // [B][C][K][4]             incoming suffix sums per chunk
    double* dA;         // [B][S][3]                (VJP)
    long long B;
    long long S;
    int C;   // chunks per instance
    int LC;  // samples per chunk (multiple of the integrator block)
    int K;   // n - 1
    int n_out, op, Q;
    double dt;
    double w[kArgMaxQ];
    double x[kArgMaxQ];
};

// forward only (vjp == false): pass 1 + combine -> out.
// vjp == true: pass 1 + combine (adjoint sources) + stream pass + reverse passes -> dA.
hipError_t launch_wb(int integ, const WbArgs& a, bool vjp, hipStream_t st);

}  // namespace magnus
