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
    double* fsum;       // [B][C][2K+1][4]  chunk summaries [Vhat_1..K, ZV_0..K]
    double* src;        // [B][C][2K+1][4]  chunk cotangents [Vhat_bar_1..K, ZV_bar_0..K]  (VJP)
    double* bases;      // [B][C][K][4]     stream values at chunk starts (VJP scratch)
    double* edge;       // [B][C][6]        dA at the chunk end / start, before the fixup (VJP)
    double* dA;         // [B][S][3]        (VJP)
    long long B;
    long long S;
    int C;   // chunks per instance
    int LC;  // samples per chunk (multiple of the integrator block)
    int K;   // n - 1
    int n_out, op, Q;
    int cw;  // lanes per instance in the combine (set by launch_wb)
    double dt;
    double w[kArgMaxQ];
    double x[kArgMaxQ];
};

// forward only (vjp == false): chunk pass + combine -> out.
// vjp == true: chunk pass + combine (cotangents per chunk) + reverse chunk pass + fixup -> dA;
// summaries_given: fsum already holds the chunk summaries of this input (no chunk pass).
hipError_t launch_wb(int integ, const WbArgs& a, bool vjp, bool summaries_given, hipStream_t st);

}  // namespace magnus
