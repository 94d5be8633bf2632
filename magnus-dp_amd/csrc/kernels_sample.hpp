// kernels_sample.hpp -- arguments of the device sampling kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace magnus {

struct SampleArgs {
    const double* coeffs;  // matrix form: [B][2*modes+1][m]
    const double* amp;     // SpaceCurve form: [B][modes][3]
    const double* phase;   // SpaceCurve form: [B][modes][3]
    double* out;           // [B][S][m] or [B][S][3]
    long long B;
    long long S;
    int m, modes;
    int sb;                // samples per workgroup (set by the launcher)
    int ib;                // instances per workgroup, sharing its sin/cos table (set by the launcher)
    double t0, tf, step;   // step = (tf - t0) / (S - 1), as numpy.linspace computes it
};

int sample_max_modes();
// sin, cos of (2 pi k) t_i on the grid of a (S, t0, tf, step, modes): out[i][k-1][0..1],
// the values the sampling kernels use
hipError_t launch_fourier_table(const SampleArgs& a, double* out, hipStream_t s);
hipError_t launch_sample_fourier(SampleArgs a, hipStream_t s);
hipError_t launch_sample_sc_fourier(const SampleArgs& a, hipStream_t s);

}  // namespace magnus
