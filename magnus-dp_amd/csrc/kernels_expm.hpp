// kernels_expm.hpp -- arguments of the batched propagator kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace magnus {

struct ExpmArgs {
    const double* in;  // dense: [B][d][d]; SpaceCurve: [B][3]
    double* out;       // dense: [B][d][d]; SpaceCurve: [B][2][2] complex (re, im)
    double* work;      // dense: [B][8][d][d] scratch
    long long B;
    int d;
};

hipError_t launch_expm(const ExpmArgs& a, hipStream_t s);
hipError_t launch_sc_expm(const ExpmArgs& a, hipStream_t s);

}  // namespace magnus
