// kernels_uneven.hpp -- packing kernels of the uneven-batch plans: instances
// given as per-instance device pointers are copied into (gather) and out of
// (scatter) the packed [count][per] layout the level-parallel kernels read.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace magnus {

struct SegmentArgs {
    const double* const* tab;  // per-instance pointers, indexed by instance id
    const uint32_t* ids;       // instance id of each packed slot
    double* packed;            // [count][per]
    long long per;             // doubles per instance
    int count;
};

hipError_t launch_gather(const SegmentArgs& a, hipStream_t s);   // packed[k] = tab[ids[k]]
hipError_t launch_scatter(const SegmentArgs& a, hipStream_t s);  // tab[ids[k]] = packed[k]

}  // namespace magnus
