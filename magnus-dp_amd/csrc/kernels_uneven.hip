// kernels_uneven.hip -- gather / scatter between per-instance buffers and the
// packed layout of one bucket of an uneven batch.  Grid: x over the segment
// (grid-stride), y over the packed slots; consecutive lanes copy consecutive
// doubles, so both sides are coalesced.
#include "kernels.hpp"
#include "kernels_uneven.hpp"

namespace magnus {

namespace {

__global__ void __launch_bounds__(256) gather_kernel(SegmentArgs a) {
    const int k = blockIdx.y;
    const double* src = a.tab[a.ids[k]];
    double* dst = a.packed + (long long)k * a.per;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < a.per; i += (long long)gridDim.x * 256)
        dst[i] = src[i];
}

__global__ void __launch_bounds__(256) scatter_kernel(SegmentArgs a) {
    const int k = blockIdx.y;
    double* dst = const_cast<double*>(a.tab[a.ids[k]]);
    const double* src = a.packed + (long long)k * a.per;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < a.per; i += (long long)gridDim.x * 256)
        dst[i] = src[i];
}

dim3 segment_grid(const SegmentArgs& a) {
    long long x = (a.per + 255) / 256;
    if (x > 1024) x = 1024;
    return dim3((unsigned)x, (unsigned)a.count);
}

}  // namespace

hipError_t launch_gather(const SegmentArgs& a, hipStream_t s) {
    if (a.count <= 0 || a.per <= 0) return hipSuccess;
    if (a.count > 65535) return hipErrorInvalidValue;
    void* h = timer_begin("uneven_gather", s);
    hipLaunchKernelGGL(gather_kernel, segment_grid(a), dim3(256), 0, s, a);
    timer_end(h);
    return hipGetLastError();
}

hipError_t launch_scatter(const SegmentArgs& a, hipStream_t s) {
    if (a.count <= 0 || a.per <= 0) return hipSuccess;
    if (a.count > 65535) return hipErrorInvalidValue;
    void* h = timer_begin("uneven_scatter", s);
    hipLaunchKernelGGL(scatter_kernel, segment_grid(a), dim3(256), 0, s, a);
    timer_end(h);
    return hipGetLastError();
}

}  // namespace magnus
