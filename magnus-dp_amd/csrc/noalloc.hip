// noalloc.hip -- the SpaceCurve VJP on a fixed device scratch array.
//
// magnus_sc_vjp_noalloc_f64 runs the same kernels as magnus_sc_vjp_f64 with
// carry == NULL on the single-pass path, but takes their scratch (chunk
// summaries, chunk cotangents, chunk-start stream values, boundary halves)
// from `scratch`, a file-scope device array of kNoAllocBytes that every GPU
// the library runs on holds until the process ends, instead of the library's
// scratch cache: the call makes no allocation call, so callers that cannot
// allocate (code under stream capture, serving loops) can use it.  Host code
// never passes the array's name to a kernel: on each device, on its first
// call there, hipGetSymbolAddress gives the array's device address, which is
// kept per device and used for every launch.  The kernels write every scratch
// value before they read it, so nothing needs clearing between calls; calls
// that overlap on one device share its array and race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "host.hpp"
#include "magnus_capi.h"

constexpr size_t kNoAllocBytes = size_t(128) << 20;

static __device__ __align__(256) char scratch[kNoAllocBytes];

namespace {

constexpr int kMaxDevices = 64;

std::mutex g_mu;
char* g_ptr[kMaxDevices] = {};

// the array's device address on the current device, looked up once per device
int scratch_address(char** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return magnus_host::device_error(e, "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) return magnus_host::fail(MAGNUS_EDEVICE, "device index out of range");
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return magnus_host::device_error(e, "hipGetSymbolAddress(scratch)");
        g_ptr[dev] = static_cast<char*>(p);
    }
    *out = g_ptr[dev];
    return MAGNUS_OK;
}

}  // namespace

extern "C" {

int magnus_sc_vjp_noalloc_f64(const magnus_problem* p, uint64_t count, const double* in3, const double* cot3,
                              double* dA3, void* stream) {
    using namespace magnus_host;
    Plan pl;
    if (int rc = make_plan(p, true, false, &pl)) return rc;
    if (!use_wb(pl, false))
        return fail(MAGNUS_EINVAL, "no-allocation VJP needs the single-pass path (su(2) input, Boole/Simpson/Riemann, "
                                   "2 <= n <= 8)");
    if (count == 0) return MAGNUS_OK;
    if (!in3) return fail(MAGNUS_EINVAL, "input buffer must not be null");
    if (!dA3) return fail(MAGNUS_EINVAL, "output buffer must not be null");
    if (!cot3) return fail(MAGNUS_EINVAL, "cotangent buffer must not be null");
    const size_t need = wb_vjp_scratch_bytes(pl, count);
    if (need > kNoAllocBytes)
        return fail(MAGNUS_ERANGE, "batch needs " + std::to_string(need) + " scratch bytes; the no-allocation VJP holds " +
                                       std::to_string(kNoAllocBytes));
    char* base = nullptr;
    if (int rc = scratch_address(&base)) return rc;
    return run_wb_vjp_on(pl, count, in3, cot3, dA3, base, kNoAllocBytes, (hipStream_t)stream);
}

}  // extern "C"
