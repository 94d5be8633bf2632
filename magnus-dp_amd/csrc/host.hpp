// host.hpp -- host-side internals shared by the C ABI (capi.cpp) and the
// uneven-batch plans (uneven.cpp): problem validation, device scratch, and the
// level-parallel forward/VJP drivers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "magnus_capi.h"

namespace magnus_host {

struct GLTable {
    uint64_t max_order;
    std::vector<double> w, x;  // flattened by order, order o at o(o-1)/2
};

// A validated problem (composer/backends.hpp:104-180, backends.cpp:26-61).
struct Plan {
    uint64_t n, S, d, m, m_in, m_cot, Q, Qj, n_out;
    int op, mat, integ;
    bool sc, pad;
    double dt;
    std::shared_ptr<const GLTable> gl;
    int tile = 0;  // level-parallel product path (LmArgs::tile); 0 is the default choice
};

int fail(int code, const std::string& msg);
int device_error(hipError_t e, const char* where);
void set_last_kernel(const char* name);
int make_plan(const magnus_problem* p, bool spacecurve, bool forward, Plan* out);

// Stop the per-kernel timers from recording events on this thread (stream
// capture records kernels only); returns the previous setting.
bool timing_suspend(bool on);

// Bump allocator over caller-owned device memory.  Allocations are released in
// stack order when the DeviceScratch that made them goes out of scope.
struct Arena {
    char* base = nullptr;
    size_t size = 0;
    size_t used = 0;
};

constexpr size_t kScratchAlign = 256;
inline size_t scratch_round(size_t bytes) { return (bytes + kScratchAlign - 1) / kScratchAlign * kScratchAlign; }

// Keep memory freed by stream-ordered frees in the current device's default
// pool (release threshold "unlimited"), once per device: with the default
// threshold 0 every synchronization returns it to the driver and the next
// call maps gigabytes of scratch again.
void keep_pool_memory();

// Temporary device buffers: stream-ordered allocations freed on the same
// stream at scope exit, or, with an arena, slices of it (no allocation calls,
// so the work can be captured into a graph).
struct DeviceScratch {
    hipStream_t stream;
    Arena* arena;
    size_t mark;
    std::vector<void*> ptrs;
    explicit DeviceScratch(hipStream_t s, Arena* a = nullptr) : stream(s), arena(a), mark(a ? a->used : 0) {}
    ~DeviceScratch() {
        if (arena) {
            arena->used = mark;
            return;
        }
        for (void* p : ptrs) (void)hipFreeAsync(p, stream);
    }
    template <class T>
    T* alloc(size_t count) {
        if (count == 0) return nullptr;
        if (arena) {
            const size_t bytes = scratch_round(count * sizeof(T));
            if (arena->used + bytes > arena->size) return nullptr;
            void* p = arena->base + arena->used;
            arena->used += bytes;
            return static_cast<T*>(p);
        }
        keep_pool_memory();
        void* p = nullptr;
        if (hipMallocAsync(&p, count * sizeof(T), stream) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        ptrs.push_back(p);
        return static_cast<T*>(p);
    }
};

// Level-parallel drivers over `count` instances of one problem, laid out
// [count][S][m_in] (in), [count][n_out][m_out] (out), [count][S][m_out] (dA).
// forward: record into `carry` when non-null; outputs=false runs VJP::none.
int run_forward(const Plan& pl, uint64_t count, const double* in, double* out, double* carry, bool outputs,
                hipStream_t stream, Arena* arena = nullptr);
int run_vjp(const Plan& pl, uint64_t count, const double* in, const double* cot, const double* carry, double* dA,
            hipStream_t stream, Arena* arena = nullptr);

// Peak arena bytes of run_forward (outputs, no carry) and of run_vjp without a
// carry (it recomputes the carry), for `count` instances.
size_t forward_scratch_bytes(const Plan& pl, uint64_t count);
size_t vjp_scratch_bytes(const Plan& pl, uint64_t count);

}  // namespace magnus_host
