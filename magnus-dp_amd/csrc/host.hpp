// host.hpp -- host-side internals shared by the C ABI (capi.cpp) and the
// uneven-batch plans (uneven.cpp): problem validation, device scratch, and the
// level-parallel forward/VJP drivers.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "magnus_capi.h"

namespace magnus {
struct WbdArgs;
}

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

// Per-(device, stream) cache of device blocks for library scratch.  A call's
// buffers are bump-allocated from it and released in stack order when the
// call returns; the next call on the same stream reuses the memory, which
// stream order makes safe.  Blocks grow geometrically and are kept until
// magnus_release_scratch(): no allocation call (and no stall inside one) on
// the steady-state path.
struct ScratchCache {
    struct Block {
        char* p;
        size_t size;
    };
    std::recursive_mutex mu;
    std::vector<Block> blocks;
    size_t bi = 0, off = 0;  // cursor: block index, offset
    void* bump(size_t bytes);
};
// Bound on a cache's cursor position (bytes of the blocks ahead of the
// cursor's block plus its offset); 0: none.  An allocation that would pass it
// fails as if the device were out of memory.
extern std::atomic<uint64_t> g_scratch_limit;
ScratchCache* scratch_cache(hipStream_t s);
void release_scratch_caches();

// Temporary device buffers of one call: slices of the stream's scratch cache,
// or of a caller-owned arena (no allocation calls either way, so the work can
// be captured into a graph once the cache has grown to size).
struct DeviceScratch {
    hipStream_t stream;
    Arena* arena;
    ScratchCache* cache = nullptr;
    size_t mark = 0, mark_bi = 0;
    explicit DeviceScratch(hipStream_t s, Arena* a = nullptr) : stream(s), arena(a) {
        if (arena) {
            mark = arena->used;
        } else {
            cache = scratch_cache(s);
            if (cache) {
                cache->mu.lock();
                mark_bi = cache->bi;
                mark = cache->off;
            }
        }
    }
    ~DeviceScratch() {
        if (arena) {
            arena->used = mark;
        } else if (cache) {
            cache->bi = mark_bi;
            cache->off = mark;
            cache->mu.unlock();
        }
    }
    DeviceScratch(const DeviceScratch&) = delete;
    DeviceScratch& operator=(const DeviceScratch&) = delete;
    template <class T>
    T* alloc(size_t count) {
        if (count == 0) return nullptr;
        const size_t bytes = scratch_round(count * sizeof(T));
        if (arena) {
            if (arena->used + bytes > arena->size) return nullptr;
            void* p = arena->base + arena->used;
            arena->used += bytes;
            return static_cast<T*>(p);
        }
        return cache ? static_cast<T*>(cache->bump(bytes)) : nullptr;
    }
};

// Level-parallel drivers over `count` instances of one problem, laid out
// [count][S][m_in] (in), [count][n_out][m_out] (out), [count][S][m_out] (dA).
// forward: record into `carry` when non-null; outputs=false runs VJP::none.
int run_forward(const Plan& pl, uint64_t count, const double* in, double* out, double* carry, bool outputs,
                hipStream_t stream, Arena* arena = nullptr);
int run_vjp(const Plan& pl, uint64_t count, const double* in, const double* cot, const double* carry, double* dA,
            hipStream_t stream, Arena* arena = nullptr);

// Single-pass SpaceCurve path (su(2) input, block integrators, 2 <= n <= 8, no
// reference-layout carry): whether a problem takes it, the scratch bytes of its
// VJP without a checkpoint for `count` instances, and that VJP on caller-given
// scratch (no allocation call).
bool use_wb(const Plan& pl, bool has_carry);
size_t wb_vjp_scratch_bytes(const Plan& pl, uint64_t count);
int run_wb_vjp_on(const Plan& pl, uint64_t count, const double* in, const double* cot, double* dA, char* scratch,
                  size_t bytes, hipStream_t stream);

// The dense W-basis GL fold's arguments for `B` instances of one problem
// (B, n, d, m, Q, op, n_out and the GL nodes; buffers left to the caller).
void fill_wbd_fold(const Plan& pl, long long B, magnus::WbdArgs& f);

// Peak arena bytes of run_forward (outputs, no carry) and of run_vjp without a
// carry (it recomputes the carry), for `count` instances.
size_t forward_scratch_bytes(const Plan& pl, uint64_t count);
size_t vjp_scratch_bytes(const Plan& pl, uint64_t count);

}  // namespace magnus_host
