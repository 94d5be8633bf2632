// capi.cpp -- host side of the C ABI (include/magnus_capi.h).
//
// Mirrors the reference composer boundary: name resolution and "Auto" rules
// (composer/dispatch.hpp:107-185, linalg/backends.hpp:11-68,
// integration/backends.hpp:53-58, util/gausslegendre_backends.hpp:11-59),
// shape and order validation with the reference's message texts
// (composer/backends.hpp:104-180, composer/backends.cpp:26-61), and the
// process-global, replaceable Gauss-Legendre table (util/gausslegendre.cpp).
// Device work is enqueued on the caller's stream; temporary device buffers
// are slices of a per-(device, stream) scratch cache: grown with hipMalloc,
// reused by the next call on the same stream (whose kernels run after this
// call's), freed by magnus_release_scratch.
#include "magnus_capi.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "host.hpp"
#include "kernels.hpp"
#include "kernels_expm.hpp"
#include "kernels_sample.hpp"
#include "kernels_wb.hpp"

namespace magnus_host {

#include "gl_table_data.inc"

thread_local std::string g_err;
thread_local const char* g_kernel = "";

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

void set_last_kernel(const char* name) { g_kernel = name; }

std::atomic<uint64_t> g_scratch_limit{0};

void* ScratchCache::bump(size_t bytes) {
    const size_t limit = (size_t)g_scratch_limit.load(std::memory_order_relaxed);
    size_t ahead = 0;  // bytes of the blocks before the cursor's block
    for (size_t i = 0; i < bi && i < blocks.size(); ++i) ahead += blocks[i].size;
    for (; bi < blocks.size(); ahead += blocks[bi].size, ++bi, off = 0) {
        if (off + bytes <= blocks[bi].size) {
            if (limit && ahead + off + bytes > limit) return nullptr;
            void* p = blocks[bi].p + off;
            off += bytes;
            return p;
        }
    }
    if (limit && ahead + bytes > limit) return nullptr;
    size_t size = bytes < (size_t(256) << 20) ? (size_t(256) << 20) : bytes;
    if (!blocks.empty() && size < 2 * blocks.back().size) size = 2 * blocks.back().size;
    if (limit && ahead + size > limit) size = limit - ahead;
    void* p = nullptr;
    if (hipMalloc(&p, size) != hipSuccess) {
        (void)hipGetLastError();
        p = nullptr;
        if (size > bytes && hipMalloc(&p, bytes) == hipSuccess) {
            size = bytes;
        } else {
            (void)hipGetLastError();
            return nullptr;
        }
    }
    blocks.push_back({static_cast<char*>(p), size});
    bi = blocks.size() - 1;
    off = bytes;
    return p;
}

namespace {
std::mutex g_cache_mu;
std::vector<std::pair<std::pair<int, hipStream_t>, std::unique_ptr<ScratchCache>>> g_caches;
}  // namespace

ScratchCache* scratch_cache(hipStream_t s) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_cache_mu);
    for (auto& c : g_caches)
        if (c.first.first == dev && c.first.second == s) return c.second.get();
    g_caches.push_back({{dev, s}, std::make_unique<ScratchCache>()});
    return g_caches.back().second.get();
}

void release_scratch_caches() {
    // Snapshot the caches under g_cache_mu and release it before taking any
    // cache's own lock: a call in flight holds its cache's lock and may take
    // g_cache_mu (a nested DeviceScratch), so holding both here in the
    // opposite order could deadlock.  Caches are never destroyed, so the
    // pointers stay valid.
    std::vector<std::pair<int, ScratchCache*>> caches;
    {
        std::lock_guard<std::mutex> lock(g_cache_mu);
        for (auto& c : g_caches) caches.push_back({c.first.first, c.second.get()});
    }
    int cur = 0;
    (void)hipGetDevice(&cur);
    for (auto& c : caches) {
        std::lock_guard<std::recursive_mutex> l2(c.second->mu);
        (void)hipSetDevice(c.first);
        (void)hipDeviceSynchronize();
        for (auto& b : c.second->blocks) (void)hipFree(b.p);
        c.second->blocks.clear();
        c.second->bi = c.second->off = 0;
    }
    (void)hipSetDevice(cur);
    (void)hipGetLastError();
}

// ---------------------------------------------------------------------------
// registries
// ---------------------------------------------------------------------------
const char* const kMatrixNames[] = {"Fixed2", "Manual", "Blas", "SpaceCurve"};
const char* const kIntegratorNames[] = {"Boole", "Simpson", "Trapezoid", "Riemann"};
const char* const kGLNames[] = {"threadpool", "openmp", "serial"};
const char* const kOpNames[] = {"one", "many", "sum"};
const char* const kNumericNames[] = {"float", "double", "std::complex<float>", "std::complex<double>"};

constexpr const char* kResolveError = "Could not resolve dispatch name option";
constexpr const char* kInvalidOption = "invalid dispatch option (based on params)";
constexpr const char* kOpError =
    "could not deduce kernel op from str in Magnus::Dispatch::op_from_str(std::string_view)";
constexpr const char* kDivisibility =
    "sample interval count violates integrator divisibility (a power of 2 is recommneded; i.e., total "
    "samples is 2^k + 1 for some k > 4)";
constexpr const char* kOrderError =
    "requested magnus order exceeds GL table. Consider providing a larger table.";

int resolve(const char* name, const char* const* names, int count, bool auto_enabled, int32_t* out) {
    if (!name || !out) return fail(MAGNUS_EINVAL, "null argument");
    if (auto_enabled && std::strcmp(name, "Auto") == 0) {
        *out = MAGNUS_AUTO;
        return MAGNUS_OK;
    }
    for (int i = 0; i < count; ++i)
        if (std::strcmp(name, names[i]) == 0) {
            *out = i;
            return MAGNUS_OK;
        }
    return fail(MAGNUS_EINVAL, kResolveError);
}

// ---------------------------------------------------------------------------
// Gauss-Legendre table
// ---------------------------------------------------------------------------
std::mutex g_gl_mu;
std::shared_ptr<const GLTable> g_gl;

std::shared_ptr<const GLTable> gl_get() {
    std::lock_guard<std::mutex> lock(g_gl_mu);
    if (!g_gl) {
        auto t = std::make_shared<GLTable>();
        t->max_order = kGLDefaultMaxOrder;
        const size_t total = kGLDefaultMaxOrder * (kGLDefaultMaxOrder + 1) / 2;
        t->w.assign(kGLDefaultWeights, kGLDefaultWeights + total);
        t->x.assign(kGLDefaultNodes, kGLDefaultNodes + total);
        g_gl = t;
    }
    return g_gl;
}

size_t gl_offset(uint64_t order) { return (size_t)(order - 1) * order / 2; }

// ---------------------------------------------------------------------------
// validation
// ---------------------------------------------------------------------------
int divisibility(int integ) {  // boole.hpp:18, simpson.hpp:18, trapezoid.hpp:18, default.hpp:15
    switch (integ) {
        case MAGNUS_INT_BOOLE: return 4;
        case MAGNUS_INT_SIMPSON: return 2;
        case MAGNUS_INT_TRAPEZOID: return 2;
        default: return 1;
    }
}

// matrix_input_shape (composer/backends.hpp:108-122) + dispatch_plan
// (composer/backends.cpp:26-61) + GL backend dispatch (backends.cpp:65-93)
int make_plan(const magnus_problem* p, bool spacecurve, bool forward, Plan* out) {
    if (!p) return fail(MAGNUS_EINVAL, "problem must not be null");
    if (p->n == 0) return fail(MAGNUS_EINVAL, "n must be at least 1");
    if (p->samples < 2) return fail(MAGNUS_EINVAL, "data must contain at least two samples");
    const uint64_t dim = spacecurve ? 2 : p->dim;
    if (dim == 0) return fail(MAGNUS_EINVAL, "data must contain square matrices");
    if (dim > std::numeric_limits<uint64_t>::max() / dim) return fail(MAGNUS_EINVAL, "matrix dimension is too large");
    if (dim * dim > std::numeric_limits<uint64_t>::max() / p->samples)
        return fail(MAGNUS_EINVAL, "input shape is too large");
    if (p->op < 0 || p->op > 2) return fail(MAGNUS_EINVAL, "invalid kernel operation");

    auto gl = gl_get();
    if (p->n > 1) {
        const uint64_t order = (p->n + 3) / 2;
        if (order > gl->max_order) return fail(MAGNUS_EINVAL, kOrderError);
    }

    // matrix backend: Auto priority Fixed2 (d==2) -> Manual (d<64) -> Blas
    int mat = spacecurve ? MAGNUS_MAT_SPACECURVE : p->matrix_backend;
    if (mat == MAGNUS_AUTO) {
        mat = (dim == 2) ? MAGNUS_MAT_FIXED2 : (dim < 64 ? MAGNUS_MAT_MANUAL : MAGNUS_MAT_BLAS);
    } else if (mat < 0 || mat > 3) {
        return fail(MAGNUS_ERANGE, "dispatch index out of range");
    } else if ((mat == MAGNUS_MAT_FIXED2 || mat == MAGNUS_MAT_SPACECURVE) && dim != 2) {
        return fail(MAGNUS_EINVAL, kInvalidOption);
    }

    int integ = p->integrator;
    if (integ == MAGNUS_AUTO) integ = MAGNUS_INT_BOOLE;
    else if (integ < 0 || integ > 3) return fail(MAGNUS_ERANGE, "dispatch index out of range");

    if ((p->samples - 1) % (uint64_t)divisibility(integ) != 0) return fail(MAGNUS_EINVAL, kDivisibility);

    if (forward) {  // GL lane executor: validated only (the GL nodes are a grid dimension)
        const int glb = p->gl_backend;
        if (glb == MAGNUS_GL_THREADPOOL) return fail(MAGNUS_EINVAL, kInvalidOption);  // no XLA pool here
        if (glb != MAGNUS_AUTO && (glb < 0 || glb > 2)) return fail(MAGNUS_ERANGE, "dispatch index out of range");
    }

    const uint64_t m = dim * dim;
    if (m > 16384) return fail(MAGNUS_EINVAL, "matrix dimension exceeds the device kernel limit (dim <= 128)");

    out->n = p->n;
    out->S = p->samples;
    out->d = dim;
    out->m = m;
    out->sc = (mat == MAGNUS_MAT_SPACECURVE);
    out->pad = spacecurve;
    out->m_in = spacecurve ? 3 : m;
    out->m_cot = spacecurve ? 3 : m;
    out->Q = (p->n + 3) / 2;
    out->Qj = (p->n > 1) ? out->Q : 1;
    out->n_out = (p->op == MAGNUS_OP_MANY) ? p->n : 1;
    out->op = p->op;
    out->mat = mat;
    out->integ = integ;
    out->dt = (p->tf - p->t0) / (double)(p->samples - 1);  // magnus.hpp:32
    out->gl = gl;
    if (out->Q > (uint64_t)magnus::kArgMaxQ && p->n > 1)
        return fail(MAGNUS_EINVAL, kOrderError);
    return MAGNUS_OK;
}

void fill_gl(const Plan& pl, double* w, double* x) {
    for (int i = 0; i < magnus::kArgMaxQ; ++i) {
        w[i] = 0.0;
        x[i] = 0.0;
    }
    if (pl.n < 2) return;
    const size_t off = gl_offset(pl.Q);
    for (uint64_t i = 0; i < pl.Q; ++i) {
        w[i] = pl.gl->w[off + i];
        x[i] = pl.gl->x[off + i];
    }
}

// ---------------------------------------------------------------------------
// kernel timing
// ---------------------------------------------------------------------------
struct TimedLaunch {
    std::string name;
    hipEvent_t start, stop;
    double bytes, flops;
};
struct TimingTotals {
    uint64_t launches = 0;
    double ms = 0.0;
    double bytes = 0.0, flops = 0.0;
};
std::mutex g_timing_mu;
bool g_timing_on = false;
thread_local bool g_timing_suspended = false;
std::vector<TimedLaunch> g_pending;
std::vector<std::pair<std::string, TimingTotals>> g_totals;

struct KernelTimer {
    hipEvent_t start = nullptr, stop = nullptr;
    hipStream_t stream;
    const char* name;
    double bytes, flops;
    KernelTimer(const char* n, hipStream_t s, double b, double f) : stream(s), name(n), bytes(b), flops(f) {
        if (g_timing_suspended) return;
        std::lock_guard<std::mutex> lock(g_timing_mu);
        if (!g_timing_on) return;
        if (hipEventCreate(&start) != hipSuccess || hipEventCreate(&stop) != hipSuccess) {
            start = stop = nullptr;
            return;
        }
        (void)hipEventRecord(start, stream);
    }
    void done() {
        if (!start) return;
        (void)hipEventRecord(stop, stream);
        std::lock_guard<std::mutex> lock(g_timing_mu);
        g_pending.push_back({name, start, stop, bytes, flops});
        start = stop = nullptr;
    }
};

bool timing_suspend(bool on) {
    const bool prev = g_timing_suspended;
    g_timing_suspended = on;
    return prev;
}

}  // namespace magnus_host

using namespace magnus_host;

namespace magnus {
void* timer_begin(const char* name, hipStream_t s, double bytes, double flops) {
    auto* t = new KernelTimer(name, s, bytes, flops);
    if (!t->start) {
        delete t;
        return nullptr;
    }
    return t;
}
void timer_end(void* h) {
    if (!h) return;
    auto* t = static_cast<KernelTimer*>(h);
    t->done();
    delete t;
}
}  // namespace magnus

namespace magnus_host {

void timing_drain_locked() {
    for (auto& t : g_pending) {
        float ms = 0.f;
        if (hipEventSynchronize(t.stop) == hipSuccess && hipEventElapsedTime(&ms, t.start, t.stop) == hipSuccess) {
            auto it = g_totals.begin();
            for (; it != g_totals.end(); ++it)
                if (it->first == t.name) break;
            if (it == g_totals.end()) {
                g_totals.push_back({t.name, TimingTotals{}});
                it = g_totals.end() - 1;
            }
            it->second.launches += 1;
            it->second.ms += ms;
            it->second.bytes += t.bytes;
            it->second.flops += t.flops;
        }
        (void)hipEventDestroy(t.start);
        (void)hipEventDestroy(t.stop);
    }
    g_pending.clear();
}

int device_error(hipError_t e, const char* where) {
    return fail(MAGNUS_EDEVICE, std::string(where) + ": " + hipGetErrorString(e));
}

// Single-pass W-basis path: su(2) SpaceCurve input, block integrators,
// 2 <= n <= kWbMaxK + 1, no reference-layout carry requested or supplied.
bool use_wb(const Plan& pl, bool has_carry) {
    if (!pl.pad || has_carry) return false;
    if (pl.integ == MAGNUS_INT_TRAPEZOID) return false;  // reference reconstruction quirk: EP path
    return pl.n >= 2 && pl.n - 1 <= (uint64_t)magnus::kWbMaxK;
}

int wb_chunk_len(const Plan& pl, uint64_t count) {
    // aim for ~128k chunk-threads (two wavefronts per SIMD), chunks of 16..512
    // samples, multiple of 4: longer chunks shrink the per-instance combine
    const double target = 131072.0;
    double lc = (double)(pl.S - 1) * (double)count / target;
    int LC = (int)lc;
    if (LC < 16) LC = 16;
    if (LC > 512) LC = 512;
    LC = (LC + 3) / 4 * 4;
    return LC;
}

uint64_t wb_summary_elems(const Plan& pl, uint64_t count) {
    const uint64_t LC = (uint64_t)wb_chunk_len(pl, count);
    const uint64_t C = (pl.S - 1 + LC - 1) / LC;
    return count * C * (2 * (pl.n - 1) + 1) * 4;
}

// Chunk layout of the single-pass path for `count` instances.
void wb_layout(const Plan& pl, uint64_t count, magnus::WbArgs& a) {
    a.B = (long long)count;
    a.S = (long long)pl.S;
    a.LC = wb_chunk_len(pl, count);
    a.C = (int)((pl.S - 1 + a.LC - 1) / a.LC);
    a.K = (int)pl.n - 1;
}

// Scratch of the single-pass VJP without a checkpoint: chunk summaries, chunk
// cotangents, chunk-start stream values and boundary halves, each 256-byte aligned.
size_t wb_vjp_scratch_bytes(const Plan& pl, uint64_t count) {
    magnus::WbArgs a{};
    wb_layout(pl, count, a);
    const size_t BC = (size_t)count * a.C, K = (size_t)a.K;
    return 2 * scratch_round(BC * (2 * K + 1) * 4 * 8) + scratch_round(BC * K * 4 * 8) + scratch_round(BC * 6 * 8);
}

// Single-pass path on given buffers (fsum always; src, bases, edge for the VJP).
int run_wb_on(const Plan& pl, uint64_t count, const double* in, const double* cot, double* out, double* dA,
              bool summaries_given, double* fsum, double* src, double* bases, double* edge, hipStream_t stream) {
    // the chunk passes pack element offsets of the input into 48 bits
    if (count > 0 && pl.S > ((uint64_t)1 << 47) / 3 / count)
        return fail(MAGNUS_EINVAL, "SpaceCurve batch exceeds 2^47 input elements");
    const bool vjp = dA != nullptr;
    magnus::WbArgs a{};
    wb_layout(pl, count, a);
    a.n_out = (int)pl.n_out;
    a.op = pl.op;
    a.Q = (int)pl.Q;
    a.dt = pl.dt;
    fill_gl(pl, a.w, a.x);
    a.in = in;
    a.cot = cot;
    a.out = out;
    a.dA = dA;
    a.fsum = fsum;
    a.src = src;
    a.bases = bases;
    a.edge = edge;
    hipError_t e = magnus::launch_wb(pl.integ, a, vjp, vjp && summaries_given, stream);
    if (e != hipSuccess) return device_error(e, "wb kernel launch");
    g_kernel = "wb";
    return MAGNUS_OK;
}

int run_wb_vjp_on(const Plan& pl, uint64_t count, const double* in, const double* cot, double* dA, char* scratch,
                  size_t bytes, hipStream_t stream) {
    magnus::WbArgs a{};
    wb_layout(pl, count, a);
    const size_t BC = (size_t)count * a.C, K = (size_t)a.K;
    if (bytes < wb_vjp_scratch_bytes(pl, count)) return fail(MAGNUS_ENOMEM, "scratch too small for this batch");
    double* fsum = reinterpret_cast<double*>(scratch);
    scratch += scratch_round(BC * (2 * K + 1) * 4 * 8);
    double* src = reinterpret_cast<double*>(scratch);
    scratch += scratch_round(BC * (2 * K + 1) * 4 * 8);
    double* bases = reinterpret_cast<double*>(scratch);
    scratch += scratch_round(BC * K * 4 * 8);
    double* edge = reinterpret_cast<double*>(scratch);
    return run_wb_on(pl, count, in, cot, nullptr, dA, false, fsum, src, bases, edge, stream);
}

// Single-pass path.  `ckpt` (optional) holds the chunk summaries: the forward
// writes them there instead of scratch; a VJP given them skips the chunk pass.
int run_wb(const Plan& pl, uint64_t count, const double* in, const double* cot, double* out, double* dA,
           double* ckpt, hipStream_t stream) {
    DeviceScratch scratch(stream);
    const bool vjp = dA != nullptr;
    magnus::WbArgs a{};
    wb_layout(pl, count, a);
    const size_t BC = (size_t)count * a.C, K = (size_t)a.K;
    double* fsum = ckpt ? ckpt : scratch.alloc<double>(BC * (2 * K + 1) * 4);
    if (!fsum) return fail(MAGNUS_ENOMEM, "device allocation failed");
    double *src = nullptr, *bases = nullptr, *edge = nullptr;
    if (vjp) {
        src = scratch.alloc<double>(BC * (2 * K + 1) * 4);
        bases = scratch.alloc<double>(BC * K * 4);
        edge = scratch.alloc<double>(BC * 6);
        if (!src || !bases || !edge) return fail(MAGNUS_ENOMEM, "device allocation failed");
    }
    return run_wb_on(pl, count, in, cot, out, dA, ckpt != nullptr, fsum, src, bases, edge, stream);
}

void fill_lm(const Plan& pl, long long jobs, magnus::LmArgs& a) {
    a.S = (long long)pl.S;
    a.jobs = jobs;
    a.n = (int)pl.n;
    a.d = (int)pl.d;
    a.m = (int)pl.m;
    a.m_in = (int)pl.m_in;
    a.m_cot = (int)pl.m_cot;
    a.Q = (int)pl.Q;
    a.Qj = (int)pl.Qj;
    a.n_out = (int)pl.n_out;
    a.op = pl.op;
    a.sc = pl.sc ? (pl.pad ? 1 : 2) : 0;
    a.bt_fused = (pl.mat == MAGNUS_MAT_FIXED2 || pl.mat == MAGNUS_MAT_SPACECURVE) ? 1 : 0;
    a.dt = pl.dt;
    a.tile = pl.tile;
    fill_gl(pl, a.w, a.x);
}

// Dense W-basis path: matrix problems of order >= 2 with no reference-layout
// carry recorded or supplied (kernels_wbd.hip).  Q times fewer products than
// the per-node recursion and a carry of n-1 levels instead of Q (n-1).  op
// "one" stays on the per-node recursion: Omega_n alone can be many orders of
// magnitude below the terms the fold combines, and the reassociated sums then
// differ from the reference's by up to ~1e-12 of Omega_n (n = 7 and 8).
bool use_wbd(const Plan& pl) { return !pl.sc && pl.n >= 2 && pl.op != MAGNUS_OP_ONE; }

// Chunk length of the sample-parallel scan (0: sequential scan).  It depends
// on the sample count only, so an instance's results do not depend on what
// it is batched with: at most 256 chunks, at least 32 samples per chunk.
int scan_chunk_len(const Plan& pl, uint64_t) {
    if (pl.integ == MAGNUS_INT_TRAPEZOID) return 0;
    uint64_t clen = (pl.S - 1 + 255) / 256;
    clen = (clen + 3) / 4 * 4;
    if (clen < 32) clen = 32;
    if (pl.S - 1 < 2 * clen) return 0;
    return (int)clen;
}

// forward scan: end values and sum shares; reverse scan (Boole): chain
// increments, pendings and raw chunk starts
size_t scan_chunk_bytes(const Plan& pl, uint64_t B) {
    const int clen = scan_chunk_len(pl, B);
    if (!clen) return 0;
    const uint64_t C = (pl.S - 1 + clen - 1) / clen;
    return scratch_round(3 * C * B * pl.m * 8);
}

// The chunked scan is part of the arithmetic (its chunk bases are summed in
// chunk order, and the d = 8 fused levels need it), so a failed allocation is
// an error, not a silent switch to the sequential scan.
bool alloc_scan_chunks(DeviceScratch& scratch, const Plan& pl, uint64_t B, magnus::LmArgs& a) {
    a.clen = scan_chunk_len(pl, B);
    a.cscan = nullptr;
    if (!a.clen) return true;
    const uint64_t C = (pl.S - 1 + a.clen - 1) / a.clen;
    a.cscan = scratch.alloc<double>(3 * C * B * pl.m);
    return a.cscan != nullptr;
}

size_t wbd_fold_bytes(const Plan& pl, uint64_t B) {  // z, totals, omega1, part, part2, work, scan chunks
    const size_t m = pl.m, n = pl.n, Q = pl.Q;
    return 2 * scratch_round(B * n * m * 8) + scratch_round(B * m * 8) + 2 * scratch_round(B * Q * n * m * 8) +
           scratch_round(B * Q * (2 * (n - 1) + 1) * m * 8) + scan_chunk_bytes(pl, B);
}

size_t forward_scratch_bytes(const Plan& pl, uint64_t count) {
    const size_t S = pl.S, m = pl.m;
    if (use_wbd(pl))
        return wbd_fold_bytes(pl, count) + scratch_round(count * 2 * S * m * 8) + scratch_round(count * S * m * 8);
    const size_t jobs = count * pl.Qj;
    size_t b = scratch_round(jobs * pl.n_out * m * 8) + scratch_round(count * m * 8);
    if (pl.n >= 2) b += scratch_round(jobs * 2 * S * m * 8) + scratch_round(jobs * S * m * 8);
    return b;
}

size_t vjp_scratch_bytes(const Plan& pl, uint64_t count) {
    if (pl.n < 2) return 0;
    const size_t S = pl.S, m = pl.m;
    if (use_wbd(pl)) {
        const size_t W = scratch_round(count * (pl.n - 1) * S * m * 8);
        const size_t fwd = scratch_round(count * S * m * 8);  // Y
        const size_t rev = 2 * scratch_round(count * pl.n * m * 8) + scratch_round(count * 2 * S * m * 8) +
                           scratch_round(count * S * m * 8);  // zbar, wbar, gbuf, dAjob
        return W + wbd_fold_bytes(pl, count) + (fwd > rev ? fwd : rev);
    }
    const size_t jobs = count * pl.Q;
    const size_t carry = scratch_round(count * pl.Q * (pl.n - 1) * S * m * 8);
    const size_t fwd = scratch_round(jobs * S * m * 8);  // VJP::none forward: Y only
    const size_t rev = scratch_round(jobs * 2 * S * m * 8) + scratch_round(jobs * S * m * 8) +
                       scratch_round(jobs * 2 * m * 8);  // gbuf, dAjob, dbar (level walkers)
    return carry + (fwd > rev ? fwd : rev);
}

// LmArgs of the W-basis sweep: one job per instance, shift 0, weight 1, all
// level sums kept (op many layout)
void fill_wbd_lm(const Plan& pl, long long B, magnus::LmArgs& a) {
    fill_lm(pl, B, a);
    a.jobs = B;
    a.Q = 1;
    a.Qj = 1;
    a.op = MAGNUS_OP_MANY;
    a.n_out = (int)pl.n;
    a.m_cot = (int)pl.m;
    for (int i = 0; i < magnus::kArgMaxQ; ++i) a.w[i] = a.x[i] = 0.0;
    a.w[0] = 1.0;
    a.x[0] = 1.0;
}

void fill_wbd_fold(const Plan& pl, long long B, magnus::WbdArgs& f) {
    f.B = B;
    f.n = (int)pl.n;
    f.d = (int)pl.d;
    f.m = (int)pl.m;
    f.Q = (int)pl.Q;
    f.op = pl.op;
    f.n_out = (int)pl.n_out;
    fill_gl(pl, f.w, f.x);
}

// d = 8 Boole problems of order 2..8: the basis sweep is one single-pass
// kernel (kernels_wbd8.hip) instead of a loop of level launches.
bool use_wbd8(const Plan& pl) { return magnus::wbd8_sweep_ok((int)pl.d, pl.integ, (int)pl.n); }

// The basis sweep of the W-basis path: sums, totals and (carry set) the levels.
int wbd_sweep(const Plan& pl, magnus::LmArgs& a, hipStream_t stream) {
    if (use_wbd8(pl)) {
        hipError_t e = magnus::launch_wbd8_sweep(a, stream);
        if (e != hipSuccess) return device_error(e, "sweep kernel launch");
        return MAGNUS_OK;
    }
    for (int j = 1; j <= (int)pl.n; ++j) {
        a.level = j;
        hipError_t e = magnus::launch_lm_forward_level(pl.integ, a, stream);
        if (e != hipSuccess) return device_error(e, "forward kernel launch");
    }
    return MAGNUS_OK;
}

// The W-basis state of `count` instances, as one buffer: the levels W_1..W_K
// [B][K][S][m] (the single-pass d = 8 sweeps keep the even levels only,
// [B][K/2][S][m]: their reverse sweep rebuilds each odd level from its total
// and the kept level below it), then the sums
// [B][n][m] (rows 1..K: Z_1..Z_K), Z_0 [B][m] and the totals [B][n][m] (rows
// 0..K-1).  It is what the reverse sweep needs besides the input and the
// cotangent.
struct WbdState {
    double *W, *z, *z0, *totals;
};
uint64_t wbd_kept_levels(const Plan& pl) {
    const uint64_t K = pl.n - 1;
    return use_wbd8(pl) ? K / 2 : K;
}
uint64_t wbd_state_elems(const Plan& pl, uint64_t count) {
    const uint64_t n = pl.n, m = pl.m;
    return count * (wbd_kept_levels(pl) * pl.S * m + (2 * n + 1) * m);
}
WbdState wbd_state_view(const Plan& pl, uint64_t count, double* base) {
    const uint64_t B = count, n = pl.n, m = pl.m;
    WbdState st;
    st.W = base;
    st.z = st.W + B * wbd_kept_levels(pl) * pl.S * m;
    st.z0 = st.z + B * n * m;
    st.totals = st.z0 + B * m;
    return st;
}

// Samples formed inside the d = 8 single-pass sweeps from Fourier coefficients
// ([count][2 modes + 1][64]) and the grid's sin/cos table ([S][modes][2])
// instead of read from an input array.
struct FourierSrc {
    const double* coef;
    const double* tab;
    int modes;
};

// The basis sweep into a state: levels (when st.W is set), sums and totals.
int wbd_sweep_state(const Plan& pl, uint64_t count, const double* in, const WbdState& st, hipStream_t stream,
                    Arena* arena, const FourierSrc* fs = nullptr) {
    DeviceScratch scratch(stream, arena);
    const long long B = (long long)count;
    const size_t S = pl.S, m = pl.m;
    const bool levels = !use_wbd8(pl);  // the level loop needs its products (and, without W, two level arrays)
    double* Y = levels ? scratch.alloc<double>(B * S * m) : nullptr;
    double* pingpong = (levels && !st.W) ? scratch.alloc<double>(B * 2 * S * m) : nullptr;
    if (levels && (!Y || (!st.W && !pingpong))) return fail(MAGNUS_ENOMEM, "device allocation failed");
    magnus::LmArgs a{};
    fill_wbd_lm(pl, B, a);
    if (levels && !alloc_scan_chunks(scratch, pl, B, a)) return fail(MAGNUS_ENOMEM, "device allocation failed");
    a.in = in;
    if (fs) {
        a.fcoef = fs->coef;
        a.ftab = fs->tab;
        a.fmodes = fs->modes;
    }
    a.carry = st.W;
    a.pingpong = pingpong;
    a.Y = Y;
    a.partial = st.z;
    a.omega1 = st.z0;
    a.totals = st.totals;
    a.outputs = 1;
    return wbd_sweep(pl, a, stream);
}

// forward (op sum / many); `state` (optional) keeps the basis state for a later VJP
int run_wbd_forward(const Plan& pl, uint64_t count, const double* in, double* out, hipStream_t stream, Arena* arena,
                    double* state = nullptr, const FourierSrc* fs = nullptr) {
    DeviceScratch scratch(stream, arena);
    const long long B = (long long)count;
    const size_t m = pl.m, n = pl.n;
    const size_t Q = pl.Q;
    WbdState st{};
    if (state) {
        st = wbd_state_view(pl, count, state);
    } else {
        st.z = scratch.alloc<double>(B * n * m);
        st.z0 = scratch.alloc<double>(B * m);
        st.totals = scratch.alloc<double>(B * n * m);
    }
    double* part = scratch.alloc<double>(B * Q * n * m);
    double* work = scratch.alloc<double>(B * Q * (2 * (n - 1) + 1) * m);
    if (!st.z || !st.z0 || !st.totals || !part || !work) return fail(MAGNUS_ENOMEM, "device allocation failed");
    if (int rc = wbd_sweep_state(pl, count, in, st, stream, arena, fs)) return rc;
    magnus::WbdArgs f{};
    fill_wbd_fold(pl, B, f);
    f.z = st.z;
    f.z0 = st.z0;
    f.totals = st.totals;
    f.out = out;
    f.part = part;
    f.work = work;
    hipError_t e = magnus::launch_wbd_fold(f, stream);
    if (e != hipSuccess) return device_error(e, "fold kernel launch");
    g_kernel = "wbd";
    return MAGNUS_OK;
}

// VJP; `state` (optional) is the basis state a forward kept for the same input,
// otherwise the sweep is rerun into scratch (the reference's VJP::none)
int run_wbd_vjp(const Plan& pl, uint64_t count, const double* in, const double* cot, double* dA, hipStream_t stream,
                Arena* arena, const double* state = nullptr, const FourierSrc* fs = nullptr) {
    DeviceScratch scratch(stream, arena);
    const long long B = (long long)count;
    const size_t S = pl.S, m = pl.m, n = pl.n;
    const size_t Q = pl.Q;
    WbdState st{};
    if (state) {
        st = wbd_state_view(pl, count, const_cast<double*>(state));  // read only below
    } else {
        const uint64_t we = B * wbd_kept_levels(pl) * S * m;  // 0 for a d = 8 problem of order 2
        st.W = we ? scratch.alloc<double>(we) : nullptr;
        st.z = scratch.alloc<double>(B * n * m);
        st.z0 = scratch.alloc<double>(B * m);
        st.totals = scratch.alloc<double>(B * n * m);
        if ((we && !st.W) || !st.z || !st.z0 || !st.totals) return fail(MAGNUS_ENOMEM, "device allocation failed");
        if (int rc = wbd_sweep_state(pl, count, in, st, stream, arena, fs)) return rc;
    }
    double* part = scratch.alloc<double>(B * Q * n * m);
    double* part2 = scratch.alloc<double>(B * Q * n * m);
    double* work = scratch.alloc<double>(B * Q * (2 * (n - 1) + 1) * m);
    if (!part || !part2 || !work) return fail(MAGNUS_ENOMEM, "device allocation failed");
    const bool rev8 = use_wbd8(pl);  // the single-pass reverse sweep writes dA itself
    double* zbar = scratch.alloc<double>(B * n * m);
    double* wbar = scratch.alloc<double>(B * n * m);
    double* gbuf = rev8 ? nullptr : scratch.alloc<double>(B * 2 * S * m);
    double* dAjob = rev8 ? nullptr : scratch.alloc<double>(B * S * m);
    if (!zbar || !wbar || (!rev8 && (!gbuf || !dAjob))) return fail(MAGNUS_ENOMEM, "device allocation failed");
    magnus::WbdArgs f{};
    fill_wbd_fold(pl, B, f);
    f.z = st.z;
    f.z0 = st.z0;
    f.totals = st.totals;
    f.cot = cot;
    f.zbar = zbar;
    f.wbar = wbar;
    f.part = part;
    f.part2 = part2;
    f.work = work;
    hipError_t e = magnus::launch_wbd_fold_vjp(f, stream);
    if (e != hipSuccess) return device_error(e, "fold kernel launch");
    if (rev8) {
        magnus::Wbd8RevArgs r{};
        r.in = in;
        if (fs) {
            r.fcoef = fs->coef;
            r.ftab = fs->tab;
            r.fmodes = fs->modes;
        }
        r.W = st.W;
        r.totals = st.totals;
        r.zbar = zbar;
        r.wbar = wbar;
        r.dA = dA;
        r.B = B;
        r.S = (long long)S;
        r.n = (int)n;
        r.dt = pl.dt;
        e = magnus::launch_wbd8_rev(r, stream);
        if (e != hipSuccess) return device_error(e, "reverse sweep launch");
        g_kernel = "wbd";
        return MAGNUS_OK;
    }
    magnus::LmArgs a{};
    fill_wbd_lm(pl, B, a);
    if (!alloc_scan_chunks(scratch, pl, B, a)) return fail(MAGNUS_ENOMEM, "device allocation failed");
    a.in = in;
    a.carry = st.W;
    a.cot = zbar;
    a.gbuf = gbuf;
    a.dAjob = dAjob;
    a.wbar = wbar;
    a.ga = 0;
    e = magnus::launch_lm_vjp_init(pl.integ, a, stream);
    for (int k = (int)n; k >= 2 && e == hipSuccess; --k) {
        a.level = k;
        e = magnus::launch_lm_vjp_level(pl.integ, a, stream);
        a.ga ^= 1;
    }
    if (e == hipSuccess) e = magnus::launch_lm_vjp_final(a, stream);
    if (e != hipSuccess) return device_error(e, "vjp kernel launch");
    magnus::VjpReduceArgs r{};
    r.dAjob = dAjob;
    r.cot = zbar;
    r.dA = dA;
    r.B = B;
    r.S = (long long)S;
    r.n = (int)n;
    r.m = (int)m;
    r.m_out = (int)m;
    r.m_cot = (int)m;
    r.n_out = (int)n;
    r.Qj = 1;
    r.op = MAGNUS_OP_MANY;
    r.sc = 0;
    r.dt = pl.dt;
    e = magnus::launch_vjp_reduce(pl.integ, r, stream);
    if (e != hipSuccess) return device_error(e, "vjp reduce launch");
    g_kernel = "wbd";
    return MAGNUS_OK;
}

// Magnus sums of every order 1..n from one basis sweep: row k-1 of out is the
// op="sum" output of order k (its own Q(k) GL nodes).
// Magnus sums of orders first..n from one basis sweep: row k - first of out is
// the op="sum" output of order k (its own Q(k) GL nodes).
int run_orders(const Plan& pl, uint64_t count, uint64_t first, const double* in, double* out, hipStream_t stream,
               const FourierSrc* fs = nullptr) {
    DeviceScratch scratch(stream);
    const long long B = (long long)count;
    const size_t S = pl.S, m = pl.m, n = pl.n, Q = pl.Q;
    double* z = scratch.alloc<double>(B * n * m);
    double* z0 = scratch.alloc<double>(B * m);
    double* totals = scratch.alloc<double>(B * n * m);
    double* part = scratch.alloc<double>(B * Q * n * m);
    double* work = scratch.alloc<double>(B * Q * (2 * (n - 1) + 1) * m);
    const bool levels = n >= 2 && !use_wbd8(pl);
    double* pingpong = levels ? scratch.alloc<double>(B * 2 * S * m) : nullptr;
    double* Y = levels ? scratch.alloc<double>(B * S * m) : nullptr;
    if (!z || !z0 || !totals || !part || !work || (levels && (!pingpong || !Y)))
        return fail(MAGNUS_ENOMEM, "device allocation failed");
    magnus::LmArgs a{};
    fill_wbd_lm(pl, B, a);
    if (levels && !alloc_scan_chunks(scratch, pl, B, a)) return fail(MAGNUS_ENOMEM, "device allocation failed");
    a.in = in;
    if (fs) {
        a.fcoef = fs->coef;
        a.ftab = fs->tab;
        a.fmodes = fs->modes;
    }
    a.pingpong = pingpong;
    a.Y = Y;
    a.partial = z;
    a.omega1 = z0;
    a.totals = totals;
    a.outputs = 1;
    if (int rc = wbd_sweep(pl, a, stream)) return rc;
    const size_t rows_out = n - first + 1;
    hipError_t e = hipSuccess;
    if (first == 1) {
        e = hipMemcpy2DAsync(out, rows_out * m * sizeof(double), z0, m * sizeof(double), m * sizeof(double), B,
                             hipMemcpyDeviceToDevice, stream);
        if (e != hipSuccess) return device_error(e, "order-1 copy");
    }
    const size_t lo = first > 2 ? first : 2;
    // orders lo..n in one fold launch: every order's GL nodes, flattened
    size_t nflat = 0;
    for (size_t k = lo; k <= n; ++k) nflat += (k + 3) / 2;
    if (n >= lo && nflat <= (size_t)magnus::kArgMaxQ) {
        double* fpart = scratch.alloc<double>(B * nflat * n * m);
        double* fwork = scratch.alloc<double>(B * nflat * (2 * (n - 1) + 1) * m);
        if (!fpart || !fwork) return fail(MAGNUS_ENOMEM, "device allocation failed");
        magnus::WbdArgs f{};
        fill_wbd_fold(pl, B, f);
        f.op = MAGNUS_OP_SUM;
        f.rows = (int)n;
        f.z = z;
        f.z0 = z0;
        f.totals = totals;
        f.out = out;
        f.ostride = (long long)(rows_out * m);
        f.part = fpart;
        f.work = fwork;
        f.lo = (int)lo;
        f.kout = (int)first;
        f.nflat = (int)nflat;
        int fi = 0;
        for (size_t k = lo; k <= n; ++k) {
            const size_t Qk = (k + 3) / 2, off = gl_offset(Qk);
            for (size_t qq = 0; qq < Qk; ++qq, ++fi) {
                f.w[fi] = pl.gl->w[off + qq];
                f.x[fi] = pl.gl->x[off + qq];
                f.korder[fi] = (int)k;
            }
        }
        e = magnus::launch_wbd_fold_orders(f, stream);
        if (e != hipSuccess) return device_error(e, "fold kernel launch");
        g_kernel = "wbd";
        return MAGNUS_OK;
    }
    for (size_t k = lo; k <= n; ++k) {
        Plan pk = pl;
        pk.n = k;
        pk.Q = (k + 3) / 2;
        pk.Qj = pk.Q;
        pk.op = MAGNUS_OP_SUM;
        pk.n_out = 1;
        magnus::WbdArgs f{};
        fill_wbd_fold(pk, B, f);
        f.rows = (int)n;
        f.z = z;
        f.z0 = z0;
        f.totals = totals;
        f.out = out + (k - first) * m;
        f.ostride = (long long)(rows_out * m);
        f.part = part;
        f.work = work;
        e = magnus::launch_wbd_fold(f, stream);
        if (e != hipSuccess) return device_error(e, "fold kernel launch");
    }
    g_kernel = "wbd";
    return MAGNUS_OK;
}

// forward: record into `carry` when non-null; outputs=false runs VJP::none
int run_forward(const Plan& pl, uint64_t count, const double* in, double* out, double* carry, bool outputs,
                hipStream_t stream, Arena* arena) {
    if (!carry && outputs && use_wbd(pl)) return run_wbd_forward(pl, count, in, out, stream, arena);
    DeviceScratch scratch(stream, arena);
    const long long B = (long long)count;
    const long long jobs = B * (long long)pl.Qj;
    const uint64_t S = pl.S, m = pl.m;
    double* partial = outputs ? scratch.alloc<double>((size_t)jobs * pl.n_out * m) : nullptr;
    double* omega1 = outputs ? scratch.alloc<double>((size_t)B * m) : nullptr;
    double* pingpong = nullptr;
    double* Y = nullptr;
    magnus::LmArgs a{};
    fill_lm(pl, jobs, a);
    const bool walk_y = magnus::walk_ok(pl.integ, a);  // the walkers form their products in registers
    if (!carry && pl.n >= 2) pingpong = scratch.alloc<double>((size_t)jobs * 2 * S * m);
    if (pl.n >= 2 && !walk_y) Y = scratch.alloc<double>((size_t)jobs * S * m);
    if ((outputs && (!partial || !omega1)) || (!carry && pl.n >= 2 && !pingpong) || (pl.n >= 2 && !walk_y && !Y))
        return fail(MAGNUS_ENOMEM, "device allocation failed");

    a.in = in;
    a.carry = carry;
    a.pingpong = pingpong;
    a.Y = Y;
    a.partial = partial;
    a.omega1 = omega1;
    a.outputs = outputs ? 1 : 0;
    const bool walk = magnus::walk_ok(pl.integ, a);  // fused level walkers (d = 16, 32, Boole)
    // without outputs the last level (sums only) has nothing to write
    const int top = outputs ? (int)pl.n : (int)pl.n - 1;
    for (int j = 1; j <= top; ++j) {
        a.level = j;
        hipError_t e = walk ? magnus::launch_walk_forward_level(a, stream)
                            : magnus::launch_lm_forward_level(pl.integ, a, stream);
        if (e != hipSuccess) return device_error(e, "forward kernel launch");
    }
    if (outputs) {
        magnus::FwdReduceArgs r{};
        r.partial = partial;
        r.omega1 = omega1;
        r.out = out;
        r.B = B;
        r.n = (int)pl.n;
        r.m = (int)m;
        r.m_out = pl.pad ? 3 : (int)m;
        r.n_out = (int)pl.n_out;
        r.Qj = (int)pl.Qj;
        r.op = pl.op;
        r.sc = pl.pad ? 1 : 0;
        hipError_t e = magnus::launch_fwd_reduce(r, stream);
        if (e != hipSuccess) return device_error(e, "reduce kernel launch");
    }
    g_kernel = "lm";
    return MAGNUS_OK;
}

// VJP from the carry on the level walkers: levels n..2, then the reduce that
// adds each job's last adjoint (with its deferred total cotangent) and the nodes
int run_walk_vjp(const Plan& pl, long long B, magnus::LmArgs& a, const double* cot, double* dA,
                 DeviceScratch& scratch, hipStream_t stream) {
    a.dbar = scratch.alloc<double>((size_t)a.jobs * 2 * pl.m);
    if (!a.dbar) return fail(MAGNUS_ENOMEM, "device allocation failed");
    hipError_t e = hipSuccess;
    for (int k = (int)pl.n; k >= 2 && e == hipSuccess; --k) {
        a.level = k;
        e = magnus::launch_walk_vjp_level(a, stream);
        a.ga ^= 1;
    }
    if (e != hipSuccess) return device_error(e, "vjp kernel launch");
    magnus::VjpReduceArgs r{};
    r.dAjob = a.dAjob;
    r.cot = cot;
    r.dA = dA;
    r.B = B;
    r.S = (long long)pl.S;
    r.n = (int)pl.n;
    r.m = (int)pl.m;
    r.m_out = (int)pl.m;
    r.m_cot = (int)pl.m_cot;
    r.n_out = (int)pl.n_out;
    r.Qj = (int)pl.Q;
    r.op = pl.op;
    r.sc = 0;
    r.dt = pl.dt;
    e = magnus::launch_walk_vjp_reduce(a, r, stream);
    if (e != hipSuccess) return device_error(e, "vjp reduce launch");
    g_kernel = "lm";  // the per-node level recursion (here on its fused walkers)
    return MAGNUS_OK;
}

bool use_walkseg(const Plan& pl);
int run_wbd_vjp_seg(const Plan& pl, uint64_t count, const double* in, const double* cot, double* dA,
                    uint64_t segments, hipStream_t stream);

int run_vjp(const Plan& pl, uint64_t count, const double* in, const double* cot, const double* carry,
            double* dA, hipStream_t stream, Arena* arena) {
    if (!carry && use_wbd(pl)) {
        const int rc = run_wbd_vjp(pl, count, in, cot, dA, stream, arena);
        // Where the n - 1 full basis levels do not fit, keep them for one time
        // segment at a time (same dA up to rounding).
        if (rc == MAGNUS_ENOMEM && !arena && use_walkseg(pl)) return run_wbd_vjp_seg(pl, count, in, cot, dA, 0, stream);
        return rc;
    }
    DeviceScratch scratch(stream, arena);
    const long long B = (long long)count;
    const uint64_t S = pl.S, m = pl.m;
    double* dAjob = nullptr;
    if (pl.n >= 2) {
        const long long jobs = B * (long long)pl.Q;
        if (!carry) {  // VJP::none recompute (grad.hpp:95-110)
            double* tmp = scratch.alloc<double>((size_t)B * pl.Q * (pl.n - 1) * S * m);
            if (!tmp) return fail(MAGNUS_ENOMEM, "device allocation failed");
            int rc = run_forward(pl, count, in, nullptr, tmp, false, stream, arena);
            if (rc) return rc;
            carry = tmp;
        }
        double* gbuf = scratch.alloc<double>((size_t)jobs * 2 * S * m);
        dAjob = scratch.alloc<double>((size_t)jobs * S * m);
        if (!gbuf || !dAjob) return fail(MAGNUS_ENOMEM, "device allocation failed");
        magnus::LmArgs a{};
        fill_lm(pl, jobs, a);
        a.in = in;
        a.cot = cot;
        a.carry = const_cast<double*>(carry);  // read only in the VJP kernels
        a.gbuf = gbuf;
        a.dAjob = dAjob;
        a.ga = 0;
        if (magnus::walk_ok(pl.integ, a)) return run_walk_vjp(pl, B, a, cot, dA, scratch, stream);
        hipError_t e = magnus::launch_lm_vjp_init(pl.integ, a, stream);
        for (int k = (int)pl.n; k >= 2 && e == hipSuccess; --k) {
            a.level = k;
            e = magnus::launch_lm_vjp_level(pl.integ, a, stream);
            a.ga ^= 1;
        }
        if (e == hipSuccess) e = magnus::launch_lm_vjp_final(a, stream);
        if (e != hipSuccess) return device_error(e, "vjp kernel launch");
    }
    magnus::VjpReduceArgs r{};
    r.dAjob = dAjob;
    r.cot = cot;
    r.dA = dA;
    r.B = B;
    r.S = (long long)S;
    r.n = (int)pl.n;
    r.m = (int)m;
    r.m_out = pl.pad ? 3 : (int)m;
    r.m_cot = (int)pl.m_cot;
    r.n_out = (int)pl.n_out;
    r.Qj = pl.n >= 2 ? (int)pl.Q : 0;
    r.op = pl.op;
    r.sc = pl.pad ? 1 : 0;
    r.dt = pl.dt;
    hipError_t e = magnus::launch_vjp_reduce(pl.integ, r, stream);
    if (e != hipSuccess) return device_error(e, "vjp reduce launch");
    if (pl.n >= 2) g_kernel = "lm";
    return MAGNUS_OK;
}

// ---------------------------------------------------------------------------
// Dense W-basis VJP with time-segment checkpoints (kernels_walkseg.hip): the
// levels are kept for one segment of L samples at a time instead of all S.
// ---------------------------------------------------------------------------
bool use_walkseg(const Plan& pl) { return !pl.sc && use_wbd(pl) && magnus::walkseg_ok((int)pl.d, pl.integ, (int)pl.n); }

// Segment length (a multiple of 4 samples).  segments = 0: the length that
// balances the per-instance segment buffers ((K + 3)(L + 1) level rows) against
// the checkpoints (2 K rows per segment), L ~ sqrt(2 K (S - 1) / (K + 3)).
uint64_t seg_length(const Plan& pl, uint64_t segments) {
    const uint64_t span = pl.S - 1, K = pl.n - 1;
    uint64_t L;
    if (segments) {
        L = (span + segments - 1) / segments;
    } else {
        L = (uint64_t)std::ceil(std::sqrt(2.0 * (double)K * (double)span / (double)(K + 3)));
    }
    L = (L + 3) / 4 * 4;
    if (L < 4) L = 4;
    if (L > span) L = span;
    return L;
}

size_t seg_scratch_bytes(const Plan& pl, uint64_t B, uint64_t L) {
    const size_t m = pl.m, n = pl.n, K = n - 1, Q = pl.Q, G = (pl.S - 1 + L - 1) / L;
    return scratch_round(B * K * (L + 1) * m * 8) + scratch_round(B * G * K * 2 * m * 8) +
           2 * scratch_round(B * n * m * 8) + scratch_round(B * m * 8) + scratch_round(B * 2 * (L + 1) * m * 8) +
           scratch_round(B * (L + 1) * m * 8) + scratch_round(B * K * 2 * m * 8) + 2 * scratch_round(B * n * m * 8) +
           2 * scratch_round(B * Q * n * m * 8) + scratch_round(B * Q * (2 * K + 1) * m * 8);
}

int run_wbd_vjp_seg(const Plan& pl, uint64_t count, const double* in, const double* cot, double* dA,
                    uint64_t segments, hipStream_t stream) {
    DeviceScratch scratch(stream, nullptr);
    const long long B = (long long)count;
    const uint64_t S = pl.S, m = pl.m, n = pl.n, K = n - 1, Q = pl.Q;
    const uint64_t L = seg_length(pl, segments), G = (S - 1 + L - 1) / L;
    magnus::WalkSegArgs a{};
    a.in = in;
    a.wseg = scratch.alloc<double>(B * K * (L + 1) * m);
    a.ckp = scratch.alloc<double>(B * G * K * 2 * m);
    a.sums = scratch.alloc<double>(B * n * m);
    a.z0 = scratch.alloc<double>(B * m);
    a.totals = scratch.alloc<double>(B * n * m);
    a.useg = scratch.alloc<double>(B * 2 * (L + 1) * m);
    a.dAseg = scratch.alloc<double>(B * (L + 1) * m);
    a.rstate = scratch.alloc<double>(B * K * 2 * m);
    double* zbar = scratch.alloc<double>(B * n * m);
    double* wbar = scratch.alloc<double>(B * n * m);
    double* part = scratch.alloc<double>(B * Q * n * m);
    double* part2 = scratch.alloc<double>(B * Q * n * m);
    double* work = scratch.alloc<double>(B * Q * (2 * K + 1) * m);
    if (!a.wseg || !a.ckp || !a.sums || !a.z0 || !a.totals || !a.useg || !a.dAseg || !a.rstate || !zbar || !wbar ||
        !part || !part2 || !work)
        return fail(MAGNUS_ENOMEM, "device allocation failed");
    a.zbar = zbar;
    a.wbar = wbar;
    a.dA = dA;
    a.B = B;
    a.S = (long long)S;
    a.L = (int)L;
    a.G = (int)G;
    a.n = (int)n;
    a.d = (int)pl.d;
    a.m = (int)m;
    a.dt = pl.dt;
    auto seg = [&](uint64_t g) {
        a.g = (int)g;
        a.s0 = (long long)(g * L);
        a.s1 = (long long)std::min<uint64_t>((g + 1) * L, S - 1);
    };
    hipError_t e = hipSuccess;
    // first pass: the segments in order, every level; checkpoints, sums, totals
    a.mode = 0;
    for (uint64_t g = 0; g < G && e == hipSuccess; ++g) {
        seg(g);
        for (uint64_t j = 1; j <= n && e == hipSuccess; ++j) {
            a.level = (int)j;
            e = magnus::launch_walkseg_forward(a, stream);
        }
    }
    if (e != hipSuccess) return device_error(e, "segment walker launch");
    magnus::WbdArgs f{};
    fill_wbd_fold(pl, B, f);
    f.z = a.sums;
    f.z0 = a.z0;
    f.totals = a.totals;
    f.cot = cot;
    f.zbar = zbar;
    f.wbar = wbar;
    f.part = part;
    f.part2 = part2;
    f.work = work;
    e = magnus::launch_wbd_fold_vjp(f, stream);
    if (e != hipSuccess) return device_error(e, "fold kernel launch");
    // reverse: the segments from the last down; rebuild the levels, then the reverse levels
    for (uint64_t gg = G; gg-- > 0 && e == hipSuccess;) {
        seg(gg);
        a.mode = 1;
        for (uint64_t j = 1; j <= K && e == hipSuccess; ++j) {
            a.level = (int)j;
            e = magnus::launch_walkseg_forward(a, stream);
        }
        a.ga = 0;
        for (uint64_t k = n; k >= 2 && e == hipSuccess; --k) {
            a.level = (int)k;
            e = magnus::launch_walkseg_reverse(a, stream);
            a.ga ^= 1;
        }
        if (e == hipSuccess) e = magnus::launch_walkseg_reduce(a, stream);
    }
    if (e != hipSuccess) return device_error(e, "segment walker launch");
    g_kernel = "walkseg";
    return MAGNUS_OK;
}

uint64_t carry_elems(const magnus_problem* p, uint64_t dim) {
    if (!p || p->n < 2) return 0;
    return ((p->n + 3) / 2) * (p->n - 1) * p->samples * dim * dim;
}

}  // namespace magnus_host

extern "C" {

int magnus_resolve_op(const char* name, int32_t* out) {
    if (!name || !out) return fail(MAGNUS_EINVAL, "null argument");
    static const char* const upper[] = {"ONE", "MANY", "SUM"};
    for (int i = 0; i < 3; ++i)
        if (std::strcmp(name, kOpNames[i]) == 0 || std::strcmp(name, upper[i]) == 0) {
            *out = i;
            return MAGNUS_OK;
        }
    return fail(MAGNUS_EINVAL, kOpError);
}
int magnus_resolve_matrix_backend(const char* name, int32_t* out) { return resolve(name, kMatrixNames, 4, true, out); }
int magnus_resolve_integrator(const char* name, int32_t* out) { return resolve(name, kIntegratorNames, 4, true, out); }
int magnus_resolve_gl_backend(const char* name, int32_t* out) { return resolve(name, kGLNames, 3, true, out); }

const char* magnus_matrix_backend_name(int32_t i) { return (i >= 0 && i < 4) ? kMatrixNames[i] : nullptr; }
const char* magnus_integrator_name(int32_t i) { return (i >= 0 && i < 4) ? kIntegratorNames[i] : nullptr; }
const char* magnus_gl_backend_name(int32_t i) { return (i >= 0 && i < 3) ? kGLNames[i] : nullptr; }
const char* magnus_op_name(int32_t i) { return (i >= 0 && i < 3) ? kOpNames[i] : nullptr; }
const char* magnus_numeric_backend_name(int32_t i) { return (i >= 0 && i < 4) ? kNumericNames[i] : nullptr; }

uint64_t magnus_max_order(void) { return gl_get()->max_order; }

int magnus_replace_gl_table(uint64_t max_order, const double* weights, const double* nodes) {
    // OwningTable constructor checks (util/gausslegendre.cpp:74-99)
    if (max_order == 0) return fail(MAGNUS_EINVAL, "OwningTable requires max_order >= 1");
    if (!weights || !nodes) return fail(MAGNUS_EINVAL, "GL data must not be null");
    auto t = std::make_shared<GLTable>();
    const size_t total = (size_t)max_order * (max_order + 1) / 2;
    t->max_order = max_order;
    t->w.assign(weights, weights + total);
    t->x.assign(nodes, nodes + total);
    std::lock_guard<std::mutex> lock(g_gl_mu);
    g_gl = t;
    return MAGNUS_OK;
}

int magnus_gl_order(uint64_t order, double* weights, double* nodes) {
    auto gl = gl_get();
    if (order == 0 || order > gl->max_order)
        return fail(MAGNUS_ERANGE, "Gauss-Legendre order is outside the available table");
    const size_t off = gl_offset(order);
    for (uint64_t i = 0; i < order; ++i) {
        if (weights) weights[i] = gl->w[off + i];
        if (nodes) nodes[i] = gl->x[off + i];
    }
    return MAGNUS_OK;
}

int magnus_validate(const magnus_problem* p, int spacecurve, int forward) {
    Plan pl;
    return make_plan(p, spacecurve != 0, forward != 0, &pl);
}

uint64_t magnus_carry_elems(const magnus_problem* p) { return p ? carry_elems(p, p->dim) : 0; }
uint64_t magnus_sc_carry_elems(const magnus_problem* p) { return carry_elems(p, 2); }
uint64_t magnus_output_elems(const magnus_problem* p) {
    if (!p) return 0;
    return (p->op == MAGNUS_OP_MANY ? p->n : 1) * p->dim * p->dim;
}
uint64_t magnus_sc_output_elems(const magnus_problem* p) {
    if (!p) return 0;
    return (p->op == MAGNUS_OP_MANY ? p->n : 1) * 3;
}

int magnus_forward_f64(const magnus_problem* p, uint64_t count, const double* in, double* out, double* carry,
                       void* stream) {
    Plan pl;
    if (int rc = make_plan(p, false, true, &pl)) return rc;
    if (count == 0) return MAGNUS_OK;
    if (!in) return fail(MAGNUS_EINVAL, "input buffer must not be null");
    if (!out) return fail(MAGNUS_EINVAL, "output buffer must not be null");
    return run_forward(pl, count, in, out, pl.n >= 2 ? carry : nullptr, true, (hipStream_t)stream);
}

int magnus_vjp_f64(const magnus_problem* p, uint64_t count, const double* in, const double* cot,
                   const double* carry, double* dA, void* stream) {
    Plan pl;
    if (int rc = make_plan(p, false, false, &pl)) return rc;
    if (count == 0) return MAGNUS_OK;
    if (!in) return fail(MAGNUS_EINVAL, "input buffer must not be null");
    if (!dA) return fail(MAGNUS_EINVAL, "output buffer must not be null");
    if (!cot) return fail(MAGNUS_EINVAL, "cotangent buffer must not be null");
    return run_vjp(pl, count, in, cot, carry, dA, (hipStream_t)stream);
}

int magnus_sc_forward_f64(const magnus_problem* p, uint64_t count, const double* in3, double* out3,
                          double* carry, void* stream) {
    Plan pl;
    if (int rc = make_plan(p, true, true, &pl)) return rc;
    if (count == 0) return MAGNUS_OK;
    if (!in3) return fail(MAGNUS_EINVAL, "input buffer must not be null");
    if (!out3) return fail(MAGNUS_EINVAL, "output buffer must not be null");
    double* c = pl.n >= 2 ? carry : nullptr;
    if (use_wb(pl, c != nullptr)) return run_wb(pl, count, in3, nullptr, out3, nullptr, nullptr, (hipStream_t)stream);
    return run_forward(pl, count, in3, out3, c, true, (hipStream_t)stream);
}

int magnus_sc_vjp_f64(const magnus_problem* p, uint64_t count, const double* in3, const double* cot3,
                      const double* carry, double* dA3, void* stream) {
    Plan pl;
    if (int rc = make_plan(p, true, false, &pl)) return rc;
    if (count == 0) return MAGNUS_OK;
    if (!in3) return fail(MAGNUS_EINVAL, "input buffer must not be null");
    if (!dA3) return fail(MAGNUS_EINVAL, "output buffer must not be null");
    if (!cot3) return fail(MAGNUS_EINVAL, "cotangent buffer must not be null");
    const double* c = pl.n >= 2 ? carry : nullptr;
    if (use_wb(pl, c != nullptr)) return run_wb(pl, count, in3, cot3, nullptr, dA3, nullptr, (hipStream_t)stream);
    return run_vjp(pl, count, in3, cot3, c, dA3, (hipStream_t)stream);
}

int magnus_forward_orders_f64(const magnus_problem* p, uint64_t count, const double* in, double* out, void* stream) {
    Plan pl;
    if (int rc = make_plan(p, false, true, &pl)) return rc;
    if (pl.sc) return fail(MAGNUS_EINVAL, "all-orders forward takes matrix problems (not the SpaceCurve backend)");
    if (count == 0) return MAGNUS_OK;
    if (!in || !out) return fail(MAGNUS_EINVAL, "null argument");
    return run_orders(pl, count, 1, in, out, (hipStream_t)stream);
}

int magnus_forward_orders_from_f64(const magnus_problem* p, uint64_t count, uint64_t first, const double* in,
                                   double* out, void* stream) {
    Plan pl;
    if (int rc = make_plan(p, false, true, &pl)) return rc;
    if (pl.sc) return fail(MAGNUS_EINVAL, "all-orders forward takes matrix problems (not the SpaceCurve backend)");
    if (first < 1 || first > pl.n) return fail(MAGNUS_EINVAL, "first order must lie in 1..n");
    if (count == 0) return MAGNUS_OK;
    if (!in || !out) return fail(MAGNUS_EINVAL, "null argument");
    return run_orders(pl, count, first, in, out, (hipStream_t)stream);
}

uint64_t magnus_basis_state_elems(const magnus_problem* p, uint64_t count) {
    Plan pl;
    if (make_plan(p, false, true, &pl)) return 0;
    if (pl.sc || !use_wbd(pl)) return 0;
    return wbd_state_elems(pl, count);
}

int magnus_forward_basis_f64(const magnus_problem* p, uint64_t count, const double* in, double* out, double* state,
                             void* stream) {
    Plan pl;
    if (int rc = make_plan(p, false, true, &pl)) return rc;
    if (pl.sc || !use_wbd(pl))
        return fail(MAGNUS_EINVAL, "basis-state forward needs the dense W-basis path (matrix problem, op many or sum, "
                                   "n >= 2)");
    if (count == 0) return MAGNUS_OK;
    if (!in || !out || !state) return fail(MAGNUS_EINVAL, "null argument");
    return run_wbd_forward(pl, count, in, out, (hipStream_t)stream, nullptr, state);
}

uint64_t magnus_vjp_segmented_scratch_bytes(const magnus_problem* p, uint64_t count, uint64_t segments) {
    Plan pl;
    if (make_plan(p, false, false, &pl) || !use_walkseg(pl)) return 0;
    return seg_scratch_bytes(pl, count, seg_length(pl, segments));
}

uint64_t magnus_vjp_segment_length(const magnus_problem* p, uint64_t segments) {
    Plan pl;
    if (make_plan(p, false, false, &pl) || !use_walkseg(pl)) return 0;
    return seg_length(pl, segments);
}

int magnus_vjp_segmented_f64(const magnus_problem* p, uint64_t count, const double* in, const double* cot, double* dA,
                             uint64_t segments, void* stream) {
    Plan pl;
    if (int rc = make_plan(p, false, false, &pl)) return rc;
    if (!use_walkseg(pl))
        return fail(MAGNUS_EINVAL, "segmented VJP needs a dense W-basis problem (op many or sum, n >= 2) with dim 16, "
                                   "32 or 64 under the Boole rule");
    if (count == 0) return MAGNUS_OK;
    if (!in || !cot || !dA) return fail(MAGNUS_EINVAL, "null argument");
    return run_wbd_vjp_seg(pl, count, in, cot, dA, segments, (hipStream_t)stream);
}

int magnus_vjp_basis_f64(const magnus_problem* p, uint64_t count, const double* in, const double* cot,
                         const double* state, double* dA, void* stream) {
    Plan pl;
    if (int rc = make_plan(p, false, false, &pl)) return rc;
    if (pl.sc || !use_wbd(pl))
        return fail(MAGNUS_EINVAL, "basis-state VJP needs the dense W-basis path (matrix problem, op many or sum, "
                                   "n >= 2)");
    if (count == 0) return MAGNUS_OK;
    if (!in || !cot || !state || !dA) return fail(MAGNUS_EINVAL, "null argument");
    return run_wbd_vjp(pl, count, in, cot, dA, (hipStream_t)stream, nullptr, state);
}

// ---- Fourier-form input: the d = 8 single-pass sweeps form the samples
// themselves.  The entry checks the problem, builds the grid's sin/cos table
// in scratch and runs the sample-array entry's work with the samples formed
// in the sweeps' loader wavefront (the same products and sums as
// magnus_sample_fourier_f64, so the results equal sampling first, bit for bit).
namespace {
int fourier_check(const Plan& pl, uint64_t modes, const char* what) {
    if (pl.sc || !use_wbd8(pl))
        return fail(MAGNUS_EINVAL, std::string(what) + " with Fourier-form input needs a dim-8 matrix problem under the "
                                                       "Boole rule with 2 <= n <= 8");
    if (modes > (uint64_t)magnus::kWbd8MaxModes) return fail(MAGNUS_ERANGE, "too many Fourier modes (max 8)");
    return MAGNUS_OK;
}
double* fourier_table(DeviceScratch& scratch, const magnus_problem* p, uint64_t modes, hipStream_t stream, int* rc) {
    *rc = MAGNUS_OK;
    if (modes == 0) return nullptr;
    double* tab = scratch.alloc<double>(p->samples * modes * 2);
    if (!tab) {
        *rc = fail(MAGNUS_ENOMEM, "device allocation failed");
        return nullptr;
    }
    magnus::SampleArgs a{};
    a.S = (long long)p->samples;
    a.modes = (int)modes;
    a.t0 = p->t0;
    a.tf = p->tf;
    a.step = (p->tf - p->t0) / (double)(p->samples - 1);
    hipError_t e = magnus::launch_fourier_table(a, tab, stream);
    if (e != hipSuccess) *rc = device_error(e, "table kernel launch");
    return tab;
}
}  // namespace

int magnus_fourier_forward_basis_f64(const magnus_problem* p, uint64_t count, const double* coeffs, uint64_t modes,
                                     double* out, double* state, void* stream) {
    Plan pl;
    if (int rc = make_plan(p, false, true, &pl)) return rc;
    if (int rc = fourier_check(pl, modes, "basis-state forward")) return rc;
    if (pl.op == MAGNUS_OP_ONE) return fail(MAGNUS_EINVAL, "basis-state forward needs op many or sum");
    if (count == 0) return MAGNUS_OK;
    if (!coeffs || !out || !state) return fail(MAGNUS_EINVAL, "null argument");
    DeviceScratch scratch((hipStream_t)stream);
    int rc;
    double* tab = fourier_table(scratch, p, modes, (hipStream_t)stream, &rc);
    if (rc) return rc;
    const FourierSrc fs{coeffs, tab, (int)modes};
    return run_wbd_forward(pl, count, nullptr, out, (hipStream_t)stream, nullptr, state, &fs);
}

int magnus_fourier_vjp_basis_f64(const magnus_problem* p, uint64_t count, const double* coeffs, uint64_t modes,
                                 const double* cot, const double* state, double* dA, void* stream) {
    Plan pl;
    if (int rc = make_plan(p, false, false, &pl)) return rc;
    if (int rc = fourier_check(pl, modes, "basis-state VJP")) return rc;
    if (pl.op == MAGNUS_OP_ONE) return fail(MAGNUS_EINVAL, "basis-state VJP needs op many or sum");
    if (count == 0) return MAGNUS_OK;
    if (!coeffs || !cot || !state || !dA) return fail(MAGNUS_EINVAL, "null argument");
    DeviceScratch scratch((hipStream_t)stream);
    int rc;
    double* tab = fourier_table(scratch, p, modes, (hipStream_t)stream, &rc);
    if (rc) return rc;
    const FourierSrc fs{coeffs, tab, (int)modes};
    return run_wbd_vjp(pl, count, nullptr, cot, dA, (hipStream_t)stream, nullptr, state, &fs);
}

int magnus_fourier_forward_orders_from_f64(const magnus_problem* p, uint64_t count, uint64_t first,
                                           const double* coeffs, uint64_t modes, double* out, void* stream) {
    Plan pl;
    if (int rc = make_plan(p, false, true, &pl)) return rc;
    if (int rc = fourier_check(pl, modes, "all-orders forward")) return rc;
    if (first < 1 || first > pl.n) return fail(MAGNUS_EINVAL, "first order must lie in 1..n");
    if (count == 0) return MAGNUS_OK;
    if (!coeffs || !out) return fail(MAGNUS_EINVAL, "null argument");
    DeviceScratch scratch((hipStream_t)stream);
    int rc;
    double* tab = fourier_table(scratch, p, modes, (hipStream_t)stream, &rc);
    if (rc) return rc;
    const FourierSrc fs{coeffs, tab, (int)modes};
    return run_orders(pl, count, first, nullptr, out, (hipStream_t)stream, &fs);
}

uint64_t magnus_sc_checkpoint_elems(const magnus_problem* p, uint64_t count) {
    Plan pl;
    if (make_plan(p, true, true, &pl)) return 0;
    if (!use_wb(pl, false)) return 0;
    return wb_summary_elems(pl, count);
}

int magnus_sc_forward_ckpt_f64(const magnus_problem* p, uint64_t count, const double* in3, double* out3,
                               double* ckpt, void* stream) {
    Plan pl;
    if (int rc = make_plan(p, true, true, &pl)) return rc;
    if (!use_wb(pl, false))
        return fail(MAGNUS_EINVAL, "checkpointed forward needs the single-pass path (su(2) input, Boole/Simpson/Riemann, "
                                   "2 <= n <= 8)");
    if (count == 0) return MAGNUS_OK;
    if (!in3 || !out3 || !ckpt) return fail(MAGNUS_EINVAL, "null argument");
    return run_wb(pl, count, in3, nullptr, out3, nullptr, ckpt, (hipStream_t)stream);
}

int magnus_sc_vjp_ckpt_f64(const magnus_problem* p, uint64_t count, const double* in3, const double* cot3,
                           const double* ckpt, double* dA3, void* stream) {
    Plan pl;
    if (int rc = make_plan(p, true, false, &pl)) return rc;
    if (!use_wb(pl, false))
        return fail(MAGNUS_EINVAL, "checkpointed VJP needs the single-pass path (su(2) input, Boole/Simpson/Riemann, "
                                   "2 <= n <= 8)");
    if (count == 0) return MAGNUS_OK;
    if (!in3 || !cot3 || !ckpt || !dA3) return fail(MAGNUS_EINVAL, "null argument");
    // the summaries are only read by the kernels of this call
    return run_wb(pl, count, in3, cot3, nullptr, dA3, const_cast<double*>(ckpt), (hipStream_t)stream);
}

int magnus_sample_fourier_f64(uint64_t count, uint64_t samples, double t0, double tf, uint64_t entries,
                              uint64_t modes, const double* coeffs, double* out, void* stream) {
    if (!coeffs || !out) return fail(MAGNUS_EINVAL, "null argument");
    if (samples < 2) return fail(MAGNUS_EINVAL, "samples must be at least 2");
    if (entries == 0) return fail(MAGNUS_EINVAL, "entries must be positive");
    if (modes > (uint64_t)magnus::sample_max_modes()) return fail(MAGNUS_ERANGE, "too many Fourier modes (max 32)");
    if (count == 0) return MAGNUS_OK;
    magnus::SampleArgs a{};
    a.coeffs = coeffs;
    a.out = out;
    a.B = (long long)count;
    a.S = (long long)samples;
    a.m = (int)entries;
    a.modes = (int)modes;
    a.t0 = t0;
    a.tf = tf;
    a.step = (tf - t0) / (double)(samples - 1);
    hipError_t e = magnus::launch_sample_fourier(a, (hipStream_t)stream);
    if (e != hipSuccess) return device_error(e, "sampling kernel launch");
    return MAGNUS_OK;
}

int magnus_sample_sc_fourier_f64(uint64_t count, uint64_t samples, double t0, double tf, uint64_t modes,
                                 const double* amp, const double* phase, double* out, void* stream) {
    if (!amp || !phase || !out) return fail(MAGNUS_EINVAL, "null argument");
    if (samples < 2) return fail(MAGNUS_EINVAL, "samples must be at least 2");
    if (modes > (uint64_t)magnus::sample_max_modes()) return fail(MAGNUS_ERANGE, "too many Fourier modes (max 32)");
    if (count == 0) return MAGNUS_OK;
    magnus::SampleArgs a{};
    a.amp = amp;
    a.phase = phase;
    a.out = out;
    a.B = (long long)count;
    a.S = (long long)samples;
    a.m = 3;
    a.modes = (int)modes;
    a.t0 = t0;
    a.tf = tf;
    a.step = (tf - t0) / (double)(samples - 1);
    hipError_t e = magnus::launch_sample_sc_fourier(a, (hipStream_t)stream);
    if (e != hipSuccess) return device_error(e, "sampling kernel launch");
    return MAGNUS_OK;
}

int magnus_expm_f64(uint64_t count, uint64_t dim, const double* in, double* out, void* stream) {
    if (!in || !out) return fail(MAGNUS_EINVAL, "null argument");
    if (dim == 0) return fail(MAGNUS_EINVAL, "dim must be positive");
    if (count == 0) return MAGNUS_OK;
    DeviceScratch scratch((hipStream_t)stream);
    magnus::ExpmArgs a{};
    a.in = in;
    a.out = out;
    a.B = (long long)count;
    a.d = (int)dim;
    a.work = scratch.alloc<double>((size_t)count * 8 * dim * dim);
    if (!a.work) return fail(MAGNUS_ENOMEM, "device allocation failed");
    hipError_t e = magnus::launch_expm(a, (hipStream_t)stream);
    if (e != hipSuccess) return device_error(e, "expm kernel launch");
    return MAGNUS_OK;
}

int magnus_sc_expm_f64(uint64_t count, const double* in3, double* out, void* stream) {
    if (!in3 || !out) return fail(MAGNUS_EINVAL, "null argument");
    if (count == 0) return MAGNUS_OK;
    magnus::ExpmArgs a{};
    a.in = in3;
    a.out = out;
    a.B = (long long)count;
    hipError_t e = magnus::launch_sc_expm(a, (hipStream_t)stream);
    if (e != hipSuccess) return device_error(e, "expm kernel launch");
    return MAGNUS_OK;
}

int magnus_timing_enable(int on) {
    std::lock_guard<std::mutex> lock(g_timing_mu);
    g_timing_on = on != 0;
    return MAGNUS_OK;
}

int magnus_timing_read(const char* name, uint64_t* launches, double* total_ms) {
    std::lock_guard<std::mutex> lock(g_timing_mu);
    timing_drain_locked();
    uint64_t n = 0;
    double ms = 0.0;
    for (auto& t : g_totals)
        if (name && t.first == name) {
            n = t.second.launches;
            ms = t.second.ms;
        }
    if (launches) *launches = n;
    if (total_ms) *total_ms = ms;
    return MAGNUS_OK;
}

int magnus_timing_counts(const char* name, double* bytes, double* flops) {
    std::lock_guard<std::mutex> lock(g_timing_mu);
    timing_drain_locked();
    double b = 0.0, f = 0.0;
    for (auto& t : g_totals)
        if (name && t.first == name) {
            b = t.second.bytes;
            f = t.second.flops;
        }
    if (bytes) *bytes = b;
    if (flops) *flops = f;
    return MAGNUS_OK;
}

void magnus_timing_reset(void) {
    std::lock_guard<std::mutex> lock(g_timing_mu);
    timing_drain_locked();
    g_totals.clear();
}

int magnus_set_scratch_limit(uint64_t bytes) {
    g_scratch_limit.store(bytes, std::memory_order_relaxed);
    return MAGNUS_OK;
}

uint64_t magnus_scratch_limit(void) { return g_scratch_limit.load(std::memory_order_relaxed); }

int magnus_release_scratch(void) {
    release_scratch_caches();
    return MAGNUS_OK;
}

const char* magnus_last_error(void) { return g_err.c_str(); }
const char* magnus_last_kernel(void) { return g_kernel; }

}  // extern "C"
