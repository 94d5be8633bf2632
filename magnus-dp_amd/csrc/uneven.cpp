// uneven.cpp -- uneven-batch plans: many matrix problems of different dim,
// samples and order solved by one enqueue (or one hipGraph launch).
//
// Schedule.  Instances are grouped into buckets of equal shape (n, samples,
// dim, op, matrix backend, integrator, dt); each bucket runs the batched
// level-parallel kernels of the one-problem path, so every instance's results
// are those of a per-instance call.  A bucket is cut into sub-batches sized to
// the per-branch workspace; sub-batches are assigned longest-first to the
// branch (HIP stream) with the least estimated work, and each branch owns one
// arena sized for its largest sub-batch, reused in stream order.  All device
// memory (pointer tables, slot lists, arenas) is allocated when the plan is
// prepared, so the enqueue makes no allocation call and can be captured.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <sstream>
#include <string>
#include <tuple>
#include <vector>

#include "host.hpp"
#include "kernels_uneven.hpp"
#include "magnus_capi.h"

using namespace magnus_host;

namespace {

constexpr int kMaxBranches = 8;

struct Bucket {
    Plan pl;
    std::vector<uint32_t> ids;  // instance ids
    size_t ids_off = 0;         // offset of ids in the device slot list
    int sub = 1;                // instances per sub-batch
    double ms_per_inst = -1.0;  // measured VJP time per instance (< 0: not measured)
};

struct Item {
    int bucket;
    uint32_t first, count;  // slots [first, first + count) of the bucket
    size_t bytes;           // arena bytes
    double est;             // estimated time (arbitrary units until tuned)
    int branch;
};

size_t fwd_item_bytes(const Plan& pl, uint64_t c) {
    return scratch_round(c * pl.S * pl.m_in * 8) + scratch_round(c * pl.n_out * pl.m * 8) +
           forward_scratch_bytes(pl, c);
}
size_t vjp_item_bytes(const Plan& pl, uint64_t c) {
    return scratch_round(c * pl.S * pl.m_in * 8) + scratch_round(c * pl.n_out * pl.m_cot * 8) +
           scratch_round(c * pl.S * pl.m * 8) + vjp_scratch_bytes(pl, c);
}
size_t item_bytes(const Plan& pl, uint64_t c) { return std::max(fwd_item_bytes(pl, c), vjp_item_bytes(pl, c)); }

// flop model of one instance's VJP with recompute, used until a bucket is measured
double model_cost(const Plan& pl) {
    const double Q = (double)pl.Q, K = (double)(pl.n - 1), S = (double)pl.S, d = (double)pl.d, m = (double)pl.m;
    const double P = 2.0 * d * d * d;
    return Q * K * S * (2.0 * P + 14.0 * m) + 2.0 * m * S + 1e5;  // + per-launch overhead
}

// product paths worth timing: at d = 32 the three MFMA splits, at d = 16 MFMA
// against the VALU row products; the per-entry VALU path is never faster at
// d >= 32 (tile 1 stays selectable through the tuning table)
std::vector<int> tile_candidates(const Plan& pl) {
    if (pl.d == 32) return {0, 2, 3};
    if (pl.d == 16) return {0, 1};
    return {0};
}

}  // namespace

struct magnus_uneven_plan {
    uint64_t count = 0;
    int branches = 1;
    size_t budget = 0;  // total workspace
    std::vector<Bucket> buckets;
    std::vector<Item> items;
    std::vector<std::vector<int>> order;  // per branch: item indices in launch order
    int device = 0;
    // device memory: tables [3][count] pointers, slot ids, arenas
    char* dmem = nullptr;
    size_t dbytes = 0;
    const double** d_tab = nullptr;
    uint32_t* d_ids = nullptr;
    std::vector<Arena> arenas;
    // pinned host table and its upload event
    void** h_tab = nullptr;
    hipEvent_t tab_done = nullptr;
    bool tab_pending = false;
    hipStream_t streams[kMaxBranches] = {};
    hipEvent_t fork = nullptr, join[kMaxBranches] = {};
    hipStream_t cap = nullptr;
    hipGraphExec_t exec[2] = {nullptr, nullptr};  // 0 forward, 1 VJP
    hipGraph_t graph[2] = {nullptr, nullptr};
};

namespace {

void drop_graphs(magnus_uneven_plan* p) {
    for (int i = 0; i < 2; ++i) {
        if (p->exec[i]) (void)hipGraphExecDestroy(p->exec[i]);
        if (p->graph[i]) (void)hipGraphDestroy(p->graph[i]);
        p->exec[i] = nullptr;
        p->graph[i] = nullptr;
    }
}

// items from the buckets' sub-batch sizes, longest-first branch assignment,
// per-branch arenas, and the device allocation that holds everything
int prepare(magnus_uneven_plan* p) {
    drop_graphs(p);
    p->items.clear();
    for (size_t b = 0; b < p->buckets.size(); ++b) {
        const Bucket& bk = p->buckets[b];
        const double per = bk.ms_per_inst >= 0 ? bk.ms_per_inst * 1e9 : model_cost(bk.pl);
        for (uint32_t f = 0; f < bk.ids.size(); f += bk.sub) {
            const uint32_t c = std::min<uint32_t>(bk.sub, (uint32_t)bk.ids.size() - f);
            p->items.push_back({(int)b, f, c, item_bytes(bk.pl, c), per * c, 0});
        }
    }
    std::vector<int> idx(p->items.size());
    for (size_t i = 0; i < idx.size(); ++i) idx[i] = (int)i;
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return p->items[a].est > p->items[b].est; });
    std::vector<double> load(p->branches, 0.0);
    std::vector<size_t> need(p->branches, 0);
    p->order.assign(p->branches, {});
    for (int i : idx) {
        int br = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        p->items[i].branch = br;
        load[br] += p->items[i].est;
        need[br] = std::max(need[br], p->items[i].bytes);
        p->order[br].push_back(i);
    }
    const size_t tab_bytes = scratch_round(3 * p->count * sizeof(void*));
    size_t ids_total = 0;
    for (auto& bk : p->buckets) {
        bk.ids_off = ids_total;
        ids_total += bk.ids.size();
    }
    const size_t ids_bytes = scratch_round(std::max<size_t>(ids_total, 1) * sizeof(uint32_t));
    size_t total = tab_bytes + ids_bytes;
    for (int b = 0; b < p->branches; ++b) total += scratch_round(need[b]);
    if (p->dmem) {
        (void)hipDeviceSynchronize();
        (void)hipFree(p->dmem);
        p->dmem = nullptr;
    }
    if (hipMalloc((void**)&p->dmem, total) != hipSuccess) {
        (void)hipGetLastError();
        return fail(MAGNUS_ENOMEM, "uneven plan: device allocation of " + std::to_string(total) + " bytes failed");
    }
    p->dbytes = total;
    p->d_tab = reinterpret_cast<const double**>(p->dmem);
    p->d_ids = reinterpret_cast<uint32_t*>(p->dmem + tab_bytes);
    std::vector<uint32_t> ids_host;
    ids_host.reserve(ids_total);
    for (auto& bk : p->buckets) ids_host.insert(ids_host.end(), bk.ids.begin(), bk.ids.end());
    if (ids_total && hipMemcpy(p->d_ids, ids_host.data(), ids_total * sizeof(uint32_t), hipMemcpyHostToDevice) !=
                         hipSuccess)
        return device_error(hipGetLastError(), "uneven plan: slot upload");
    p->arenas.assign(p->branches, Arena{});
    char* at = p->dmem + tab_bytes + ids_bytes;
    for (int b = 0; b < p->branches; ++b) {
        p->arenas[b].base = at;
        p->arenas[b].size = need[b];
        p->arenas[b].used = 0;
        at += scratch_round(need[b]);
    }
    return MAGNUS_OK;
}

// one sub-batch on its branch stream
int enqueue_item(magnus_uneven_plan* p, const Item& it, bool vjp, hipStream_t s) {
    const Bucket& bk = p->buckets[it.bucket];
    const Plan& pl = bk.pl;
    Arena* ar = &p->arenas[it.branch];
    DeviceScratch ds(s, ar);
    const uint32_t* ids = p->d_ids + bk.ids_off + it.first;
    const uint64_t c = it.count;
    double* in = ds.alloc<double>(c * pl.S * pl.m_in);
    magnus::SegmentArgs g{};
    g.ids = ids;
    g.count = (int)c;
    g.tab = p->d_tab;
    g.packed = in;
    g.per = (long long)(pl.S * pl.m_in);
    if (!in) return fail(MAGNUS_ENOMEM, "uneven plan: arena exhausted");
    hipError_t e = magnus::launch_gather(g, s);
    if (e != hipSuccess) return device_error(e, "uneven gather");
    if (!vjp) {
        double* out = ds.alloc<double>(c * pl.n_out * pl.m);
        if (!out) return fail(MAGNUS_ENOMEM, "uneven plan: arena exhausted");
        if (int rc = run_forward(pl, c, in, out, nullptr, true, s, ar)) return rc;
        g.tab = p->d_tab + p->count;
        g.packed = out;
        g.per = (long long)(pl.n_out * pl.m);
        e = magnus::launch_scatter(g, s);
        if (e != hipSuccess) return device_error(e, "uneven scatter");
        return MAGNUS_OK;
    }
    double* cot = ds.alloc<double>(c * pl.n_out * pl.m_cot);
    double* dA = ds.alloc<double>(c * pl.S * pl.m);
    if (!cot || !dA) return fail(MAGNUS_ENOMEM, "uneven plan: arena exhausted");
    g.tab = p->d_tab + p->count;
    g.packed = cot;
    g.per = (long long)(pl.n_out * pl.m_cot);
    e = magnus::launch_gather(g, s);
    if (e != hipSuccess) return device_error(e, "uneven gather");
    if (int rc = run_vjp(pl, c, in, cot, nullptr, dA, s, ar)) return rc;
    g.tab = p->d_tab + 2 * p->count;
    g.packed = dA;
    g.per = (long long)(pl.S * pl.m);
    e = magnus::launch_scatter(g, s);
    if (e != hipSuccess) return device_error(e, "uneven scatter");
    return MAGNUS_OK;
}

// fork from `origin` onto the branch streams, every sub-batch, join back
int enqueue_all(magnus_uneven_plan* p, bool vjp, hipStream_t origin) {
    hipError_t e = hipEventRecord(p->fork, origin);
    if (e != hipSuccess) return device_error(e, "uneven fork");
    for (int b = 0; b < p->branches; ++b) {
        if (p->order[b].empty()) continue;
        e = hipStreamWaitEvent(p->streams[b], p->fork, 0);
        if (e != hipSuccess) return device_error(e, "uneven fork");
        for (int i : p->order[b])
            if (int rc = enqueue_item(p, p->items[i], vjp, p->streams[b])) return rc;
        e = hipEventRecord(p->join[b], p->streams[b]);
        if (e == hipSuccess) e = hipStreamWaitEvent(origin, p->join[b], 0);
        if (e != hipSuccess) return device_error(e, "uneven join");
    }
    return MAGNUS_OK;
}

// pointer tables: rows in, (out | cot), dA; uploaded on `s` ahead of the work
int upload_tables(magnus_uneven_plan* p, const double* const* r0, const double* const* r1, const double* const* r2,
                  hipStream_t s) {
    if (p->tab_pending) {
        hipError_t e = hipEventSynchronize(p->tab_done);  // the previous upload has read the host table
        if (e != hipSuccess) return device_error(e, "uneven table");
    }
    const double* const* rows[3] = {r0, r1, r2};
    for (int r = 0; r < 3; ++r)
        for (uint64_t i = 0; i < p->count; ++i) p->h_tab[r * p->count + i] = rows[r] ? (void*)rows[r][i] : nullptr;
    hipError_t e = hipMemcpyAsync(p->d_tab, p->h_tab, 3 * p->count * sizeof(void*), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipEventRecord(p->tab_done, s);
    if (e != hipSuccess) return device_error(e, "uneven table");
    p->tab_pending = true;
    return MAGNUS_OK;
}

int run(magnus_uneven_plan* p, bool vjp, bool use_graph, hipStream_t s) {
    if (!use_graph) return enqueue_all(p, vjp, s);
    const int g = vjp ? 1 : 0;
    if (!p->exec[g]) {
        hipError_t e = hipStreamBeginCapture(p->cap, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return device_error(e, "uneven capture");
        const bool prev = timing_suspend(true);
        int rc = enqueue_all(p, vjp, p->cap);
        timing_suspend(prev);
        hipGraph_t graph = nullptr;
        e = hipStreamEndCapture(p->cap, &graph);
        if (rc) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        if (e != hipSuccess) return device_error(e, "uneven capture");
        e = hipGraphInstantiate(&p->exec[g], graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            (void)hipGraphDestroy(graph);
            p->exec[g] = nullptr;
            return device_error(e, "uneven graph instantiate");
        }
        p->graph[g] = graph;
    }
    hipError_t e = hipGraphLaunch(p->exec[g], s);
    if (e != hipSuccess) return device_error(e, "uneven graph launch");
    return MAGNUS_OK;
}

void destroy(magnus_uneven_plan* p) {
    if (!p) return;
    (void)hipDeviceSynchronize();
    drop_graphs(p);
    for (int b = 0; b < kMaxBranches; ++b) {
        if (p->streams[b]) (void)hipStreamDestroy(p->streams[b]);
        if (p->join[b]) (void)hipEventDestroy(p->join[b]);
    }
    if (p->cap) (void)hipStreamDestroy(p->cap);
    if (p->fork) (void)hipEventDestroy(p->fork);
    if (p->tab_done) (void)hipEventDestroy(p->tab_done);
    if (p->h_tab) (void)hipHostFree(p->h_tab);
    if (p->dmem) (void)hipFree(p->dmem);
    delete p;
}

}  // namespace

extern "C" {

int magnus_uneven_create(const magnus_problem* probs, uint64_t count, uint64_t workspace_bytes, int branches,
                         magnus_uneven_plan** out) {
    if (!out) return fail(MAGNUS_EINVAL, "null argument");
    *out = nullptr;
    if (!probs && count) return fail(MAGNUS_EINVAL, "null argument");
    if (branches < 1 || branches > kMaxBranches) return fail(MAGNUS_EINVAL, "branches must be in 1..8");
    if (count > 0xffffffffull) return fail(MAGNUS_EINVAL, "too many instances");
    std::vector<Plan> plans(count);
    for (uint64_t i = 0; i < count; ++i) {
        if (int rc = make_plan(&probs[i], false, true, &plans[i])) {
            std::string msg = "instance " + std::to_string(i) + ": " + magnus_last_error();
            return fail(rc, msg);
        }
    }
    auto* p = new magnus_uneven_plan();
    p->count = count;
    p->branches = branches;
    if (hipGetDevice(&p->device) != hipSuccess) {
        delete p;
        return device_error(hipGetLastError(), "uneven plan: no device");
    }
    if (workspace_bytes == 0) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
            delete p;
            return device_error(hipGetLastError(), "uneven plan: memory query");
        }
        workspace_bytes = fr / 2;
    }
    p->budget = workspace_bytes;
    const size_t per_branch = workspace_bytes / branches;
    std::map<std::tuple<uint64_t, uint64_t, uint64_t, int, int, int, uint64_t>, int> keys;
    for (uint64_t i = 0; i < count; ++i) {
        const Plan& pl = plans[i];
        uint64_t dtb;
        std::memcpy(&dtb, &pl.dt, sizeof dtb);
        auto key = std::make_tuple(pl.d, pl.S, pl.n, pl.op, pl.mat, pl.integ, dtb);
        auto it = keys.find(key);
        if (it == keys.end()) {
            it = keys.emplace(key, (int)p->buckets.size()).first;
            p->buckets.push_back(Bucket{});
            p->buckets.back().pl = pl;
        }
        p->buckets[it->second].ids.push_back((uint32_t)i);
    }
    for (auto& bk : p->buckets) {
        if (item_bytes(bk.pl, 1) > per_branch) {
            destroy(p);
            return fail(MAGNUS_ENOMEM, "uneven plan: one instance of dim " + std::to_string(bk.pl.d) + ", " +
                                           std::to_string(bk.pl.S) + " samples needs " +
                                           std::to_string(item_bytes(bk.pl, 1)) + " workspace bytes per branch");
        }
        uint64_t sub = bk.ids.size();
        while (sub > 1 && item_bytes(bk.pl, sub) > per_branch) sub = (sub + 1) / 2;
        bk.sub = (int)std::min<uint64_t>(sub, 65535);
    }
    hipError_t e = hipSuccess;
    for (int b = 0; b < branches && e == hipSuccess; ++b) {
        e = hipStreamCreateWithFlags(&p->streams[b], hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&p->join[b], hipEventDisableTiming);
    }
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&p->cap, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&p->fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&p->tab_done, hipEventDisableTiming);
    if (e == hipSuccess) e = hipHostMalloc((void**)&p->h_tab, std::max<uint64_t>(3 * count, 1) * sizeof(void*));
    if (e != hipSuccess) {
        destroy(p);
        return device_error(e, "uneven plan: stream/event/host table");
    }
    if (int rc = prepare(p)) {
        destroy(p);
        return rc;
    }
    *out = p;
    return MAGNUS_OK;
}

void magnus_uneven_destroy(magnus_uneven_plan* plan) { destroy(plan); }

int magnus_uneven_forward_f64(magnus_uneven_plan* p, const double* const* in, double* const* out, int use_graph,
                              void* stream) {
    if (!p) return fail(MAGNUS_EINVAL, "null plan");
    if (p->count == 0) return MAGNUS_OK;
    if (!in || !out) return fail(MAGNUS_EINVAL, "null argument");
    for (uint64_t i = 0; i < p->count; ++i)
        if (!in[i] || !out[i]) return fail(MAGNUS_EINVAL, "instance " + std::to_string(i) + ": null buffer");
    hipStream_t s = (hipStream_t)stream;
    if (int rc = upload_tables(p, in, const_cast<const double* const*>(out), nullptr, s)) return rc;
    if (int rc = run(p, false, use_graph != 0, s)) return rc;
    set_last_kernel("lm");
    return MAGNUS_OK;
}

int magnus_uneven_vjp_f64(magnus_uneven_plan* p, const double* const* in, const double* const* cot,
                          double* const* dA, int use_graph, void* stream) {
    if (!p) return fail(MAGNUS_EINVAL, "null plan");
    if (p->count == 0) return MAGNUS_OK;
    if (!in || !cot || !dA) return fail(MAGNUS_EINVAL, "null argument");
    for (uint64_t i = 0; i < p->count; ++i)
        if (!in[i] || !cot[i] || !dA[i]) return fail(MAGNUS_EINVAL, "instance " + std::to_string(i) + ": null buffer");
    hipStream_t s = (hipStream_t)stream;
    if (int rc = upload_tables(p, in, cot, const_cast<const double* const*>(dA), s)) return rc;
    if (int rc = run(p, true, use_graph != 0, s)) return rc;
    set_last_kernel("lm");
    return MAGNUS_OK;
}

int magnus_uneven_autotune(magnus_uneven_plan* p, const double* const* in, const double* const* cot,
                           double* const* dA, void* stream) {
    if (!p) return fail(MAGNUS_EINVAL, "null plan");
    if (p->count == 0) return MAGNUS_OK;
    if (!in || !cot || !dA) return fail(MAGNUS_EINVAL, "null argument");
    hipStream_t s = (hipStream_t)stream;
    // earlier calls may still use the arenas on the branch streams
    if (hipDeviceSynchronize() != hipSuccess) return device_error(hipGetLastError(), "uneven autotune");
    if (int rc = upload_tables(p, in, cot, const_cast<const double* const*>(dA), s)) return rc;
    hipEvent_t t0, t1;
    if (hipEventCreate(&t0) != hipSuccess || hipEventCreate(&t1) != hipSuccess)
        return device_error(hipGetLastError(), "uneven autotune events");
    int rc = MAGNUS_OK;
    const bool prev = timing_suspend(true);
    for (size_t b = 0; b < p->buckets.size() && rc == MAGNUS_OK; ++b) {
        Bucket& bk = p->buckets[b];
        // the bucket's first sub-batch, on the arena of the branch that holds it
        const Item* first = nullptr;
        for (auto& it : p->items)
            if (it.bucket == (int)b && it.first == 0) first = &it;
        if (!first) continue;
        Item probe = *first;
        double best_ms = -1.0;
        int best_tile = bk.pl.tile;
        for (int tile : tile_candidates(bk.pl)) {
            bk.pl.tile = tile;
            rc = enqueue_item(p, probe, true, s);  // warm-up
            float ms = 0.f;
            if (!rc) rc = hipEventRecord(t0, s) == hipSuccess ? MAGNUS_OK : device_error(hipGetLastError(), "event");
            for (int r = 0; r < 2 && !rc; ++r) rc = enqueue_item(p, probe, true, s);
            if (!rc) rc = hipEventRecord(t1, s) == hipSuccess ? MAGNUS_OK : device_error(hipGetLastError(), "event");
            if (!rc && hipEventSynchronize(t1) == hipSuccess && hipEventElapsedTime(&ms, t0, t1) == hipSuccess) {
                const double per = ms / 2.0 / probe.count;
                if (best_ms < 0 || per < best_ms) {
                    best_ms = per;
                    best_tile = tile;
                }
            }
            if (rc) break;
        }
        bk.pl.tile = best_tile;
        if (best_ms >= 0) bk.ms_per_inst = best_ms;
    }
    timing_suspend(prev);
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    if (rc) return rc;
    if (hipStreamSynchronize(s) != hipSuccess) return device_error(hipGetLastError(), "uneven autotune");
    return prepare(p);
}

uint64_t magnus_uneven_tuning(const magnus_uneven_plan* p, char* buf, uint64_t len) {
    if (!p) return 0;
    std::ostringstream os;
    for (auto& bk : p->buckets) {
        char line[128];
        std::snprintf(line, sizeof line, "%llu %llu %llu %d %.6g\n", (unsigned long long)bk.pl.d,
                      (unsigned long long)bk.pl.S, (unsigned long long)bk.pl.n, bk.pl.tile, bk.ms_per_inst);
        os << line;
    }
    const std::string t = os.str();
    if (buf && len) {
        const size_t n = std::min<size_t>(t.size(), len - 1);
        std::memcpy(buf, t.data(), n);
        buf[n] = '\0';
    }
    return t.size();
}

int magnus_uneven_set_tuning(magnus_uneven_plan* p, const char* text) {
    if (!p || !text) return fail(MAGNUS_EINVAL, "null argument");
    std::istringstream is(text);
    std::string line;
    bool changed = false;
    while (std::getline(is, line)) {
        unsigned long long d, S, n;
        int tile;
        double ms = -1.0;
        const int got = std::sscanf(line.c_str(), "%llu %llu %llu %d %lf", &d, &S, &n, &tile, &ms);
        if (got < 4) continue;
        if (tile < 0 || tile > 3) return fail(MAGNUS_EINVAL, "tile must be in 0..3");
        for (auto& bk : p->buckets)
            if (bk.pl.d == d && bk.pl.S == S && bk.pl.n == n) {
                bk.pl.tile = tile;
                if (got == 5) bk.ms_per_inst = ms;
                changed = true;
            }
    }
    if (!changed) return MAGNUS_OK;
    (void)hipDeviceSynchronize();
    return prepare(p);
}

uint64_t magnus_uneven_info(const magnus_uneven_plan* p, int what) {
    if (!p) return 0;
    switch (what) {
        case 0: return p->buckets.size();
        case 1: return p->items.size();
        case 2: {
            uint64_t used = 0;
            for (auto& o : p->order) used += o.empty() ? 0 : 1;
            return used;
        }
        case 3: return p->dbytes;
        default: return 0;
    }
}

}  // extern "C"
