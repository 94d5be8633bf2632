// uneven.cpp -- uneven-batch plans: many matrix problems of different dim,
// samples and order solved by one enqueue (or one hipGraph launch).
//
// Schedule.  Dense W-basis problems of dim 16, 32 or 64 under the Boole rule (op
// many or sum, n >= 2) run as one grouped pipeline: every such instance, of
// any samples and order, is a set of tasks (instance, segment of samples,
// level) of the grouped walkers, and one launch per pipeline step and dim
// runs every task of that step (kernels_walkseg.hip); the GL fold between the
// passes runs once per (dim, order, op) group.  Instances that do not fit the
// pipeline's workspace together run in consecutive waves.
//
// The other instances are grouped into buckets of equal shape (n, samples,
// dim, op, matrix backend, integrator, dt); each bucket runs the batched
// level-parallel kernels of the one-problem path, so every such instance's
// results are those of a per-instance call.  A bucket is cut into sub-batches
// sized to the per-branch workspace; sub-batches are assigned longest-first to
// the branch (HIP stream) with the least estimated work, and each branch owns
// one arena sized for its largest sub-batch, reused in stream order.  The
// pipeline of each dim runs on a stream of its own beside the branches (the
// dims' instances share nothing).  All device memory
// (pointer tables, slot lists, arenas, task tables, the pipeline's workspace)
// is allocated when the plan is prepared, so the enqueue makes no allocation
// call and can be captured.
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
#include "kernels.hpp"
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

// ---- grouped walker pipeline -------------------------------------------------
constexpr int kGwDims = 3;  // dim 16, 32, 64
constexpr int kGwDim[kGwDims] = {16, 32, 64};
constexpr int kGwL = 128;   // default segment length in samples (a multiple of the Boole block)
constexpr int kGwLCandidates[3] = {64, 128, 256};  // the autotuner's choices; sizes are planned for the largest
constexpr int kGwLMax = 256;

int gw_dim_index(uint64_t d) { return d == 16 ? 0 : (d == 32 ? 1 : 2); }
bool gw_eligible(const Plan& pl) {
    return !pl.sc && pl.op != MAGNUS_OP_ONE && magnus::gw_ok((int)pl.d, pl.integ, (int)pl.n);
}

// instances of one (dim, order, op): one fold launch, packed per-slot arrays
struct GwGroup {
    Plan pl;
    std::vector<uint32_t> ids;  // the instances' table rows, in slot order
    size_t ids_off = 0;         // offset of ids in the device slot list
    size_t z = 0, z0 = 0, totals = 0, zbar = 0, wbar = 0, cot = 0, out = 0, part = 0, part2 = 0, work = 0;
};

struct GwWave {
    std::vector<GwGroup> groups;
    size_t inst_off = 0, ninst = 0;  // in the device GwInst array
    size_t bytes = 0;                // workspace bytes
    // per dim: task offsets of each step (size steps + 1), into the device task table at *_base
    std::vector<size_t> fwd_off[kGwDims], rev_off[kGwDims];
    size_t fwd_base[kGwDims] = {}, rev_base[kGwDims] = {};
    std::vector<double> fwd_bytes[kGwDims], fwd_flops[kGwDims], rev_bytes[kGwDims], rev_flops[kGwDims];
};

size_t gw_inst_bytes(const Plan& pl, int L) {
    const size_t K = pl.n - 1, m = pl.m, n = pl.n, S = pl.S, Q = pl.Q, no = pl.n_out;
    return scratch_round(K * S * m * 8) + 2 * scratch_round(K * 2 * m * 8) + scratch_round(K * 2 * (L + 1) * m * 8) +
           scratch_round(n * m * 8) * 5 + scratch_round(m * 8) + 2 * scratch_round(no * m * 8) +
           2 * scratch_round(Q * n * m * 8) + scratch_round(Q * (2 * K + 1) * m * 8);
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
    // grouped walker pipeline
    std::vector<Plan> gw_plans;  // per instance (empty Plan where the instance is not in the pipeline)
    std::vector<uint32_t> gw_ids;
    std::vector<GwWave> gw;
    size_t gw_budget = 0;
    int gw_L = kGwL;            // segment length of the pipeline's tasks
    double gw_ms = -1.0;        // measured pipeline VJP time (< 0: not measured)
    magnus::GwInst* d_inst = nullptr;
    magnus::GwTask* d_tasks = nullptr;
    char* gw_ws = nullptr;
    hipStream_t gw_stream[kGwDims] = {};  // one pipeline per dim: the dims' instances are independent
    hipEvent_t gw_join[kGwDims] = {};
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

// Waves of pipeline instances (largest first) within the pipeline's budget,
// their (dim, order, op) groups, workspace offsets and the task tables of every
// step.  Group slot lists are appended to the device slot list at ids_total.
int gw_layout(magnus_uneven_plan* p, size_t& ids_total, std::vector<magnus::GwInst>& inst,
              std::vector<magnus::GwTask>& tasks, std::vector<size_t>& slot_ws) {
    p->gw.clear();
    if (p->gw_ids.empty()) return MAGNUS_OK;
    const int L = p->gw_L;
    std::vector<uint32_t> order = p->gw_ids;
    auto work = [&](uint32_t i) {
        const Plan& pl = p->gw_plans[i];
        return (double)pl.S * (double)(pl.n - 1) * (double)pl.d * pl.d * pl.d;
    };
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return work(a) > work(b); });
    std::vector<std::vector<uint32_t>> waves;
    size_t cur = 0;
    for (uint32_t i : order) {
        const size_t b = gw_inst_bytes(p->gw_plans[i], L);
        if (b > p->gw_budget)
            return fail(MAGNUS_ENOMEM, "uneven plan: one instance of dim " + std::to_string(p->gw_plans[i].d) + ", " +
                                           std::to_string(p->gw_plans[i].S) + " samples needs " + std::to_string(b) +
                                           " pipeline workspace bytes");
        if (waves.empty() || cur + b > p->gw_budget) {
            waves.push_back({});
            cur = 0;
        }
        waves.back().push_back(i);
        cur += b;
    }
    for (auto& members : waves) {
        GwWave wv;
        std::map<std::tuple<uint64_t, uint64_t, int>, int> gkey;
        for (uint32_t i : members) {
            const Plan& pl = p->gw_plans[i];
            auto key = std::make_tuple(pl.d, pl.n, pl.op);
            auto it = gkey.find(key);
            if (it == gkey.end()) {
                it = gkey.emplace(key, (int)wv.groups.size()).first;
                wv.groups.push_back(GwGroup{});
                wv.groups.back().pl = pl;
            }
            wv.groups[it->second].ids.push_back(i);
        }
        size_t off = 0;
        auto take = [&](size_t bytes) {
            const size_t at = off;
            off += scratch_round(bytes);
            return at;
        };
        wv.inst_off = inst.size();
        std::vector<int> steps_f(kGwDims, 0), steps_r(kGwDims, 0);
        struct Pending {
            int slot, G, n, dim;
        };
        std::vector<Pending> pend;
        for (auto& grp : wv.groups) {
            const Plan& pl = grp.pl;
            const size_t B = grp.ids.size(), m = pl.m, n = pl.n, K = n - 1, Q = pl.Q, no = pl.n_out;
            grp.ids_off = ids_total;
            ids_total += B;
            grp.z = take(B * n * m * 8);
            grp.z0 = take(B * m * 8);
            grp.totals = take(B * n * m * 8);
            grp.zbar = take(B * n * m * 8);
            grp.wbar = take(B * n * m * 8);
            grp.cot = take(B * no * m * 8);
            grp.out = take(B * no * m * 8);
            grp.part = take(B * Q * n * m * 8);
            grp.part2 = take(B * Q * n * m * 8);
            grp.work = take(B * Q * (2 * K + 1) * m * 8);
            for (size_t b = 0; b < B; ++b) {
                const uint32_t i = grp.ids[b];
                const Plan& pi = p->gw_plans[i];
                magnus::GwInst gi{};
                gi.S = (long long)pi.S;
                gi.n = (int)pi.n;
                gi.id = (int)i;
                gi.dt = pi.dt;
                // device pointers are bound once the workspace is allocated: keep offsets
                // in the pointer fields' place through slot_ws (arrays below) and the group
                const size_t own = take(K * pi.S * m * 8);
                take(K * 2 * m * 8);
                take(K * 2 * (L + 1) * m * 8);
                take(K * 2 * m * 8);
                slot_ws.push_back(own);
                const int G = (int)((pi.S - 1 + L - 1) / L);
                const int di = gw_dim_index(pi.d);
                pend.push_back({(int)(inst.size() - wv.inst_off), G, (int)pi.n, di});
                inst.push_back(gi);
                steps_f[di] = std::max(steps_f[di], G + (int)pi.n - 1);
                steps_r[di] = std::max(steps_r[di], G + (int)pi.n - 2 + 1);
            }
        }
        wv.ninst = inst.size() - wv.inst_off;
        wv.bytes = off;
        // task tables: per dim, steps in order
        for (int di = 0; di < kGwDims; ++di) {
            const double d = kGwDim[di], m = d * d;
            std::vector<std::vector<magnus::GwTask>> fs(steps_f[di]), rs(steps_r[di]);
            std::vector<double> fb(steps_f[di], 0.0), ff(steps_f[di], 0.0), rb(steps_r[di], 0.0), rf(steps_r[di], 0.0);
            for (auto& q : pend) {
                if (q.dim != di) continue;
                const long long S = (long long)inst[wv.inst_off + q.slot].S;
                for (int g = 0; g < q.G; ++g) {
                    const long long s0 = (long long)g * L, s1 = std::min<long long>(s0 + L, S - 1);
                    const double len = (double)(s1 - s0 + (g == 0 ? 1 : 0));
                    const double own = (double)((s1 == S - 1 ? S - 1 : s1 - 1) - s0 + 1);
                    for (int j = 1; j <= q.n; ++j) {
                        const int t = g + j - 1;
                        fs[t].push_back({q.slot, g, j, 0});
                        fb[t] += 8.0 * len * m * ((j >= 2 ? 2 : 1) + (j < q.n ? 1 : 0));
                        ff[t] += len * ((j >= 2 ? 2.0 * d * d * d : 0.0) + 8.0 * m);
                    }
                    for (int k = q.n; k >= 2; --k) {
                        const int t = (q.G - 1 - g) + (q.n - k);
                        rs[t].push_back({q.slot, g, k, 1});
                        rb[t] += 8.0 * own * m * (k == q.n ? 4 : 6);
                        rf[t] += own * (4.0 * d * d * d + 16.0 * m);
                    }
                    const int tf = (q.G - 1 - g) + (q.n - 2) + 1;
                    rs[tf].push_back({q.slot, g, 0, 2});
                    rb[tf] += 8.0 * own * m * 3;
                }
            }
            wv.fwd_base[di] = tasks.size();
            wv.fwd_off[di].assign(1, 0);
            for (int t = 0; t < steps_f[di]; ++t) {
                tasks.insert(tasks.end(), fs[t].begin(), fs[t].end());
                wv.fwd_off[di].push_back(tasks.size() - wv.fwd_base[di]);
            }
            wv.rev_base[di] = tasks.size();
            wv.rev_off[di].assign(1, 0);
            for (int t = 0; t < steps_r[di]; ++t) {
                tasks.insert(tasks.end(), rs[t].begin(), rs[t].end());
                wv.rev_off[di].push_back(tasks.size() - wv.rev_base[di]);
            }
            wv.fwd_bytes[di] = fb;
            wv.fwd_flops[di] = ff;
            wv.rev_bytes[di] = rb;
            wv.rev_flops[di] = rf;
        }
        p->gw.push_back(std::move(wv));
    }
    return MAGNUS_OK;
}

// device pointers of every pipeline instance, from the workspace offsets
void gw_bind(magnus_uneven_plan* p, std::vector<magnus::GwInst>& inst, const std::vector<size_t>& slot_ws) {
    char* ws = p->gw_ws;
    for (auto& wv : p->gw) {
        size_t slot = wv.inst_off;
        for (auto& grp : wv.groups) {
            const Plan& pl = grp.pl;
            const size_t m = pl.m, n = pl.n, K = n - 1;
            for (size_t b = 0; b < grp.ids.size(); ++b, ++slot) {
                magnus::GwInst& gi = inst[slot];
                const size_t S = (size_t)gi.S;
                size_t at = slot_ws[slot];
                gi.W = reinterpret_cast<double*>(ws + at);
                at += scratch_round(K * S * m * 8);
                gi.fst = reinterpret_cast<double*>(ws + at);
                at += scratch_round(K * 2 * m * 8);
                gi.U = reinterpret_cast<double*>(ws + at);
                at += scratch_round(K * 2 * (p->gw_L + 1) * m * 8);
                gi.rst = reinterpret_cast<double*>(ws + at);
                gi.sums = reinterpret_cast<double*>(ws + grp.z) + b * n * m;
                gi.z0 = reinterpret_cast<double*>(ws + grp.z0) + b * m;
                gi.totals = reinterpret_cast<double*>(ws + grp.totals) + b * n * m;
                gi.zbar = reinterpret_cast<double*>(ws + grp.zbar) + b * n * m;
                gi.wbar = reinterpret_cast<double*>(ws + grp.wbar) + b * n * m;
            }
        }
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
    // grouped pipeline: waves of instances within the budget, groups, tasks
    std::vector<magnus::GwInst> gw_inst;
    std::vector<magnus::GwTask> gw_tasks;
    std::vector<size_t> gw_slot_ws;  // per GwInst: workspace offset of its own arrays
    if (int rc = gw_layout(p, ids_total, gw_inst, gw_tasks, gw_slot_ws)) return rc;
    const size_t ids_bytes = scratch_round(std::max<size_t>(ids_total, 1) * sizeof(uint32_t));
    const size_t inst_bytes = scratch_round(std::max<size_t>(gw_inst.size(), 1) * sizeof(magnus::GwInst));
    const size_t task_bytes = scratch_round(std::max<size_t>(gw_tasks.size(), 1) * sizeof(magnus::GwTask));
    size_t gw_ws_bytes = 0;
    for (auto& wv : p->gw) gw_ws_bytes = std::max(gw_ws_bytes, wv.bytes);
    size_t total = tab_bytes + ids_bytes + inst_bytes + task_bytes + scratch_round(gw_ws_bytes);
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
    for (auto& wv : p->gw)
        for (auto& grp : wv.groups) ids_host.insert(ids_host.end(), grp.ids.begin(), grp.ids.end());
    if (ids_total && hipMemcpy(p->d_ids, ids_host.data(), ids_total * sizeof(uint32_t), hipMemcpyHostToDevice) !=
                         hipSuccess)
        return device_error(hipGetLastError(), "uneven plan: slot upload");
    p->d_inst = reinterpret_cast<magnus::GwInst*>(p->dmem + tab_bytes + ids_bytes);
    p->d_tasks = reinterpret_cast<magnus::GwTask*>(p->dmem + tab_bytes + ids_bytes + inst_bytes);
    p->gw_ws = p->dmem + tab_bytes + ids_bytes + inst_bytes + task_bytes;
    if (!gw_inst.empty()) {
        gw_bind(p, gw_inst, gw_slot_ws);
        if (hipMemcpy(p->d_inst, gw_inst.data(), gw_inst.size() * sizeof(magnus::GwInst), hipMemcpyHostToDevice) !=
                hipSuccess ||
            hipMemcpy(p->d_tasks, gw_tasks.data(), gw_tasks.size() * sizeof(magnus::GwTask), hipMemcpyHostToDevice) !=
                hipSuccess)
            return device_error(hipGetLastError(), "uneven plan: task upload");
    }
    p->arenas.assign(p->branches, Arena{});
    char* at = p->gw_ws + scratch_round(gw_ws_bytes);
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

// the grouped pipeline, wave by wave; each dim's tasks, folds and copies on
// that dim's stream (instances of different dims share nothing), the dims
// joined between waves (a wave reuses the workspace of the one before)
int enqueue_gw(magnus_uneven_plan* p, bool vjp) {
    for (size_t wi = 0; wi < p->gw.size(); ++wi) {
        GwWave& wv = p->gw[wi];
        if (wi > 0) {
            for (int di = 0; di < kGwDims; ++di) {
                hipError_t e = hipEventRecord(p->gw_join[di], p->gw_stream[di]);
                if (e != hipSuccess) return device_error(e, "uneven pipeline join");
            }
            for (int di = 0; di < kGwDims; ++di)
                for (int dj = 0; dj < kGwDims; ++dj) {
                    if (di == dj) continue;
                    hipError_t e = hipStreamWaitEvent(p->gw_stream[di], p->gw_join[dj], 0);
                    if (e != hipSuccess) return device_error(e, "uneven pipeline join");
                }
        }
        magnus::GwArgs ga{};
        ga.inst = p->d_inst + wv.inst_off;
        ga.tab = p->d_tab;
        ga.count = (long long)p->count;
        ga.L = p->gw_L;
        hipError_t e = hipSuccess;
        for (int di = 0; di < kGwDims; ++di) {
            hipStream_t s = p->gw_stream[di];
            if (vjp) {
                for (auto& grp : wv.groups) {
                    if ((int)grp.pl.d != kGwDim[di]) continue;
                    magnus::SegmentArgs g{};
                    g.ids = p->d_ids + grp.ids_off;
                    g.count = (int)grp.ids.size();
                    g.tab = p->d_tab + p->count;
                    g.packed = reinterpret_cast<double*>(p->gw_ws + grp.cot);
                    g.per = (long long)(grp.pl.n_out * grp.pl.m_cot);
                    e = magnus::launch_gather(g, s);
                    if (e != hipSuccess) return device_error(e, "uneven pipeline gather");
                }
            }
            for (size_t t = 0; t + 1 < wv.fwd_off[di].size(); ++t) {
                ga.tasks = p->d_tasks + wv.fwd_base[di] + wv.fwd_off[di][t];
                const int cnt = (int)(wv.fwd_off[di][t + 1] - wv.fwd_off[di][t]);
                e = magnus::launch_gw_forward(kGwDim[di], ga, cnt, wv.fwd_bytes[di][t], wv.fwd_flops[di][t], s);
                if (e != hipSuccess) return device_error(e, "uneven pipeline forward");
            }
            for (auto& grp : wv.groups) {
                if ((int)grp.pl.d != kGwDim[di]) continue;
                magnus::WbdArgs f{};
                fill_wbd_fold(grp.pl, (long long)grp.ids.size(), f);
                f.z = reinterpret_cast<double*>(p->gw_ws + grp.z);
                f.z0 = reinterpret_cast<double*>(p->gw_ws + grp.z0);
                f.totals = reinterpret_cast<double*>(p->gw_ws + grp.totals);
                f.part = reinterpret_cast<double*>(p->gw_ws + grp.part);
                f.work = reinterpret_cast<double*>(p->gw_ws + grp.work);
                if (vjp) {
                    f.cot = reinterpret_cast<double*>(p->gw_ws + grp.cot);
                    f.zbar = reinterpret_cast<double*>(p->gw_ws + grp.zbar);
                    f.wbar = reinterpret_cast<double*>(p->gw_ws + grp.wbar);
                    f.part2 = reinterpret_cast<double*>(p->gw_ws + grp.part2);
                    e = magnus::launch_wbd_fold_vjp(f, s);
                    if (e != hipSuccess) return device_error(e, "uneven pipeline fold");
                } else {
                    f.out = reinterpret_cast<double*>(p->gw_ws + grp.out);
                    e = magnus::launch_wbd_fold(f, s);
                    if (e != hipSuccess) return device_error(e, "uneven pipeline fold");
                    magnus::SegmentArgs g{};
                    g.ids = p->d_ids + grp.ids_off;
                    g.count = (int)grp.ids.size();
                    g.tab = p->d_tab + p->count;
                    g.packed = f.out;
                    g.per = (long long)(grp.pl.n_out * grp.pl.m);
                    e = magnus::launch_scatter(g, s);
                    if (e != hipSuccess) return device_error(e, "uneven pipeline scatter");
                }
            }
            if (!vjp) continue;
            for (size_t t = 0; t + 1 < wv.rev_off[di].size(); ++t) {
                ga.tasks = p->d_tasks + wv.rev_base[di] + wv.rev_off[di][t];
                const int cnt = (int)(wv.rev_off[di][t + 1] - wv.rev_off[di][t]);
                e = magnus::launch_gw_reverse(kGwDim[di], ga, cnt, wv.rev_bytes[di][t], wv.rev_flops[di][t], s);
                if (e != hipSuccess) return device_error(e, "uneven pipeline reverse");
            }
        }
    }
    return MAGNUS_OK;
}

// fork from `origin` onto the branch streams and the pipeline's stream, every
// sub-batch and pipeline step, join back
int enqueue_all(magnus_uneven_plan* p, bool vjp, hipStream_t origin) {
    hipError_t e = hipEventRecord(p->fork, origin);
    if (e != hipSuccess) return device_error(e, "uneven fork");
    if (!p->gw.empty()) {
        for (int di = 0; di < kGwDims && e == hipSuccess; ++di) e = hipStreamWaitEvent(p->gw_stream[di], p->fork, 0);
        if (e != hipSuccess) return device_error(e, "uneven fork");
        if (int rc = enqueue_gw(p, vjp)) return rc;
        for (int di = 0; di < kGwDims && e == hipSuccess; ++di) {
            e = hipEventRecord(p->gw_join[di], p->gw_stream[di]);
            if (e == hipSuccess) e = hipStreamWaitEvent(origin, p->gw_join[di], 0);
        }
        if (e != hipSuccess) return device_error(e, "uneven join");
    }
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
    for (int di = 0; di < kGwDims; ++di) {
        if (p->gw_stream[di]) (void)hipStreamDestroy(p->gw_stream[di]);
        if (p->gw_join[di]) (void)hipEventDestroy(p->gw_join[di]);
    }
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
    // The pipeline runs in one wave (its steps are efficient only when they
    // hold many tasks): eligible instances join it in order of their chain of
    // steps (about S + (n - 1) L samples, each costing more at a larger dim),
    // shortest first, while their full levels fit in the workspace left after
    // the branches' reserve (an eighth, and at least one instance of the largest
    // remaining bucket per branch); the others, the longest chains, run in
    // buckets on the level kernels, parallel over samples.
    p->gw_plans.assign(count, Plan{});
    std::vector<uint32_t> cand;
    for (uint64_t i = 0; i < count; ++i)
        if (gw_eligible(plans[i])) cand.push_back((uint32_t)i);
    auto chain = [&](uint32_t i) {
        const Plan& pl = plans[i];
        return (double)(pl.S + (pl.n - 1) * kGwL) * (double)pl.d * pl.d * pl.d;
    };
    std::stable_sort(cand.begin(), cand.end(), [&](uint32_t a, uint32_t b) { return chain(a) < chain(b); });
    std::vector<char> in_gw(count, 0);
    {
        size_t noncand_max = 0;
        bool noncand = false;
        for (uint64_t k = 0; k < count; ++k)
            if (!gw_eligible(plans[k])) {
                noncand = true;
                noncand_max = std::max(noncand_max, item_bytes(plans[k], 1));
            }
        std::vector<size_t> suffix(cand.size() + 1, 0);  // largest bucket instance among cand[c..]
        for (size_t c = cand.size(); c-- > 0;) suffix[c] = std::max(suffix[c + 1], item_bytes(plans[cand[c]], 1));
        size_t used = 0;
        for (size_t c = 0; c < cand.size(); ++c) {
            const bool left = noncand || c + 1 < cand.size();  // instances left for the branches
            const size_t res = left ? std::max(workspace_bytes / 8,
                                               std::max(noncand_max, suffix[c + 1]) * (size_t)branches)
                                    : 0;
            const size_t b = gw_inst_bytes(plans[cand[c]], kGwLMax);
            if (workspace_bytes < res || used + b > workspace_bytes - res) break;
            used += b;
            in_gw[cand[c]] = 1;
        }
        p->gw_budget = used;
    }
    bool others = false;
    for (uint64_t i = 0; i < count; ++i) {
        if (in_gw[i]) {
            p->gw_plans[i] = plans[i];
            p->gw_ids.push_back((uint32_t)i);
        } else {
            others = true;
        }
    }
    if (!others) p->gw_budget = workspace_bytes;
    const size_t per_branch = (workspace_bytes - p->gw_budget) / branches;
    std::map<std::tuple<uint64_t, uint64_t, uint64_t, int, int, int, uint64_t>, int> keys;
    for (uint64_t i = 0; i < count; ++i) {
        if (in_gw[i]) continue;
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
            const std::string msg = "uneven plan: one instance of dim " + std::to_string(bk.pl.d) + ", " +
                                    std::to_string(bk.pl.S) + " samples needs " +
                                    std::to_string(item_bytes(bk.pl, 1)) + " workspace bytes per branch";
            destroy(p);
            return fail(MAGNUS_ENOMEM, msg);
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
    for (int di = 0; di < kGwDims && e == hipSuccess; ++di) {
        e = hipStreamCreateWithFlags(&p->gw_stream[di], hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&p->gw_join[di], hipEventDisableTiming);
    }
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
    // the pipeline's segment length: the whole pipeline's VJP under each candidate
    if (rc == MAGNUS_OK && !p->gw_ids.empty()) {
        auto once = [&]() -> int {
            hipError_t e = hipEventRecord(p->fork, s);
            for (int di = 0; di < kGwDims && e == hipSuccess; ++di) e = hipStreamWaitEvent(p->gw_stream[di], p->fork, 0);
            if (e != hipSuccess) return device_error(e, "uneven autotune");
            if (int r = enqueue_gw(p, true)) return r;
            for (int di = 0; di < kGwDims && e == hipSuccess; ++di) {
                e = hipEventRecord(p->gw_join[di], p->gw_stream[di]);
                if (e == hipSuccess) e = hipStreamWaitEvent(s, p->gw_join[di], 0);
            }
            return e == hipSuccess ? MAGNUS_OK : device_error(e, "uneven autotune");
        };
        double best = -1.0;
        int best_L = p->gw_L;
        for (int L : kGwLCandidates) {
            p->gw_L = L;
            if ((rc = prepare(p))) break;  // reallocates the tables: upload them again
            if ((rc = upload_tables(p, in, cot, const_cast<const double* const*>(dA), s))) break;
            if ((rc = once())) break;  // warm-up
            if (hipEventRecord(t0, s) != hipSuccess) {
                rc = device_error(hipGetLastError(), "event");
                break;
            }
            for (int r = 0; r < 2 && !rc; ++r) rc = once();
            if (rc) break;
            float ms = 0.f;
            if (hipEventRecord(t1, s) == hipSuccess && hipEventSynchronize(t1) == hipSuccess &&
                hipEventElapsedTime(&ms, t0, t1) == hipSuccess) {
                const double per = ms / 2.0;
                if (best < 0 || per < best) {
                    best = per;
                    best_L = L;
                }
            }
        }
        p->gw_L = best_L;
        if (best >= 0) p->gw_ms = best;
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
    if (!p->gw_ids.empty()) {
        char line[128];
        std::snprintf(line, sizeof line, "pipeline %d %.6g\n", p->gw_L, p->gw_ms);
        os << line;
    }
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
        int L = 0;
        double pms = -1.0;
        const int pg = std::sscanf(line.c_str(), "pipeline %d %lf", &L, &pms);
        if (pg >= 1) {
            if (L < 4 || L > kGwLMax || L % 4) return fail(MAGNUS_EINVAL, "pipeline segment length must be a multiple of 4 in 4..256");
            if (!p->gw_ids.empty()) {
                p->gw_L = L;
                if (pg == 2) p->gw_ms = pms;
                changed = true;
            }
            continue;
        }
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

int magnus_uneven_in_pipeline(const magnus_uneven_plan* p, uint64_t i) {
    if (!p || i >= p->count || p->gw_plans.empty()) return 0;
    return p->gw_plans[i].n ? 1 : 0;
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
        case 4: return p->gw_ids.size();
        case 5: return p->gw.size();
        default: return 0;
    }
}

}  // extern "C"
