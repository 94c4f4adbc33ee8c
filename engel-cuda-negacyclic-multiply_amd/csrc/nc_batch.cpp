// Batches of mixed ring degree as one HIP graph (include/negacyclic_batch.h).
//
// Each degree keeps one resident plan on the batch's one table set (the
// tables of every degree; each plan owns only its N-dependent constants).  A capture forks
// one stream per distinct degree from the capture stream, enqueues every
// group of that degree on it with the configuration its plan holds for the
// group's count, and joins them back: the graph has one branch per degree,
// so kernels of different degrees overlap on the device, and one graph
// launch replays the whole batch.  The graph owns its workspace, so later
// calls on the plans (which may regrow theirs) never move memory under it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <new>
#include <string>
#include <vector>

#include "../../include/negacyclic_batch.h"
#include "nc_internal.hpp"

struct nc_batch {
    uint32_t L = 0;
    std::vector<uint32_t> q;
    nc_tableset* tables = nullptr;  // one table set for every degree, made with the first plan
    std::map<uint32_t, nc_plan*> plans;
    std::map<uint32_t, hipStream_t> branch;  // capture branch per degree
    hipStream_t capture = nullptr;
};

struct nc_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<void*> workspace;
};

namespace {

using nc::set_error;

int hip_err(hipError_t e, const char* what) { return set_error(NC_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e)); }

int check_groups(const nc_group* groups, uint32_t n) {
    if (!groups || n == 0) return set_error(NC_ERR_INVALID, "no groups");
    for (uint32_t i = 0; i < n; ++i) {
        const nc_group& g = groups[i];
        if (!g.d_a || !g.d_b || !g.d_c) return set_error(NC_ERR_INVALID, "group " + std::to_string(i) + ": null buffer");
        if (g.count == 0) return set_error(NC_ERR_INVALID, "group " + std::to_string(i) + ": count is 0");
    }
    return NC_OK;
}

void destroy_graph(nc_graph* g) {
    if (!g) return;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    for (void* p : g->workspace) (void)hipFree(p);
    delete g;
}

}  // namespace

extern "C" {

int nc_batch_create(uint32_t L, const uint32_t* q, nc_batch** out) {
    if (!out) return set_error(NC_ERR_INVALID, "out is null");
    *out = nullptr;
    if (!q) return set_error(NC_ERR_INVALID, "q is null");
    if (L == 0 || L > 32) return set_error(NC_ERR_INVALID, "L must be in [1, 32]");
    nc_batch* b = new (std::nothrow) nc_batch();
    if (!b) return set_error(NC_ERR_NOMEM, "batch");
    b->L = L;
    b->q.assign(q, q + L);
    const hipError_t e = hipStreamCreateWithFlags(&b->capture, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete b;
        return hip_err(e, "capture stream");
    }
    *out = b;
    return NC_OK;
}

void nc_batch_destroy(nc_batch* b) {
    if (!b) return;
    for (auto& kv : b->branch) (void)hipStreamDestroy(kv.second);
    if (b->capture) (void)hipStreamDestroy(b->capture);
    for (auto& kv : b->plans) nc_plan_destroy(kv.second);
    nc_tableset_destroy(b->tables);
    delete b;
}

int nc_batch_plan(nc_batch* b, uint32_t N, nc_plan** out) {
    if (!b || !out) return set_error(NC_ERR_INVALID, "null argument");
    auto it = b->plans.find(N);
    if (it == b->plans.end()) {
        if (!b->tables)
            if (int rc = nc_tableset_create(b->L, b->q.data(), &b->tables)) return rc;
        nc_plan* p = nullptr;
        if (int rc = nc_plan_create_shared(b->tables, N, &p)) return rc;
        it = b->plans.emplace(N, p).first;
    }
    *out = it->second;
    return NC_OK;
}

int nc_batch_autotune(nc_batch* b, const nc_group* groups, uint32_t n, void* stream) {
    if (!b) return set_error(NC_ERR_INVALID, "null batch");
    if (int rc = check_groups(groups, n)) return rc;
    for (uint32_t i = 0; i < n; ++i) {
        nc_plan* p = nullptr;
        if (int rc = nc_batch_plan(b, groups[i].N, &p)) return rc;
        if (int rc = nc_plan_autotune(p, groups[i].count, stream, nullptr, nullptr)) return rc;
    }
    return NC_OK;
}

int nc_batch_run(nc_batch* b, const nc_group* groups, uint32_t n, void* stream) {
    if (!b) return set_error(NC_ERR_INVALID, "null batch");
    if (int rc = check_groups(groups, n)) return rc;
    for (uint32_t i = 0; i < n; ++i) {
        nc_plan* p = nullptr;
        if (int rc = nc_batch_plan(b, groups[i].N, &p)) return rc;
        if (int rc = nc_multiply_device(p, groups[i].d_a, groups[i].d_b, groups[i].d_c, groups[i].count, stream))
            return rc;
    }
    return NC_OK;
}

int nc_batch_capture(nc_batch* b, const nc_group* groups, uint32_t n, nc_graph** out) {
    if (!b || !out) return set_error(NC_ERR_INVALID, "null argument");
    *out = nullptr;
    if (int rc = check_groups(groups, n)) return rc;
    // everything that allocates or synchronises happens before the capture
    std::map<uint32_t, std::vector<uint32_t>> by_degree;  // degree -> group indices, in order
    std::vector<uint32_t> order;                          // degrees by first appearance
    for (uint32_t i = 0; i < n; ++i) {
        nc_plan* p = nullptr;
        if (int rc = nc_batch_plan(b, groups[i].N, &p)) return rc;
        if (int rc = nc::validate_batch(p, groups[i].count)) return rc;  // the graph owns its workspace
        if (!by_degree.count(groups[i].N)) order.push_back(groups[i].N);
        by_degree[groups[i].N].push_back(i);
    }
    nc_graph* g = new (std::nothrow) nc_graph();
    if (!g) return set_error(NC_ERR_NOMEM, "graph");
    struct Branch {
        hipStream_t s;
        void* ws_a;
        void* ws_b;
    };
    std::map<uint32_t, Branch> br;
    for (uint32_t N : order) {
        uint32_t most = 0;
        for (uint32_t i : by_degree[N]) most = std::max(most, groups[i].count);
        const size_t bytes = (size_t)most * nc::workspace_bytes_per_product(b->plans[N]);
        void* wa = nullptr;
        void* wb = nullptr;
        hipError_t e = hipMalloc(&wa, bytes);
        if (e == hipSuccess) g->workspace.push_back(wa);
        if (e == hipSuccess) e = hipMalloc(&wb, bytes);
        if (e == hipSuccess) g->workspace.push_back(wb);
        if (e == hipSuccess && !b->branch.count(N)) {
            hipStream_t s = nullptr;
            e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
            if (e == hipSuccess) b->branch[N] = s;
        }
        if (e != hipSuccess) {
            destroy_graph(g);
            return hip_err(e, "graph workspace");
        }
        br[N] = Branch{b->branch[N], wa, wb};
    }
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> joins(order.size(), nullptr);
    hipError_t e = hipEventCreateWithFlags(&fork, hipEventDisableTiming);
    for (size_t i = 0; i < joins.size() && e == hipSuccess; ++i) e = hipEventCreateWithFlags(&joins[i], hipEventDisableTiming);
    int rc = NC_OK;
    bool capturing = false;
    if (e == hipSuccess) e = hipStreamBeginCapture(b->capture, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        capturing = true;
        e = hipEventRecord(fork, b->capture);
    }
    for (size_t bi = 0; bi < order.size() && e == hipSuccess && rc == NC_OK; ++bi) {
        const uint32_t N = order[bi];
        const Branch& x = br[N];
        nc_plan* p = b->plans[N];
        e = hipStreamWaitEvent(x.s, fork, 0);
        for (uint32_t i : by_degree[N]) {
            if (e != hipSuccess || rc != NC_OK) break;
            const nc_group& gr = groups[i];
            rc = nc::run_pipeline_ws(p, x.ws_a, x.ws_b, gr.d_a, gr.d_b, gr.d_c, gr.count, x.s,
                                     nc::choose_for(p, gr.count));
        }
        if (e == hipSuccess && rc == NC_OK) e = hipEventRecord(joins[bi], x.s);
        if (e == hipSuccess && rc == NC_OK) e = hipStreamWaitEvent(b->capture, joins[bi], 0);
    }
    if (capturing) {
        hipGraph_t graph = nullptr;
        const hipError_t e2 = hipStreamEndCapture(b->capture, &graph);
        if (e == hipSuccess) e = e2;
        g->graph = graph;
    }
    if (e == hipSuccess && rc == NC_OK) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (fork) (void)hipEventDestroy(fork);
    for (hipEvent_t j : joins)
        if (j) (void)hipEventDestroy(j);
    if (rc == NC_OK && e != hipSuccess) rc = hip_err(e, "graph capture");
    if (rc != NC_OK) {
        destroy_graph(g);
        return rc;
    }
    *out = g;
    return NC_OK;
}

int nc_graph_launch(nc_graph* g, void* stream) {
    if (!g || !g->exec) return set_error(NC_ERR_INVALID, "null graph");
    const hipError_t e = hipGraphLaunch(g->exec, static_cast<hipStream_t>(stream));
    return e == hipSuccess ? NC_OK : hip_err(e, "graph launch");
}

void nc_graph_destroy(nc_graph* g) { destroy_graph(g); }

}  // extern "C"
