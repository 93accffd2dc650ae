// Serving: a request loop over device-resident graphs with continuous
// batching.  There is no counterpart in the reference (one graph per process,
// src/scc_main.cpp:7-57); each request's labels are exactly those of a single
// solve of its graph.
//
// * Graphs are registered once and stay in device memory until removed.
//   Registration is admitted against an HBM budget that also covers the
//   largest batch's working state.
// * A request names a graph.  The worker thread takes every pending request
//   that fits the batch limits and the budget, solves them as one union (one
//   cooperative launch), and starts again with whatever arrived meanwhile:
//   requests that arrive while a batch runs join the next one.
// * Labels come back through two device/pinned slots: the copy of batch k's
//   labels (second stream) overlaps batch k+1's solve; a completion thread
//   hands them to the waiting requests.
#include <string.h>

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/ispan_hip.h"
#include "internal.h"

namespace {

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

struct Graph {
    int32_t V = 0;
    int64_t E = 0;
    int32_t* d[4] = {nullptr, nullptr, nullptr, nullptr};
    int64_t bytes = 0;
    int pending = 0;  // requests not yet finished
    bool removed = false;
};

struct Request {
    int64_t graph = 0;
    int32_t V = 0;
    double t_submit = 0, t_start = 0, t_done = 0;
    bool done = false;
    int status = ISPAN_OK;
    int64_t batch = -1;
    int32_t batch_graphs = 0;
    std::vector<int32_t> labels;
};

// One batch in flight between the worker and the completion thread.
struct Flight {
    std::vector<int64_t> reqs;
    std::vector<int64_t> v_off;
    int slot = 0;
    int status = ISPAN_OK;
    hipEvent_t copied = nullptr;
    int64_t id = 0;
};

}  // namespace

struct ispan_server {
    ispan_ctx* ctx = nullptr;
    int device = 0;
    ispan_params params{};
    int64_t budget = 0;
    int32_t max_graphs = 0;
    int64_t max_edges = 0;
    hipStream_t compute_stream = nullptr;  // the solver context's stream
    hipStream_t copy_stream = nullptr;

    std::mutex mu;
    std::condition_variable cv_work, cv_done, cv_flight, cv_slot;
    std::unordered_map<int64_t, Graph> graphs;
    std::unordered_map<int64_t, Request> reqs;
    std::deque<int64_t> pending;
    std::deque<Flight> flights;
    int64_t next_graph = 1, next_req = 1, next_batch = 0;
    int64_t bytes_graphs = 0, bytes_ws = 0;
    int64_t ws_need = 0;  // working-state bytes the largest admitted graph needs alone
    bool stop = false;
    bool slot_busy[2] = {false, false};
    int32_t* d_labels[2] = {nullptr, nullptr};
    int32_t* h_labels[2] = {nullptr, nullptr};  // pinned
    int64_t cap_labels[2] = {0, 0};
    std::vector<double> latencies;
    int64_t batches = 0, batch_graphs_total = 0;
    std::string err;
    std::thread worker, completer;
};

namespace {

void free_graph_mem(Graph& g) {
    for (auto& p : g.d)
        if (p) {
            (void)hipFree(p);
            p = nullptr;
        }
}

// Fills a batch from the front of the queue: FIFO, bounded by the graph and
// edge limits, the int32 union and the working-state budget.
void take_batch(ispan_server* S, std::vector<int64_t>& out) {
    int64_t V = 0, E = 0;
    while (!S->pending.empty()) {
        const int64_t rid = S->pending.front();
        const Request& r = S->reqs[rid];
        const Graph& g = S->graphs[r.graph];
        const int64_t V2 = V + g.V, E2 = E + g.E;
        const bool first = out.empty();
        if (!first) {
            if (S->max_graphs > 0 && (int32_t)out.size() >= S->max_graphs) break;
            if (E2 > S->max_edges || V2 > INT32_MAX - 64) break;
            if (S->bytes_graphs + ispan::batch_workspace_bytes(V2, E2) > S->budget) break;
        }
        out.push_back(rid);
        V = V2;
        E = E2;
        S->pending.pop_front();
    }
}

int ensure_slot(ispan_server* S, int slot, int64_t n) {
    if (n <= S->cap_labels[slot]) return ISPAN_OK;
    if (S->d_labels[slot]) (void)hipFree(S->d_labels[slot]);
    if (S->h_labels[slot]) (void)hipHostFree(S->h_labels[slot]);
    S->d_labels[slot] = nullptr;
    S->h_labels[slot] = nullptr;
    S->cap_labels[slot] = 0;
    const int64_t cap = std::max<int64_t>(n, 1 << 20);
    if (hipMalloc(&S->d_labels[slot], (size_t)cap * 4) != hipSuccess) return ISPAN_ERR_NOMEM;
    if (hipHostMalloc(&S->h_labels[slot], (size_t)cap * 4, hipHostMallocDefault) != hipSuccess) return ISPAN_ERR_NOMEM;
    S->cap_labels[slot] = cap;
    return ISPAN_OK;
}

void worker_loop(ispan_server* S) {
    (void)hipSetDevice(S->device);
    int slot = 0;
    for (;;) {
        std::vector<int64_t> batch;
        std::vector<int64_t> v_off(1, 0), e_off(1, 0);
        std::vector<const int32_t*> t[4];
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_work.wait(lk, [S] { return S->stop || !S->pending.empty(); });
            if (S->stop && S->pending.empty()) break;
            take_batch(S, batch);
            const double t0 = now_ms();
            for (int64_t rid : batch) {
                Request& r = S->reqs[rid];
                const Graph& g = S->graphs[r.graph];
                r.t_start = t0;
                v_off.push_back(v_off.back() + g.V);
                e_off.push_back(e_off.back() + g.E);
                for (int k = 0; k < 4; ++k) t[k].push_back(g.d[k]);
            }
            // a label slot is reused only after its previous copy has been consumed
            S->cv_slot.wait(lk, [S, slot] { return !S->slot_busy[slot]; });
            S->slot_busy[slot] = true;
        }
        Flight f;
        f.reqs = batch;
        f.v_off = v_off;
        f.slot = slot;
        int rc = ensure_slot(S, slot, v_off.back());
        ispan_stats st{};
        if (!rc)
            rc = ispan::batch_solve_tables(S->ctx, (int32_t)batch.size(), v_off.data(), e_off.data(), t[0].data(),
                                           t[1].data(), t[2].data(), t[3].data(), &S->params, S->d_labels[slot], &st);
        if (!rc) {
            // labels to pinned memory on the copy stream, after the solve's
            // last kernel on the compute stream; the next batch's solve runs
            // meanwhile
            hipEvent_t solved = nullptr;
            if (hipEventCreateWithFlags(&solved, hipEventDisableTiming) != hipSuccess ||
                hipEventRecord(solved, S->compute_stream) != hipSuccess ||
                hipStreamWaitEvent(S->copy_stream, solved, 0) != hipSuccess ||
                hipEventCreateWithFlags(&f.copied, hipEventDisableTiming) != hipSuccess)
                rc = ISPAN_ERR_HIP;
            if (!rc && v_off.back() > 0 &&
                hipMemcpyAsync(S->h_labels[slot], S->d_labels[slot], (size_t)v_off.back() * 4,
                               hipMemcpyDeviceToHost, S->copy_stream) != hipSuccess)
                rc = ISPAN_ERR_HIP;
            if (!rc && hipEventRecord(f.copied, S->copy_stream) != hipSuccess) rc = ISPAN_ERR_HIP;
            if (solved) (void)hipEventDestroy(solved);
        }
        f.status = rc;
        {
            std::lock_guard<std::mutex> lk(S->mu);
            if (rc) S->err = ispan_ctx_last_error(S->ctx);
            f.id = S->next_batch++;
            S->flights.push_back(std::move(f));
        }
        S->cv_flight.notify_one();
        slot ^= 1;
    }
    {
        std::lock_guard<std::mutex> lk(S->mu);
        Flight end;
        end.id = -1;
        S->flights.push_back(std::move(end));
    }
    S->cv_flight.notify_one();
}

void completer_loop(ispan_server* S) {
    (void)hipSetDevice(S->device);
    for (;;) {
        Flight f;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_flight.wait(lk, [S] { return !S->flights.empty(); });
            f = std::move(S->flights.front());
            S->flights.pop_front();
        }
        if (f.id < 0) break;
        if (f.copied) {
            if (hipEventSynchronize(f.copied) != hipSuccess && f.status == ISPAN_OK) f.status = ISPAN_ERR_HIP;
            (void)hipEventDestroy(f.copied);
        }
        const double t = now_ms();
        {
            std::lock_guard<std::mutex> lk(S->mu);
            for (size_t i = 0; i < f.reqs.size(); ++i) {
                Request& r = S->reqs[f.reqs[i]];
                if (f.status == ISPAN_OK) {
                    r.labels.assign(S->h_labels[f.slot] + f.v_off[i], S->h_labels[f.slot] + f.v_off[i + 1]);
                }
                r.status = f.status;
                r.t_done = t;
                r.batch = f.id;
                r.batch_graphs = (int32_t)f.reqs.size();
                r.done = true;
                S->latencies.push_back(t - r.t_submit);
                auto it = S->graphs.find(r.graph);
                if (it != S->graphs.end() && --it->second.pending == 0 && it->second.removed) {
                    S->bytes_graphs -= it->second.bytes;
                    free_graph_mem(it->second);
                    S->graphs.erase(it);
                }
            }
            S->batches += 1;
            S->batch_graphs_total += (int64_t)f.reqs.size();
            S->slot_busy[f.slot] = false;
        }
        S->cv_slot.notify_all();
        S->cv_done.notify_all();
    }
}

int add_graph_impl(ispan_server* S, int32_t V, int64_t E, const int32_t* const* src, bool device, int64_t* id) {
    if (!S || V < 0 || E < 0 || E > INT32_MAX || !id || !src[0] || !src[2] || (E > 0 && (!src[1] || !src[3])))
        return ISPAN_ERR_ARG;
    if (!device) {
        if (src[0][0] != 0 || src[2][0] != 0 || (int64_t)src[0][V] != E || (int64_t)src[2][V] != E)
            return ISPAN_ERR_CSR;
        // the host arrays are checked in full before they are admitted
        if (ispan_csr_check(V, E, src[0], src[1], src[2], src[3]) != ISPAN_OK) return ISPAN_ERR_CSR;
    }
    const int64_t bytes = 2 * ((int64_t)V + 1) * 4 + 2 * std::max<int64_t>(E, 1) * 4;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        const int64_t need = std::max(S->ws_need, ispan::batch_workspace_bytes(V, E));
        if (S->bytes_graphs + bytes + need > S->budget) return ISPAN_ERR_BUDGET;
        S->bytes_graphs += bytes;  // reserved while copying
        S->ws_need = need;
    }
    Graph g;
    g.V = V;
    g.E = E;
    g.bytes = bytes;
    const size_t sz[4] = {((size_t)V + 1) * 4, (size_t)std::max<int64_t>(E, 1) * 4, ((size_t)V + 1) * 4,
                          (size_t)std::max<int64_t>(E, 1) * 4};
    const size_t n[4] = {((size_t)V + 1) * 4, (size_t)E * 4, ((size_t)V + 1) * 4, (size_t)E * 4};
    int rc = ISPAN_OK;
    hipStream_t s = S->copy_stream;
    for (int k = 0; k < 4 && !rc; ++k) {
        if (hipMalloc(&g.d[k], sz[k]) != hipSuccess) rc = ISPAN_ERR_NOMEM;
        else if (n[k] && hipMemcpyAsync(g.d[k], src[k], n[k], device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                                        s) != hipSuccess)
            rc = ISPAN_ERR_HIP;
    }
    if (!rc && hipStreamSynchronize(s) != hipSuccess) rc = ISPAN_ERR_HIP;
    if (!rc && device) rc = ispan::check_csr_device(V, E, g.d[0], g.d[1], g.d[2], g.d[3], s);  // the copy, on the device
    std::lock_guard<std::mutex> lk(S->mu);
    if (rc) {
        free_graph_mem(g);
        S->bytes_graphs -= bytes;
        return rc;
    }
    *id = S->next_graph++;
    S->graphs.emplace(*id, std::move(g));
    return ISPAN_OK;
}

double percentile(std::vector<double> v, double q) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double pos = q * (double)(v.size() - 1);
    const size_t i = (size_t)pos;
    const double f = pos - (double)i;
    return i + 1 < v.size() ? v[i] * (1 - f) + v[i + 1] * f : v[i];
}

}  // namespace

extern "C" {

int ispan_server_create(ispan_server** out, int device, const ispan_server_config* cfg, const ispan_params* params) {
    if (!out) return ISPAN_ERR_ARG;
    *out = nullptr;
    auto* S = new ispan_server();
    S->device = device;
    ispan_default_params(&S->params);
    if (params) S->params = *params;
    S->budget = cfg && cfg->hbm_budget > 0 ? cfg->hbm_budget : (int64_t)200 << 30;
    S->max_graphs = cfg ? std::max(cfg->max_batch_graphs, 0) : 0;
    S->max_edges = cfg && cfg->max_batch_edges > 0 ? std::min<int64_t>(cfg->max_batch_edges, INT32_MAX) : INT32_MAX;
    int rc = ispan_ctx_create(&S->ctx, device);
    if (!rc && (hipStreamCreateWithFlags(&S->compute_stream, hipStreamNonBlocking) != hipSuccess ||
                hipStreamCreateWithFlags(&S->copy_stream, hipStreamNonBlocking) != hipSuccess))
        rc = ISPAN_ERR_HIP;
    if (!rc) rc = ispan_ctx_set_stream(S->ctx, S->compute_stream);
    if (rc) {
        if (S->ctx) ispan_ctx_destroy(S->ctx);
        if (S->copy_stream) (void)hipStreamDestroy(S->copy_stream);
        if (S->compute_stream) (void)hipStreamDestroy(S->compute_stream);
        delete S;
        return rc;
    }
    S->worker = std::thread(worker_loop, S);
    S->completer = std::thread(completer_loop, S);
    *out = S;
    return ISPAN_OK;
}

int ispan_server_destroy(ispan_server* S) {
    if (!S) return ISPAN_ERR_ARG;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    if (S->worker.joinable()) S->worker.join();
    if (S->completer.joinable()) S->completer.join();
    (void)hipSetDevice(S->device);
    for (auto& kv : S->graphs) free_graph_mem(kv.second);
    for (int k = 0; k < 2; ++k) {
        if (S->d_labels[k]) (void)hipFree(S->d_labels[k]);
        if (S->h_labels[k]) (void)hipHostFree(S->h_labels[k]);
    }
    ispan_ctx_destroy(S->ctx);
    if (S->copy_stream) (void)hipStreamDestroy(S->copy_stream);
    if (S->compute_stream) (void)hipStreamDestroy(S->compute_stream);
    delete S;
    return ISPAN_OK;
}

int ispan_server_add_graph(ispan_server* S, int32_t V, int64_t E, const int32_t* fw_beg, const int32_t* fw_adj,
                           const int32_t* bw_beg, const int32_t* bw_adj, int64_t* graph_id) {
    const int32_t* src[4] = {fw_beg, fw_adj, bw_beg, bw_adj};
    if (!S) return ISPAN_ERR_ARG;
    (void)hipSetDevice(S->device);
    return add_graph_impl(S, V, E, src, false, graph_id);
}

int ispan_server_add_graph_device(ispan_server* S, int32_t V, int64_t E, const int32_t* d_fw_beg,
                                  const int32_t* d_fw_adj, const int32_t* d_bw_beg, const int32_t* d_bw_adj,
                                  int64_t* graph_id) {
    const int32_t* src[4] = {d_fw_beg, d_fw_adj, d_bw_beg, d_bw_adj};
    if (!S) return ISPAN_ERR_ARG;
    (void)hipSetDevice(S->device);
    return add_graph_impl(S, V, E, src, true, graph_id);
}

int ispan_server_remove_graph(ispan_server* S, int64_t graph_id) {
    if (!S) return ISPAN_ERR_ARG;
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->graphs.find(graph_id);
    if (it == S->graphs.end() || it->second.removed) return ISPAN_ERR_NOTFOUND;
    it->second.removed = true;
    if (it->second.pending == 0) {
        S->bytes_graphs -= it->second.bytes;
        free_graph_mem(it->second);
        S->graphs.erase(it);
    }
    return ISPAN_OK;
}

int ispan_server_submit(ispan_server* S, int64_t graph_id, int64_t* request_id) {
    if (!S || !request_id) return ISPAN_ERR_ARG;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        auto it = S->graphs.find(graph_id);
        if (it == S->graphs.end() || it->second.removed) return ISPAN_ERR_NOTFOUND;
        it->second.pending += 1;
        const int64_t id = S->next_req++;
        Request& r = S->reqs[id];
        r.graph = graph_id;
        r.V = it->second.V;
        r.t_submit = now_ms();
        S->pending.push_back(id);
        *request_id = id;
    }
    S->cv_work.notify_one();
    return ISPAN_OK;
}

int ispan_server_wait(ispan_server* S, int64_t request_id, int32_t* labels, ispan_request_info* info,
                      int64_t timeout_ms) {
    if (!S) return ISPAN_ERR_ARG;
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->reqs.find(request_id);
    if (it == S->reqs.end()) return ISPAN_ERR_NOTFOUND;
    auto ready = [S, request_id] { return S->reqs[request_id].done; };
    if (timeout_ms < 0) S->cv_done.wait(lk, ready);
    else if (!S->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready)) return ISPAN_ERR_TIMEOUT;
    Request& r = S->reqs[request_id];
    if (info) {
        info->submit_ms = r.t_submit;
        info->start_ms = r.t_start;
        info->done_ms = r.t_done;
        info->latency_ms = r.t_done - r.t_submit;
        info->batch = r.batch;
        info->batch_graphs = r.batch_graphs;
        info->status = r.status;
    }
    const int status = r.status;
    if (status == ISPAN_OK && labels && !r.labels.empty())
        memcpy(labels, r.labels.data(), r.labels.size() * 4);
    S->reqs.erase(request_id);
    return status;
}

int ispan_server_stats_get(ispan_server* S, ispan_server_stats* out) {
    if (!S || !out) return ISPAN_ERR_ARG;
    std::lock_guard<std::mutex> lk(S->mu);
    memset(out, 0, sizeof(*out));
    out->requests_done = (int64_t)S->latencies.size();
    out->batches = S->batches;
    out->graphs_resident = (int64_t)S->graphs.size();
    out->bytes_resident = S->bytes_graphs;
    out->bytes_budget = S->budget;
    out->p50_ms = percentile(S->latencies, 0.50);
    out->p99_ms = percentile(S->latencies, 0.99);
    double sum = 0, mx = 0;
    for (double x : S->latencies) {
        sum += x;
        mx = std::max(mx, x);
    }
    out->mean_ms = S->latencies.empty() ? 0.0 : sum / (double)S->latencies.size();
    out->max_ms = mx;
    out->mean_batch_graphs = S->batches ? (double)S->batch_graphs_total / (double)S->batches : 0.0;
    return ISPAN_OK;
}

int ispan_server_reset_stats(ispan_server* S) {
    if (!S) return ISPAN_ERR_ARG;
    std::lock_guard<std::mutex> lk(S->mu);
    S->latencies.clear();
    S->batches = 0;
    S->batch_graphs_total = 0;
    return ISPAN_OK;
}

const char* ispan_server_last_error(const ispan_server* S) { return S ? S->err.c_str() : "null server"; }

}  // extern "C"
