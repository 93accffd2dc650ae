// Serving: a request loop over device-resident graphs with continuous
// batching.  There is no counterpart in the reference (one graph per process,
// src/scc_main.cpp:7-57); each request's labels are exactly those of a single
// solve of its graph.
//
// * Registered graphs are appended once to an arena: one CSR pair in device
//   memory holding every resident graph side by side with global ids, with
//   its first-neighbour arrays.  Registration is admitted against an HBM
//   budget that covers the arena, the solve state sized for it and the
//   outstanding label buffers, and against the device's free memory at the
//   time (hipMemGetInfo) when the arena has to grow.
// * A request names a graph.  The worker thread takes every pending request
//   within the batch limits, solves the distinct graphs among them in one pass
//   over the arena (a selection bitmap keeps every other vertex inactive, so
//   nothing is copied per batch), and starts again with whatever arrived
//   meanwhile.  Requests that arrive while a batch runs join the next one:
//   a batch is one cooperative launch whose blocks fill every CU and pass
//   grid barriers level by level, so no second launch can run beside it, and
//   a graph admitted into the running launch could only start at the next
//   phase boundary (both traversals are seeded before their first level),
//   which is no earlier than the next batch in practice.
// * Labels stay in device memory (a pool of batch buffers) until the client
//   takes them: copied to the host by wait(), or read in place through
//   wait_device() and released.  Latency is submit -> labels ready on the
//   device.
#include <string.h>

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/ispan_hip.h"
#include "common.h"
#include "internal.h"

namespace {

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

struct Graph {
    int64_t vb = 0;  // first arena vertex
    int64_t eb = 0;  // first arena edge
    int32_t V = 0;
    int64_t E = 0;
    int pending = 0;  // requests not yet finished
    bool removed = false;
};

struct LabelBuf {
    int32_t* d = nullptr;
    int64_t cap = 0;  // int32 entries
    int refs = 0;     // requests still holding part of it
};

struct Request {
    int64_t graph = 0;
    int32_t V = 0;
    double t_submit = 0, t_start = 0, t_done = 0;
    bool done = false;
    int status = ISPAN_OK;
    int64_t batch = -1;
    int32_t batch_graphs = 0;
    int buf = -1;            // label buffer
    int32_t* d_labels = nullptr;
};

// One batch between the worker and the completion thread.
struct Flight {
    std::vector<int64_t> reqs;
    int buf = -1;
    int status = ISPAN_OK;
    hipEvent_t done = nullptr;
    int64_t id = 0;
};

}  // namespace

struct ispan_server {
    ispan_ctx* ctx = nullptr;
    int device = 0;
    ispan_params params{};
    bool tuned = false;               // batches take params by size class from `tuning`
    ispan_tuning_table tuning{};
    int64_t budget = 0;
    int32_t max_graphs = 0;
    int64_t max_vertices = 0;
    hipStream_t compute_stream = nullptr;  // the solver context's stream
    hipStream_t copy_stream = nullptr;     // registration

    std::mutex mu;
    std::condition_variable cv_work, cv_done, cv_flight, cv_mem, cv_arena;
    // arena (grown only while no batch is in flight)
    int32_t* csr[4] = {nullptr, nullptr, nullptr, nullptr};
    int32_t* first[2] = {nullptr, nullptr};
    int64_t capV = 0, capE = 0, usedV = 0, usedE = 0;
    int inflight = 0;     // batches whose kernels may still read the arena
    bool growing = false;
    std::map<int64_t, Graph> graphs;  // ordered by id = arena order
    std::unordered_map<int64_t, Request> reqs;
    std::deque<int64_t> pending;
    std::deque<Flight> flights;
    std::vector<LabelBuf> bufs;
    int64_t next_graph = 1, next_req = 1, next_batch = 0;
    int64_t bytes_live = 0;  // live graphs' arena share
    bool stop = false;
    std::vector<double> latencies;
    int64_t batches = 0, batch_graphs_total = 0;
    std::string err;
    std::thread worker, completer;
};

namespace {

// Device memory left free when the arena grows (runtime and other users).
constexpr int64_t kFreeReserve = (int64_t)256 << 20;

int64_t arena_bytes(int64_t V, int64_t E) { return 2 * (V + 1) * 4 + 2 * std::max<int64_t>(E, 1) * 4 + 2 * V * 4; }

int64_t label_bytes(const ispan_server* S) {
    int64_t b = 0;
    for (const auto& x : S->bufs) b += x.cap * 4;
    return b;
}

// Device bytes in use or reserved: arena capacity, solve state for it, label buffers.
int64_t bytes_used(const ispan_server* S) {
    return arena_bytes(S->capV, S->capE) + ispan::solve_state_bytes(S->capV, S->capE) + label_bytes(S);
}

// A label buffer of at least n entries: a free one, else a new one within the
// budget (-1: none fits right now).  Called with mu held.
int acquire_buf(ispan_server* S, int64_t n) {
    int best = -1;
    for (int i = 0; i < (int)S->bufs.size(); ++i)
        if (S->bufs[i].refs == 0 && S->bufs[i].cap >= n && (best < 0 || S->bufs[i].cap < S->bufs[best].cap)) best = i;
    if (best >= 0) return best;
    int64_t cap = std::max<int64_t>(n, 1 << 20);  // small batches share a reusable size
    if (bytes_used(S) + cap * 4 > S->budget) cap = n;
    if (bytes_used(S) + cap * 4 > S->budget) {
        // free idle buffers that are too small, then retry once
        for (auto& b : S->bufs)
            if (b.refs == 0 && b.d) {
                (void)hipFree(b.d);
                b.d = nullptr;
                b.cap = 0;
            }
        if (bytes_used(S) + cap * 4 > S->budget) return -1;
    }
    LabelBuf b;
    if (hipMalloc(&b.d, (size_t)cap * 4) != hipSuccess) return -1;
    b.cap = cap;
    for (int i = 0; i < (int)S->bufs.size(); ++i)
        if (!S->bufs[i].d) {
            S->bufs[i] = b;
            return i;
        }
    S->bufs.push_back(b);
    return (int)S->bufs.size() - 1;
}

void drop_ref(ispan_server* S, Request& r) {
    if (r.buf >= 0 && --S->bufs[r.buf].refs == 0) S->cv_mem.notify_all();
    r.buf = -1;
    r.d_labels = nullptr;
}

// Arena space of removed graphs at its end is reused.  Called with mu held.
void trim_arena(ispan_server* S) {
    while (!S->graphs.empty()) {
        auto it = std::prev(S->graphs.end());
        if (!it->second.removed || it->second.pending) break;
        S->usedV = it->second.vb;
        S->usedE = it->second.eb;
        S->graphs.erase(it);
    }
    if (S->graphs.empty()) S->usedV = S->usedE = 0;
}

// Fills a batch from the front of the queue, FIFO, within the graph and
// vertex limits and the label memory the budget leaves.
void take_batch(ispan_server* S, std::vector<int64_t>& out) {
    std::unordered_map<int64_t, int> seen;
    int64_t V = 0, busy = 0;
    for (const auto& b : S->bufs) busy += b.refs ? b.cap * 4 : 0;
    const int64_t room = S->budget - arena_bytes(S->capV, S->capE) - ispan::solve_state_bytes(S->capV, S->capE) - busy;
    while (!S->pending.empty()) {
        const int64_t rid = S->pending.front();
        const Request& r = S->reqs[rid];
        const bool is_new = !seen.count(r.graph);
        const int64_t V2 = V + (is_new ? r.V : 0);
        if (!out.empty() && is_new) {
            if (S->max_graphs > 0 && (int32_t)seen.size() >= S->max_graphs) break;
            if (S->max_vertices > 0 && V2 > S->max_vertices) break;
            if (4 * V2 > room) break;
        }
        if (is_new) seen[r.graph] = 1;
        out.push_back(rid);
        V = V2;
        S->pending.pop_front();
    }
}

void worker_loop(ispan_server* S) {
    (void)hipSetDevice(S->device);
    for (;;) {
        std::vector<int64_t> batch;
        std::vector<int64_t> vb, ve;
        std::vector<int32_t*> out;
        std::unordered_map<int64_t, int64_t> offset;  // graph -> offset in the label buffer
        Flight f;
        ispan::ArenaView av{};
        ispan_params bp{};
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_work.wait(lk, [S] { return S->stop || !S->pending.empty(); });
            if (S->stop && S->pending.empty()) break;
            take_batch(S, batch);
            int64_t n = 0;
            for (int64_t rid : batch) {
                const Request& r = S->reqs[rid];
                if (!offset.count(r.graph)) {
                    offset[r.graph] = n;
                    n += r.V;
                }
            }
            // a label buffer for the batch, once memory allows; if it cannot
            // fit even with every buffer released, the batch fails
            int b = -1;
            bool never = false;
            S->cv_mem.wait(lk, [&] {
                if ((b = acquire_buf(S, std::max<int64_t>(n, 1))) >= 0) return true;
                bool busy = false;
                for (const auto& x : S->bufs) busy |= x.refs > 0;
                never = !busy;
                return never || S->stop;
            });
            if (b < 0) {
                const double t = now_ms();
                for (int64_t rid : batch) {
                    Request& r = S->reqs[rid];
                    r.status = ISPAN_ERR_BUDGET;
                    r.t_start = r.t_done = t;
                    r.done = true;
                    auto it = S->graphs.find(r.graph);
                    if (it != S->graphs.end() && --it->second.pending == 0 && it->second.removed) trim_arena(S);
                }
                S->err = "the labels of a batch do not fit the HBM budget";
                S->cv_done.notify_all();
                continue;
            }
            S->cv_arena.wait(lk, [S] { return !S->growing; });
            S->inflight += 1;
            f.buf = b;
            LabelBuf& lb = S->bufs[b];
            lb.refs = (int)batch.size();
            const double t0 = now_ms();
            for (int64_t rid : batch) {
                Request& r = S->reqs[rid];
                r.t_start = t0;
                r.buf = b;
                r.d_labels = lb.d + offset[r.graph];
            }
            int64_t bigV = 0;
            for (const auto& kv : offset) {
                const Graph& g = S->graphs[kv.first];
                if (g.V == 0) continue;
                vb.push_back(g.vb);
                ve.push_back(g.vb + g.V);
                out.push_back(lb.d + kv.second);
                bigV = std::max<int64_t>(bigV, g.V);
            }
            // parameters of the batch: by the size class of its largest graph
            bp = S->params;
            if (S->tuned) ispan_tuning_lookup(&S->tuning, bigV, &bp);
            av = ispan::ArenaView{S->csr[0], S->csr[1], S->csr[2], S->csr[3], S->first[0], S->first[1],
                                  (int32_t)S->usedV, S->usedE};
        }
        f.reqs = batch;
        int rc = ISPAN_OK;
        if (!vb.empty()) {
            ispan_stats st{};
            rc = ispan::solve_arena(S->ctx, av, (int32_t)vb.size(), vb.data(), ve.data(), out.data(), &bp, &st);
        }
        if (!rc && (hipEventCreateWithFlags(&f.done, hipEventDisableTiming) != hipSuccess ||
                    hipEventRecord(f.done, S->compute_stream) != hipSuccess))
            rc = ISPAN_ERR_HIP;
        f.status = rc;
        {
            std::lock_guard<std::mutex> lk(S->mu);
            if (rc) S->err = ispan_ctx_last_error(S->ctx);
            f.id = S->next_batch++;
            S->flights.push_back(std::move(f));
        }
        S->cv_flight.notify_one();
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
        if (f.done) {
            if (hipEventSynchronize(f.done) != hipSuccess && f.status == ISPAN_OK) f.status = ISPAN_ERR_HIP;
            (void)hipEventDestroy(f.done);
        } else if (f.status == ISPAN_OK) {
            (void)hipStreamSynchronize(S->compute_stream);
        }
        const double t = now_ms();
        {
            std::lock_guard<std::mutex> lk(S->mu);
            for (int64_t rid : f.reqs) {
                Request& r = S->reqs[rid];
                r.status = f.status;
                r.t_done = t;
                r.batch = f.id;
                r.batch_graphs = (int32_t)f.reqs.size();
                r.done = true;
                if (f.status != ISPAN_OK) drop_ref(S, r);
                S->latencies.push_back(t - r.t_submit);
                auto it = S->graphs.find(r.graph);
                if (it != S->graphs.end() && --it->second.pending == 0 && it->second.removed) trim_arena(S);
            }
            S->inflight -= 1;
            S->batches += 1;
            S->batch_graphs_total += (int64_t)f.reqs.size();
        }
        S->cv_arena.notify_all();
        S->cv_done.notify_all();
    }
}

// Grows the arena to hold V more vertices and E more edges (called with the
// lock held and no batch in flight; the lock is dropped during the copy).
int grow_arena(ispan_server* S, std::unique_lock<std::mutex>& lk, int64_t needV, int64_t needE) {
    int64_t nV = std::max<int64_t>(needV, std::min<int64_t>(2 * S->capV, INT32_MAX - 64));
    int64_t nE = std::max<int64_t>(needE, std::min<int64_t>(2 * S->capE, INT32_MAX));
    // doubling only while the budget allows it
    // room is kept for the labels of one batch as large as the arena.  Besides
    // the configured budget, the growth must fit the device's free memory
    // now: the new arena is allocated while the old one is still alive, and
    // the solve state and the label room grow to the new size.
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return ISPAN_ERR_HIP;
    const int64_t cur = arena_bytes(S->capV, S->capE) + ispan::solve_state_bytes(S->capV, S->capE) +
                        4 * (S->capV + 1);
    auto fits = [&](int64_t v, int64_t e) {
        const int64_t need = arena_bytes(v, e) + ispan::solve_state_bytes(v, e) + 4 * (v + 1);
        const int64_t extra = need - cur + arena_bytes(S->capV, S->capE);
        return need <= S->budget && extra + kFreeReserve <= (int64_t)free_b;
    };
    if (!fits(nV, nE)) {
        nV = needV;
        nE = needE;
        if (!fits(nV, nE)) return ISPAN_ERR_BUDGET;
    }
    S->growing = true;
    lk.unlock();
    int32_t* n[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    const size_t sz[6] = {((size_t)nV + 1) * 4, (size_t)std::max<int64_t>(nE, 1) * 4, ((size_t)nV + 1) * 4,
                          (size_t)std::max<int64_t>(nE, 1) * 4, (size_t)std::max<int64_t>(nV, 1) * 4,
                          (size_t)std::max<int64_t>(nV, 1) * 4};
    const size_t used[6] = {((size_t)S->usedV + 1) * 4, (size_t)S->usedE * 4, ((size_t)S->usedV + 1) * 4,
                            (size_t)S->usedE * 4, (size_t)S->usedV * 4, (size_t)S->usedV * 4};
    int32_t* old[6] = {S->csr[0], S->csr[1], S->csr[2], S->csr[3], S->first[0], S->first[1]};
    int rc = ISPAN_OK;
    for (int k = 0; k < 6 && !rc; ++k) {
        if (hipMalloc(&n[k], sz[k]) != hipSuccess) rc = ISPAN_ERR_NOMEM;
        else if (old[k] && used[k] && hipMemcpyAsync(n[k], old[k], used[k], hipMemcpyDeviceToDevice, S->copy_stream))
            rc = ISPAN_ERR_HIP;
    }
    if (!rc && hipStreamSynchronize(S->copy_stream) != hipSuccess) rc = ISPAN_ERR_HIP;
    if (!rc && !old[0]) {  // empty arena: offset 0 at vertex 0
        const int32_t zero = 0;
        if (hipMemcpy(n[0], &zero, 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(n[2], &zero, 4, hipMemcpyHostToDevice) != hipSuccess)
            rc = ISPAN_ERR_HIP;
    }
    lk.lock();
    if (rc) {
        for (auto* p : n)
            if (p) (void)hipFree(p);
    } else {
        for (auto* p : old)
            if (p) (void)hipFree(p);
        for (int k = 0; k < 4; ++k) S->csr[k] = n[k];
        S->first[0] = n[4];
        S->first[1] = n[5];
        S->capV = nV;
        S->capE = nE;
    }
    S->growing = false;
    S->cv_arena.notify_all();
    return rc;
}

// Registers a graph whose CSR pair is in device memory (checked here).
int register_device(ispan_server* S, int32_t V, int64_t E, const int32_t* const* d, int64_t* id) {
    hipStream_t s = S->copy_stream;
    int rc = ispan::check_csr_device(V, E, d[0], d[1], d[2], d[3], s);
    if (rc) return rc;
    std::unique_lock<std::mutex> lk(S->mu);
    // growth and appends wait until no batch reads the arena
    S->cv_arena.wait(lk, [S] { return S->inflight == 0 && !S->growing; });
    const int64_t vb = S->usedV, eb = S->usedE;
    if (vb + V > INT32_MAX - 64 || eb + E > INT32_MAX) return ISPAN_ERR_BUDGET;
    if (vb + V > S->capV || eb + E > S->capE) {
        rc = grow_arena(S, lk, vb + V, eb + E);
        if (rc) return rc;
        S->cv_arena.wait(lk, [S] { return S->inflight == 0 && !S->growing; });
    }
    S->growing = true;  // keeps batches out while the graph is appended
    lk.unlock();
    if (ispan::launch_arena_append(d[0], d[1], V, E, vb, eb, S->csr[0], S->csr[1], s) != hipSuccess ||
        ispan::launch_arena_append(d[2], d[3], V, E, vb, eb, S->csr[2], S->csr[3], s) != hipSuccess ||
        (V > 0 && (ispan::launch_first(S->csr[0] + vb, S->csr[1], V, S->first[0] + vb, 1024, s) != hipSuccess ||
                   ispan::launch_first(S->csr[2] + vb, S->csr[3], V, S->first[1] + vb, 1024, s) != hipSuccess)) ||
        hipStreamSynchronize(s) != hipSuccess)
        rc = ISPAN_ERR_HIP;
    lk.lock();
    S->growing = false;
    S->cv_arena.notify_all();
    if (rc) return rc;
    Graph g;
    g.vb = vb;
    g.eb = eb;
    g.V = V;
    g.E = E;
    S->usedV = vb + V;
    S->usedE = eb + E;
    S->bytes_live += arena_bytes(V, E);
    *id = S->next_graph++;
    S->graphs.emplace(*id, g);
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

int wait_done(ispan_server* S, std::unique_lock<std::mutex>& lk, int64_t rid, int64_t timeout_ms) {
    auto it = S->reqs.find(rid);
    if (it == S->reqs.end()) return ISPAN_ERR_NOTFOUND;
    auto ready = [S, rid] { return S->reqs[rid].done; };
    if (timeout_ms < 0) S->cv_done.wait(lk, ready);
    else if (!S->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready)) return ISPAN_ERR_TIMEOUT;
    return ISPAN_OK;
}

void fill_info(const Request& r, ispan_request_info* info) {
    if (!info) return;
    info->submit_ms = r.t_submit;
    info->start_ms = r.t_start;
    info->done_ms = r.t_done;
    info->latency_ms = r.t_done - r.t_submit;
    info->batch = r.batch;
    info->batch_graphs = r.batch_graphs;
    info->status = r.status;
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
    S->max_vertices = cfg && cfg->max_batch_vertices > 0 ? cfg->max_batch_vertices : 0;
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
    S->cv_mem.notify_all();
    if (S->worker.joinable()) S->worker.join();
    if (S->completer.joinable()) S->completer.join();
    (void)hipSetDevice(S->device);
    ispan_ctx_destroy(S->ctx);  // waits for its stream
    for (auto* p : S->csr)
        if (p) (void)hipFree(p);
    for (auto* p : S->first)
        if (p) (void)hipFree(p);
    for (auto& b : S->bufs)
        if (b.d) (void)hipFree(b.d);
    if (S->copy_stream) (void)hipStreamDestroy(S->copy_stream);
    if (S->compute_stream) (void)hipStreamDestroy(S->compute_stream);
    delete S;
    return ISPAN_OK;
}

int ispan_server_add_graph(ispan_server* S, int32_t V, int64_t E, const int32_t* fw_beg, const int32_t* fw_adj,
                           const int32_t* bw_beg, const int32_t* bw_adj, int64_t* graph_id) {
    if (!S || V < 0 || E < 0 || E > INT32_MAX || !graph_id || !fw_beg || !bw_beg || (E > 0 && (!fw_adj || !bw_adj)))
        return ISPAN_ERR_ARG;
    if (ispan_csr_check(V, E, fw_beg, fw_adj, bw_beg, bw_adj) != ISPAN_OK) return ISPAN_ERR_CSR;
    (void)hipSetDevice(S->device);
    // staged through temporary device arrays
    int32_t* d[4] = {nullptr, nullptr, nullptr, nullptr};
    const int32_t* h[4] = {fw_beg, fw_adj, bw_beg, bw_adj};
    const size_t n[4] = {((size_t)V + 1) * 4, (size_t)E * 4, ((size_t)V + 1) * 4, (size_t)E * 4};
    int rc = ISPAN_OK;
    for (int k = 0; k < 4 && !rc; ++k) {
        if (hipMalloc(&d[k], std::max<size_t>(n[k], 4)) != hipSuccess) rc = ISPAN_ERR_NOMEM;
        else if (n[k] && hipMemcpy(d[k], h[k], n[k], hipMemcpyHostToDevice) != hipSuccess) rc = ISPAN_ERR_HIP;
    }
    if (!rc) {
        const int32_t* dc[4] = {d[0], d[1], d[2], d[3]};
        rc = register_device(S, V, E, dc, graph_id);
    }
    for (auto* p : d)
        if (p) (void)hipFree(p);
    return rc;
}

int ispan_server_add_graph_device(ispan_server* S, int32_t V, int64_t E, const int32_t* d_fw_beg,
                                  const int32_t* d_fw_adj, const int32_t* d_bw_beg, const int32_t* d_bw_adj,
                                  int64_t* graph_id) {
    if (!S || V < 0 || E < 0 || E > INT32_MAX || !graph_id || !d_fw_beg || !d_bw_beg ||
        (E > 0 && (!d_fw_adj || !d_bw_adj)))
        return ISPAN_ERR_ARG;
    (void)hipSetDevice(S->device);
    const int32_t* d[4] = {d_fw_beg, d_fw_adj, d_bw_beg, d_bw_adj};
    return register_device(S, V, E, d, graph_id);
}

int ispan_server_remove_graph(ispan_server* S, int64_t graph_id) {
    if (!S) return ISPAN_ERR_ARG;
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->graphs.find(graph_id);
    if (it == S->graphs.end() || it->second.removed) return ISPAN_ERR_NOTFOUND;
    it->second.removed = true;
    S->bytes_live -= arena_bytes(it->second.V, it->second.E);
    if (it->second.pending == 0) trim_arena(S);
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

int ispan_server_submit_many(ispan_server* S, int32_t n, const int64_t* graph_ids, int64_t* request_ids) {
    if (!S || n < 0 || (n > 0 && (!graph_ids || !request_ids))) return ISPAN_ERR_ARG;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        for (int32_t i = 0; i < n; ++i) {
            auto it = S->graphs.find(graph_ids[i]);
            if (it == S->graphs.end() || it->second.removed) return ISPAN_ERR_NOTFOUND;
        }
        const double t = now_ms();
        for (int32_t i = 0; i < n; ++i) {
            Graph& g = S->graphs[graph_ids[i]];
            g.pending += 1;
            const int64_t id = S->next_req++;
            Request& r = S->reqs[id];
            r.graph = graph_ids[i];
            r.V = g.V;
            r.t_submit = t;
            S->pending.push_back(id);
            request_ids[i] = id;
        }
    }
    S->cv_work.notify_one();
    return ISPAN_OK;
}

int ispan_server_wait(ispan_server* S, int64_t request_id, int32_t* labels, ispan_request_info* info,
                      int64_t timeout_ms) {
    if (!S) return ISPAN_ERR_ARG;
    std::unique_lock<std::mutex> lk(S->mu);
    int rc = wait_done(S, lk, request_id, timeout_ms);
    if (rc) return rc;
    Request r = S->reqs[request_id];
    S->reqs.erase(request_id);
    fill_info(r, info);
    rc = r.status;
    if (rc == ISPAN_OK && labels && r.V > 0) {
        lk.unlock();  // the buffer stays referenced until the copy is done
        if (hipSetDevice(S->device) != hipSuccess ||
            hipMemcpy(labels, r.d_labels, (size_t)r.V * 4, hipMemcpyDeviceToHost) != hipSuccess)
            rc = ISPAN_ERR_HIP;
        lk.lock();
    }
    drop_ref(S, r);
    return rc;
}

int ispan_server_wait_device(ispan_server* S, int64_t request_id, const int32_t** d_labels, ispan_request_info* info,
                             int64_t timeout_ms) {
    if (!S || !d_labels) return ISPAN_ERR_ARG;
    std::unique_lock<std::mutex> lk(S->mu);
    const int rc = wait_done(S, lk, request_id, timeout_ms);
    if (rc) return rc;
    const Request& r = S->reqs[request_id];
    fill_info(r, info);
    *d_labels = r.d_labels;
    return r.status;
}

int ispan_server_release(ispan_server* S, int64_t request_id) {
    if (!S) return ISPAN_ERR_ARG;
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(request_id);
    if (it == S->reqs.end() || !it->second.done) return ISPAN_ERR_NOTFOUND;
    drop_ref(S, it->second);
    S->reqs.erase(it);
    return ISPAN_OK;
}

int ispan_server_stats_get(ispan_server* S, ispan_server_stats* out) {
    if (!S || !out) return ISPAN_ERR_ARG;
    std::lock_guard<std::mutex> lk(S->mu);
    memset(out, 0, sizeof(*out));
    out->requests_done = (int64_t)S->latencies.size();
    out->batches = S->batches;
    int64_t live = 0;
    for (const auto& kv : S->graphs) live += !kv.second.removed;
    out->graphs_resident = live;
    out->bytes_resident = S->bytes_live;
    out->bytes_device = bytes_used(S);
    // the budget in effect: the configured one, capped by what the device
    // could still give this server
    size_t free_b = 0, total_b = 0;
    out->bytes_budget = S->budget;
    if (hipSetDevice(S->device) == hipSuccess && hipMemGetInfo(&free_b, &total_b) == hipSuccess)
        out->bytes_budget = std::min<int64_t>(S->budget, out->bytes_device + (int64_t)free_b);
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

int ispan_server_set_tuning(ispan_server* S, const ispan_tuning_table* table) {
    if (!S) return ISPAN_ERR_ARG;
    std::lock_guard<std::mutex> lk(S->mu);
    S->tuned = table != nullptr;
    if (table) S->tuning = *table;
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
