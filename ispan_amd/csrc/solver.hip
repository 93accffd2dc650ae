// Host side of libispan_hip.so: the context (device buffers, stream, pinned
// readback block), the SCC phase sequence of scc_detection
// (src/scc_core.cpp:27-1094) driven over the kernels in kernels.hip, and the
// C ABI declared in include/ispan_hip.h.
#include <string.h>

#include <cstring>
#include <hip/hip_runtime.h>
#include <rocprim/rocprim.hpp>

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/time.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/ispan_hip.h"
#include "common.h"
#include "internal.h"

using namespace ispan;

struct ispan_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t own_stream = nullptr;
    int cu = 0;
    int grid = 0;
    std::string err;

    // bound graph
    int32_t V = -1;
    int64_t E = 0;
    const int32_t* fwb = nullptr;
    const int32_t* fwa = nullptr;
    const int32_t* bwb = nullptr;
    const int32_t* bwa = nullptr;
    int32_t* owned[4] = {nullptr, nullptr, nullptr, nullptr};
    int64_t nwords = 0;
    int32_t* fw_first = nullptr;  // first out-/in-neighbour per vertex, built at bind
    int32_t* bw_first = nullptr;

    // solve state (capacity cap_V / cap_E)
    int32_t cap_V = -1;
    int64_t cap_E = -1;
    int32_t* lab = nullptr;
    uint64_t* act = nullptr;
    uint64_t* fvis = nullptr;
    uint64_t* bvis = nullptr;
    int32_t* fq[2] = {nullptr, nullptr};
    int32_t* q[2] = {nullptr, nullptr};
    HubItem* hub[2] = {nullptr, nullptr};
    int32_t* parent = nullptr;
    int32_t* stamp = nullptr;
    uint64_t* keys = nullptr;
    uint64_t* keys_alt = nullptr;
    int32_t* starts = nullptr;
    int32_t* qbuf = nullptr;
    PivotPart* part = nullptr;
    Ctl* ctl = nullptr;
    void* sort_tmp = nullptr;
    size_t sort_tmp_bytes = 0;
    Ctl* h_ctl = nullptr;  // pinned
    // cooperative giant-phase kernel
    int coop_grid = 0;
    bool level_trace = false;  // ISPAN_LEVEL_TRACE=1: per-level lines of the giant phase on stderr
    double clock_khz = 0.0;
    uint64_t* snap = nullptr;
    HubItem* gh[2] = {nullptr, nullptr};   // forward hub work items
    HubItem* ghb[2] = {nullptr, nullptr};  // backward
    uint64_t* snapb = nullptr;             // backward bottom-up snapshot
    int64_t* gpart = nullptr;
    int32_t* gdbg = nullptr;
    GiantState* gstate = nullptr;
    GiantState* h_gstate = nullptr;  // pinned
    hipEvent_t ev[16] = {};
    bool have_solution = false;
    // batched solves: consecutive graphs grouped so that each group's
    // disjoint union fits int32 offsets; each group owns its union CSR
    struct BatchGroup {
        int32_t g0 = 0, ng = 0;  // first graph (in bind order) and graph count
        int32_t cls = 0;         // size class of its largest graph
        int32_t V = 0;
        int64_t E = 0;
        int32_t* csr[4] = {nullptr, nullptr, nullptr, nullptr};  // fw_beg, fw_adj, bw_beg, bw_adj
        int32_t* first[2] = {nullptr, nullptr};
        BatchRanges rb;  // its graphs' ranges inside the union, labels into batch_labels
    };
    std::vector<BatchGroup> groups;
    int32_t batch_n = -1;
    std::vector<int64_t> batch_v_off;
    int32_t* batch_labels = nullptr;  // [sum V_g], each graph's local canonical ids
    bool batch_solved = false;
    const int32_t* batch_roots = nullptr;  // set while one group is solved
    int32_t batch_nroots = 0;
    const uint64_t* batch_sel = nullptr;  // set while a selection of an arena is solved
    // ranges and selection bitmap of solve_arena (the server's batches)
    BatchRanges ws_rb;
    int32_t ws_capN = -1;
    int64_t ws_capC = -1;
    uint64_t* sel = nullptr;
    int64_t sel_cap = -1;
    void* stage = nullptr;  // pinned staging of range uploads
    size_t stage_bytes = 0;
    // kernel profiling (params.timing >= 2)
    std::vector<hipEvent_t> pool;
    size_t pool_used = 0;
    struct Rec {
        int cls;
        hipEvent_t a, b;
    };
    std::vector<Rec> recs;
    ispan_kernel_profile prof{};
    int64_t* hist = nullptr;  // kHistSlots results of ispan_scc_histogram
};

namespace {

const char* const kVersion = "ispan_hip 0.1 (gfx950)";

double wall_ms() {
    struct timeval tv;
    gettimeofday(&tv, nullptr);
    return tv.tv_sec * 1e3 + tv.tv_usec * 1e-3;
}

#define CK(call)                                                                                  \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess) {                                                                   \
            ctx->err = std::string(#call) + ": " + hipGetErrorString(e_);                         \
            return e_ == hipErrorOutOfMemory ? ISPAN_ERR_NOMEM : ISPAN_ERR_HIP;                   \
        }                                                                                         \
    } while (0)

template <class T>
void dfree(T*& p) {
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

void free_solve_buffers(ispan_ctx* ctx) {
    dfree(ctx->lab);
    dfree(ctx->act);
    dfree(ctx->fvis);
    dfree(ctx->bvis);
    for (int i = 0; i < 2; ++i) {
        dfree(ctx->fq[i]);
        dfree(ctx->q[i]);
        dfree(ctx->hub[i]);
    }
    dfree(ctx->parent);
    dfree(ctx->stamp);
    dfree(ctx->keys);
    dfree(ctx->keys_alt);
    dfree(ctx->starts);
    dfree(ctx->qbuf);
    dfree(ctx->part);
    dfree(ctx->snap);
    dfree(ctx->gh[0]);
    dfree(ctx->gh[1]);
    dfree(ctx->ghb[0]);
    dfree(ctx->ghb[1]);
    dfree(ctx->snapb);
    dfree(ctx->gpart);
    dfree(ctx->gdbg);
    if (ctx->sort_tmp) (void)hipFree(ctx->sort_tmp);
    ctx->sort_tmp = nullptr;
    ctx->sort_tmp_bytes = 0;
    ctx->cap_V = -1;
    ctx->cap_E = -1;
    ctx->have_solution = false;
}

void free_ranges(BatchRanges& r) {
    dfree(r.vb);  // one block: vb, ve, cstart, out
    r.ve = r.cstart = nullptr;
    r.out = nullptr;
    dfree(r.cp);
    dfree(r.ci);
    dfree(r.roots);
    dfree(r.gmin);
    r.n = 0;
    r.nchunks = 0;
}

void free_batch(ispan_ctx* ctx) {
    for (auto& g : ctx->groups) {
        for (auto& p : g.csr) dfree(p);
        for (auto& p : g.first) dfree(p);
        free_ranges(g.rb);
    }
    ctx->groups.clear();
    dfree(ctx->batch_labels);
    ctx->batch_n = -1;
    ctx->batch_v_off.clear();
    ctx->batch_solved = false;
    ctx->batch_roots = nullptr;
    ctx->batch_nroots = 0;
}

void free_workspace(ispan_ctx* ctx) {
    free_ranges(ctx->ws_rb);
    ctx->ws_capN = -1;
    ctx->ws_capC = -1;
    dfree(ctx->sel);
    ctx->sel_cap = -1;
    if (ctx->stage) (void)hipHostFree(ctx->stage);
    ctx->stage = nullptr;
    ctx->stage_bytes = 0;
}

void free_graph(ispan_ctx* ctx) {
    free_batch(ctx);
    for (auto& p : ctx->owned) dfree(p);
    dfree(ctx->fw_first);
    dfree(ctx->bw_first);
    ctx->fwb = ctx->fwa = ctx->bwb = ctx->bwa = nullptr;
    ctx->V = -1;
    ctx->E = 0;
    ctx->have_solution = false;
}

int64_t hub_capacity(int64_t E) { return 2 * (E / kHubChunk) + 64; }

int ensure_buffers(ispan_ctx* ctx, int32_t V, int64_t E) {
    if (V <= ctx->cap_V && E <= ctx->cap_E) return ISPAN_OK;
    free_solve_buffers(ctx);
    const size_t nv = (size_t)std::max(V, 1) + 1;
    const size_t nw = (size_t)((V + 63) / 64) + 1;
    CK(hipMalloc(&ctx->lab, nv * 4));
    CK(hipMalloc(&ctx->act, nw * 8));
    CK(hipMalloc(&ctx->fvis, nw * 8));
    CK(hipMalloc(&ctx->bvis, nw * 8));
    for (int i = 0; i < 2; ++i) {
        // 8 bytes a vertex: the giant kernel's top-down queue entries
        // (first edge, degree); the other phases queue 4-byte ids in them
        CK(hipMalloc(&ctx->fq[i], nv * 8));
        CK(hipMalloc(&ctx->q[i], nv * 8));
        CK(hipMalloc(&ctx->hub[i], (size_t)hub_capacity(E) * sizeof(HubItem)));
    }
    CK(hipMalloc(&ctx->parent, nv * 4));
    CK(hipMalloc(&ctx->stamp, nv * 4));
    CK(hipMalloc(&ctx->keys, nv * 8));
    CK(hipMalloc(&ctx->keys_alt, nv * 8));
    CK(hipMalloc(&ctx->starts, nv * 4));
    CK(hipMalloc(&ctx->qbuf, nv * 4));
    CK(hipMalloc(&ctx->part, (size_t)ctx->grid * sizeof(PivotPart)));
    CK(hipMalloc(&ctx->snap, nw * 8));
    for (int i = 0; i < 2; ++i) {
        CK(hipMalloc(&ctx->gh[i], (size_t)giant_hub_capacity(E) * sizeof(HubItem)));
        CK(hipMalloc(&ctx->ghb[i], (size_t)giant_hub_capacity(E) * sizeof(HubItem)));
    }
    CK(hipMalloc(&ctx->snapb, nw * 8));
    CK(hipMalloc(&ctx->gpart, (size_t)kGiantSlots * std::max(ctx->coop_grid, 1) * 8));
    CK(hipMalloc(&ctx->gdbg, (size_t)2 * std::max(ctx->coop_grid, 1) * 4));
    size_t bytes = 0;
    CK(rocprim::radix_sort_keys(nullptr, bytes, ctx->keys, ctx->keys_alt, (size_t)nv, 0, 64, ctx->stream));
    CK(hipMalloc(&ctx->sort_tmp, bytes + 256));
    ctx->sort_tmp_bytes = bytes + 256;
    ctx->cap_V = V;
    ctx->cap_E = E;
    return ISPAN_OK;
}

hipEvent_t pool_event(ispan_ctx* ctx) {
    if (ctx->pool_used == ctx->pool.size()) {
        hipEvent_t e = nullptr;
        // timing only, read after a stream synchronize: no system-scope fence
        // (it would write the L2s back to memory at every record)
        if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return nullptr;
        ctx->pool.push_back(e);
    }
    return ctx->pool[ctx->pool_used++];
}

// Brackets of one traversal level when profiling: returns the events to
// record before / between / after its kernels (null when not profiling).
struct LevelEv {
    hipEvent_t a = nullptr, mid = nullptr, b = nullptr;
};

LevelEv level_events(ispan_ctx* ctx, const ispan_params* P, bool td) {
    LevelEv e;
    if (P->timing < 2) return e;
    e.a = pool_event(ctx);
    e.b = pool_event(ctx);
    if (td) e.mid = pool_event(ctx);
    if (!e.a || !e.b || (td && !e.mid)) return LevelEv{};
    if (td) {
        ctx->recs.push_back({0, e.a, e.mid});
        ctx->recs.push_back({1, e.mid, e.b});
    } else {
        ctx->recs.push_back({2, e.a, e.b});
    }
    return e;
}

int readback(ispan_ctx* ctx, int64_t* syncs) {
    CK(hipMemcpyAsync(ctx->h_ctl, ctx->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    if (syncs) ++*syncs;
    return ISPAN_OK;
}

// One direction-optimising traversal (fw_bfs_fq_queue, src/fw_bw.h:396-698;
// bw_bfs_fq_queue, src/fw_bw.h:944-1248) over the kernels: sync top-down
// (mode 0), Rsync bottom-up (mode 1), then top-down to exhaustion (mode 2),
// with the reference's switch rules on alpha, beta and gamma.  set_size is
// the reference's fq_size for this traversal (the active queue for FW, the
// FW set for BW).  Top-down levels of mode 2 are enqueued async_batch at a
// time between host checks (a level with an empty input does nothing).
int traverse(ispan_ctx* ctx, bool bw, int64_t set_size, const ispan_params* P, double avg_degree, int64_t* syncs) {
    hipStream_t s = ctx->stream;
    const int32_t* td_beg = bw ? ctx->bwb : ctx->fwb;
    const int32_t* td_adj = bw ? ctx->bwa : ctx->fwa;
    const int32_t* bu_beg = bw ? ctx->fwb : ctx->bwb;
    const int32_t* bu_adj = bw ? ctx->fwa : ctx->bwa;
    const uint64_t* allow = bw ? ctx->fvis : ctx->act;
    uint64_t* vis = bw ? ctx->bvis : ctx->fvis;
    int par = 0, mode = 0;
    int64_t visited = 0;
    int32_t level = 0;
    const int batch = P->async_batch > 0 ? P->async_batch : 8;
    const int64_t level_cap = (int64_t)ctx->V + 2;
    while (true) {
        if (mode == 2) {
            for (int b = 0; b < batch; ++b) {
                LevelEv le = level_events(ctx, P, true);
                if (le.a) CK(hipEventRecord(le.a, s));
                CK(launch_td_level(bw, td_beg, td_adj, allow, vis, ctx->q, ctx->hub, ctx->ctl, par, ctx->grid, s,
                                   le.mid));
                if (le.b) CK(hipEventRecord(le.b, s));
                par ^= 1;
            }
            level += batch;
            int rc = readback(ctx, syncs);
            if (rc) return rc;
            if (ctx->h_ctl->q_n[par] == 0) break;
            if (level > level_cap) {
                ctx->err = "traversal exceeded V levels";
                return ISPAN_ERR_ITER;
            }
            continue;
        }
        LevelEv le = level_events(ctx, P, mode == 0);
        if (le.a) CK(hipEventRecord(le.a, s));
        if (mode == 0)
            CK(launch_td_level(bw, td_beg, td_adj, allow, vis, ctx->q, ctx->hub, ctx->ctl, par, ctx->grid, s,
                               le.mid));
        else
            CK(launch_bu_level(bw, bu_beg, bu_adj, allow, vis, ctx->q, ctx->hub, ctx->ctl, par, ctx->nwords,
                               ctx->grid, s));
        if (le.b) CK(hipEventRecord(le.b, s));
        int rc = readback(ctx, syncs);
        if (rc) return rc;
        const int64_t frontier = ctx->h_ctl->q_n[par ^ 1];
        visited += frontier;
        if (frontier == 0) break;
        if (mode == 0) {
            const double ef = (double)frontier * avg_degree;
            const double er = (double)(set_size - visited) * avg_degree;
            if (er / P->alpha < ef) mode = 1;
        } else if ((set_size * 1.0 / P->beta) > frontier || level > P->gamma) {
            mode = 2;  // BU pushed its finds into q[par^1]: the next input
        }
        par ^= 1;
        ++level;
        if (level > level_cap) {
            ctx->err = "traversal exceeded V levels";
            return ISPAN_ERR_ITER;
        }
    }
    return ISPAN_OK;
}

void rec(ispan_ctx* ctx, const ispan_params* P, int i) {
    if (P->timing) (void)hipEventRecord(ctx->ev[i], ctx->stream);
}

float ev_ms(ispan_ctx* ctx, int a, int b) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, ctx->ev[a], ctx->ev[b]) != hipSuccess) return 0.f;
    return ms;
}

int check_bound(ispan_ctx* ctx, const ispan_params* P) {
    if (!P || !P->validate) return ISPAN_OK;
    CK(launch_reset_ctl(ctx->ctl, ctx->stream));
    if (ctx->V > 0) {
        CK(launch_check_csr(ctx->fwb, ctx->fwa, ctx->V, ctx->E, ctx->ctl, ctx->grid, ctx->stream));
        CK(launch_check_csr(ctx->bwb, ctx->bwa, ctx->V, ctx->E, ctx->ctl, ctx->grid, ctx->stream));
    }
    int rc = readback(ctx, nullptr);
    if (rc) return rc;
    if (ctx->h_ctl->bad_csr) {
        ctx->err = "CSR check failed (" + std::to_string(ctx->h_ctl->bad_csr) + " violations)";
        return ISPAN_ERR_CSR;
    }
    return ISPAN_OK;
}

int solve_impl(ispan_ctx* ctx, const ispan_params* P, int32_t* d_labels, ispan_stats* st) {
    const double t0 = wall_ms();
    hipStream_t s = ctx->stream;
    const int32_t V = ctx->V;
    const int64_t E = ctx->E;
    ispan_stats local{};
    int64_t syncs = 0;
    int rc = ensure_buffers(ctx, V, E);
    if (rc) return rc;
    ctx->have_solution = false;
    const double avg_degree = V > 0 ? (double)E / V : 0.0;
    const int G = ctx->grid;

    ctx->recs.clear();
    ctx->pool_used = 0;
    ctx->prof = ispan_kernel_profile{};
    // Engine 0 needs a cooperative launch (all blocks resident at once).  The
    // device grants it to one process at a time; when it is refused (or the
    // device has none) this solve runs engine 1, and stats.engine_used says so.
    bool coop = P->engine == 0 && ctx->coop_grid > 0 && V > 0;
    if (!coop) CK(launch_reset_ctl(ctx->ctl, s));  // the cooperative launch clears it with its own state
    // Phase events bracket the host-driven phases; the cooperative kernel
    // times its own phases with the device clock, so it records none.
    if (!coop) rec(ctx, P, 0);
    if (V == 0) {
        CK(hipStreamSynchronize(s));
        ctx->have_solution = true;
        if (st) *st = local;
        return ISPAN_OK;
    }
    int64_t fq0 = 0, n = 0, fw_count = 0;
    int32_t root = -1;
    bool rest_done = false;
    double giant_phase_ms[4] = {0, 0, 0, 0};  // engine 0: trim-1-first+pivot, FW, BW+relabel, trims+WCC
    bool counted = false;  // ctl->unlabelled holds the count of the finished labelling
    GiantArgs ga_tail{};   // engine 0: the giant kernel's arguments, for a colouring tail
    if (coop) {
        // P1-P6 + relabel in one cooperative kernel
        GiantArgs ga{};
        ga.fwb = ctx->fwb;
        ga.fwa = ctx->fwa;
        ga.bwb = ctx->bwb;
        ga.bwa = ctx->bwa;
        ga.fw_first = ctx->fw_first;
        ga.bw_first = ctx->bw_first;
        ga.V = V;
        ga.nwords = ctx->nwords;
        ga.lab = ctx->lab;
        ga.act = ctx->act;
        ga.fvis = ctx->fvis;
        ga.bvis = ctx->bvis;
        ga.snap = ctx->snap;
        ga.q0 = reinterpret_cast<uint64_t*>(ctx->q[0]);
        ga.q1 = reinterpret_cast<uint64_t*>(ctx->q[1]);
        ga.h0 = ctx->gh[0];
        ga.h1 = ctx->gh[1];
        ga.snapb = ctx->snapb;
        // the backward direction runs beside the forward one: its queues are
        // the active-queue buffers, idle until the post phase
        ga.qb0 = reinterpret_cast<uint64_t*>(ctx->fq[0]);
        ga.qb1 = reinterpret_cast<uint64_t*>(ctx->fq[1]);
        ga.hb0 = ctx->ghb[0];
        ga.hb1 = ctx->ghb[1];
        ga.fq = ctx->fq[1];
        ga.gs = ctx->gstate;
        ga.part = ctx->gpart;
        ga.fq_n_out = &ctx->ctl->fq_n[1];
        ga.hub_cap = giant_hub_capacity(E);
        ga.alpha = P->alpha;
        ga.beta = P->beta;
        ga.gamma = P->gamma;
        ga.avg_degree = avg_degree;
        ga.wrap32 = P->pivot_wrap32;
        ga.switch_rule = P->switch_rule;
        ga.roots = ctx->batch_roots;
        ga.n_roots = ctx->batch_nroots;
        ga.sel = ctx->batch_sel;
        ga.rest = 1;
        ga.trim_times = P->trim_times;
        ga.theta = P->theta;
        ga.fq_alt = ctx->fq[0];
        ga.stamp = ctx->stamp;
        ga.parent = ctx->parent;
        ga.fq_n_ctl = ctx->ctl->fq_n;
        ga.dbg = ctx->gdbg;
        ga.trace = ctx->level_trace ? 1 : 0;
        ga.bu_mult = giant_bu_mult(ctx->coop_grid);
        ga.wcc_mode = P->wcc_mode;
        ga.trim_scope = P->trim_scope;
        ga.td_local = std::max(P->td_local, 0);
        ga.ckey = ctx->keys;
        ga.aux = ctx->qbuf;
        ga_tail = ga;
        hipEvent_t ga0 = nullptr, ga1 = nullptr;
        if (P->timing >= 2 && (ga0 = pool_event(ctx)) && (ga1 = pool_event(ctx))) ctx->recs.push_back({3, ga0, ga1});
        const hipError_t le = launch_giant(ga, ctx->ctl, ctx->coop_grid, s, ga0);
        if (le == hipErrorCooperativeLaunchTooLarge || le == hipErrorIllegalState || le == hipErrorNotSupported ||
            le == hipErrorLaunchOutOfResources) {
            (void)hipGetLastError();
            if (ga0) ctx->recs.pop_back();
            coop = false;
            rec(ctx, P, 0);  // engine 1 times its phases with events
        } else {
            CK(le);
            if (ga1) CK(hipEventRecord(ga1, s));
            CK(hipMemcpyAsync(ctx->h_gstate, ctx->gstate, ctx->level_trace ? sizeof(GiantState) : kGiantStateHead,
                              hipMemcpyDeviceToHost, s));
            // the kernel normally labels every vertex and then checks, itself,
            // that none is left at -1 (unl_checked): one host wait covers the
            // solve and its check.  Everything the host reads now is in the
            // state; the control block (cleared by the launch) is copied only
            // by the readbacks of later launches.
            CK(hipStreamSynchronize(s));
            ++syncs;
            *ctx->h_ctl = Ctl{};
            ctx->h_ctl->unlabelled = (int32_t)ctx->h_gstate->unlabelled;
            bool tail = false;
            if (!ctx->h_gstate->err.v && ctx->h_gstate->n_final > 0 && (P->wcc_mode == 1 || P->wcc_mode == 2)) {
                tail = true;
                // active vertices are left: the WCC by label propagation, or
                // the colouring SCC method, on the device (one more launch
                // and wait, as the mice phase would take)
                CK(launch_tail(ga, ctx->coop_grid, s));
                CK(hipMemcpyAsync(ctx->h_gstate, ctx->gstate,
                                  ctx->level_trace ? sizeof(GiantState) : kGiantStateHead, hipMemcpyDeviceToHost, s));
                CK(launch_fill_i32(&ctx->ctl->unlabelled, 0, 1, 1, s));
                CK(launch_count_unlabelled(ctx->lab, V, ctx->batch_sel, ctx->ctl, G, s));
                if ((rc = readback(ctx, &syncs))) return rc;
            }
            const GiantState& gs = *ctx->h_gstate;
            if (gs.err.v) {
                if (gs.err.v & 1) {  // where did the blocks stop?
                    std::vector<int32_t> d((size_t)2 * ctx->coop_grid);
                    if (hipMemcpy(d.data(), ctx->gdbg, d.size() * 4, hipMemcpyDeviceToHost) == hipSuccess) {
                        int amin = INT32_MAX, amax = 0, pmin = INT32_MAX, pmax = 0, zero = 0;
                        for (int b = 0; b < ctx->coop_grid; ++b) {
                            amin = std::min(amin, d[b]);
                            amax = std::max(amax, d[b]);
                            pmin = std::min(pmin, d[ctx->coop_grid + b]);
                            pmax = std::max(pmax, d[ctx->coop_grid + b]);
                            zero += d[b] == 0;
                        }
                        fprintf(stderr, "[ispan] barrier timeout: grid %d, arrived min %d max %d (blocks at 0: %d), "
                                "passed min %d max %d, gen %d\n", ctx->coop_grid, amin, amax, zero, pmin, pmax, gs.gen.v);
                    }
                }
                ctx->err = gs.err.v & 1   ? "giant phase: grid barrier timed out"
                           : gs.err.v & 2 ? "giant phase: level cap exceeded"
                                          : "giant phase: queue count out of range";
                return gs.err.v & 2 ? ISPAN_ERR_ITER : ISPAN_ERR_HIP;
            }
            fq0 = gs.fq0;
            root = (int32_t)gs.root;
            fw_count = gs.fw_count;
            n = gs.fq_n;
            local.fq0 = fq0;
            local.pivot = root;
            local.fw_size = fw_count;
            local.giant_size = gs.giant;
            local.bw_reach = gs.bw_reach;
            local.fw_levels = gs.fw_levels;
            local.bw_levels = gs.bw_levels;
            if (ctx->clock_khz > 0) {
                const double k = 1.0 / ctx->clock_khz;  // ms per tick
                giant_phase_ms[0] = (gs.ts[1] - gs.ts[0]) * k;
                const bool traversed = root >= 0 || ctx->batch_nroots > 0;
                giant_phase_ms[1] = traversed ? (gs.ts[2] - gs.ts[1]) * k : 0.0;
                giant_phase_ms[2] = (gs.ts[4] - (traversed ? gs.ts[2] : gs.ts[1])) * k;
                giant_phase_ms[3] = gs.ts[5] > gs.ts[4] ? (gs.ts[5] - gs.ts[4]) * k : 0.0;
            }
            if (ctx->level_trace && ctx->clock_khz > 0) {
                const double k0 = 1.0 / ctx->clock_khz;
                fprintf(stderr,
                        "[ispan] blocks started %.3f..%.3f ms, trim-1-first done %.3f..%.3f ms, pivot known %.3f ms\n",
                        ((int64_t)~gs.t_start[0] - gs.ts[0]) * k0, (gs.t_start[1] - gs.ts[0]) * k0,
                        ((int64_t)~gs.t_arrive[0] - gs.ts[0]) * k0, (gs.t_arrive[1] - gs.ts[0]) * k0,
                        (gs.ts[1] - gs.ts[0]) * k0);
                const double k = 1.0 / ctx->clock_khz;
                if (gs.post_cp[0])
                    fprintf(stderr, "[ispan] post after BW: counts %.3f ms, barrier %.3f ms, labels and queue %.3f ms\n",
                            (gs.post_cp[0] - gs.ts[3]) * k, (gs.post_cp[1] - gs.ts[3]) * k,
                            (gs.post_cp[2] - gs.ts[3]) * k);
                if (gs.bu1_scan[0]) {  // first forward bottom-up level, per block and by XCC (block % 8)
                    const int nb = std::min(ctx->coop_grid, 1024);
                    double xs[8] = {0}, xw[8] = {0}, xmax[8] = {0};
                    int xn[8] = {0};
                    for (int b = 0; b < nb; ++b) {
                        const int x = b % 8;
                        xs[x] += (double)(gs.bu1_scan[b] - 1);
                        xw[x] += gs.bu1_work[b] * k;
                        xmax[x] = std::max(xmax[x], gs.bu1_work[b] * k);
                        ++xn[x];
                    }
                    for (int x = 0; x < 8; ++x)
                        if (xn[x])
                            fprintf(stderr, "[ispan] BU1 xcc %d: mean work %.3f ms max %.3f ms, mean scan %.0f\n", x,
                                    xw[x] / xn[x], xmax[x], xs[x] / xn[x]);
                    fprintf(stderr, "[ispan] BU1 work by block:");
                    for (int b = 0; b < nb; ++b) fprintf(stderr, "%s%.0f", b % 32 ? " " : "\n  ", gs.bu1_work[b] * k * 1e3);
                    fprintf(stderr, "\n");
                }
                for (int d = 0; d < 2; ++d) {
                    int64_t t = gs.ts[1];  // both directions start after the pivot
                    const int64_t nl = std::min<int64_t>(d == 0 ? gs.fw_levels : gs.bw_levels, kLevelTrace);
                    for (int64_t l = 0; l < nl; ++l) {
                        fprintf(stderr,
                                "[ispan] %s level %2lld %s found %10lld edges %11lld  %8.3f ms  work %.3f..%.3f "
                                "max block scan %.0f marks dense %.3f probed %.3f misses %.3f\n",
                                d ? "BW" : "FW", (long long)l, gs.lv_mode[d][l] ? "BU" : "TD",
                                (long long)gs.lv_found[d][l], (long long)gs.lv_edges[d][l], (gs.lv_ts[d][l] - t) * k,
                                (int64_t)~(uint64_t)gs.lv_wmin[d][l] * k, gs.lv_wmax[d][l] * k, (double)gs.lv_cp[d][l][0],
                                gs.lv_cp[d][l][3] * k, gs.lv_cp[d][l][1] * k, gs.lv_cp[d][l][2] * k);
                        t = gs.lv_ts[d][l];
                    }
                }
            }
            rest_done = true;
            // k_tail's labels are counted by the scan after it; otherwise the
            // kernel's own check counts when nothing was left for the host
            counted = tail ? gs.n_final == 0 : gs.unl_checked != 0;
            ctx->prof.edges[0] = gs.work[0];
            ctx->prof.edges[2] = gs.work[1];
            ctx->prof.found[0] = gs.work[2];
            ctx->prof.found[2] = gs.work[3];
            ctx->prof.found[3] = fq0;
        }
    }
    if (!coop && ctx->batch_nroots > 0) {
        ctx->err = "batched solves need the cooperative engine (the device refused a cooperative launch)";
        return ISPAN_ERR_HIP;
    }
    if (!coop) {
        // P1-P3: trim-1-first, active queue, pivot
        hipEvent_t t1a = nullptr, t1b = nullptr;
        if (P->timing >= 2 && (t1a = pool_event(ctx)) && (t1b = pool_event(ctx))) {
            ctx->recs.push_back({3, t1a, t1b});
            CK(hipEventRecord(t1a, s));
        }
        CK(launch_trim1_first(ctx->fwb, ctx->bwb, V, ctx->lab, ctx->act, ctx->fvis, ctx->bvis, ctx->fq[0],
                              ctx->q[0], ctx->ctl, ctx->part, P->pivot_wrap32, G, s));
        if (t1b) CK(hipEventRecord(t1b, s));
        rec(ctx, P, 1);
        if ((rc = readback(ctx, &syncs))) return rc;
        fq0 = ctx->h_ctl->fq_n[0];
        root = ctx->h_ctl->root;
        local.fq0 = fq0;
        local.pivot = root;
        rec(ctx, P, 2);
        if (root >= 0) {
            // P4 forward traversal
            if ((rc = traverse(ctx, false, fq0, P, avg_degree, &syncs))) return rc;
            local.fw_levels = ctx->h_ctl->levels;
            CK(launch_popcount(ctx->fvis, ctx->nwords, &ctx->ctl->fw_count, G, s));
            rec(ctx, P, 3);
            if ((rc = readback(ctx, &syncs))) return rc;
            fw_count = ctx->h_ctl->fw_count;
            // P5-P6 backward traversal inside the FW set
            CK(launch_bw_init(ctx->bvis, ctx->q[0], ctx->ctl, s));
            if ((rc = traverse(ctx, true, fw_count, P, avg_degree, &syncs))) return rc;
            local.bw_levels = ctx->h_ctl->levels;
        } else {
            rec(ctx, P, 3);
        }
        local.fw_size = fw_count;
        CK(launch_post_giant(V, ctx->lab, ctx->act, ctx->bvis, ctx->fq[1], ctx->ctl, G, s));
        rec(ctx, P, 4);
        if ((rc = readback(ctx, &syncs))) return rc;
        n = ctx->h_ctl->fq_n[1];
        local.giant_size = root >= 0 ? fq0 - n : 0;  // giant = removed by post_giant
        local.bw_reach = local.giant_size;           // engine 1: backward inside the forward set
    }
    int sel = 1;
    int64_t trim1 = 0, pat2 = 0, pat3 = 0;
    bool wcc_ready = false;
    if (rest_done) {
        const GiantState& gs = *ctx->h_gstate;
        sel = (int)gs.final_sel;
        n = gs.n_final;
        trim1 = gs.trim1;
        pat2 = gs.pat2;
        pat3 = gs.pat3;
        wcc_ready = gs.wcc_done != 0;
        if (n > 0) rec(ctx, P, 7);  // host-side WCC and mice follow
    } else {
        // P7 trim-1 sweeps with the theta stop rule (src/scc_core.cpp:562-602):
        // the first comparison is against the pre-giant queue, as there.
        int64_t prev = fq0;
        int it = 0;
        while (prev > 0 && it < P->trim_times) {
            if (n == 0) break;
            CK(launch_trim1_sweep(ctx->fwb, ctx->fwa, ctx->bwb, ctx->bwa, ctx->lab, ctx->act, ctx->fq, ctx->ctl, sel,
                                  G, s));
            sel ^= 1;
            if ((rc = readback(ctx, &syncs))) return rc;
            const int64_t nn = ctx->h_ctl->fq_n[sel];
            trim1 += ctx->h_ctl->trimmed;
            n = nn;
            if (prev - nn < prev * P->theta) break;
            prev = nn;
            ++it;
        }
        rec(ctx, P, 5);
        if (n > 0) {
            // P8-P10 size-2/3 pattern trims
            CK(launch_pattern_trim(ctx->fwb, ctx->fwa, ctx->bwb, ctx->bwa, ctx->lab, ctx->stamp, ctx->act,
                                   ctx->fq[sel], ctx->ctl, sel, G, s));
            CK(launch_pattern_commit(ctx->lab, ctx->act, ctx->stamp, ctx->fq[sel], ctx->ctl, sel, G, s));
            rec(ctx, P, 6);
            // P11 one trim-1 sweep
            CK(launch_trim1_sweep(ctx->fwb, ctx->fwa, ctx->bwb, ctx->bwa, ctx->lab, ctx->act, ctx->fq, ctx->ctl, sel,
                                  G, s));
            sel ^= 1;
            rec(ctx, P, 7);
            if ((rc = readback(ctx, &syncs))) return rc;
            n = ctx->h_ctl->fq_n[sel];
            trim1 += ctx->h_ctl->trimmed;
            pat2 = ctx->h_ctl->pattern2;
            pat3 = ctx->h_ctl->pattern3;
        } else {
            rec(ctx, P, 6);
            rec(ctx, P, 7);
        }
    }
    const bool host_tail = n > 0;
    if (n > 0) {
        // P12 WCC of the remaining subgraph (engine 0 has done it on the
        // device; engine 1 runs label propagation with a host check per
        // round, and the union-find otherwise, also for wcc_mode 2)
        if (wcc_ready) {
        } else if (P->wcc_mode == 1) {
            CK(launch_uf_init(ctx->fq[sel], ctx->parent, ctx->ctl, sel, G, s));
            int it = 0;
            for (;; ++it) {
                if (it > 4 * 1024 * 1024) {
                    ctx->err = "label propagation did not converge";
                    return ISPAN_ERR_ITER;
                }
                CK(launch_lp_round(ctx->fwb, ctx->fwa, ctx->bwb, ctx->bwa, ctx->act, ctx->fq[sel], ctx->parent,
                                   ctx->ctl, sel, G, s));
                if ((rc = readback(ctx, &syncs))) return rc;
                if (!ctx->h_ctl->changed) break;
            }
        } else {
            CK(launch_wcc_uf(ctx->fwb, ctx->fwa, ctx->act, ctx->fq[sel], ctx->parent, ctx->ctl, sel, G, s));
        }
        rec(ctx, P, 8);
        // P13-P14 mice: group by WCC, FW-BW per WCC
        CK(launch_mice_keys(ctx->fq[sel], ctx->parent, ctx->stamp, ctx->keys, ctx->ctl, sel, G, s));
        int end_bit = 32;
        while (end_bit < 64 && ((int64_t)1 << (end_bit - 32)) < (int64_t)V) ++end_bit;
        size_t bytes = ctx->sort_tmp_bytes;
        CK(rocprim::radix_sort_keys(ctx->sort_tmp, bytes, ctx->keys, ctx->keys_alt, (size_t)n, 0, (unsigned)end_bit,
                                    s));
        CK(launch_mice_starts(ctx->keys_alt, (int32_t)n, ctx->starts, ctx->ctl, G, s));
        // FW-BW per WCC costs a long WCC of small SCCs its size per SCC (each
        // search runs down the rest of it): on engine 0, past
        // ISPAN_MICE_MAX_WCC vertices k_mice leaves everything to the
        // colouring method, launched after the readback below shows it
        CK(launch_mice(ctx->fwb, ctx->fwa, ctx->bwb, ctx->bwa, ctx->lab, ctx->stamp, ctx->keys_alt, (int32_t)n,
                       ctx->starts, ctx->ctl, ctx->qbuf, coop ? ISPAN_MICE_MAX_WCC : 0, G, s));
        local.mice = n;
    } else if (!coop) {
        rec(ctx, P, 8);
    }
    if (!coop || host_tail) rec(ctx, P, 9);
    if (!counted) {
        CK(launch_fill_i32(&ctx->ctl->unlabelled, 0, 1, 1, s));
        CK(launch_count_unlabelled(ctx->lab, V, ctx->batch_sel, ctx->ctl, G, s));
        if ((rc = readback(ctx, &syncs))) return rc;
    }
    if (host_tail) local.max_wcc = ctx->h_ctl->max_wcc;
    if (coop && host_tail && ctx->h_ctl->max_wcc > ISPAN_MICE_MAX_WCC) {
        // k_mice left this rest alone: the colouring method labels it
        GiantArgs gc = ga_tail;
        gc.wcc_mode = 2;
        CK(launch_tail(gc, ctx->coop_grid, s));
        CK(launch_fill_i32(&ctx->ctl->unlabelled, 0, 1, 1, s));
        CK(launch_count_unlabelled(ctx->lab, V, ctx->batch_sel, ctx->ctl, G, s));
        if ((rc = readback(ctx, &syncs))) return rc;
        local.coloured = local.mice;
        local.mice = 0;
    }
    if (ctx->h_ctl->unlabelled) {
        ctx->err = std::to_string(ctx->h_ctl->unlabelled) + " vertices left unlabelled";
        return ISPAN_ERR_UNLABELLED;
    }
    if (d_labels && d_labels != ctx->lab) {
        CK(hipMemcpyAsync(d_labels, ctx->lab, (size_t)V * 4, hipMemcpyDeviceToDevice, s));
        CK(hipStreamSynchronize(s));
    }
    local.n_wcc = ctx->h_ctl->n_wcc;
    local.trim1 = trim1 + (V - fq0);
    local.trim2 = pat2;
    local.trim3 = pat3;
    local.host_syncs = syncs;
    local.engine_used = coop ? 0 : 1;
    local.ms[3] = wall_ms() - t0;
    if (P->timing) {
        // event map: 0 start, 1 trim-1-first+queue+pivot, 2 pivot readback,
        // 3 FW done, 4 BW + giant relabel done, 5 trim-1 sweeps, 6 size-2/3,
        // 7 last trim-1 sweep, 8 WCC, 9 mice
        double first, fw, bw, rest_ms = 0.0, t1 = 0.0, t23 = 0.0, t1b = 0.0, wcc = 0.0, mice = 0.0;
        if (coop) {
            first = giant_phase_ms[0];
            fw = giant_phase_ms[1];
            bw = giant_phase_ms[2];
            rest_ms = giant_phase_ms[3];
        } else {
            first = ev_ms(ctx, 0, 1);
            fw = ev_ms(ctx, 2, 3);
            bw = ev_ms(ctx, 3, 4);
            t1 = ev_ms(ctx, 4, 5);
            t23 = ev_ms(ctx, 5, 6);
            t1b = ev_ms(ctx, 6, 7);
        }
        if (!coop || host_tail) {
            wcc = ev_ms(ctx, 7, 8);
            mice = ev_ms(ctx, 8, 9);
        }
        local.ms[11] = first;
        local.ms[14] = t1 + t1b + rest_ms;  // engine 0: in-kernel sweeps, patterns and WCC
        local.ms[4] = first + t1 + t1b + rest_ms;
        local.ms[5] = t23;
        local.ms[13] = 0.0;  // size-3 patterns run fused with size-2 (ms[5])
        local.ms[6] = 0.0;   // pivot reduction is fused into trim-1-first (ms[11])
        local.ms[7] = fw;
        local.ms[8] = bw;
        local.ms[9] = wcc;
        local.ms[10] = mice;
        local.ms[0] = local.ms[4] + local.ms[5] + local.ms[13];
        local.ms[1] = fw + bw;
        local.ms[2] = wcc + mice;
    }
    if (P->timing >= 2) {
        ispan_kernel_profile& pf = ctx->prof;
        for (const auto& r : ctx->recs) {
            float ms = 0.f;
            CK(hipEventElapsedTime(&ms, r.a, r.b));
            pf.launches[r.cls] += 1;
            pf.ms[r.cls] += ms;
        }
        if (!coop) {
            for (int c = 0; c < 3; ++c) {
                pf.edges[c] = ctx->h_ctl->work_edges[c];
                pf.found[c] = ctx->h_ctl->work_found[c];
            }
            pf.edges[3] = 0;
            pf.found[3] = fq0;
        }
    }
    ctx->have_solution = true;
    if (st) *st = local;
    return ISPAN_OK;
}

int bind_common(ispan_ctx* ctx, int32_t V, int64_t E) {
    ctx->V = V;
    ctx->E = E;
    ctx->nwords = ((int64_t)V + 63) / 64;
    const size_t nv = (size_t)std::max(V, 1);
    CK(hipMalloc(&ctx->fw_first, nv * 4));
    CK(hipMalloc(&ctx->bw_first, nv * 4));
    if (V > 0) {
        CK(launch_first(ctx->fwb, ctx->fwa, V, ctx->fw_first, ctx->grid, ctx->stream));
        CK(launch_first(ctx->bwb, ctx->bwa, V, ctx->bw_first, ctx->grid, ctx->stream));
    }
    CK(hipStreamSynchronize(ctx->stream));
    return ISPAN_OK;
}

int64_t file_size(const char* path) {
    struct stat sb;
    if (stat(path, &sb) != 0) return -1;
    return (int64_t)sb.st_size;
}

int read_i32(const char* path, int64_t count, int32_t* out) {
    FILE* f = fopen(path, "rb");
    if (!f) return ISPAN_ERR_IO;
    size_t got = fread(out, 4, (size_t)count, f);
    fclose(f);
    return got == (size_t)count ? ISPAN_OK : ISPAN_ERR_IO;
}

}  // namespace

// ============================================================== C ABI

extern "C" {

void ispan_default_params(ispan_params* p) {
    if (!p) return;
    memset(p, 0, sizeof(*p));
    p->alpha = 30;
    p->beta = 200;
    p->gamma = 10;
    p->theta = 0.1;
    p->trim_times = 3;
    p->thread_count = 1;
    p->validate = 1;
    p->async_batch = 8;
    p->engine = 0;
    p->switch_rule = 1;
    p->td_local = 64;
}

const char* ispan_strerror(int code) {
    switch (code) {
        case ISPAN_OK: return "ok";
        case ISPAN_ERR_ARG: return "invalid argument";
        case ISPAN_ERR_CSR: return "malformed CSR";
        case ISPAN_ERR_NOMEM: return "out of memory";
        case ISPAN_ERR_HIP: return "HIP runtime error";
        case ISPAN_ERR_ITER: return "iteration cap reached";
        case ISPAN_ERR_UNLABELLED: return "vertices left unlabelled";
        case ISPAN_ERR_IO: return "file I/O error";
        case ISPAN_ERR_NOGRAPH: return "no graph bound";
        case ISPAN_ERR_BUDGET: return "HBM budget exceeded";
        case ISPAN_ERR_TIMEOUT: return "timed out";
        case ISPAN_ERR_NOTFOUND: return "unknown id";
        case ISPAN_ERR_SCRATCH: return "graph too large for the allocation-free solve's fixed device scratch";
        default: return "unknown error";
    }
}

const char* ispan_version(void) { return kVersion; }

int ispan_ctx_create(ispan_ctx** out, int device) {
    if (!out) return ISPAN_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return ISPAN_ERR_ARG;
    ispan_ctx* ctx = new ispan_ctx();
    ctx->device = device;
    auto fail = [&](int code) {
        ispan_ctx_destroy(ctx);
        return code;
    };
    if (hipSetDevice(device) != hipSuccess) return fail(ISPAN_ERR_HIP);
    if (hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking) != hipSuccess) return fail(ISPAN_ERR_HIP);
    ctx->stream = ctx->own_stream;
    if (hipDeviceGetAttribute(&ctx->cu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        return fail(ISPAN_ERR_HIP);
    ctx->grid = std::max(ctx->cu, 1) * 8;
    {
        int coop = 0, rate = 0;
        (void)hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, device);
        (void)hipDeviceGetAttribute(&rate, hipDeviceAttributeWallClockRate, device);
        ctx->clock_khz = rate;
        int per_cu = coop ? std::min(giant_max_blocks_per_cu(), 8) : 0;
        if (const char* env = getenv("ISPAN_COOP_BLOCKS_PER_CU")) per_cu = std::min(per_cu, std::max(atoi(env), 1));
        ctx->coop_grid = per_cu * ctx->cu;
        if (const char* env = getenv("ISPAN_LEVEL_TRACE")) ctx->level_trace = atoi(env) != 0;
    }
    if (hipMalloc(&ctx->gstate, sizeof(GiantState)) != hipSuccess) return fail(ISPAN_ERR_NOMEM);
    // epoch 0: no launch of this process has cleared it (the kernel clears the rest)
    if (hipMemset(ctx->gstate, 0, sizeof(GiantState)) != hipSuccess) return fail(ISPAN_ERR_HIP);
    if (hipHostMalloc(&ctx->h_gstate, sizeof(GiantState), hipHostMallocDefault) != hipSuccess)
        return fail(ISPAN_ERR_NOMEM);
    if (hipMalloc(&ctx->ctl, sizeof(Ctl)) != hipSuccess) return fail(ISPAN_ERR_NOMEM);
    if (hipHostMalloc(&ctx->h_ctl, sizeof(Ctl), hipHostMallocDefault) != hipSuccess) return fail(ISPAN_ERR_NOMEM);
    for (auto& e : ctx->ev)  // phase timing events, read after a stream synchronize
        if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return fail(ISPAN_ERR_HIP);
    *out = ctx;
    return ISPAN_OK;
}

int ispan_ctx_destroy(ispan_ctx* ctx) {
    if (!ctx) return ISPAN_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    free_solve_buffers(ctx);
    free_graph(ctx);
    free_workspace(ctx);
    dfree(ctx->ctl);
    dfree(ctx->gstate);
    dfree(ctx->hist);
    if (ctx->h_ctl) (void)hipHostFree(ctx->h_ctl);
    if (ctx->h_gstate) (void)hipHostFree(ctx->h_gstate);
    for (auto& e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : ctx->pool) (void)hipEventDestroy(e);
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
    delete ctx;
    return ISPAN_OK;
}

int ispan_ctx_set_stream(ispan_ctx* ctx, void* hip_stream) {
    if (!ctx) return ISPAN_ERR_ARG;
    ctx->stream = hip_stream ? (hipStream_t)hip_stream : ctx->own_stream;
    return ISPAN_OK;
}

int ispan_ctx_synchronize(ispan_ctx* ctx) {
    if (!ctx) return ISPAN_ERR_ARG;
    const hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        ctx->err = std::string("hipStreamSynchronize: ") + hipGetErrorString(e);
        return ISPAN_ERR_HIP;
    }
    return ISPAN_OK;
}

const char* ispan_ctx_last_error(const ispan_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int ispan_graph_upload(ispan_ctx* ctx, int32_t V, int64_t E, const int32_t* fw_beg, const int32_t* fw_adj,
                       const int32_t* bw_beg, const int32_t* bw_adj, const ispan_params* params) {
    if (!ctx || V < 0 || E < 0 || E > INT32_MAX || !fw_beg || !bw_beg || (E > 0 && (!fw_adj || !bw_adj)))
        return ISPAN_ERR_ARG;
    if (fw_beg[0] != 0 || bw_beg[0] != 0 || (int64_t)fw_beg[V] != E || (int64_t)bw_beg[V] != E) {
        ctx->err = "beg[0] must be 0 and beg[V] must equal E";
        return ISPAN_ERR_CSR;
    }
    CK(hipSetDevice(ctx->device));
    free_graph(ctx);
    const size_t nb = ((size_t)V + 1) * 4, na = (size_t)std::max<int64_t>(E, 1) * 4;
    CK(hipMalloc(&ctx->owned[0], nb));
    CK(hipMalloc(&ctx->owned[1], na));
    CK(hipMalloc(&ctx->owned[2], nb));
    CK(hipMalloc(&ctx->owned[3], na));
    CK(hipMemcpyAsync(ctx->owned[0], fw_beg, nb, hipMemcpyHostToDevice, ctx->stream));
    CK(hipMemcpyAsync(ctx->owned[2], bw_beg, nb, hipMemcpyHostToDevice, ctx->stream));
    if (E > 0) {
        CK(hipMemcpyAsync(ctx->owned[1], fw_adj, (size_t)E * 4, hipMemcpyHostToDevice, ctx->stream));
        CK(hipMemcpyAsync(ctx->owned[3], bw_adj, (size_t)E * 4, hipMemcpyHostToDevice, ctx->stream));
    }
    CK(hipStreamSynchronize(ctx->stream));
    ctx->fwb = ctx->owned[0];
    ctx->fwa = ctx->owned[1];
    ctx->bwb = ctx->owned[2];
    ctx->bwa = ctx->owned[3];
    ispan_params d;
    ispan_default_params(&d);
    ctx->V = V;
    ctx->E = E;
    int rc = check_bound(ctx, params ? params : &d);
    if (!rc) rc = bind_common(ctx, V, E);
    if (rc) free_graph(ctx);
    return rc;
}

int ispan_graph_bind_device(ispan_ctx* ctx, int32_t V, int64_t E, const int32_t* d_fw_beg, const int32_t* d_fw_adj,
                            const int32_t* d_bw_beg, const int32_t* d_bw_adj, const ispan_params* params) {
    if (!ctx || V < 0 || E < 0 || E > INT32_MAX || !d_fw_beg || !d_bw_beg || (E > 0 && (!d_fw_adj || !d_bw_adj)))
        return ISPAN_ERR_ARG;
    CK(hipSetDevice(ctx->device));
    free_graph(ctx);
    ctx->fwb = d_fw_beg;
    ctx->fwa = d_fw_adj;
    ctx->bwb = d_bw_beg;
    ctx->bwa = d_bw_adj;
    ctx->V = V;
    ctx->E = E;
    ispan_params d;
    ispan_default_params(&d);
    int rc = check_bound(ctx, params ? params : &d);
    if (!rc) rc = bind_common(ctx, V, E);
    if (rc) free_graph(ctx);
    return rc;
}

int ispan_solve(ispan_ctx* ctx, const ispan_params* params, int32_t* d_labels, ispan_stats* stats) {
    if (!ctx) return ISPAN_ERR_ARG;
    if (ctx->V < 0) return ISPAN_ERR_NOGRAPH;
    ispan_params d;
    ispan_default_params(&d);
    const ispan_params* P = params ? params : &d;
    if (P->alpha <= 0 || P->beta <= 0 || P->trim_times < 0) return ISPAN_ERR_ARG;
    CK(hipSetDevice(ctx->device));
    return solve_impl(ctx, P, d_labels, stats);
}

int ispan_labels_download(ispan_ctx* ctx, int32_t* labels) {
    if (!ctx || (!labels && ctx->V > 0)) return ISPAN_ERR_ARG;
    if (!ctx->have_solution) return ISPAN_ERR_NOGRAPH;
    if (ctx->V == 0) return ISPAN_OK;
    CK(hipMemcpyAsync(labels, ctx->lab, (size_t)ctx->V * 4, hipMemcpyDeviceToHost, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    return ISPAN_OK;
}

const int32_t* ispan_labels_device(const ispan_ctx* ctx) {
    return (ctx && ctx->have_solution) ? ctx->lab : nullptr;
}

int ispan_solve_stats(ispan_ctx* ctx, ispan_stats* st) {
    if (!ctx || !st) return ISPAN_ERR_ARG;
    if (!ctx->have_solution) return ISPAN_ERR_NOGRAPH;
    if (ctx->V == 0) return ISPAN_OK;
    CK(launch_reset_ctl(ctx->ctl, ctx->stream));
    CK(launch_stats(ctx->fwb, ctx->bwb, ctx->lab, ctx->fvis, ctx->bvis, ctx->V, ctx->ctl, ctx->grid, ctx->stream));
    int rc = readback(ctx, nullptr);
    if (rc) return rc;
    st->fw_edges = ctx->h_ctl->fw_edges;
    st->bw_edges = ctx->h_ctl->bw_edges;
    st->n_scc = ctx->h_ctl->n_scc;
    return ISPAN_OK;
}

namespace {

// The CSR a batch solve runs on (a group's union or the server's arena).
struct View {
    const int32_t* csr[4];
    const int32_t* first[2];
    int32_t V;
    int64_t E;
};

void use_view(ispan_ctx* ctx, const View& v) {
    ctx->fwb = v.csr[0];
    ctx->fwa = v.csr[1];
    ctx->bwb = v.csr[2];
    ctx->bwa = v.csr[3];
    ctx->fw_first = const_cast<int32_t*>(v.first[0]);
    ctx->bw_first = const_cast<int32_t*>(v.first[1]);
    ctx->V = v.V;
    ctx->E = v.E;
    ctx->nwords = ((int64_t)v.V + 63) / 64;
}

void drop_group_view(ispan_ctx* ctx) {
    ctx->fwb = ctx->fwa = ctx->bwb = ctx->bwa = nullptr;
    ctx->fw_first = ctx->bw_first = nullptr;
    ctx->V = -1;
    ctx->E = 0;
    ctx->nwords = 0;
    ctx->batch_roots = nullptr;
    ctx->batch_nroots = 0;
    ctx->batch_sel = nullptr;
}

View group_view(const ispan_ctx::BatchGroup& g) {
    return View{{g.csr[0], g.csr[1], g.csr[2], g.csr[3]}, {g.first[0], g.first[1]}, g.V, g.E};
}

// Device arrays of n graph ranges (grown to capacity *capN graphs / *capC
// chunks when given, else sized exactly) and their upload: vb, ve, cstart
// and out packed into one pinned staging area, sent with one copy on the
// context's stream (the solve that follows orders every later use of it).
int ranges_upload(ispan_ctx* ctx, BatchRanges& r, int32_t* capN, int64_t* capC, int32_t n, const int64_t* vb,
                  const int64_t* ve, int32_t* const* out) {
    const size_t m = (size_t)std::max(n, 1) + 1;
    const size_t bytes = 4 * m * 8;  // vb, ve, cstart (int64) and out (pointers), m entries each
    if (bytes > ctx->stage_bytes) {
        CK(hipStreamSynchronize(ctx->stream));  // a pending copy may still read the old area
        if (ctx->stage) (void)hipHostFree(ctx->stage);
        ctx->stage = nullptr;
        ctx->stage_bytes = 0;
        CK(hipHostMalloc(&ctx->stage, bytes, hipHostMallocDefault));
        ctx->stage_bytes = bytes;
    } else {
        CK(hipStreamSynchronize(ctx->stream));  // the previous upload has been read
    }
    int64_t* h = static_cast<int64_t*>(ctx->stage);
    int64_t* hcs = h + 2 * m;
    hcs[0] = 0;
    for (int32_t g = 0; g < n; ++g) {
        h[g] = vb[g];
        h[m + g] = ve[g];
        hcs[g + 1] = hcs[g] + (ve[g] - vb[g] + kBatchChunk - 1) / kBatchChunk;
        reinterpret_cast<int32_t**>(h + 3 * m)[g] = out[g];
    }
    const int64_t nc = hcs[n];
    const bool growN = !capN || n > *capN, growC = !capC || nc > *capC;
    if (growN) {
        dfree(r.vb);  // vb owns the packed block: vb, ve, cstart, out
        dfree(r.roots);
        dfree(r.gmin);
        CK(hipMalloc(&r.vb, bytes));
        CK(hipMalloc(&r.roots, m * 4));
        CK(hipMalloc(&r.gmin, m * 4));
        if (capN) *capN = n;
    }
    r.ve = r.vb + m;
    r.cstart = r.vb + 2 * m;
    r.out = reinterpret_cast<int32_t**>(r.vb + 3 * m);
    if (growC) {
        dfree(r.cp);
        dfree(r.ci);
        const size_t c = (size_t)std::max<int64_t>(nc, 1);
        CK(hipMalloc(&r.cp, c * 8));
        CK(hipMalloc(&r.ci, c * 4));
        if (capC) *capC = nc;
    }
    CK(hipMemcpyAsync(r.vb, h, bytes, hipMemcpyHostToDevice, ctx->stream));
    r.n = n;
    r.nchunks = nc;
    return ISPAN_OK;
}

// One pass of the pipeline over v, seeded with the pivot of every range in
// r, each range's local labels into its output.  sel (optional) limits the
// solve to the ranges' vertices.
int solve_view(ispan_ctx* ctx, const View& v, const BatchRanges& r, const uint64_t* sel, const ispan_params* P,
               ispan_stats* st) {
    if (v.V == 0 || r.n == 0) return ISPAN_OK;
    use_view(ctx, v);
    ctx->batch_roots = r.roots;
    ctx->batch_nroots = r.n;
    ctx->batch_sel = sel;
    int rc = ISPAN_OK;
    hipError_t e = launch_batch_roots(v.csr[0], v.csr[2], r, P->pivot_wrap32, ctx->stream);
    if (e != hipSuccess) {
        ctx->err = std::string("launch_batch_roots: ") + hipGetErrorString(e);
        rc = ISPAN_ERR_HIP;
    }
    if (!rc) rc = solve_impl(ctx, P, nullptr, st);
    if (!rc) {
        e = launch_batch_finish(ctx->lab, ctx->bvis, r, ctx->stream);
        if (e != hipSuccess) {
            ctx->err = std::string("launch_batch_finish: ") + hipGetErrorString(e);
            rc = ISPAN_ERR_HIP;
        }
    }
    drop_group_view(ctx);
    ctx->have_solution = false;
    return rc;
}

// Size class of a graph of V vertices: floor(log2 V), 0 for V <= 1.
int size_class(int64_t V) {
    int k = 0;
    while (k < ISPAN_SIZE_CLASSES - 1 && (V >> (k + 1)) > 0) ++k;
    return k;
}

// Builds every group of a batch; on failure the caller frees the batch.
// Groups the batch's graphs into unions of at most cap edges (and int32
// vertex ids).  Without by_class a group is a run of consecutive graphs;
// with it the graphs are ordered by size class first (stably) and then
// packed the same way, so each group spans a narrow range of sizes (its
// class is that of its largest graph) and its union is built from scattered
// arrays.  Groups are not split at class boundaries: every group costs a
// launch and a full level structure (measured ~0.37 ms per group on the
// 4096-graph batch), which splitting 3 groups into 12 did not pay back.
int batch_build(ispan_ctx* ctx, int32_t n, const int64_t* v_off, const int64_t* e_off, const int32_t* fwb_cat,
                const int32_t* fwa_cat, const int32_t* bwb_cat, const int32_t* bwa_cat, int64_t cap,
                const ispan_params* P, bool by_class) {
    hipStream_t s = ctx->stream;
    int64_t* d_off = nullptr;
    int32_t* d_bad = nullptr;
    int rc = ISPAN_OK;
    int32_t maxV = 0;
    int64_t maxE = 0;
    auto hip = [&](hipError_t e, const char* what) {
        if (e != hipSuccess && rc == ISPAN_OK) {
            ctx->err = std::string(what) + ": " + hipGetErrorString(e);
            rc = e == hipErrorOutOfMemory ? ISPAN_ERR_NOMEM : ISPAN_ERR_HIP;
        }
        return rc == ISPAN_OK;
    };
    if (!hip(hipMalloc(&d_off, (size_t)2 * (n + 1) * 8), "hipMalloc") ||
        !hip(hipMalloc(&d_bad, 4), "hipMalloc") ||
        !hip(hipMemcpyAsync(d_off, v_off, (size_t)(n + 1) * 8, hipMemcpyHostToDevice, s), "hipMemcpyAsync") ||
        !hip(hipMemcpyAsync(d_off + n + 1, e_off, (size_t)(n + 1) * 8, hipMemcpyHostToDevice, s), "hipMemcpyAsync") ||
        !hip(hipMemsetAsync(d_bad, 0, 4, s), "hipMemsetAsync")) {
        dfree(d_off);
        dfree(d_bad);
        return rc;
    }
    hip(hipMalloc(&ctx->batch_labels, (size_t)std::max<int64_t>(v_off[n], 1) * 4), "hipMalloc");
    std::vector<int32_t> order(n);
    for (int32_t g = 0; g < n; ++g) order[g] = g;
    auto cls_of = [&](int32_t g) { return size_class(v_off[g + 1] - v_off[g]); };
    if (by_class) std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) { return cls_of(x) < cls_of(y); });
    int64_t* d_perm = nullptr;  // a scattered group's graph table: lv_off, le_off, sv, se (int64), sg (int32)
    for (int32_t p0 = 0; p0 < n && rc == ISPAN_OK;) {
        int32_t p1 = p0 + 1;
        int64_t sV = v_off[order[p0] + 1] - v_off[order[p0]], sE = e_off[order[p0] + 1] - e_off[order[p0]];
        while (p1 < n) {
            const int32_t g = order[p1];
            const int64_t Vg = v_off[g + 1] - v_off[g], Eg = e_off[g + 1] - e_off[g];
            if (sE + Eg > cap || sV + Vg > INT32_MAX - 64) break;
            sV += Vg;
            sE += Eg;
            ++p1;
        }
        bool contiguous = true;
        for (int32_t p = p0 + 1; p < p1; ++p) contiguous &= order[p] == order[p - 1] + 1;
        ctx->groups.emplace_back();
        ispan_ctx::BatchGroup& G = ctx->groups.back();
        G.g0 = order[p0];
        G.ng = p1 - p0;
        G.V = (int32_t)sV;
        G.E = sE;
        maxV = std::max(maxV, G.V);
        maxE = std::max(maxE, G.E);
        const size_t nb = ((size_t)G.V + 1) * 4, na = (size_t)std::max<int64_t>(G.E, 1) * 4;
        std::vector<int64_t> vb(G.ng), ve(G.ng);
        std::vector<int32_t*> outp(G.ng);
        int64_t acc = 0, big = 0;
        for (int32_t i = 0; i < G.ng; ++i) {
            const int32_t g = order[p0 + i];
            vb[i] = acc;
            acc += v_off[g + 1] - v_off[g];
            ve[i] = acc;
            outp[i] = ctx->batch_labels + v_off[g];
            big = std::max(big, v_off[g + 1] - v_off[g]);
        }
        G.cls = size_class(big);
        if (!hip(hipMalloc(&G.csr[0], nb), "hipMalloc") || !hip(hipMalloc(&G.csr[1], na), "hipMalloc") ||
            !hip(hipMalloc(&G.csr[2], nb), "hipMalloc") || !hip(hipMalloc(&G.csr[3], na), "hipMalloc") ||
            !hip(hipMalloc(&G.first[0], (size_t)std::max(G.V, 1) * 4), "hipMalloc") ||
            !hip(hipMalloc(&G.first[1], (size_t)std::max(G.V, 1) * 4), "hipMalloc"))
            break;
        if (contiguous) {
            if (!hip(launch_batch_union(fwb_cat, fwa_cat, bwb_cat, bwa_cat, d_off, d_off + n + 1, G.g0, G.ng, G.V,
                                        G.E, G.csr[0], G.csr[1], G.csr[2], G.csr[3], d_bad, s),
                     "launch_batch_union"))
                break;
        } else {
            const int32_t m = G.ng;
            std::vector<int64_t> tab((size_t)4 * (m + 1) + (m + 2) / 2, 0);
            int64_t* lv = tab.data();
            int64_t* le = lv + m + 1;
            int64_t* sv = le + m + 1;
            int64_t* se = sv + m + 1;
            int32_t* sg = reinterpret_cast<int32_t*>(se + m + 1);
            for (int32_t i = 0; i < m; ++i) {
                const int32_t g = order[p0 + i];
                lv[i + 1] = lv[i] + (v_off[g + 1] - v_off[g]);
                le[i + 1] = le[i] + (e_off[g + 1] - e_off[g]);
                sv[i] = v_off[g];
                se[i] = e_off[g];
                sg[i] = g;
            }
            // the previous group's union is built (its check below waited)
            dfree(d_perm);
            if (!hip(hipMalloc(&d_perm, tab.size() * 8), "hipMalloc") ||
                !hip(hipMemcpyAsync(d_perm, tab.data(), tab.size() * 8, hipMemcpyHostToDevice, s), "hipMemcpyAsync") ||
                !hip(launch_batch_union_perm(fwb_cat, fwa_cat, bwb_cat, bwa_cat, d_perm, d_perm + m + 1,
                                             d_perm + 2 * (m + 1), d_perm + 3 * (m + 1),
                                             reinterpret_cast<const int32_t*>(d_perm + 4 * (m + 1)), m, G.V, G.E,
                                             G.csr[0], G.csr[1], G.csr[2], G.csr[3], d_bad, s),
                     "launch_batch_union_perm"))
                break;
        }
        int32_t bad = 0;
        if (!hip(hipMemcpyAsync(&bad, d_bad, 4, hipMemcpyDeviceToHost, s), "hipMemcpyAsync") ||
            !hip(hipStreamSynchronize(s), "hipStreamSynchronize"))
            break;
        if (bad) {
            ctx->err = (bad & 1) ? "batch: a graph's offsets do not run from 0 to its edge count"
                                 : "batch: an adjacency id lies outside its graph";
            rc = ISPAN_ERR_CSR;
            break;
        }
        use_view(ctx, group_view(G));
        rc = check_bound(ctx, P);
        drop_group_view(ctx);
        if (rc) break;
        rc = ranges_upload(ctx, G.rb, nullptr, nullptr, G.ng, vb.data(), ve.data(), outp.data());
        if (rc) break;
        // first arrays only once the offsets are known to be in range
        if (G.V > 0 && (!hip(launch_first(G.csr[0], G.csr[1], G.V, G.first[0], ctx->grid, s), "launch_first") ||
                        !hip(launch_first(G.csr[2], G.csr[3], G.V, G.first[1], ctx->grid, s), "launch_first")))
            break;
        p0 = p1;
    }
    if (rc == ISPAN_OK) rc = ensure_buffers(ctx, maxV, maxE);
    (void)hipStreamSynchronize(s);
    dfree(d_perm);
    dfree(d_off);
    dfree(d_bad);
    return rc;
}

void add_stats(ispan_stats& t, const ispan_stats& st) {
    for (int k = 0; k < 16; ++k) t.ms[k] += st.ms[k];
    t.fq0 += st.fq0;
    t.fw_size += st.fw_size;
    t.giant_size += st.giant_size;
    t.trim1 += st.trim1;
    t.trim2 += st.trim2;
    t.trim3 += st.trim3;
    t.n_wcc += st.n_wcc;
    t.mice += st.mice;
    t.fw_levels = std::max(t.fw_levels, st.fw_levels);
    t.bw_levels = std::max(t.bw_levels, st.bw_levels);
    t.host_syncs += st.host_syncs;
}

}  // namespace

}  // extern "C"

namespace ispan {

int check_csr_device(int32_t V, int64_t E, const int32_t* fwb, const int32_t* fwa, const int32_t* bwb,
                     const int32_t* bwa, hipStream_t s) {
    Ctl* d = nullptr;
    Ctl h{};
    int rc = ISPAN_OK;
    if (hipMalloc(&d, sizeof(Ctl)) != hipSuccess) return ISPAN_ERR_NOMEM;
    int dev = 0, cu = 1;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev);
    if (launch_reset_ctl(d, s) != hipSuccess ||
        (V > 0 && (launch_check_csr(fwb, fwa, V, E, d, cu * 8, s) != hipSuccess ||
                   launch_check_csr(bwb, bwa, V, E, d, cu * 8, s) != hipSuccess)) ||
        hipMemcpyAsync(&h, d, sizeof(Ctl), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        rc = ISPAN_ERR_HIP;
    else if (h.bad_csr)
        rc = ISPAN_ERR_CSR;
    (void)hipFree(d);
    return rc;
}

int64_t solve_state_bytes(int64_t V, int64_t E) {
    // per-vertex solve state (labels, bitmaps, queues, WCC and mice arrays)
    // and the hub work items of both kinds of traversal
    return 88 * (V + 1) + 2 * (2 * giant_hub_capacity(E) + 2 * (E / kHubChunk) + 64) * (int64_t)sizeof(HubItem);
}

int solve_arena(ispan_ctx* ctx, const ArenaView& av, int32_t n, const int64_t* vb, const int64_t* ve,
                int32_t* const* out, const ispan_params* P, ispan_stats* st) {
    if (!ctx || n < 1 || av.V < 0) return ISPAN_ERR_ARG;
    if (ctx->batch_n >= 0 || ctx->V >= 0) free_graph(ctx);
    int rc = ranges_upload(ctx, ctx->ws_rb, &ctx->ws_capN, &ctx->ws_capC, n, vb, ve, out);
    if (rc) return rc;
    const int64_t nwords = ((int64_t)av.V + 63) / 64;
    if (nwords > ctx->sel_cap) {
        dfree(ctx->sel);
        CK(hipMalloc(&ctx->sel, (size_t)std::max<int64_t>(nwords, 1) * 8));
        ctx->sel_cap = nwords;
    }
    CK(launch_batch_select(ctx->ws_rb, ctx->sel, nwords, ctx->stream));
    const View v{{av.fwb, av.fwa, av.bwb, av.bwa}, {av.fw_first, av.bw_first}, av.V, av.E};
    ispan_stats local{};
    rc = solve_view(ctx, v, ctx->ws_rb, ctx->sel, P, &local);
    if (st) *st = local;
    return rc;
}

}  // namespace ispan

extern "C" {

int ispan_batch_bind_device(ispan_ctx* ctx, int32_t n_graphs, const int64_t* h_v_off, const int64_t* h_e_off,
                            const int32_t* d_fw_beg, const int32_t* d_fw_adj, const int32_t* d_bw_beg,
                            const int32_t* d_bw_adj, int64_t max_group_edges, const ispan_params* params) {
    if (!ctx || n_graphs < 1 || !h_v_off || !h_e_off || !d_fw_beg || !d_bw_beg || max_group_edges < 0)
        return ISPAN_ERR_ARG;
    if (h_v_off[0] != 0 || h_e_off[0] != 0) return ISPAN_ERR_ARG;
    for (int32_t g = 0; g < n_graphs; ++g) {
        const int64_t Vg = h_v_off[g + 1] - h_v_off[g], Eg = h_e_off[g + 1] - h_e_off[g];
        if (Vg < 0 || Eg < 0 || Vg > INT32_MAX - 64 || Eg > INT32_MAX) return ISPAN_ERR_ARG;
    }
    if (h_e_off[n_graphs] > 0 && (!d_fw_adj || !d_bw_adj)) return ISPAN_ERR_ARG;
    CK(hipSetDevice(ctx->device));
    free_graph(ctx);
    ispan_params d;
    ispan_default_params(&d);
    const int64_t cap = max_group_edges > 0 ? std::min<int64_t>(max_group_edges, INT32_MAX) : INT32_MAX;
    const ispan_params* P = params ? params : &d;
    int rc = batch_build(ctx, n_graphs, h_v_off, h_e_off, d_fw_beg, d_fw_adj, d_bw_beg, d_bw_adj, cap, P,
                         P->batch_by_class != 0);
    if (rc) {
        free_graph(ctx);
        return rc;
    }
    ctx->batch_n = n_graphs;
    ctx->batch_v_off.assign(h_v_off, h_v_off + n_graphs + 1);
    return ISPAN_OK;
}

int ispan_batch_groups(const ispan_ctx* ctx) { return ctx && ctx->batch_n >= 0 ? (int)ctx->groups.size() : 0; }

int ispan_batch_solve(ispan_ctx* ctx, const ispan_params* params, ispan_stats* stats) {
    if (!ctx) return ISPAN_ERR_ARG;
    if (ctx->batch_n < 0) return ISPAN_ERR_NOGRAPH;
    ispan_params d;
    ispan_default_params(&d);
    const ispan_params* P = params ? params : &d;
    if (P->alpha <= 0 || P->beta <= 0 || P->trim_times < 0) return ISPAN_ERR_ARG;
    CK(hipSetDevice(ctx->device));
    const double t0 = wall_ms();
    ispan_stats tot{};
    ctx->batch_solved = false;
    for (const auto& G : ctx->groups) {
        ispan_stats st{};
        const int rc = solve_view(ctx, group_view(G), G.rb, nullptr, P, &st);
        if (rc) return rc;
        add_stats(tot, st);
    }
    CK(hipStreamSynchronize(ctx->stream));
    tot.ms[3] = wall_ms() - t0;
    tot.pivot = -1;
    tot.engine_used = 0;
    ctx->batch_solved = true;
    if (stats) *stats = tot;
    return ISPAN_OK;
}

int ispan_autotune(ispan_ctx* ctx, const ispan_params* candidates, int32_t n, int32_t reps, int32_t* best,
                   double* ms) {
    if (!ctx || !candidates || n < 1 || reps < 1 || !best) return ISPAN_ERR_ARG;
    const bool batch = ctx->batch_n >= 0;
    if (!batch && ctx->V < 0) return ISPAN_ERR_NOGRAPH;
    CK(hipSetDevice(ctx->device));
    double best_ms = 0.0;
    *best = -1;
    for (int32_t i = 0; i < n; ++i) {
        const ispan_params& P = candidates[i];
        if (P.alpha <= 0 || P.beta <= 0 || P.trim_times < 0) return ISPAN_ERR_ARG;
        std::vector<double> t;
        for (int32_t r = 0; r <= reps; ++r) {
            const double t0 = wall_ms();
            const int rc = batch ? ispan_batch_solve(ctx, &P, nullptr) : solve_impl(ctx, &P, nullptr, nullptr);
            if (rc) return rc;
            CK(hipStreamSynchronize(ctx->stream));
            if (r > 0) t.push_back(wall_ms() - t0);  // the first run warms up
        }
        std::sort(t.begin(), t.end());
        const double med = t[t.size() / 2];
        if (ms) ms[i] = med;
        if (*best < 0 || med < best_ms) {
            *best = i;
            best_ms = med;
        }
    }
    return ISPAN_OK;
}

void ispan_tuning_lookup(const ispan_tuning_table* table, int64_t V, ispan_params* out) {
    if (!out) return;
    ispan_default_params(out);
    if (!table) return;
    const int k = size_class(V);
    if (table->tuned[k]) *out = table->params[k];
}

int ispan_autotune_classes(ispan_ctx* ctx, const ispan_params* candidates, int32_t n, int32_t reps,
                           ispan_tuning_table* table) {
    if (!ctx || !candidates || n < 1 || reps < 1 || !table) return ISPAN_ERR_ARG;
    const bool batch = ctx->batch_n >= 0;
    if (!batch && ctx->V < 0) return ISPAN_ERR_NOGRAPH;
    for (int32_t i = 0; i < n; ++i)
        if (candidates[i].alpha <= 0 || candidates[i].beta <= 0 || candidates[i].trim_times < 0) return ISPAN_ERR_ARG;
    CK(hipSetDevice(ctx->device));
    // t[k][i]: candidate i's median time over class k (summed over its groups)
    std::vector<std::vector<double>> t(ISPAN_SIZE_CLASSES, std::vector<double>(n, 0.0));
    std::vector<bool> seen(ISPAN_SIZE_CLASSES, false);
    auto time_one = [&](const ispan_params& P, auto&& solve_once, double& med) -> int {
        std::vector<double> x;
        for (int32_t r = 0; r <= reps; ++r) {
            const double t0 = wall_ms();
            const int rc = solve_once(P);
            if (rc) return rc;
            CK(hipStreamSynchronize(ctx->stream));
            if (r > 0) x.push_back(wall_ms() - t0);  // the first run warms up
        }
        std::sort(x.begin(), x.end());
        med = x[x.size() / 2];
        return ISPAN_OK;
    };
    if (batch) {
        for (const auto& G : ctx->groups) {
            seen[G.cls] = true;
            for (int32_t i = 0; i < n; ++i) {
                double med = 0.0;
                const int rc = time_one(candidates[i], [&](const ispan_params& P) {
                    return solve_view(ctx, group_view(G), G.rb, nullptr, &P, nullptr);
                }, med);
                if (rc) return rc;
                t[G.cls][i] += med;
            }
        }
        ctx->batch_solved = false;  // only some groups were solved last
    } else {
        const int k = size_class(ctx->V);
        seen[k] = true;
        for (int32_t i = 0; i < n; ++i) {
            double med = 0.0;
            const int rc = time_one(candidates[i], [&](const ispan_params& P) {
                return solve_impl(ctx, &P, nullptr, nullptr);
            }, med);
            if (rc) return rc;
            t[k][i] = med;
        }
    }
    for (int k = 0; k < ISPAN_SIZE_CLASSES; ++k) {
        if (!seen[k]) continue;
        const int32_t b = (int32_t)(std::min_element(t[k].begin(), t[k].end()) - t[k].begin());
        table->tuned[k] = 1;
        table->params[k] = candidates[b];
        table->params[k].timing = 0;
        table->ms[k] = t[k][b];
    }
    return ISPAN_OK;
}

int ispan_batch_solve_tuned(ispan_ctx* ctx, const ispan_tuning_table* table, ispan_stats* stats) {
    if (!ctx || !table) return ISPAN_ERR_ARG;
    if (ctx->batch_n < 0) return ISPAN_ERR_NOGRAPH;
    CK(hipSetDevice(ctx->device));
    const double t0 = wall_ms();
    ispan_stats tot{};
    ctx->batch_solved = false;
    ispan_params d;
    ispan_default_params(&d);
    for (const auto& G : ctx->groups) {
        const ispan_params* P = table->tuned[G.cls] ? &table->params[G.cls] : &d;
        if (P->alpha <= 0 || P->beta <= 0 || P->trim_times < 0) return ISPAN_ERR_ARG;
        ispan_stats st{};
        const int rc = solve_view(ctx, group_view(G), G.rb, nullptr, P, &st);
        if (rc) return rc;
        add_stats(tot, st);
    }
    CK(hipStreamSynchronize(ctx->stream));
    tot.ms[3] = wall_ms() - t0;
    tot.pivot = -1;
    tot.engine_used = 0;
    ctx->batch_solved = true;
    if (stats) *stats = tot;
    return ISPAN_OK;
}

// One line per tuned class: class, then the fields a tuning run chooses.
int ispan_tuning_save(const ispan_tuning_table* table, const char* path) {
    if (!table || !path) return ISPAN_ERR_ARG;
    FILE* f = fopen(path, "w");
    if (!f) return ISPAN_ERR_IO;
    fprintf(f, "# ispan tuning table: class alpha beta gamma theta trim_times wcc_mode trim_scope switch_rule "
               "engine ms\n");
    for (int k = 0; k < ISPAN_SIZE_CLASSES; ++k) {
        if (!table->tuned[k]) continue;
        const ispan_params& p = table->params[k];
        fprintf(f, "%d %d %d %d %.17g %d %d %d %d %d %.6f\n", k, p.alpha, p.beta, p.gamma, p.theta, p.trim_times,
                p.wcc_mode, p.trim_scope, p.switch_rule, p.engine, table->ms[k]);
    }
    const bool ok = fclose(f) == 0;
    return ok ? ISPAN_OK : ISPAN_ERR_IO;
}

int ispan_tuning_load(const char* path, ispan_tuning_table* table) {
    if (!table || !path) return ISPAN_ERR_ARG;
    FILE* f = fopen(path, "r");
    if (!f) return ISPAN_ERR_IO;
    ispan_tuning_table t{};
    char line[512];
    int rc = ISPAN_OK;
    while (fgets(line, sizeof line, f)) {
        if (line[0] == '#' || line[0] == '\n') continue;
        int k = -1;
        ispan_params p;
        ispan_default_params(&p);
        double ms = 0.0;
        if (sscanf(line, "%d %d %d %d %lf %d %d %d %d %d %lf", &k, &p.alpha, &p.beta, &p.gamma, &p.theta,
                   &p.trim_times, &p.wcc_mode, &p.trim_scope, &p.switch_rule, &p.engine, &ms) != 11 ||
            k < 0 || k >= ISPAN_SIZE_CLASSES || p.alpha <= 0 || p.beta <= 0 || p.trim_times < 0) {
            rc = ISPAN_ERR_IO;
            break;
        }
        t.tuned[k] = 1;
        t.params[k] = p;
        t.ms[k] = ms;
    }
    fclose(f);
    if (rc == ISPAN_OK) *table = t;
    return rc;
}

int ispan_batch_labels_download(ispan_ctx* ctx, int32_t* labels) {
    if (!ctx || !labels) return ISPAN_ERR_ARG;
    if (!ctx->batch_solved) return ISPAN_ERR_NOGRAPH;
    const int64_t n = ctx->batch_v_off.back();
    if (n == 0) return ISPAN_OK;
    CK(hipMemcpyAsync(labels, ctx->batch_labels, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    return ISPAN_OK;
}

const int32_t* ispan_batch_labels_device(const ispan_ctx* ctx) {
    return (ctx && ctx->batch_solved) ? ctx->batch_labels : nullptr;
}

int ispan_scc_detection(int32_t V, int64_t E, const int32_t* fw_beg, const int32_t* fw_adj, const int32_t* bw_beg,
                        const int32_t* bw_adj, const ispan_params* params, int32_t* labels, ispan_stats* stats) {
    if (!labels && V > 0) return ISPAN_ERR_ARG;
    ispan_ctx* ctx = nullptr;
    int rc = ispan_ctx_create(&ctx, 0);
    if (rc) return rc;
    rc = ispan_graph_upload(ctx, V, E, fw_beg, fw_adj, bw_beg, bw_adj, params);
    if (!rc) rc = ispan_solve(ctx, params, nullptr, stats);
    if (!rc) rc = ispan_labels_download(ctx, labels);
    ispan_ctx_destroy(ctx);
    return rc;
}

// ---- allocation-free one-shot solve ---------------------------------------
// Every buffer of ispan_scc_detection_noalloc's solve lives in this array:
// one copy per GPU, held until the process ends.  In host code the name is
// a host-side stand-in, never a device address: the device address comes
// from hipGetSymbolAddress, taken once per device and kept below.
constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxScratchDevices = 16;

struct ScratchDevice {
    bool ready = false;
    unsigned char* ptr = nullptr;  // this GPU's `scratch` (device address)
    int cu = 0, grid = 0, coop_grid = 0;
    double clock_khz = 0.0;
    Ctl h_ctl;                     // host copies of the control block and the
    GiantState h_gstate;           // giant kernel's state, read after each wait
};
std::mutex g_scratch_mu;
ScratchDevice g_scratch[kMaxScratchDevices];

// Set-up of device `dev` (current): the address of its copy of `scratch`
// and its launch geometry, on its first call.
int scratch_device(int dev, ScratchDevice** out) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    ScratchDevice& sd = g_scratch[dev];
    if (!sd.ready) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return ISPAN_ERR_HIP;
        int cu = 0, coop = 0, rate = 0;
        if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return ISPAN_ERR_HIP;
        (void)hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev);
        (void)hipDeviceGetAttribute(&rate, hipDeviceAttributeWallClockRate, dev);
        sd.cu = cu;
        sd.grid = std::max(cu, 1) * 8;  // as ispan_ctx_create
        sd.coop_grid = (coop ? std::min(giant_max_blocks_per_cu(), 8) : 0) * cu;
        sd.clock_khz = rate;
        sd.ptr = static_cast<unsigned char*>(p);
        sd.ready = true;
    }
    *out = &sd;
    return ISPAN_OK;
}

// Places the buffers ensure_buffers would allocate (same sizes), the control
// block, the giant kernel's state and the first-neighbour arrays of bind
// one after another from base, 256-byte aligned; with base == nullptr only
// sizes them.  Returns the bytes used.
size_t carve_solve_state(ispan_ctx* c, unsigned char* base, int32_t V, int64_t E, size_t sort_bytes) {
    size_t off = 0;
    auto take = [&](auto*& p, size_t bytes) {
        using T = std::remove_reference_t<decltype(p)>;
        p = base ? reinterpret_cast<T>(base + off) : nullptr;
        off += (bytes + 255) & ~(size_t)255;
    };
    const size_t nv = (size_t)std::max(V, 1) + 1;
    const size_t nw = (size_t)((V + 63) / 64) + 1;
    take(c->lab, nv * 4);
    take(c->act, nw * 8);
    take(c->fvis, nw * 8);
    take(c->bvis, nw * 8);
    for (int i = 0; i < 2; ++i) {
        take(c->fq[i], nv * 8);
        take(c->q[i], nv * 8);
        take(c->hub[i], (size_t)hub_capacity(E) * sizeof(HubItem));
    }
    take(c->parent, nv * 4);
    take(c->stamp, nv * 4);
    take(c->keys, nv * 8);
    take(c->keys_alt, nv * 8);
    take(c->starts, nv * 4);
    take(c->qbuf, nv * 4);
    take(c->part, (size_t)c->grid * sizeof(PivotPart));
    take(c->snap, nw * 8);
    for (int i = 0; i < 2; ++i) {
        take(c->gh[i], (size_t)giant_hub_capacity(E) * sizeof(HubItem));
        take(c->ghb[i], (size_t)giant_hub_capacity(E) * sizeof(HubItem));
    }
    take(c->snapb, nw * 8);
    take(c->gpart, (size_t)kGiantSlots * std::max(c->coop_grid, 1) * 8);
    take(c->gdbg, (size_t)2 * std::max(c->coop_grid, 1) * 4);
    unsigned char* st = nullptr;
    take(st, sort_bytes + 256);
    c->sort_tmp = st;
    c->sort_tmp_bytes = sort_bytes + 256;
    take(c->ctl, sizeof(Ctl));
    take(c->gstate, sizeof(GiantState));
    take(c->fw_first, (size_t)std::max(V, 1) * 4);
    take(c->bw_first, (size_t)std::max(V, 1) * 4);
    return off;
}

}  // namespace

int ispan_scc_detection_noalloc(int32_t V, int64_t E, const int32_t* d_fw_beg, const int32_t* d_fw_adj,
                                const int32_t* d_bw_beg, const int32_t* d_bw_adj, const ispan_params* params,
                                int32_t* d_labels, void* hip_stream, ispan_stats* stats) {
    if (V < 0 || V == INT32_MAX || E < 0 || E > INT32_MAX) return ISPAN_ERR_ARG;
    if (V > 0 && (!d_fw_beg || !d_bw_beg || !d_labels)) return ISPAN_ERR_ARG;
    if (E > 0 && (!d_fw_adj || !d_bw_adj)) return ISPAN_ERR_ARG;
    ispan_params P;
    ispan_default_params(&P);
    if (params) P = *params;
    if (P.alpha <= 0 || P.beta <= 0 || P.trim_times < 0) return ISPAN_ERR_ARG;
    P.timing = 0;  // phase timing would need events created
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxScratchDevices) return ISPAN_ERR_HIP;
    ScratchDevice* sd = nullptr;
    int rc = scratch_device(dev, &sd);
    if (rc) return rc;
    ispan_ctx c;
    c.device = dev;
    c.stream = static_cast<hipStream_t>(hip_stream);
    c.cu = sd->cu;
    c.grid = sd->grid;
    c.coop_grid = sd->coop_grid;
    c.clock_khz = sd->clock_khz;
    size_t sort_bytes = 0;
    if (rocprim::radix_sort_keys(nullptr, sort_bytes, static_cast<uint64_t*>(nullptr), static_cast<uint64_t*>(nullptr),
                                 (size_t)std::max(V, 1) + 1, 0, 64, c.stream) != hipSuccess)
        return ISPAN_ERR_HIP;
    if (carve_solve_state(&c, nullptr, V, E, sort_bytes) > kScratchBytes) return ISPAN_ERR_SCRATCH;
    carve_solve_state(&c, sd->ptr, V, E, sort_bytes);
    c.cap_V = V;  // the buffers are in place: ensure_buffers allocates nothing
    c.cap_E = E;
    c.h_ctl = &sd->h_ctl;
    c.h_gstate = &sd->h_gstate;
    c.V = V;
    c.E = E;
    c.nwords = ((int64_t)V + 63) / 64;
    c.fwb = d_fw_beg;
    c.fwa = d_fw_adj;
    c.bwb = d_bw_beg;
    c.bwa = d_bw_adj;
    if ((rc = check_bound(&c, &P))) return rc;
    if (V > 0) {
        // the kernel's state sits wherever this call's layout puts it: its
        // epoch word must not hold a value a launch is still to publish
        if (launch_fill_i32(reinterpret_cast<int32_t*>(&c.gstate->epoch), 0, 2, 1, c.stream) != hipSuccess ||
            launch_first(c.fwb, c.fwa, V, c.fw_first, c.grid, c.stream) != hipSuccess ||
            launch_first(c.bwb, c.bwa, V, c.bw_first, c.grid, c.stream) != hipSuccess)
            return ISPAN_ERR_HIP;
    }
    return solve_impl(&c, &P, d_labels, stats);
}

int ispan_graph_load(const char* fw_beg_file, const char* fw_adj_file, const char* bw_beg_file,
                     const char* bw_adj_file, ispan_host_graph* out) {
    if (!fw_beg_file || !fw_adj_file || !bw_beg_file || !bw_adj_file || !out) return ISPAN_ERR_ARG;
    memset(out, 0, sizeof(*out));
    const int64_t sb = file_size(fw_beg_file), sa = file_size(fw_adj_file);
    const int64_t sb2 = file_size(bw_beg_file), sa2 = file_size(bw_adj_file);
    if (sb < 4 || sa < 0 || sb2 != sb || sa2 != sa || sb % 4 || sa % 4) return ISPAN_ERR_IO;
    const int64_t V = sb / 4 - 1, E = sa / 4;
    if (V > INT32_MAX - 1 || E > INT32_MAX) return ISPAN_ERR_IO;
    out->V = (int32_t)V;
    out->E = E;
    out->fw_beg = (int32_t*)malloc((size_t)(V + 1) * 4);
    out->bw_beg = (int32_t*)malloc((size_t)(V + 1) * 4);
    out->fw_adj = (int32_t*)malloc((size_t)std::max<int64_t>(E, 1) * 4);
    out->bw_adj = (int32_t*)malloc((size_t)std::max<int64_t>(E, 1) * 4);
    if (!out->fw_beg || !out->bw_beg || !out->fw_adj || !out->bw_adj) {
        ispan_graph_free(out);
        return ISPAN_ERR_NOMEM;
    }
    int rc = read_i32(fw_beg_file, V + 1, out->fw_beg);
    if (!rc) rc = read_i32(fw_adj_file, E, out->fw_adj);
    if (!rc) rc = read_i32(bw_beg_file, V + 1, out->bw_beg);
    if (!rc) rc = read_i32(bw_adj_file, E, out->bw_adj);
    if (rc) ispan_graph_free(out);
    return rc;
}

void ispan_graph_free(ispan_host_graph* g) {
    if (!g) return;
    free(g->fw_beg);
    free(g->fw_adj);
    free(g->bw_beg);
    free(g->bw_adj);
    memset(g, 0, sizeof(*g));
}

int ispan_csr_check(int32_t V, int64_t E, const int32_t* fw_beg, const int32_t* fw_adj, const int32_t* bw_beg,
                    const int32_t* bw_adj) {
    if (V < 0 || E < 0 || !fw_beg || !bw_beg || (E > 0 && (!fw_adj || !bw_adj))) return ISPAN_ERR_ARG;
    for (int d = 0; d < 2; ++d) {
        const int32_t* b = d ? bw_beg : fw_beg;
        const int32_t* a = d ? bw_adj : fw_adj;
        if (b[0] != 0 || (int64_t)b[V] != E) return ISPAN_ERR_CSR;
        for (int32_t v = 0; v < V; ++v)
            if (b[v] > b[v + 1]) return ISPAN_ERR_CSR;
        for (int64_t k = 0; k < E; ++k)
            if (a[k] < 0 || a[k] >= V) return ISPAN_ERR_CSR;
    }
    return ISPAN_OK;
}

int ispan_summarize(int32_t V, const int32_t* labels, ispan_summary* out) {
    if (V < 0 || !out || (V > 0 && !labels)) return ISPAN_ERR_ARG;
    memset(out, 0, sizeof(*out));
    std::vector<int32_t> size((size_t)V, 0);
    for (int32_t v = 0; v < V; ++v) {
        const int32_t l = labels[v];
        if (l < 0 || l > v) return ISPAN_ERR_ARG;  // canonical: label <= own id
        ++size[l];
    }
    int64_t largest = 0;
    for (int32_t v = 0; v < V; ++v) largest = std::max<int64_t>(largest, size[v]);
    bool largest_taken = false;
    for (int32_t v = 0; v < V; ++v) {
        const int32_t c = size[v];
        if (!c) continue;
        ++out->total;
        if (c == largest && !largest_taken && largest > 3) {
            largest_taken = true;
            continue;
        }
        if (c == 1) ++out->size_1;
        else if (c == 2) ++out->size_2;
        else if (c == 3) ++out->size_3;
        else ++out->others;
    }
    out->largest = largest;
    return ISPAN_OK;
}

int ispan_scc_histogram(ispan_ctx* ctx, int64_t* bins, int32_t n_bins, ispan_summary* summary) {
    if (!ctx || n_bins < 0 || n_bins > 64 || (n_bins > 0 && !bins)) return ISPAN_ERR_ARG;
    if (!ctx->have_solution) return ISPAN_ERR_NOGRAPH;
    int64_t h[kHistSlots] = {0};
    if (ctx->V > 0) {
        CK(hipSetDevice(ctx->device));
        if (!ctx->hist) CK(hipMalloc(&ctx->hist, sizeof(h)));
        // the mice queue buffer (int32[V + 1]) holds the sizes: the solve is over
        CK(launch_scc_histogram(ctx->lab, ctx->V, ctx->qbuf, ctx->hist, ctx->grid, ctx->stream));
        CK(hipMemcpyAsync(h, ctx->hist, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
        CK(hipStreamSynchronize(ctx->stream));
    }
    for (int k = 0; k < n_bins; ++k) bins[k] = h[k];
    for (int k = n_bins; k < 64 && n_bins > 0; ++k) bins[n_bins - 1] += h[k];  // the last bin takes every larger size
    if (summary) {
        // the classes of ispan_summarize: the largest SCC is counted apart
        // when it has more than 3 vertices
        summary->largest = h[64];
        summary->size_1 = h[65];
        summary->size_2 = h[66];
        summary->size_3 = h[67];
        summary->total = h[68];
        summary->others = h[68] - h[65] - h[66] - h[67] - (h[64] > 3 ? 1 : 0);
    }
    return ISPAN_OK;
}

static int gen_rc(int rc) {
    return rc == 0 ? ISPAN_OK : rc == -1 ? ISPAN_ERR_ARG : rc == -2 ? ISPAN_ERR_CSR : rc == -3 ? ISPAN_ERR_NOMEM
                                                                                               : ISPAN_ERR_HIP;
}

int ispan_gen_rmat_device(ispan_ctx* ctx, int32_t V, double edge_factor, double a, double b, double c, uint64_t seed,
                          int32_t** d_fw_beg, int32_t** d_fw_adj, int32_t** d_bw_beg, int32_t** d_bw_adj, int64_t* E) {
    if (!ctx || !d_fw_beg || !d_fw_adj || !d_bw_beg || !d_bw_adj || !E) return ISPAN_ERR_ARG;
    CK(hipSetDevice(ctx->device));
    return gen_rc(gen_rmat_device(ctx->stream, V, edge_factor, a, b, c, seed, d_fw_beg, d_fw_adj, d_bw_beg, d_bw_adj,
                                  E, ctx->err));
}

int ispan_gen_grid_device(ispan_ctx* ctx, int32_t n, double p_bidir, uint64_t seed, int32_t** d_fw_beg,
                          int32_t** d_fw_adj, int32_t** d_bw_beg, int32_t** d_bw_adj, int64_t* E) {
    if (!ctx || !d_fw_beg || !d_fw_adj || !d_bw_beg || !d_bw_adj || !E) return ISPAN_ERR_ARG;
    CK(hipSetDevice(ctx->device));
    return gen_rc(
        gen_grid_device(ctx->stream, n, p_bidir, seed, d_fw_beg, d_fw_adj, d_bw_beg, d_bw_adj, E, ctx->err));
}

int ispan_build_csr_device(ispan_ctx* ctx, int32_t V, int64_t m, const int32_t* d_src, const int32_t* d_dst,
                           int32_t** d_fw_beg, int32_t** d_fw_adj, int32_t** d_bw_beg, int32_t** d_bw_adj,
                           int64_t* E) {
    if (!ctx || (m > 0 && (!d_src || !d_dst)) || !d_fw_beg || !d_fw_adj || !d_bw_beg || !d_bw_adj || !E)
        return ISPAN_ERR_ARG;
    CK(hipSetDevice(ctx->device));
    return gen_rc(build_csr_device(ctx->stream, V, m, d_src, d_dst, d_fw_beg, d_fw_adj, d_bw_beg, d_bw_adj, E,
                                   ctx->err));
}

int ispan_kernel_profile_get(const ispan_ctx* ctx, ispan_kernel_profile* out) {
    if (!ctx || !out) return ISPAN_ERR_ARG;
    *out = ctx->prof;
    return ISPAN_OK;
}

// Text edge list -> the reference's binary files (graph_converter/
// txt_to_bin_fw_int.cpp:31-217 and txt_to_bin_bw_int.cpp).  Leading lines
// starting with '%' are skipped; ids are shifted by the smallest id.
static int parse_edge_list(const char* path, std::vector<int32_t>& src, std::vector<int32_t>& dst, int32_t* V,
                           std::string& err) {
    FILE* f = fopen(path, "rb");
    if (!f) {
        err = std::string("cannot open ") + path;
        return ISPAN_ERR_IO;
    }
    std::vector<char> buf;
    char chunk[1 << 16];
    size_t got;
    while ((got = fread(chunk, 1, sizeof(chunk), f)) > 0) buf.insert(buf.end(), chunk, chunk + got);
    fclose(f);
    size_t i = 0;
    const size_t n = buf.size();
    while (i < n && buf[i] == '%') {
        while (i < n && buf[i] != '\n') ++i;
        ++i;
    }
    std::vector<int64_t> ids;
    ids.reserve(n / 4);
    while (i < n) {
        while (i < n && (buf[i] == ' ' || buf[i] == '\t' || buf[i] == '\n' || buf[i] == '\r')) ++i;
        if (i >= n) break;
        bool neg = false;
        if (buf[i] == '-') {
            neg = true;
            ++i;
        }
        if (i >= n || buf[i] < '0' || buf[i] > '9') {
            err = "edge list: malformed token at byte " + std::to_string(i);
            return ISPAN_ERR_IO;
        }
        int64_t x = 0;
        while (i < n && buf[i] >= '0' && buf[i] <= '9') {
            x = x * 10 + (buf[i] - '0');
            if (x > INT32_MAX) {
                err = "edge list: id exceeds int32";
                return ISPAN_ERR_IO;
            }
            ++i;
        }
        ids.push_back(neg ? -x : x);
    }
    if (ids.size() % 2) {
        err = "edge list: odd number of ids";
        return ISPAN_ERR_IO;
    }
    int64_t lo = INT64_MAX, hi = INT64_MIN;
    for (int64_t x : ids) {
        lo = std::min(lo, x);
        hi = std::max(hi, x);
    }
    if (ids.empty()) lo = hi = 0;
    if (hi - lo + 1 > INT32_MAX - 1) {
        err = "edge list: id range exceeds int32";
        return ISPAN_ERR_IO;
    }
    *V = ids.empty() ? 0 : (int32_t)(hi - lo + 1);
    const size_t m = ids.size() / 2;
    src.resize(m);
    dst.resize(m);
    for (size_t k = 0; k < m; ++k) {
        src[k] = (int32_t)(ids[2 * k] - lo);
        dst[k] = (int32_t)(ids[2 * k + 1] - lo);
    }
    return ISPAN_OK;
}

static int write_i32(const std::string& path, const int32_t* p, size_t count) {
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) return ISPAN_ERR_IO;
    const size_t w = count ? fwrite(p, 4, count, f) : 0;
    fclose(f);
    return w == count ? ISPAN_OK : ISPAN_ERR_IO;
}

int ispan_convert_edge_list(ispan_ctx* ctx, const char* txt_path, const char* out_dir, int32_t clean, int32_t* V_out,
                            int64_t* E_out) {
    if (!ctx || !txt_path || !out_dir) return ISPAN_ERR_ARG;
    CK(hipSetDevice(ctx->device));
    std::vector<int32_t> src, dst;
    int32_t V = 0;
    int rc = parse_edge_list(txt_path, src, dst, &V, ctx->err);
    if (rc) return rc;
    const int64_t m = (int64_t)src.size();
    int32_t *d_src = nullptr, *d_dst = nullptr, *fb = nullptr, *fa = nullptr, *bb = nullptr, *ba = nullptr;
    int64_t E = m;
    std::vector<int32_t> h[4];
    CK(hipMalloc(&d_src, (size_t)std::max<int64_t>(m, 1) * 4));
    CK(hipMalloc(&d_dst, (size_t)std::max<int64_t>(m, 1) * 4));
    if (m) {
        CK(hipMemcpy(d_src, src.data(), (size_t)m * 4, hipMemcpyHostToDevice));
        CK(hipMemcpy(d_dst, dst.data(), (size_t)m * 4, hipMemcpyHostToDevice));
    }
    if (clean) {
        rc = gen_rc(build_csr_device(ctx->stream, V, m, d_src, d_dst, &fb, &fa, &bb, &ba, &E, ctx->err));
    } else {
        rc = gen_rc(csr_stable_one(ctx->stream, V, m, d_src, d_dst, &fb, &fa, ctx->err));
        if (!rc) rc = gen_rc(csr_stable_one(ctx->stream, V, m, d_dst, d_src, &bb, &ba, ctx->err));
    }
    (void)hipFree(d_src);
    (void)hipFree(d_dst);
    if (!rc) {
        int32_t* dp[4] = {fb, fa, bb, ba};
        const size_t cnt[4] = {(size_t)V + 1, (size_t)E, (size_t)V + 1, (size_t)E};
        for (int k = 0; k < 4 && !rc; ++k) {
            h[k].resize(std::max<size_t>(cnt[k], 1));
            if (cnt[k] && hipMemcpy(h[k].data(), dp[k], cnt[k] * 4, hipMemcpyDeviceToHost) != hipSuccess)
                rc = ISPAN_ERR_HIP;
        }
    }
    for (int32_t* p : {fb, fa, bb, ba})
        if (p) (void)hipFree(p);
    if (rc) return rc;
    // side files of the reference converters: source/destination per edge and degrees
    std::vector<int32_t> head((size_t)std::max<int64_t>(E, 1)), bhead((size_t)std::max<int64_t>(E, 1));
    std::vector<int32_t> odeg((size_t)std::max(V, 1)), ideg((size_t)std::max(V, 1));
    for (int32_t v = 0; v < V; ++v) {
        odeg[v] = h[0][v + 1] - h[0][v];
        ideg[v] = h[2][v + 1] - h[2][v];
        for (int32_t k = h[0][v]; k < h[0][v + 1]; ++k) head[k] = v;
        for (int32_t k = h[2][v]; k < h[2][v + 1]; ++k) bhead[k] = v;
    }
    const std::string d(out_dir);
    const struct {
        const char* name;
        const int32_t* p;
        size_t n;
    } files[] = {{"fw_begin.bin", h[0].data(), (size_t)V + 1}, {"fw_adjacent.bin", h[1].data(), (size_t)E},
                 {"head.bin", head.data(), (size_t)E},          {"out_degree.bin", odeg.data(), (size_t)V},
                 {"bw_begin.bin", h[2].data(), (size_t)V + 1}, {"bw_adjacent.bin", h[3].data(), (size_t)E},
                 {"bw_head.bin", bhead.data(), (size_t)E},      {"in_degree.bin", ideg.data(), (size_t)V}};
    for (const auto& fl : files) {
        if (write_i32(d + "/" + fl.name, fl.p, fl.n)) {
            ctx->err = std::string("cannot write ") + d + "/" + fl.name;
            return ISPAN_ERR_IO;
        }
    }
    if (V_out) *V_out = V;
    if (E_out) *E_out = E;
    return ISPAN_OK;
}

void ispan_device_free(void* d_ptr) {
    if (d_ptr) (void)hipFree(d_ptr);
}

}  // extern "C"
