/*
 * ispan_hip.h -- C ABI of libispan_hip.so, the MI355X (gfx950) strongly
 * connected components pipeline of iSpan (SC'18), built for HIP.
 *
 * The reference exposes its shared-memory path as C++ functions compiled
 * into the `ispan` executable (src/scc_common.h:7-33):
 *     graph_load(fw_beg, fw_csr, bw_beg, bw_csr, avg_time)   src/scc_common.h:7-13
 *     scc_detection(g, alpha, beta, gamma, theta, T, avg)    src/scc_common.h:15-23
 *     get_scc_result(scc_id, vert_count)                     src/scc_common.h:25-28
 *     print_time_result(run_times, avg_time)                 src/scc_common.h:30-33
 * This header replaces them with plain-pointer entry points.  Inputs are the
 * reference's int32 CSR pair (src/graph.h:9-31, src/util.h:17-18): fw_beg[V+1],
 * fw_adj[E] (out-edges) and bw_beg[V+1], bw_adj[E] (in-edges), headerless
 * little-endian int32 on disk (src/graph.cpp:13-14).
 *
 * Output: canonical labels, label[v] = smallest vertex id of v's SCC.  The
 * reference's raw scc_id encoding (1 giant, -1/-2/-3/-4 trimmed, v >= 2 mice
 * pivot) is not reproduced; its partition is.
 *
 * Errors are negative return codes (the reference calls exit(-1) or asserts,
 * src/scc_main.cpp:13, src/graph.cpp:21-74).  A context is not thread-safe;
 * distinct contexts may be used from distinct threads.  Inputs stay owned by
 * the caller; the library never frees them.
 */
#ifndef ISPAN_HIP_H
#define ISPAN_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ISPAN_OK 0
#define ISPAN_ERR_ARG (-1)        /* null pointer, bad size or bad parameter     */
#define ISPAN_ERR_CSR (-2)        /* malformed CSR (offsets, ids out of range)   */
#define ISPAN_ERR_NOMEM (-3)      /* device or host allocation failed            */
#define ISPAN_ERR_HIP (-4)        /* HIP runtime error                           */
#define ISPAN_ERR_ITER (-5)       /* a fixed-point loop hit its iteration cap    */
#define ISPAN_ERR_UNLABELLED (-6) /* internal: a vertex left without a label     */
#define ISPAN_ERR_IO (-7)         /* file missing, short or of the wrong size    */
#define ISPAN_ERR_NOGRAPH (-8)    /* solve called before a graph was bound       */
#define ISPAN_ERR_BUDGET (-9)     /* server: the HBM budget cannot admit it      */
#define ISPAN_ERR_TIMEOUT (-10)   /* server: the wait timed out                  */
#define ISPAN_ERR_NOTFOUND (-11)  /* server: unknown graph or request id         */
#define ISPAN_ERR_SCRATCH (-12)   /* too large for ispan_scc_detection_noalloc   */

/* Tuning knobs.  alpha/beta/gamma/theta keep the meaning of the `ispan`
 * argv (src/scc_main.cpp:20-25; used at src/fw_bw.h:581,599 and
 * src/scc_core.cpp:597).  thread_count is accepted for drop-in argv
 * compatibility and recorded; the GPU path does not use it. */
typedef struct ispan_params {
    int32_t alpha;        /* top-down -> bottom-up switch (default 30)              */
    int32_t beta;         /* bottom-up -> top-down-queue switch (default 200)       */
    int32_t gamma;        /* level cap for bottom-up (default 10)                   */
    double theta;         /* trim-1 sweep stop ratio (default 0.1)                  */
    int32_t trim_times;   /* max trim-1 sweeps after the giant SCC (TRIM_TIMES, 3)  */
    int32_t thread_count; /* recorded only                                          */
    int32_t pivot_wrap32; /* 1: reproduce the reference's int32 out*in product      */
    int32_t timing;       /* 1: per-phase HIP event timing into ispan_stats.ms      */
    int32_t validate;     /* 1: check CSR structure when a graph is bound           */
    int32_t wcc_mode;     /* small SCCs after the trims: 0 WCC by union-find
                             hooking (openmp_wcc's union, src/openmp_wcc.hpp),
                             1 WCC by label propagation + jumping (coloring_wcc,
                             src/color_propagation.h:468-593), each followed by
                             FW-BW per WCC (mice); 2 the colouring SCC method
                             (graph_color, src/color_propagation.h:9-467), no
                             mice phase.  Engine 0 runs all three on the device
                             with no host round trip; engine 1 checks label
                             propagation from the host each round and runs 2
                             as 0.  Chooses the work, never the labels.       */
    int32_t async_batch;  /* engine 1: top-down levels enqueued per host check (8)  */
    int32_t engine;       /* giant phase: 0 one cooperative persistent kernel with
                             grid barriers (default), 1 one launch per level with
                             host-side switch decisions                              */
    int32_t switch_rule;  /* top-down -> bottom-up test, (set - visited) * avg_deg
                             / alpha < frontier work: 1 (default) the frontier's
                             work is its exact out-degree sum; 0 it is found *
                             avg_deg, as src/fw_bw.h:578-581.  Engine 1 always
                             uses 0.  Chooses the work, never the labels.       */
    int32_t trim_scope;   /* size-2/3 pattern trims walk: 0 the active queue
                             (trim_*_from_fq, src/trim_2_3.h:146-1080), 1 every
                             active vertex of the graph (trim_*_from_graph,
                             src/trim_2_3.h:9-144, 287-480, 682-879); engine 0
                             only.  Same detections, different work.          */
    int32_t batch_by_class; /* ispan_batch_bind_device: 1 orders the graphs by
                             size class (see ispan_tuning_table) before packing
                             them into groups, so a group spans a narrow range
                             of sizes and is solved with the tuned parameters
                             of its largest graph's class; 0 packs consecutive
                             graphs                                           */
    int32_t td_local;     /* engine 0, top-down to exhaustion (after bottom-up):
                             a wave expands what it claims itself, in the same
                             level, for up to this many steps of 64 vertices
                             before handing the rest to the next level (the
                             reference's asynchronous top-down, src/fw_bw.h:
                             506-544); 0 keeps every level synchronous.
                             Default 64.  Chooses the work, never the labels.  */
} ispan_params;

/* Per-solve statistics.  ms[] uses the reference's avg_time slot map
 * (src/load_print.cpp:67-68, src/scc_core.cpp:1038-1052):
 *   0 trim, 1 elephant SCC (pivot+fw+bw), 2 mice SCC (wcc+mice), 3 total,
 *   4 trim size_1, 5 trim size_2, 6 pivot, 7 fw bfs, 8 bw bfs, 9 wcc,
 *   10 mice fw-bw, 13 trim size_3; plus 11 trim-1-first (with queue and
 *   pivot), 14 trim-1 sweeps.  Filled when params.timing = 1, except ms[3],
 *   always the host-measured solve time. */
typedef struct ispan_stats {
    double ms[16];
    int64_t fq0;          /* vertices left after trim-1-first                */
    int64_t pivot;        /* giant-SCC pivot (-1: none)                      */
    int64_t fw_size;      /* |FW reach set of the pivot|                     */
    int64_t giant_size;   /* |giant SCC|                                     */
    int64_t fw_edges;     /* sum of out-degrees over the FW set              */
    int64_t bw_edges;     /* sum of in-degrees over the giant SCC            */
    int64_t trim1;        /* vertices labelled by trim-1 (all sweeps)        */
    int64_t trim2;        /* vertices labelled by the size-2 pattern         */
    int64_t trim3;        /* vertices labelled by the size-3 patterns        */
    int64_t n_wcc;        /* WCCs entering the mice phase                    */
    int64_t mice;         /* vertices entering the mice phase                */
    int64_t fw_levels;    /* FW traversal levels                             */
    int64_t bw_levels;    /* BW traversal levels                             */
    int64_t n_scc;        /* number of SCCs (computed on request)            */
    int64_t host_syncs;   /* host<->device round trips in this solve         */
    int64_t engine_used;  /* giant-phase engine that ran (see ispan_params)  */
    int64_t bw_reach;     /* vertices the backward traversal reached: engine
                             0 runs it beside the forward one over every
                             active vertex (the giant is the intersection);
                             engine 1 inside the forward set (= giant_size) */
    int64_t max_wcc;      /* largest WCC left for the mice phase             */
    int64_t coloured;     /* engine 0: vertices the colouring method took
                             instead of the mice phase, because a WCC had
                             more than ISPAN_MICE_MAX_WCC vertices           */
    int64_t reserved[5];
} ispan_stats;

/* Engine 0 hands the vertices left after the trims to the colouring method
 * (as wcc_mode 2) instead of the mice phase when one of their WCCs has more
 * vertices than this: FW-BW per WCC searches forward from each smallest
 * unlabelled vertex in turn, which on a long WCC of small SCCs costs its size
 * per SCC.  Labels do not change. */
#define ISPAN_MICE_MAX_WCC 1024

typedef struct ispan_ctx ispan_ctx;

/* Defaults: alpha 30, beta 200, gamma 10, theta 0.1, trim_times 3
 * (src/bash_one.sh:3, src/util.h:34), validate 1, async_batch 8,
 * switch_rule 1, td_local 64, everything else 0. */
void ispan_default_params(ispan_params* p);

/* Human-readable text for an error code. */
const char* ispan_strerror(int code);

/* Context: device buffers, a HIP stream, pinned readback area.  `device` is
 * a HIP ordinal.  Replaces the per-call allocation at src/scc_core.cpp:56-92. */
int ispan_ctx_create(ispan_ctx** out, int device);
int ispan_ctx_destroy(ispan_ctx* ctx);
/* Use a caller-owned hipStream_t (passed as void*) for all work; NULL = the
 * context's own stream. */
int ispan_ctx_set_stream(ispan_ctx* ctx, void* hip_stream);
/* Wait until all work queued by this context has finished (the reference's
 * calls are synchronous; this is the point where its timer would stop). */
int ispan_ctx_synchronize(ispan_ctx* ctx);
/* Message of the last error seen by this context. */
const char* ispan_ctx_last_error(const ispan_ctx* ctx);

/* Copy a host CSR pair to the device (the device-side half of graph::graph,
 * src/graph.cpp:2-91).  Validates when params->validate (or params NULL). */
int ispan_graph_upload(ispan_ctx* ctx, int32_t V, int64_t E, const int32_t* fw_beg, const int32_t* fw_adj,
                       const int32_t* bw_beg, const int32_t* bw_adj, const ispan_params* params);
/* Bind caller-owned device arrays (no copy).  They must stay valid and
 * unchanged until the next bind/upload or ctx destruction. */
int ispan_graph_bind_device(ispan_ctx* ctx, int32_t V, int64_t E, const int32_t* d_fw_beg, const int32_t* d_fw_adj,
                            const int32_t* d_bw_beg, const int32_t* d_bw_adj, const ispan_params* params);

/* One SCC solve of the bound graph (scc_detection, src/scc_core.cpp:27-1094).
 * Labels stay on the device; d_labels (int32[V], device) receives them when
 * not NULL.  stats may be NULL. */
int ispan_solve(ispan_ctx* ctx, const ispan_params* params, int32_t* d_labels, ispan_stats* stats);
/* Copy the labels of the last solve to host memory (int32[V]). */
int ispan_labels_download(ispan_ctx* ctx, int32_t* labels);
/* Device pointer of the labels of the last solve (int32[V], owned by ctx). */
const int32_t* ispan_labels_device(const ispan_ctx* ctx);
/* Fill stats->n_scc, fw_edges and bw_edges for the last solve (extra pass). */
int ispan_solve_stats(ispan_ctx* ctx, ispan_stats* stats);

/* Per-kernel profile of the last solve run with params.timing >= 2: every
 * traversal launch is bracketed by HIP events on the solve stream.
 * Engine 1 classes: 0 top-down level (k_td), 1 top-down hub items
 * (k_td_hub), 2 bottom-up sweep (k_bu), 3 trim-1-first + queue + pivot.
 * Engine 0: class 3 is the whole giant-phase kernel (k_giant: trim-1-first,
 * pivot, both traversals, relabel); edges[0]/[2] and found[0]/[2] hold its
 * top-down and bottom-up work.  edges = adjacency entries read, found =
 * vertices reached. */
typedef struct ispan_kernel_profile {
    int64_t launches[4];
    double ms[4];
    int64_t edges[4];
    int64_t found[4];
} ispan_kernel_profile;
int ispan_kernel_profile_get(const ispan_ctx* ctx, ispan_kernel_profile* out);

/* One-shot host-pointer convenience: upload, solve, download (the whole of
 * scc_detection + the label hook of get_scc_result).  labels: int32[V]. */
int ispan_scc_detection(int32_t V, int64_t E, const int32_t* fw_beg, const int32_t* fw_adj, const int32_t* bw_beg,
                        const int32_t* bw_adj, const ispan_params* params, int32_t* labels, ispan_stats* stats);

/* Allocation-free variant of ispan_scc_detection, on device memory.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * allocation).  It calls no hipMalloc/hipFree, creates no stream or event and
 * allocates no host memory on success: every buffer of the solve comes from
 * one fixed device array of 64 MiB that the library holds on each GPU it runs
 * on until the process ends.  The solve state takes about 85 bytes per vertex
 * plus 0.075 bytes per edge plus ~0.4 MB, so graphs of up to ~750,000 vertices
 * fit at edge factor 16 (R-MAT scale 19 does; scale 20 does not); a larger one
 * returns ISPAN_ERR_SCRATCH before any device work.  Inputs: the CSR pair in
 * device memory (as ispan_graph_bind_device); output: d_labels (int32[V],
 * device).  Runs on the current HIP device, on hip_stream (NULL: the null
 * stream), and returns when the labels are in d_labels.  params->timing is
 * ignored (stats->ms[3] is still the wall time).  Every call on one GPU uses
 * that GPU's one array, which keeps the last call's contents (each solve
 * clears what it reads): calls that overlap on one GPU race, so callers must
 * serialise them per device. */
int ispan_scc_detection_noalloc(int32_t V, int64_t E, const int32_t* d_fw_beg, const int32_t* d_fw_adj,
                                const int32_t* d_bw_beg, const int32_t* d_bw_adj, const ispan_params* params,
                                int32_t* d_labels, void* hip_stream, ispan_stats* stats);

/* ---- batched solves ----------------------------------------------------
 * Many graphs in one pass of the pipeline over their disjoint union.  There
 * is no counterpart in the reference, which solves one graph per process
 * (src/scc_main.cpp:7-57); per graph the labels are those a single solve
 * gives.  Layout (device memory, caller-owned): graph g has local ids
 * 0..V_g-1; h_v_off / h_e_off (host, int64[n_graphs + 1], starting at 0) give
 * V_g = v_off[g+1] - v_off[g] and E_g = e_off[g+1] - e_off[g]; its V_g + 1
 * offsets (0 .. E_g) start at d_fw_beg[v_off[g] + g], its E_g adjacency
 * entries at d_fw_adj[e_off[g]]; the bw pair likewise.
 * Binding groups consecutive graphs whose union fits int32 offsets (at most
 * max_group_edges edges per group; 0 = 2^31 - 1), builds each group's union
 * CSR once and checks every graph (ISPAN_ERR_CSR).  The inputs may be freed
 * after the call.  Binding replaces any bound graph or batch. */
int ispan_batch_bind_device(ispan_ctx* ctx, int32_t n_graphs, const int64_t* h_v_off, const int64_t* h_e_off,
                            const int32_t* d_fw_beg, const int32_t* d_fw_adj, const int32_t* d_bw_beg,
                            const int32_t* d_bw_adj, int64_t max_group_edges, const ispan_params* params);
/* Number of groups of the bound batch (0 if none). */
int ispan_batch_groups(const ispan_ctx* ctx);
/* Solve every graph of the bound batch: one cooperative launch per group,
 * every graph's own pivot seeding the traversals.  Needs the cooperative
 * engine.  stats: totals over the batch, ms[3] = wall time of the batch. */
int ispan_batch_solve(ispan_ctx* ctx, const ispan_params* params, ispan_stats* stats);
/* Labels of the last batch solve: int32[v_off[n_graphs]], graph g's at
 * v_off[g], each the smallest local id of the vertex's SCC. */
int ispan_batch_labels_download(ispan_ctx* ctx, int32_t* labels);
const int32_t* ispan_batch_labels_device(const ispan_ctx* ctx);

/* Autotuning: times every candidate parameter set on the bound graph (or
 * the bound batch) `reps` times after one warm-up solve and reports the
 * fastest (its index in *best, median ms of each in ms[n] if not NULL).
 * Parameters choose the work, never the labels, so any candidate is valid.
 * No counterpart in the reference, whose alpha/beta/gamma come from argv. */
int ispan_autotune(ispan_ctx* ctx, const ispan_params* candidates, int32_t n, int32_t reps, int32_t* best,
                   double* ms);

/* ---- tuning per size class ---------------------------------------------
 * No counterpart in the reference, whose alpha/beta/gamma/theta come from
 * argv (src/scc_main.cpp:20-25) for every graph alike.  A graph of V vertices
 * is in size class floor(log2 V) (class 0 also takes V = 0); the table holds
 * one parameter set per class.  Parameters choose the work, never the
 * labels. */
#define ISPAN_SIZE_CLASSES 32
typedef struct ispan_tuning_table {
    int32_t tuned[ISPAN_SIZE_CLASSES];       /* 1: params[k] was measured          */
    ispan_params params[ISPAN_SIZE_CLASSES];
    double ms[ISPAN_SIZE_CLASSES];           /* the choice's median time per class:
                                                per graph for a single graph, per
                                                pass over the class for a batch   */
} ispan_tuning_table;

/* Times every candidate on the bound graph (its class) or on each size class
 * of the bound batch (a group counts for the class of its largest graph;
 * bind with params->batch_by_class = 1 so that groups span narrow size
 * ranges), `reps` times after one warm-up, and writes the fastest per
 * measured class into table; other classes keep what table held. */
int ispan_autotune_classes(ispan_ctx* ctx, const ispan_params* candidates, int32_t n, int32_t reps,
                           ispan_tuning_table* table);
/* ispan_batch_solve with each group solved by the parameters of its size
 * class (a group's class: that of its largest graph); untuned classes use
 * the defaults. */
int ispan_batch_solve_tuned(ispan_ctx* ctx, const ispan_tuning_table* table, ispan_stats* stats);
/* The parameters the table gives a graph of V vertices (defaults when its
 * class is untuned). */
void ispan_tuning_lookup(const ispan_tuning_table* table, int64_t V, ispan_params* out);
/* A table as text (one line per tuned class) and back; ISPAN_ERR_IO when the
 * file cannot be written, read or parsed. */
int ispan_tuning_save(const ispan_tuning_table* table, const char* path);
int ispan_tuning_load(const char* path, ispan_tuning_table* table);

/* ---- serving --------------------------------------------------------------
 * A request loop over device-resident graphs with continuous batching; no
 * counterpart in the reference.  Registered graphs are appended once to an
 * arena (one CSR pair in device memory holding every resident graph side by
 * side); registration is checked and admitted against an HBM budget that
 * covers the arena, the solve state sized for it and the label buffers, and
 * against the device memory actually free when the arena has to grow.  A
 * request names a graph.  A worker thread takes every pending request within
 * the batch limits, solves the distinct graphs among them in one pass over
 * the arena (nothing is copied per batch) and starts over with what arrived
 * meanwhile.  Labels stay in device memory until the client takes them.
 * Each request's labels equal those of a single solve of its graph.
 * Latency is submit -> labels ready on the device.  All functions are
 * thread-safe. */
typedef struct ispan_server ispan_server;
typedef struct {
    int64_t hbm_budget;         /* device bytes for the arena, solve state, labels (0: 200 GiB) */
    int32_t max_batch_graphs;   /* distinct graphs per batch (0: no limit)                      */
    int32_t reserved0;
    int64_t max_batch_vertices; /* vertices per batch (0: no limit)                             */
    int64_t reserved[5];
} ispan_server_config;
typedef struct {
    double submit_ms, start_ms, done_ms; /* steady clock, ms                       */
    double latency_ms;                   /* done - submit                          */
    int64_t batch;                       /* batch that solved it                   */
    int32_t batch_graphs;                /* requests in that batch                 */
    int32_t status;                      /* ISPAN_OK or the batch's error          */
} ispan_request_info;
typedef struct {
    int64_t requests_done, batches, graphs_resident;
    int64_t bytes_resident;             /* the live graphs' share of the arena     */
    int64_t bytes_budget, bytes_device; /* budget in effect (the configured one,
                                           capped by device use + free memory);
                                           arena + solve state + labels             */
    double p50_ms, p99_ms, mean_ms, max_ms; /* latency since the last reset        */
    double mean_batch_graphs;
    int64_t reserved[5];
} ispan_server_stats;

int ispan_server_create(ispan_server** out, int device, const ispan_server_config* cfg, const ispan_params* params);
/* Finishes every submitted request, then stops. */
int ispan_server_destroy(ispan_server* srv);
/* Register a graph from host or device memory (checked in full first);
 * ISPAN_ERR_BUDGET when the arena cannot take it. */
int ispan_server_add_graph(ispan_server* srv, int32_t V, int64_t E, const int32_t* fw_beg, const int32_t* fw_adj,
                           const int32_t* bw_beg, const int32_t* bw_adj, int64_t* graph_id);
int ispan_server_add_graph_device(ispan_server* srv, int32_t V, int64_t E, const int32_t* d_fw_beg,
                                  const int32_t* d_fw_adj, const int32_t* d_bw_beg, const int32_t* d_bw_adj,
                                  int64_t* graph_id);
/* Unregister (pending requests still finish).  Arena space is reused once
 * every graph after it is gone too. */
int ispan_server_remove_graph(ispan_server* srv, int64_t graph_id);
/* Queue a solve of graph_id; returns at once. */
int ispan_server_submit(ispan_server* srv, int64_t graph_id, int64_t* request_id);
/* Queue n solves at once (they are in the queue together, so they can share
 * a batch); nothing is queued if an id is unknown. */
int ispan_server_submit_many(ispan_server* srv, int32_t n, const int64_t* graph_ids, int64_t* request_ids);
/* Wait for a request (timeout_ms < 0: no limit), copy its labels (int32[V])
 * to host memory if labels is not NULL, and forget it.  Returns its status. */
int ispan_server_wait(ispan_server* srv, int64_t request_id, int32_t* labels, ispan_request_info* info,
                      int64_t timeout_ms);
/* Wait and leave the labels in device memory: *d_labels stays valid until
 * ispan_server_release(request_id). */
int ispan_server_wait_device(ispan_server* srv, int64_t request_id, const int32_t** d_labels, ispan_request_info* info,
                             int64_t timeout_ms);
int ispan_server_release(ispan_server* srv, int64_t request_id);
int ispan_server_stats_get(ispan_server* srv, ispan_server_stats* out);
int ispan_server_reset_stats(ispan_server* srv);
/* Solve each batch with the table's parameters for the size class of the
 * batch's largest graph (NULL: back to the server's own parameters). */
int ispan_server_set_tuning(ispan_server* srv, const ispan_tuning_table* table);
const char* ispan_server_last_error(const ispan_server* srv);

/* ---- reference file I/O and reporting ---------------------------------- */

/* Host graph as loaded from the four iSpan binaries (graph_load,
 * src/load_print.cpp:6-22 / graph::graph, src/graph.cpp:2-91):
 * V = size(fw_begin)/4 - 1, E = size(fw_adjacent)/4. */
typedef struct ispan_host_graph {
    int32_t V;
    int64_t E;
    int32_t* fw_beg;
    int32_t* fw_adj;
    int32_t* bw_beg;
    int32_t* bw_adj;
} ispan_host_graph;

int ispan_graph_load(const char* fw_beg_file, const char* fw_adj_file, const char* bw_beg_file,
                     const char* bw_adj_file, ispan_host_graph* out);
void ispan_graph_free(ispan_host_graph* g);
/* Structural check of a host CSR pair (offsets monotone, beg[V] == E, ids in
 * range, bw the transpose-size of fw).  Returns ISPAN_OK or ISPAN_ERR_CSR. */
int ispan_csr_check(int32_t V, int64_t E, const int32_t* fw_beg, const int32_t* fw_adj, const int32_t* bw_beg,
                    const int32_t* bw_adj);

/* SCC size summary of canonical labels (get_scc_result, src/load_print.cpp:
 * 24-60, without its phantom entry): largest SCC size, counts of SCCs of
 * size 1, 2, 3 and larger-than-3-but-not-largest, total SCC count. */
typedef struct ispan_summary {
    int64_t largest;
    int64_t size_1;
    int64_t size_2;
    int64_t size_3;
    int64_t others;
    int64_t total;
} ispan_summary;
int ispan_summarize(int32_t V, const int32_t* labels, ispan_summary* out);
/* SCC-size histogram of the last ispan_solve's labels, computed on the
 * device: bins[k] (host, n_bins <= 64 entries) = number of SCCs with size in
 * [2^k, 2^(k+1)); the last bin also counts every larger size.  summary (may
 * be NULL) receives what ispan_summarize gives for the same labels.  Extends
 * get_scc_result (src/load_print.cpp:24-60), which tallies the sizes of the
 * SCCs it cannot classify in a std::map but prints only count classes. */
int ispan_scc_histogram(ispan_ctx* ctx, int64_t* bins, int32_t n_bins, ispan_summary* summary);

/* ---- synthetic inputs built on the device ------------------------------ */

/* R-MAT (a, b, c, 1-a-b-c) draws m = round(edge_factor * V) at
 * ceil(log2 V) bit levels from a counter hash of (seed, draw); draws with an
 * endpoint >= V are dropped, then self-loops and duplicates are removed and
 * both CSR directions are built on the device.  The caller frees the four
 * device arrays with ispan_device_free.  (Format producer of
 * graph_converter/txt_to_bin_{fw,bw}_int.cpp, with dedup.) */
int ispan_gen_rmat_device(ispan_ctx* ctx, int32_t V, double edge_factor, double a, double b, double c, uint64_t seed,
                          int32_t** d_fw_beg, int32_t** d_fw_adj, int32_t** d_bw_beg, int32_t** d_bw_adj, int64_t* E);
/* n x n lattice; each lattice edge bidirectional with probability p_bidir,
 * else one random direction; plus L/100 random directed edges. */
int ispan_gen_grid_device(ispan_ctx* ctx, int32_t n, double p_bidir, uint64_t seed, int32_t** d_fw_beg,
                          int32_t** d_fw_adj, int32_t** d_bw_beg, int32_t** d_bw_adj, int64_t* E);
/* Edge list (device int32 src[m], dst[m]) -> deduplicated CSR pair. */
int ispan_build_csr_device(ispan_ctx* ctx, int32_t V, int64_t m, const int32_t* d_src, const int32_t* d_dst,
                           int32_t** d_fw_beg, int32_t** d_fw_adj, int32_t** d_bw_beg, int32_t** d_bw_adj,
                           int64_t* E);
void ispan_device_free(void* d_ptr);

/* Text edge list ("src dst" pairs, leading '%' lines skipped, ids shifted by
 * the smallest id) -> the reference's binary files in out_dir: fw_begin.bin,
 * fw_adjacent.bin, head.bin, out_degree.bin (graph_converter/
 * txt_to_bin_fw_int.cpp:31-217) and bw_begin.bin, bw_adjacent.bin,
 * bw_head.bin, in_degree.bin (txt_to_bin_bw_int.cpp).  clean = 0 keeps the
 * converters' layout (rows in input order, duplicates and self-loops kept);
 * clean = 1 drops self-loops and duplicates and sorts every row.  The CSR is
 * built on the device by stable radix sorts. */
int ispan_convert_edge_list(ispan_ctx* ctx, const char* txt_path, const char* out_dir, int32_t clean,
                            int32_t* V_out, int64_t* E_out);

/* Library version string. */
const char* ispan_version(void);

#ifdef __cplusplus
}
#endif
#endif /* ISPAN_HIP_H */
