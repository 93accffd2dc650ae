// Internal declarations shared by the kernels and the host-side solver.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <string>

namespace ispan {

constexpr int kBlock = 256;          // threads per workgroup (4 wave64)
constexpr int kWave = 64;            // gfx950 wavefront width
constexpr int kTrimTile = 4096;      // vertices per tile in the first trim sweep
constexpr int kHubDeg = 4096;        // top-down: degree at which a vertex is split
constexpr int kHubChunk = 4096;      // edges per hub work item
constexpr int kBuTileWords = 32;     // bottom-up: bitmap words (x64 vertices) per tile
constexpr int kWaveBuf = 512;        // per-wave LDS staging for queue pushes

// Device control block: every count a kernel needs to size its work lives
// here, so kernels launch on fixed grids and the host reads it back only
// where a decision needs it.
struct alignas(128) Ctl {
    int32_t fq_n[2];        // active-vertex queue lengths (ping-pong)
    int32_t q_n[2];         // traversal frontier lengths (ping-pong by level parity)
    int32_t hub_n[2];       // hub work items (ping-pong by level parity)
    int32_t root;           // giant-SCC pivot, -1 if none
    int32_t giant_min;      // smallest id in the giant SCC
    int32_t levels;         // traversal levels that had a non-empty frontier
    int32_t trimmed;        // vertices labelled by the current trim sweep
    int32_t pattern2;       // vertices labelled by the size-2 pattern
    int32_t pattern3;       // vertices labelled by the size-3 patterns
    int32_t n_wcc;          // WCCs in the mice phase
    int32_t changed;        // fixed-point flag (label propagation)
    int32_t unlabelled;     // final check
    uint32_t done_blocks;   // last-block-done counter (pivot reduction)
    int32_t bad_csr;        // CSR validation failures
    int32_t max_wcc;        // largest WCC entering the mice phase
    int64_t fw_count;       // |FW set| (popcount of fvis)
    int64_t fw_edges;       // statistics pass
    int64_t bw_edges;
    int64_t n_scc;
    int64_t work_edges[3];  // adjacency entries read: 0 top-down, 1 hub items, 2 bottom-up
    int64_t work_found[3];  // vertices reached, same classes
    int32_t pad[2];
};

struct PivotPart {
    int64_t prod;
    int32_t id;
    int32_t pad;
};

struct HubItem {
    int32_t v;
    int32_t start;  // first edge offset of this chunk
};

// ---- cooperative giant-phase kernel (giant.hip) --------------------------
struct alignas(128) GLine {
    int32_t v;
    int32_t pad[31];
};

// Reduction carried by one grid barrier: three sums and a max over keys
// (identity 0 for all four).
struct alignas(128) RLine {
    int64_t sum[3];
    int64_t key;
    int64_t pad[12];
};

// Device state of one k_giant launch; zeroed before every launch.
constexpr int kLevelTrace = 1024;
struct GiantState {
    GLine sub[32];      // barrier arrival counters (two-level barrier)
    GLine top;
    GLine gen;          // barrier generation
    GLine qn[2][3];     // [direction][level % 3] queue lengths
    GLine hn[2][3];     // [direction][level % 3] hub work items
    GLine err;          // 1 barrier timeout, 2 level cap, 4 queue count out of range
    int64_t root, fq0, fw_count, giant, gmin, fq_n, fw_levels, bw_levels;
    int64_t work[4];    // adjacency entries read top-down / bottom-up, vertices found top-down / bottom-up
    int64_t ts[6];      // wall-clock stamps: start, pivot known, FW done, BW done (the two run in the
                        // same levels), giant relabelled, end
    GLine fqn[3];       // active-queue lengths of the in-kernel trim sweeps (rotating)
    int64_t n_final, final_sel, trim1, pat2, pat3, sweeps, wcc_done;
    int64_t tail_rounds;  // label-propagation sweeps (wcc_mode 1) or colouring rounds (2)
    int64_t tail_sccs;    // colouring: SCCs it labelled
    int64_t bw_reach;     // vertices the backward traversal reached (the giant and what reaches it)
    int64_t unlabelled;   // with unl_checked: vertices the kernel left at -1 (its final check)
    int64_t unl_checked;  // 1: the kernel labelled everything it was given and counted `unlabelled`
    RLine gacc[32];     // barrier reductions: per arrival group
    RLine tacc;         // barrier reductions: over groups
    RLine res[2];       // barrier reductions: result of generation g in res[g & 1]
    // blocks' start and first-barrier arrival clocks: min, max of each
    int64_t t_start[2], t_arrive[2];
    // per level of each direction (first kLevelTrace levels): end time,
    // vertices found, adjacency entries read, 0 top-down / 1 bottom-up.
    // Last on purpose: a solve clears and copies back only what precedes
    // them, unless the level trace is on.
    int64_t lv_ts[2][kLevelTrace], lv_found[2][kLevelTrace], lv_edges[2][kLevelTrace], lv_mode[2][kLevelTrace];
    // with GiantArgs::trace: per level, the longest and (as ~t) the shortest
    // block work time (barrier exit to next arrival); bottom-up levels: [0]
    // the most adjacency entries one block scanned, and the latest time any
    // block's first wave had [1] probed its candidates, [2] scanned its
    // misses, [3] finished its dense words (clock ticks from its barrier exit)
    int64_t lv_wmax[2][kLevelTrace], lv_wmin[2][kLevelTrace], lv_cp[2][kLevelTrace][4];
    // with GiantArgs::trace: each block's work time in the first bottom-up
    // level of the forward traversal, and the adjacency entries it scanned
    int64_t bu1_work[1024], bu1_scan[1024];
    // with GiantArgs::trace: latest clock of any block after the post
    // phase's count pass, its barrier, and the labels and active queue
    int64_t post_cp[3];
    // the launch that last cleared this state (GiantArgs::epoch); never
    // cleared itself, and 0 when the state is allocated
    int64_t epoch;
};
constexpr size_t kGiantStateHead = offsetof(GiantState, lv_ts);

struct GiantArgs {
    const int32_t* fwb;
    const int32_t* fwa;
    const int32_t* bwb;
    const int32_t* bwa;
    const int32_t* fw_first;  // first out-neighbour per vertex (-1: none)
    const int32_t* bw_first;  // first in-neighbour per vertex
    int32_t V;
    int64_t nwords;
    int32_t* lab;
    uint64_t* act;
    uint64_t* fvis;
    uint64_t* bvis;
    uint64_t* snap;           // forward bottom-up: visited words before the level
    uint64_t* snapb;          // backward
    uint64_t* q0;             // forward top-down queues (ping-pong by level): (first edge << 32) | degree
    uint64_t* q1;
    HubItem* h0;              // forward hub work items
    HubItem* h1;
    uint64_t* qb0;            // backward queues (the same entries) and hub work items
    uint64_t* qb1;
    HubItem* hb0;
    HubItem* hb1;
    int32_t* fq;              // active queue after the giant phase
    GiantState* gs;
    int64_t* part;            // kGiantSlots x grid partial counts
    int32_t* fq_n_out;        // receives the active queue length
    int64_t hub_cap;          // capacity of h0 / h1
    // after the giant (when rest != 0): trim-1 sweeps, size-2/3 patterns,
    // one more sweep and the WCC union-find run in the same launch
    int32_t rest;
    int32_t trim_times;
    double theta;
    int32_t* fq_alt;          // second active-queue buffer
    int32_t* stamp;           // pattern size per labelled vertex
    int32_t* parent;          // WCC union-find forest
    int32_t* fq_n_ctl;        // Ctl::fq_n: receives the final queue length at [final buffer]
    int32_t* dbg;             // 2 x grid: barriers each block arrived at / passed
    int32_t alpha, beta, gamma;
    double avg_degree;
    int32_t wrap32;
    // batched solve over a disjoint union of graphs: every graph's pivot
    // (-1 entries are skipped) seeds both traversals; giant vertices keep
    // their own id as label, to be replaced per graph afterwards
    const int32_t* roots;
    int32_t n_roots;          // 0: single graph, pivot chosen in the kernel
    const uint64_t* sel;      // optional: only these vertices take part (the rest is never active)
    int32_t trace;            // nonzero: per-level block timings (GiantState::lv_w*, lv_cp, bu1_*)
    int64_t bu_mult;          // coprime to grid * waves per block: bottom-up run permutation
    int32_t switch_rule;      // top-down -> bottom-up estimate of the frontier's work: 0 found x average
                              // degree (the reference's), 1 the frontier's exact out-degree sum
    int32_t wcc_mode;         // after the trims: 0 union-find WCC, 1 label-propagation WCC, 2 the
                              // colouring SCC method (no mice phase left)
    int32_t trim_scope;       // size-2/3 pattern walk: 0 the active queue, 1 every active vertex
    int32_t td_local;         // top-down to exhaustion: steps a wave expands its own finds in place (0: off)
    uint64_t* ckey;           // colouring: (priority, colour) per vertex
    int32_t* aux;             // colouring: smallest member per SCC root
    // set by launch_giant: block 0 clears `state_words` words of gs and
    // `ctl_words` of ctl, then publishes `epoch` (unique per launch, never 0)
    // in gs->epoch; the other blocks start once they read it
    Ctl* ctl;
    int64_t epoch;
    int32_t state_words, ctl_words;
};

constexpr int kGiantSlots = 16;
constexpr int kGiantHubDeg = 512;   // degree from which a claimed vertex is split into work items
constexpr int kGiantChunk = 8192;   // edges per work item
// A vertex of degree d >= kGiantHubDeg becomes ceil(d / kGiantChunk) <=
// d / kGiantChunk + 1 items, and at most E / kGiantHubDeg vertices qualify.
inline int64_t giant_hub_capacity(int64_t E) { return E / kGiantChunk + E / kGiantHubDeg + 64; }

// 0 on error
int giant_max_blocks_per_cu();
// GiantArgs::bu_mult for a grid of `grid` blocks
int64_t giant_bu_mult(int grid);
// records `before` (when given), then launches k_giant cooperatively; the
// kernel clears a.gs and ctl itself (a kernel queued in front of a
// cooperative launch costs a cross-queue wait).  gs->epoch must not hold an
// epoch of this process that is still to come: 0 at allocation does.
hipError_t launch_giant(const GiantArgs& a, Ctl* ctl, int grid, hipStream_t s, hipEvent_t before);
// after launch_giant on the same stream, for a.wcc_mode 1 or 2 when the giant
// phase left active vertices: the WCC by label propagation, or the colouring
// SCC method, over them
hipError_t launch_tail(const GiantArgs& a, int grid, hipStream_t s);
// batched solves (batch.hip)
hipError_t launch_batch_union(const int32_t* fwb_cat, const int32_t* fwa_cat, const int32_t* bwb_cat,
                              const int32_t* bwa_cat, const int64_t* d_v_off, const int64_t* d_e_off, int32_t g0,
                              int32_t ng, int64_t Vu, int64_t Eu, int32_t* ufwb, int32_t* ufwa, int32_t* ubwb,
                              int32_t* ubwa, int32_t* bad, hipStream_t s);
// The same for a group of graphs that are not consecutive in the caller's
// layout: group graph i is the caller's graph sg[i] with the caller's offsets
// sv[i] = v_off[sg[i]], se[i] = e_off[sg[i]]; lv_off / le_off are the
// group's own prefix sums (ng + 1 entries from 0).
hipError_t launch_batch_union_perm(const int32_t* fwb_cat, const int32_t* fwa_cat, const int32_t* bwb_cat,
                                   const int32_t* bwa_cat, const int64_t* d_lv_off, const int64_t* d_le_off,
                                   const int64_t* d_sv, const int64_t* d_se, const int32_t* d_sg, int32_t ng,
                                   int64_t Vu, int64_t Eu, int32_t* ufwb, int32_t* ufwa, int32_t* ubwb, int32_t* ubwa,
                                   int32_t* bad, hipStream_t s);
// Graph ranges of one batch solve (device arrays, see batch.hip): graph g
// owns vertices [vb[g], ve[g]) of the CSR being solved; its local labels go
// to out[g]; cstart is the prefix of its kBatchChunk-vertex chunk counts.
struct BatchRanges {
    int32_t n = 0;
    int64_t nchunks = 0;
    int64_t* vb = nullptr;
    int64_t* ve = nullptr;
    int64_t* cstart = nullptr;
    int32_t** out = nullptr;
    int64_t* cp = nullptr;  // per-chunk pivot candidates
    int32_t* ci = nullptr;
    int32_t* roots = nullptr;
    int32_t* gmin = nullptr;
};
constexpr int64_t kBatchChunk = 4096;  // vertices per chunk: one block each
hipError_t launch_batch_roots(const int32_t* fwb, const int32_t* bwb, const BatchRanges& r, int32_t wrap32,
                              hipStream_t s);
hipError_t launch_batch_finish(const int32_t* lab, const uint64_t* bvis, const BatchRanges& r, hipStream_t s);
hipError_t launch_batch_select(const BatchRanges& r, uint64_t* sel, int64_t nwords, hipStream_t s);
hipError_t launch_arena_append(const int32_t* beg, const int32_t* adj, int32_t V, int64_t E, int64_t vb, int64_t eb,
                               int32_t* abeg, int32_t* aadj, hipStream_t s);
hipError_t launch_first(const int32_t* beg, const int32_t* adj, int32_t V, int32_t* first, int grid, hipStream_t s);

// Grid sizes for grid-stride kernels.
struct Launch {
    int grid_v;     // sweeps over V
    int grid_q;     // sweeps over queues
    int cu;         // compute units
};

// ---- kernel launchers (kernels.hip) --------------------------------------
// All take the stream last and return the hipError_t of the launch.
hipError_t launch_reset_ctl(Ctl* ctl, hipStream_t s);
hipError_t launch_fill_i32(int32_t* p, int32_t v, int64_t n, int grid, hipStream_t s);
hipError_t launch_trim1_first(const int32_t* fwb, const int32_t* bwb, int32_t V, int32_t* lab, uint64_t* act,
                              uint64_t* fvis, uint64_t* bvis, int32_t* fq, int32_t* q0, Ctl* ctl, PivotPart* part,
                              int pivot_wrap32, int grid, hipStream_t s);
// One top-down level: expand q[par] into q[par^1] (forward, bw == false:
// out-edges, claim act & ~fvis; backward: in-edges, claim fvis & ~bvis).
hipError_t launch_td_level(bool bw, const int32_t* beg, const int32_t* adj, const uint64_t* allow, uint64_t* vis,
                           int32_t* const q[2], HubItem* const hub[2], Ctl* ctl, int par, int grid, hipStream_t s,
                           hipEvent_t ev_mid = nullptr);
// One bottom-up sweep over the candidates allow & ~vis; pushes every vertex
// it visits into q[par^1].
hipError_t launch_bu_level(bool bw, const int32_t* beg, const int32_t* adj, const uint64_t* allow, uint64_t* vis,
                           int32_t* const q[2], HubItem* const hub[2], Ctl* ctl, int par, int64_t nwords, int grid,
                           hipStream_t s);
hipError_t launch_bw_init(uint64_t* bvis, int32_t* q0, Ctl* ctl, hipStream_t s);
hipError_t launch_popcount(const uint64_t* bm, int64_t nwords, int64_t* out, int grid, hipStream_t s);
// act &= ~bvis, giant labels, fq[1] rebuilt from act.
hipError_t launch_post_giant(int32_t V, int32_t* lab, uint64_t* act, const uint64_t* bvis, int32_t* fq, Ctl* ctl,
                             int grid, hipStream_t s);
// trim-1 sweep fq[sel] -> fq[sel^1]
hipError_t launch_trim1_sweep(const int32_t* fwb, const int32_t* fwa, const int32_t* bwb, const int32_t* bwa,
                              int32_t* lab, uint64_t* act, int32_t* const fq[2], Ctl* ctl, int sel, int grid,
                              hipStream_t s);
hipError_t launch_pattern_trim(const int32_t* fwb, const int32_t* fwa, const int32_t* bwb, const int32_t* bwa,
                               int32_t* lab, int32_t* stamp, const uint64_t* act, const int32_t* fq, Ctl* ctl,
                               int sel, int grid, hipStream_t s);
hipError_t launch_pattern_commit(int32_t* lab, uint64_t* act, const int32_t* stamp, const int32_t* fq, Ctl* ctl,
                                 int sel, int grid, hipStream_t s);
hipError_t launch_wcc_uf(const int32_t* fwb, const int32_t* fwa, const uint64_t* act, const int32_t* fq,
                         int32_t* parent, Ctl* ctl, int sel, int grid, hipStream_t s);
hipError_t launch_uf_init(const int32_t* fq, int32_t* parent, Ctl* ctl, int sel, int grid, hipStream_t s);
hipError_t launch_lp_round(const int32_t* fwb, const int32_t* fwa, const int32_t* bwb, const int32_t* bwa,
                           const uint64_t* act, const int32_t* fq, int32_t* color, Ctl* ctl, int sel, int grid,
                           hipStream_t s);
hipError_t launch_mice_keys(const int32_t* fq, const int32_t* wcc, int32_t* stamp, uint64_t* keys, Ctl* ctl, int sel,
                            int grid, hipStream_t s);
hipError_t launch_mice_starts(const uint64_t* keys, int32_t n, int32_t* starts, Ctl* ctl, int grid, hipStream_t s);
// max_wcc > 0: does nothing when a WCC has more vertices than that
// (ctl->max_wcc, from launch_mice_starts); the caller labels them otherwise
hipError_t launch_mice(const int32_t* fwb, const int32_t* fwa, const int32_t* bwb, const int32_t* bwa, int32_t* lab,
                       int32_t* stamp, const uint64_t* keys, int32_t n, const int32_t* starts, Ctl* ctl,
                       int32_t* qbuf, int32_t max_wcc, int grid, hipStream_t s);
hipError_t launch_count_unlabelled(const int32_t* lab, int32_t V, const uint64_t* sel, Ctl* ctl, int grid,
                                   hipStream_t s);
hipError_t launch_stats(const int32_t* fwb, const int32_t* bwb, const int32_t* lab, const uint64_t* fvis,
                        const uint64_t* bvis, int32_t V, Ctl* ctl, int grid, hipStream_t s);
// SCC sizes of canonical labels into size[V] (cleared first), then
// out[kHistSlots]: 64 log2 bins, largest, counts of sizes 1-3, SCC count.
constexpr int kHistSlots = 69;
hipError_t launch_scc_histogram(const int32_t* lab, int32_t V, int32_t* size, int64_t* out, int grid, hipStream_t s);
hipError_t launch_check_csr(const int32_t* beg, const int32_t* adj, int32_t V, int64_t E, Ctl* ctl, int grid,
                            hipStream_t s);

// ---- graph building (gen.hip).  Return 0, -1 bad argument, -2 bad edge,
// -3 out of memory, -4 HIP error (message in err).
int gen_rmat_device(hipStream_t s, int32_t V, double ef, double a, double b, double c, uint64_t seed, int32_t** fwb,
                    int32_t** fwa, int32_t** bwb, int32_t** bwa, int64_t* E, std::string& err);
int gen_grid_device(hipStream_t s, int32_t n, double p_bidir, uint64_t seed, int32_t** fwb, int32_t** fwa,
                    int32_t** bwb, int32_t** bwa, int64_t* E, std::string& err);
int build_csr_device(hipStream_t s, int32_t V, int64_t m, const int32_t* src, const int32_t* dst, int32_t** fwb,
                     int32_t** fwa, int32_t** bwb, int32_t** bwa, int64_t* E, std::string& err);
int csr_stable_one(hipStream_t s, int32_t V, int64_t m, const int32_t* key, const int32_t* val, int32_t** beg,
                   int32_t** adj, std::string& err);

}  // namespace ispan
