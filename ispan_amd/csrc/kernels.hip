// gfx950 kernels of the iSpan SCC pipeline.
//
// Per-vertex state in HBM:
//   lab[V]   int32   -1 while unassigned, else the canonical SCC label
//   act[W]   uint64  bit v: v is still unassigned (W = ceil(V/64))
//   fvis[W]  uint64  bit v: reached by the forward traversal from the pivot
//   bvis[W]  uint64  bit v: reached by the backward traversal inside fvis
// The bitmaps are 2 MB each at V = 2^24, so the random probes of the
// traversals stay in L2/MALL while adjacency is streamed from HBM.
#include "common.h"
#include "device_util.h"

namespace ispan {

using namespace dev;


// ------------------------------------------------------------------ reset

__global__ void k_reset_ctl(Ctl* ctl) {
    int32_t* p = reinterpret_cast<int32_t*>(ctl);
    for (int i = threadIdx.x; i < (int)(sizeof(Ctl) / 4); i += blockDim.x) p[i] = 0;
}

__global__ void k_fill_i32(int32_t* p, int32_t v, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// ---------------------------------------- P1-P3: trim-1-first, queue, pivot
// trim_1_first (src/trim_1_gfq.h:9-33) + generate_frontier_queue
// (src/trim_1_gfq.h:486-529) + pivot_selection_from_fq (src/scc_common.h:
// 92-150) in one sweep over the two offset arrays.  Each 4096-vertex tile
// compacts its survivors in ascending order with one global atomic; the
// pivot is the largest out*in, lowest id on ties, reduced by the last block.
__global__ __launch_bounds__(kBlock) void k_trim1_first(const int32_t* __restrict__ fwb, const int32_t* __restrict__ bwb,
                                                        int32_t V, int32_t* __restrict__ lab, uint64_t* act,
                                                        uint64_t* fvis, uint64_t* bvis, int32_t* __restrict__ fq,
                                                        int32_t* q0, Ctl* ctl, PivotPart* part, int wrap32) {
    constexpr int J = kTrimTile / kBlock;
    __shared__ int s_off[J * (kBlock / kWave)];
    __shared__ int64_t s_bp[kBlock / kWave];
    __shared__ int32_t s_bi[kBlock / kWave];
    __shared__ int s_last;
    const int tid = threadIdx.x, lane = lane_id(), wave = wave_id();
    int64_t bestp = 0;
    int32_t besti = INT32_MAX;
    const int64_t ntiles = ((int64_t)V + kTrimTile - 1) / kTrimTile;
    for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int64_t base = tile * kTrimTile;
        uint64_t masks[J];
#pragma unroll
        for (int j = 0; j < J; ++j) {
            const int64_t v = base + j * kBlock + tid;
            bool surv = false;
            if (v < V) {
                const int32_t od = fwb[v + 1] - fwb[v];
                const int32_t id = bwb[v + 1] - bwb[v];
                surv = od > 0 && id > 0;
                lab[v] = surv ? -1 : (int32_t)v;
                if (surv) {
                    int64_t p = wrap32 ? (int64_t)(int32_t)((uint32_t)od * (uint32_t)id) : (int64_t)od * id;
                    if (p > 0 && pivot_better(p, (int32_t)v, bestp, besti)) {
                        bestp = p;
                        besti = (int32_t)v;
                    }
                }
            }
            const uint64_t m = __ballot(surv);
            masks[j] = m;
            const int64_t first = base + j * kBlock + wave * kWave;
            if (lane == 0) {
                if (first < V) {
                    act[first >> 6] = m;
                    fvis[first >> 6] = 0;
                    bvis[first >> 6] = 0;
                }
                s_off[j * (kBlock / kWave) + wave] = __popcll(m);
            }
        }
        __syncthreads();
        if (wave == 0) {
            const int c = s_off[lane];
            const int incl = wave_incl_scan(c);
            const int total = lane_get(incl, kWave - 1);
            int b = 0;
            if (lane == 0 && total) b = atomicAdd(&ctl->fq_n[0], total);
            b = lane_get(b, 0);
            s_off[lane] = b + incl - c;
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < J; ++j) {
            if ((masks[j] >> lane) & 1ull)
                fq[s_off[j * (kBlock / kWave) + wave] + __popcll(masks[j] & lanemask_lt())] =
                    (int32_t)(base + j * kBlock + tid);
        }
        __syncthreads();
    }
    // block reduction of the pivot candidate
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) {
        int64_t op = __shfl_xor(bestp, d, kWave);
        int32_t oi = __shfl_xor(besti, d, kWave);
        if (pivot_better(op, oi, bestp, besti)) {
            bestp = op;
            besti = oi;
        }
    }
    if (lane == 0) {
        s_bp[wave] = bestp;
        s_bi[wave] = besti;
    }
    __syncthreads();
    if (tid == 0) {
        for (int w = 1; w < kBlock / kWave; ++w)
            if (pivot_better(s_bp[w], s_bi[w], bestp, besti)) {
                bestp = s_bp[w];
                besti = s_bi[w];
            }
        part[blockIdx.x].prod = bestp;
        part[blockIdx.x].id = besti;
        __threadfence();
        const unsigned prev = atomicAdd(&ctl->done_blocks, 1u);
        s_last = prev == gridDim.x - 1;
    }
    __syncthreads();
    if (!s_last) return;
    // last block: global pivot and traversal start state
    __threadfence();
    bestp = 0;
    besti = INT32_MAX;
    for (int b = tid; b < (int)gridDim.x; b += kBlock) {
        const int64_t p = load_agent(&part[b].prod);
        const int32_t i = load_agent(&part[b].id);
        if (pivot_better(p, i, bestp, besti)) {
            bestp = p;
            besti = i;
        }
    }
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) {
        int64_t op = __shfl_xor(bestp, d, kWave);
        int32_t oi = __shfl_xor(besti, d, kWave);
        if (pivot_better(op, oi, bestp, besti)) {
            bestp = op;
            besti = oi;
        }
    }
    if (lane == 0) {
        s_bp[wave] = bestp;
        s_bi[wave] = besti;
    }
    __syncthreads();
    if (tid == 0) {
        for (int w = 1; w < kBlock / kWave; ++w)
            if (pivot_better(s_bp[w], s_bi[w], bestp, besti)) {
                bestp = s_bp[w];
                besti = s_bi[w];
            }
        const int32_t root = (bestp > 0 && besti < V) ? besti : -1;
        ctl->root = root;
        ctl->done_blocks = 0;
        ctl->giant_min = root;
        ctl->q_n[0] = root >= 0 ? 1 : 0;
        ctl->q_n[1] = 0;
        ctl->hub_n[0] = ctl->hub_n[1] = 0;
        ctl->levels = 0;
        if (root >= 0) {
            atomicOr(reinterpret_cast<unsigned long long*>(fvis + (root >> 6)), 1ull << (root & 63));
            q0[0] = root;
        }
    }
}

// ------------------------------------------------------- traversal levels
// Top-down level (fw_bfs_fq_queue / bw_bfs_fq_queue sync and async top-down
// branches, src/fw_bw.h:453-478,506-544 and :992-1026,1066-1108).  A block
// stages 256 frontier vertices, scans their degrees in LDS and walks the
// block's combined edge range with all 256 threads (coalesced adjacency
// reads).  Vertices of degree >= kHubDeg are split into kHubChunk-edge work
// items for k_td_hub, which every level runs last.
template <bool BW>
__global__ __launch_bounds__(kBlock) void k_td(const int32_t* __restrict__ beg, const int32_t* __restrict__ adj,
                                               const uint64_t* __restrict__ allow, uint64_t* vis,
                                               const int32_t* __restrict__ qin, int32_t* qout, HubItem* hub_out,
                                               Ctl* ctl, int par) {
    __shared__ int32_t s_off[kBlock];
    __shared__ int32_t s_start[kBlock];
    __shared__ int s_tmp[kBlock / kWave];
    __shared__ int32_t s_wbuf[kBlock / kWave][kWaveBuf];
    const int tid = threadIdx.x;
    const int n = ctl->q_n[par];
    int32_t* qn_out = &ctl->q_n[par ^ 1];
    int32_t* hub_n = &ctl->hub_n[par];
    WaveQueue wq{s_wbuf[wave_id()], 0};
    int gmin = INT32_MAX;
    int64_t edges = 0, found = 0;
    for (int64_t tb = (int64_t)blockIdx.x * kBlock; tb < n; tb += (int64_t)gridDim.x * kBlock) {
        const int64_t i = tb + tid;
        int32_t start = 0, deg = 0;
        if (i < n) {
            const int32_t v = qin[i];
            start = beg[v];
            deg = beg[v + 1] - start;
            if (deg >= kHubDeg) {
                const int nch = (deg + kHubChunk - 1) / kHubChunk;
                const int h = atomicAdd(hub_n, nch);
                for (int c = 0; c < nch; ++c) hub_out[h + c] = HubItem{v, start + c * kHubChunk};
                deg = 0;
            }
        }
        int excl, total;
        block_excl_scan(deg, excl, total, s_tmp);
        s_off[tid] = excl;
        s_start[tid] = start;
        if (tid == 0) edges += total;
        __syncthreads();
        for (int eb = 0; eb < total; eb += kBlock) {
            const int e = eb + tid;
            bool claim = false;
            int32_t w = 0;
            if (e < total) {
                int lo = 0, hi = kBlock - 1;  // last j with s_off[j] <= e
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (s_off[mid] <= e) lo = mid;
                    else hi = mid - 1;
                }
                w = adj[s_start[lo] + (e - s_off[lo])];
                claim = try_claim(w, allow, vis);
            }
            wq_push(wq, claim, w, qout, qn_out);
            found += claim;
            if (BW && claim) gmin = min(gmin, w);
        }
        __syncthreads();
    }
    wq_flush(wq, qout, qn_out);
    wave_add64(&ctl->work_edges[0], edges);
    wave_add64(&ctl->work_found[0], found);
    if (BW) {
        gmin = wave_min(gmin);
        if (lane_id() == 0 && gmin != INT32_MAX) atomicMin(&ctl->giant_min, gmin);
    }
}

// Hub work items of this level; also the level's bookkeeping: the input
// count q_n[par] is consumed (k_td has completed) and reset, and the next
// level's hub counter is cleared.
template <bool BW>
__global__ __launch_bounds__(kBlock) void k_td_hub(const int32_t* __restrict__ beg, const int32_t* __restrict__ adj,
                                                   const uint64_t* __restrict__ allow, uint64_t* vis,
                                                   const HubItem* __restrict__ hub_in, int32_t* qout, Ctl* ctl,
                                                   int par) {
    __shared__ int32_t s_wbuf[kBlock / kWave][kWaveBuf];
    const int tid = threadIdx.x;
    if (blockIdx.x == 0 && tid == 0) {
        if (ctl->q_n[par] > 0) ctl->levels += 1;
        ctl->q_n[par] = 0;
        ctl->hub_n[par ^ 1] = 0;
    }
    const int nh = ctl->hub_n[par];
    int32_t* qn_out = &ctl->q_n[par ^ 1];
    WaveQueue wq{s_wbuf[wave_id()], 0};
    int gmin = INT32_MAX;
    int64_t edges = 0, found = 0;
    for (int h = blockIdx.x; h < nh; h += gridDim.x) {
        const HubItem it = hub_in[h];
        const int32_t end = min(it.start + kHubChunk, beg[it.v + 1]);
        if (tid == 0) edges += end - it.start;
        for (int32_t eb = it.start; eb < end; eb += kBlock) {
            const int32_t e = eb + tid;
            bool claim = false;
            int32_t w = 0;
            if (e < end) {
                w = adj[e];
                claim = try_claim(w, allow, vis);
            }
            wq_push(wq, claim, w, qout, qn_out);
            found += claim;
            if (BW && claim) gmin = min(gmin, w);
        }
    }
    wq_flush(wq, qout, qn_out);
    wave_add64(&ctl->work_edges[1], edges);
    wave_add64(&ctl->work_found[1], found);
    if (BW) {
        gmin = wave_min(gmin);
        if (lane_id() == 0 && gmin != INT32_MAX) atomicMin(&ctl->giant_min, gmin);
    }
}

// Bottom-up sweep (the Rsync branches, src/fw_bw.h:480-505 and :1040-1064).
// Candidates come from the bitmaps (allow & ~vis), 2048 vertices per tile
// compacted into LDS; each candidate scans its reverse-direction edges and
// stops at the first visited neighbour.  Bits set during the sweep are seen
// by later checks (asynchronous, as in the reference).  Every vertex found
// is also pushed, so a switch to top-down needs no gather pass.
template <bool BW>
__global__ __launch_bounds__(kBlock) void k_bu(const int32_t* __restrict__ beg, const int32_t* __restrict__ adj,
                                               const uint64_t* __restrict__ allow, uint64_t* vis, int32_t* qout,
                                               Ctl* ctl, int par, int64_t nwords) {
    __shared__ int32_t s_list[kBuTileWords * 64];
    __shared__ int s_n;
    __shared__ int32_t s_wbuf[kBlock / kWave][kWaveBuf];
    const int tid = threadIdx.x, lane = lane_id();
    if (blockIdx.x == 0 && tid == 0) {
        ctl->levels += 1;
        ctl->q_n[par] = 0;
        ctl->hub_n[par ^ 1] = 0;
    }
    int32_t* qn_out = &ctl->q_n[par ^ 1];
    WaveQueue wq{s_wbuf[wave_id()], 0};
    int gmin = INT32_MAX;
    int scanned = 0, nfound = 0;
    auto visited = [vis](int32_t u, int32_t) { return bit_of(vis, u); };
    for (int64_t t = blockIdx.x; t * kBuTileWords < nwords; t += gridDim.x) {
        if (wave_id() == 0) {
            const int64_t wi = t * kBuTileWords + lane;
            uint64_t m = 0;
            if (lane < kBuTileWords && wi < nwords) m = allow[wi] & ~vis[wi];
            const int c = __popcll(m);
            const int incl = wave_incl_scan(c);
            int o = incl - c;
            while (m) {
                const int b = __ffsll((unsigned long long)m) - 1;
                s_list[o++] = (int32_t)((wi << 6) + b);
                m &= m - 1;
            }
            if (lane == kWave - 1) s_n = incl;
        }
        __syncthreads();
        const int n = s_n;
        for (int kb = 0; kb < n; kb += kBlock) {
            const int k = kb + tid;
            const bool have = k < n;
            const int32_t v = have ? s_list[k] : 0;
            const int32_t p = have ? beg[v] : 0;
            const int32_t end = have ? beg[v + 1] : 0;
            const bool found = wave_any_edge(have, p, end, v, adj, visited, scanned);
            if (found)
                atomicOr(reinterpret_cast<unsigned long long*>(vis + (v >> 6)), 1ull << (v & 63));
            wq_push(wq, found, v, qout, qn_out);
            nfound += found;
            if (BW && found) gmin = min(gmin, v);
        }
        __syncthreads();
    }
    wq_flush(wq, qout, qn_out);
    wave_add64(&ctl->work_edges[2], scanned);
    wave_add64(&ctl->work_found[2], nfound);
    if (BW) {
        gmin = wave_min(gmin);
        if (lane_id() == 0 && gmin != INT32_MAX) atomicMin(&ctl->giant_min, gmin);
    }
}

// Backward traversal start: the pivot (src/fw_bw.h:972-976).
__global__ void k_bw_init(uint64_t* bvis, int32_t* q0, Ctl* ctl) {
    const int32_t root = ctl->root;
    ctl->q_n[0] = root >= 0 ? 1 : 0;
    ctl->q_n[1] = 0;
    ctl->hub_n[0] = ctl->hub_n[1] = 0;
    ctl->levels = 0;
    if (root >= 0) {
        bvis[root >> 6] |= 1ull << (root & 63);
        q0[0] = root;
        ctl->giant_min = root;
    }
}

__global__ __launch_bounds__(kBlock) void k_popcount(const uint64_t* __restrict__ bm, int64_t nwords, int64_t* out) {
    int c = 0;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < nwords; i += (int64_t)gridDim.x * kBlock)
        c += __popcll(bm[i]);
    c = wave_sum(c);
    if (lane_id() == 0 && c) atomicAdd(reinterpret_cast<unsigned long long*>(out), (unsigned long long)c);
}

// ------------------------------------------------------------ after giant
// Giant members get the giant's smallest id; act loses them; the active
// queue fq is rebuilt from the bitmap (ascending within each wave).
__global__ __launch_bounds__(kBlock) void k_post_giant(int32_t V, int32_t* __restrict__ lab, uint64_t* act,
                                                       const uint64_t* __restrict__ bvis, int32_t* fq, Ctl* ctl) {
    __shared__ int32_t s_wbuf[kBlock / kWave][kWaveBuf];
    const int32_t gmin = ctl->giant_min;
    WaveQueue wq{s_wbuf[wave_id()], 0};
    int32_t* cnt = &ctl->fq_n[1];
    for (int64_t tb = (int64_t)blockIdx.x * kBlock; tb < V; tb += (int64_t)gridDim.x * kBlock) {
        const int64_t v = tb + threadIdx.x;
        const bool in = v < V;
        uint64_t g = 0, a = 0;
        if (in) {
            g = bvis[v >> 6];
            a = act[v >> 6];
        }
        const bool giant = in && ((g >> (v & 63)) & 1ull);
        const bool keep = in && ((a >> (v & 63)) & 1ull) && !giant;
        if (giant) lab[v] = gmin;
        const uint64_t mk = __ballot(keep);
        if (lane_id() == 0 && in && g) act[v >> 6] = mk;
        wq_push(wq, keep, (int32_t)v, fq, cnt);
    }
    wq_flush(wq, fq, cnt);
}

// ---------------------------------------------------------------- trim-1
// trim_1_from_fq_gfq (src/trim_1_gfq.h:365-458): an active vertex with no
// active non-self out-neighbour or no active non-self in-neighbour is a
// singleton SCC.  act is updated in place, so later checks in the same sweep
// see it; survivors are compacted into fq[sel^1].
__global__ void k_prep_sweep(Ctl* ctl, int sel) {
    ctl->fq_n[sel ^ 1] = 0;
    ctl->trimmed = 0;
}

__global__ __launch_bounds__(kBlock) void k_trim1(const int32_t* __restrict__ fwb, const int32_t* __restrict__ fwa,
                                                  const int32_t* __restrict__ bwb, const int32_t* __restrict__ bwa,
                                                  int32_t* __restrict__ lab, uint64_t* act,
                                                  const int32_t* __restrict__ fq_in, int32_t* fq_out, Ctl* ctl,
                                                  int sel) {
    __shared__ int32_t s_wbuf[kBlock / kWave][kWaveBuf];
    const int n = ctl->fq_n[sel];
    int32_t* cnt = &ctl->fq_n[sel ^ 1];
    WaveQueue wq{s_wbuf[wave_id()], 0};
    int trimmed = 0, scanned = 0;
    auto live = [act](int32_t w, int32_t self) { return w != self && bit_of(act, w); };
    for (int64_t tb = (int64_t)blockIdx.x * kBlock; tb < n; tb += (int64_t)gridDim.x * kBlock) {
        const int64_t i = tb + threadIdx.x;
        int32_t v = 0;
        bool alive = false;
        if (i < n) {
            v = fq_in[i];
            alive = bit_of(act, v);
        }
        const bool out = wave_any_edge(alive, alive ? fwb[v] : 0, alive ? fwb[v + 1] : 0, v, fwa, live, scanned);
        const bool in =
            wave_any_edge(alive && out, alive ? bwb[v] : 0, alive ? bwb[v + 1] : 0, v, bwa, live, scanned);
        const bool trim = alive && !(out && in);
        if (trim) {
            lab[v] = v;
            atomicAnd(reinterpret_cast<unsigned long long*>(act + (v >> 6)), ~(1ull << (v & 63)));
            ++trimmed;
        }
        wq_push(wq, alive && !trim, v, fq_out, cnt);
    }
    wq_flush(wq, fq_out, cnt);
    trimmed = wave_sum(trimmed);
    if (lane_id() == 0 && trimmed) atomicAdd(&ctl->trimmed, trimmed);
}

// ---------------------------------------------------------- size-2/3 trims
// trim_2_from_fq (src/trim_2_3.h:146-283), trim_3_1_from_fq (:482-677) and
// trim_3_2_from_fq (:881-1080), checked against act as it stands when the
// kernel starts: detections only write lab (and the pattern size into
// stamp); act is cleared afterwards by k_pattern_commit.  With a fixed
// active set every detected group is an SCC of the active subgraph, so two
// detections of one vertex always agree on its label.

__global__ __launch_bounds__(kBlock) void k_pattern(const int32_t* __restrict__ fwb, const int32_t* __restrict__ fwa,
                                                    const int32_t* __restrict__ bwb, const int32_t* __restrict__ bwa,
                                                    int32_t* lab, int32_t* stamp, const uint64_t* __restrict__ act,
                                                    const int32_t* __restrict__ fq, const Ctl* ctl, int sel) {
    const int n = ctl->fq_n[sel];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t a = fq[i];
        if (bit_of(act, a)) pattern_check(fwb, fwa, bwb, bwa, lab, stamp, act, a);
    }
}

__global__ __launch_bounds__(kBlock) void k_pattern_commit(const int32_t* __restrict__ lab, uint64_t* act,
                                                           const int32_t* __restrict__ stamp,
                                                           const int32_t* __restrict__ fq, Ctl* ctl, int sel) {
    const int n = ctl->fq_n[sel];
    int c2 = 0, c3 = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t v = fq[i];
        if (lab[v] >= 0 && bit_of(act, v)) {
            atomicAnd(reinterpret_cast<unsigned long long*>(act + (v >> 6)), ~(1ull << (v & 63)));
            if (stamp[v] == 2) ++c2;
            else ++c3;
        }
    }
    c2 = wave_sum(c2);
    c3 = wave_sum(c3);
    if (lane_id() == 0) {
        if (c2) atomicAdd(&ctl->pattern2, c2);
        if (c3) atomicAdd(&ctl->pattern3, c3);
    }
}

// ------------------------------------------------------------------- WCC
// Weakly connected components of the active subgraph (the role of
// coloring_wcc, src/color_propagation.h:468-593, and openmp_wcc's CAS union,
// src/openmp_wcc.hpp:86-114).  Union-find: each active edge links the larger
// root under the smaller one with a CAS, then every vertex jumps to its
// root, so parent[v] = smallest id of v's WCC.  Out-edges of active vertices
// cover every active edge once.
__device__ __forceinline__ int32_t uf_root(int32_t* parent, int32_t v) {
    int32_t nx = load_agent(&parent[v]);
    while (nx != v) {
        v = nx;
        nx = load_agent(&parent[v]);
    }
    return v;
}

__global__ __launch_bounds__(kBlock) void k_uf_init(const int32_t* __restrict__ fq, int32_t* parent, const Ctl* ctl,
                                                    int sel) {
    const int n = ctl->fq_n[sel];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        parent[fq[i]] = fq[i];
}

__global__ __launch_bounds__(kBlock) void k_uf_hook(const int32_t* __restrict__ fwb, const int32_t* __restrict__ fwa,
                                                    const uint64_t* __restrict__ act, const int32_t* __restrict__ fq,
                                                    int32_t* parent, const Ctl* ctl, int sel) {
    const int n = ctl->fq_n[sel];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t v = fq[i];
        for (int32_t k = fwb[v]; k < fwb[v + 1]; ++k) {
            const int32_t w = fwa[k];
            if (w == v || !bit_of(act, w)) continue;
            int32_t a = v, b = w;
            while (true) {
                a = uf_root(parent, a);
                b = uf_root(parent, b);
                if (a == b) break;
                if (a < b) {
                    const int32_t t = a;
                    a = b;
                    b = t;
                }
                if (atomicCAS(&parent[a], a, b) == a) break;
            }
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_uf_compress(const int32_t* __restrict__ fq, int32_t* parent,
                                                        const Ctl* ctl, int sel) {
    const int n = ctl->fq_n[sel];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t v = fq[i];
        parent[v] = uf_root(parent, v);
    }
}

// Label propagation variant (coloring_wcc's scheme with the smallest id):
// color[v] = min over active in/out neighbours, then pointer jumping, until
// a sweep changes nothing.
__global__ __launch_bounds__(kBlock) void k_lp_sweep(const int32_t* __restrict__ fwb, const int32_t* __restrict__ fwa,
                                                     const int32_t* __restrict__ bwb, const int32_t* __restrict__ bwa,
                                                     const uint64_t* __restrict__ act, const int32_t* __restrict__ fq,
                                                     int32_t* color, Ctl* ctl, int sel) {
    const int n = ctl->fq_n[sel];
    int changed = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t v = fq[i];
        int32_t c = load_agent(&color[v]);
        const int32_t c0 = c;
        for (int dir = 0; dir < 2; ++dir) {
            const int32_t* beg = dir == 0 ? bwb : fwb;
            const int32_t* adj = dir == 0 ? bwa : fwa;
            for (int32_t k = beg[v]; k < beg[v + 1]; ++k) {
                const int32_t w = adj[k];
                if (w == v || !bit_of(act, w)) continue;
                c = min(c, load_agent(&color[w]));
            }
        }
        if (c < c0) {
            atomicMin(&color[v], c);
            changed = 1;
        }
    }
    if (__ballot(changed) && lane_id() == 0) atomicOr(&ctl->changed, 1);
}

__global__ __launch_bounds__(kBlock) void k_lp_jump(const int32_t* __restrict__ fq, int32_t* color, const Ctl* ctl,
                                                    int sel) {
    const int n = ctl->fq_n[sel];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t v = fq[i];
        int32_t r = load_agent(&color[v]);
        for (int hop = 0; hop < 100; ++hop) {
            const int32_t nr = load_agent(&color[r]);
            if (nr == r) break;
            r = nr;
        }
        atomicMin(&color[v], r);
    }
}

__global__ void k_clear_changed(Ctl* ctl) { ctl->changed = 0; }

// ------------------------------------------------------------------ mice
// init_fw_sa + mice_fw_bw (src/fw_bw.h:1548-1691) regrouped: (wcc, v) keys
// are sorted so each WCC is a contiguous ascending run; one wave owns a WCC
// and repeatedly takes its smallest unlabelled member as pivot, runs a FW
// search (visit stamp = pivot) and a BW search inside it, and labels that
// SCC with the pivot -- its smallest id.
__global__ __launch_bounds__(kBlock) void k_mice_keys(const int32_t* __restrict__ fq, const int32_t* __restrict__ wcc,
                                                      int32_t* stamp, uint64_t* keys, const Ctl* ctl, int sel) {
    const int n = ctl->fq_n[sel];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t v = fq[i];
        keys[i] = ((uint64_t)(uint32_t)wcc[v] << 32) | (uint32_t)v;
        stamp[v] = -1;
    }
}

// The start of every WCC in the sorted (WCC, vertex) keys, and the largest
// WCC's size: a start finds its WCC's end by binary search (keys ascend by
// WCC in their high half).
__global__ __launch_bounds__(kBlock) void k_mice_starts(const uint64_t* __restrict__ keys, int32_t n, int32_t* starts,
                                                        Ctl* ctl) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t w = keys[i] >> 32;
        if (i == 0 || w != (keys[i - 1] >> 32)) {
            starts[atomicAdd(&ctl->n_wcc, 1)] = (int32_t)i;
            int64_t lo = i + 1, hi = n;  // first index past the WCC
            while (lo < hi) {
                const int64_t mid = (lo + hi) / 2;
                if ((keys[mid] >> 32) == w) lo = mid + 1;
                else hi = mid;
            }
            atomicMax(&ctl->max_wcc, (int32_t)(lo - i));
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_mice(const int32_t* __restrict__ fwb, const int32_t* __restrict__ fwa,
                                                 const int32_t* __restrict__ bwb, const int32_t* __restrict__ bwa,
                                                 int32_t* lab, int32_t* stamp, const uint64_t* __restrict__ keys,
                                                 int32_t n, const int32_t* __restrict__ starts, const Ctl* ctl,
                                                 int32_t* qbuf, int32_t max_wcc) {
    __shared__ int s_tail[kBlock / kWave];
    if (max_wcc > 0 && ctl->max_wcc > max_wcc) return;  // left to the caller's other method
    const int lane = lane_id(), wave = wave_id();
    const int nw = ctl->n_wcc;
    const int waves = gridDim.x * (kBlock / kWave);
    for (int idx = blockIdx.x * (kBlock / kWave) + wave; idx < nw; idx += waves) {
        const int32_t s = starts[idx];
        const uint32_t root = (uint32_t)(keys[s] >> 32);
        // segment end
        int32_t e = s;
        for (int32_t b = s; b < n; b += kWave) {
            const int32_t j = b + lane;
            const bool other = j >= n || (uint32_t)(keys[j] >> 32) != root;
            const uint64_t m = __ballot(other);
            if (m) {
                e = b + __ffsll((unsigned long long)m) - 1;
                break;
            }
            e = b + kWave;
        }
        int32_t* q = qbuf + s;
        int32_t cursor = s;
        while (true) {
            // next pivot: smallest member still unlabelled
            int32_t pj = -1;
            for (int32_t b = cursor; b < e; b += kWave) {
                const int32_t j = b + lane;
                const bool fr = j < e && load_agent(&lab[(int32_t)(uint32_t)keys[j]]) == -1;
                const uint64_t m = __ballot(fr);
                if (m) {
                    pj = b + __ffsll((unsigned long long)m) - 1;
                    break;
                }
            }
            if (pj < 0) break;
            cursor = pj + 1;
            const int32_t pv = (int32_t)(uint32_t)keys[pj];
            for (int phase = 0; phase < 2; ++phase) {
                const int32_t* beg = phase == 0 ? fwb : bwb;
                const int32_t* adj = phase == 0 ? fwa : bwa;
                if (lane == 0) {
                    if (phase == 0) stamp[pv] = pv;
                    else lab[pv] = pv;
                    q[0] = pv;
                    s_tail[wave] = 1;
                }
                wave_sync();
                int32_t head = 0, tail = 1;
                while (head < tail) {
                    const int32_t lo = head, hi = tail;
                    head = hi;
                    for (int32_t b = lo; b < hi; b += kWave) {
                        const int32_t i = b + lane;
                        if (i < hi) {
                            const int32_t x = load_agent(&q[i]);
                            for (int32_t k = beg[x]; k < beg[x + 1]; ++k) {
                                const int32_t w = adj[k];
                                if (load_agent(&lab[w]) != -1) continue;
                                bool claim;
                                if (phase == 0)
                                    claim = load_agent(&stamp[w]) != pv && atomicExch(&stamp[w], pv) != pv;
                                else
                                    claim = load_agent(&stamp[w]) == pv && atomicCAS(&lab[w], -1, pv) == -1;
                                if (claim) q[atomicAdd(&s_tail[wave], 1)] = w;
                            }
                        }
                    }
                    wave_sync();
                    tail = s_tail[wave];
                    wave_sync();
                }
            }
        }
    }
}

// ---------------------------------------------------------------- checks
// sel (optional): only the selected vertices count.
__global__ __launch_bounds__(kBlock) void k_count_unlabelled(const int32_t* __restrict__ lab, int32_t V,
                                                             const uint64_t* __restrict__ sel, Ctl* ctl) {
    int c = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < V; i += (int64_t)gridDim.x * blockDim.x)
        c += (!sel || ((sel[i >> 6] >> (i & 63)) & 1ull)) && lab[i] < 0;
    c = wave_sum(c);
    if (lane_id() == 0 && c) atomicAdd(&ctl->unlabelled, c);
}

// Work statistics of the last solve: edges incident to the FW set (out) and
// to the giant (in), and the SCC count.
__global__ __launch_bounds__(kBlock) void k_stats(const int32_t* __restrict__ fwb, const int32_t* __restrict__ bwb,
                                                  const int32_t* __restrict__ lab, const uint64_t* __restrict__ fvis,
                                                  const uint64_t* __restrict__ bvis, int32_t V, Ctl* ctl) {
    int64_t fe = 0, be = 0;
    int ns = 0;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (int64_t)gridDim.x * blockDim.x) {
        if (bit_of(fvis, (int32_t)v)) fe += fwb[v + 1] - fwb[v];
        if (bit_of(bvis, (int32_t)v)) be += bwb[v + 1] - bwb[v];
        ns += lab[v] == v;
    }
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) {
        fe += __shfl_xor(fe, d, kWave);
        be += __shfl_xor(be, d, kWave);
    }
    ns = wave_sum(ns);
    if (lane_id() == 0) {
        if (fe) atomicAdd(reinterpret_cast<unsigned long long*>(&ctl->fw_edges), (unsigned long long)fe);
        if (be) atomicAdd(reinterpret_cast<unsigned long long*>(&ctl->bw_edges), (unsigned long long)be);
        if (ns) atomicAdd(reinterpret_cast<unsigned long long*>(&ctl->n_scc), (unsigned long long)ns);
    }
}

// Structural CSR check: beg[0] == 0, beg non-decreasing, beg[V] == E,
// every adjacency id in [0, V).
__global__ __launch_bounds__(kBlock) void k_check_csr(const int32_t* __restrict__ beg, const int32_t* __restrict__ adj,
                                                      int32_t V, int64_t E, Ctl* ctl) {
    int bad = 0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) bad += beg[0] != 0 || (int64_t)beg[V] != E;
    for (int64_t v = t; v < V; v += stride) bad += beg[v] > beg[v + 1];
    for (int64_t k = t; k < E; k += stride) bad += adj[k] < 0 || adj[k] >= V;
    bad = wave_sum(bad);
    if (lane_id() == 0 && bad) atomicAdd(&ctl->bad_csr, bad);
}

// ---------------------------------------------------- SCC size histogram
// Extends get_scc_result (src/load_print.cpp:24-60), which tallies SCC
// sizes in a std::map but prints only count classes.
//
// size[lab[v]] += 1 over canonical labels.  A wave takes 64 consecutive
// vertices; lanes that share a label add together (one leader per round,
// up to kSizeRounds rounds), the rest one atomic each, so a label carried by
// a third of the vertices (R-MAT's giant) costs one atomic per wave step
// instead of one per vertex.
constexpr int kSizeRounds = 4;
__global__ __launch_bounds__(kBlock) void k_scc_sizes(const int32_t* __restrict__ lab, int32_t V, int32_t* size) {
    const int lane = lane_id();
    const int64_t waves = (int64_t)gridDim.x * (kBlock / kWave);
    for (int64_t base = ((int64_t)blockIdx.x * (kBlock / kWave) + wave_id()) * kWave; base < V;
         base += waves * kWave) {
        const int64_t v = base + lane;
        const int32_t l = v < V ? lab[v] : -1;
        bool todo = v < V;
        for (int r = 0; r < kSizeRounds; ++r) {
            const uint64_t m = __ballot(todo);
            if (!m) break;
            const int lead = __ffsll((unsigned long long)m) - 1;
            const int32_t L = lane_get(l, lead);
            const uint64_t same = __ballot(todo && l == L);
            if (lane == lead) atomicAdd(&size[L], __popcll(same));
            if (l == L) todo = false;
        }
        if (todo) atomicAdd(&size[l], 1);
    }
}

// Log2 bins of the SCC sizes (bin k: sizes in [2^k, 2^(k+1))), per block in
// LDS, plus the largest size, the counts of sizes 1, 2, 3 and the SCC
// count: out[0..63] bins, out[64] largest, out[65..67] sizes 1-3, out[68]
// total.
__global__ __launch_bounds__(kBlock) void k_scc_hist(const int32_t* __restrict__ size, int32_t V, int64_t* out) {
    __shared__ int s_bin[64];
    if (threadIdx.x < 64) s_bin[threadIdx.x] = 0;
    __syncthreads();
    int largest = 0, c1 = 0, c2 = 0, c3 = 0, tot = 0;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (int64_t)gridDim.x * blockDim.x) {
        const int c = size[v];
        if (c <= 0) continue;
        atomicAdd(&s_bin[31 - __clz(c)], 1);
        largest = max(largest, c);
        c1 += c == 1;
        c2 += c == 2;
        c3 += c == 3;
        ++tot;
    }
    c1 = wave_sum(c1);
    c2 = wave_sum(c2);
    c3 = wave_sum(c3);
    tot = wave_sum(tot);
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) largest = max(largest, __shfl_xor(largest, d, kWave));
    if (lane_id() == 0) {
        if (largest) atomicMax(reinterpret_cast<unsigned long long*>(&out[64]), (unsigned long long)largest);
        if (c1) atomicAdd(reinterpret_cast<unsigned long long*>(&out[65]), (unsigned long long)c1);
        if (c2) atomicAdd(reinterpret_cast<unsigned long long*>(&out[66]), (unsigned long long)c2);
        if (c3) atomicAdd(reinterpret_cast<unsigned long long*>(&out[67]), (unsigned long long)c3);
        if (tot) atomicAdd(reinterpret_cast<unsigned long long*>(&out[68]), (unsigned long long)tot);
    }
    __syncthreads();
    if (threadIdx.x < 64 && s_bin[threadIdx.x])
        atomicAdd(reinterpret_cast<unsigned long long*>(&out[threadIdx.x]), (unsigned long long)s_bin[threadIdx.x]);
}

// ------------------------------------------------------------- launchers

#define LAUNCH(kern, grid, ...)                                  \
    do {                                                         \
        hipLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), 0, s, __VA_ARGS__); \
        return hipGetLastError();                                \
    } while (0)

hipError_t launch_reset_ctl(Ctl* ctl, hipStream_t s) { LAUNCH(k_reset_ctl, 1, ctl); }

hipError_t launch_fill_i32(int32_t* p, int32_t v, int64_t n, int grid, hipStream_t s) {
    LAUNCH(k_fill_i32, grid, p, v, n);
}

hipError_t launch_trim1_first(const int32_t* fwb, const int32_t* bwb, int32_t V, int32_t* lab, uint64_t* act,
                              uint64_t* fvis, uint64_t* bvis, int32_t* fq, int32_t* q0, Ctl* ctl, PivotPart* part,
                              int pivot_wrap32, int grid, hipStream_t s) {
    LAUNCH(k_trim1_first, grid, fwb, bwb, V, lab, act, fvis, bvis, fq, q0, ctl, part, pivot_wrap32);
}

hipError_t launch_td_level(bool bw, const int32_t* beg, const int32_t* adj, const uint64_t* allow, uint64_t* vis,
                           int32_t* const q[2], HubItem* const hub[2], Ctl* ctl, int par, int grid, hipStream_t s,
                           hipEvent_t ev_mid) {
    if (bw)
        hipLaunchKernelGGL(k_td<true>, dim3(grid), dim3(kBlock), 0, s, beg, adj, allow, vis, q[par], q[par ^ 1],
                           hub[par], ctl, par);
    else
        hipLaunchKernelGGL(k_td<false>, dim3(grid), dim3(kBlock), 0, s, beg, adj, allow, vis, q[par], q[par ^ 1],
                           hub[par], ctl, par);
    if (ev_mid) (void)hipEventRecord(ev_mid, s);
    if (bw)
        hipLaunchKernelGGL(k_td_hub<true>, dim3(grid), dim3(kBlock), 0, s, beg, adj, allow, vis, hub[par],
                           q[par ^ 1], ctl, par);
    else
        hipLaunchKernelGGL(k_td_hub<false>, dim3(grid), dim3(kBlock), 0, s, beg, adj, allow, vis, hub[par],
                           q[par ^ 1], ctl, par);
    return hipGetLastError();
}

hipError_t launch_bu_level(bool bw, const int32_t* beg, const int32_t* adj, const uint64_t* allow, uint64_t* vis,
                           int32_t* const q[2], HubItem* const hub[2], Ctl* ctl, int par, int64_t nwords, int grid,
                           hipStream_t s) {
    (void)hub;
    if (bw) {
        LAUNCH(k_bu<true>, grid, beg, adj, allow, vis, q[par ^ 1], ctl, par, nwords);
    } else {
        LAUNCH(k_bu<false>, grid, beg, adj, allow, vis, q[par ^ 1], ctl, par, nwords);
    }
}

hipError_t launch_bw_init(uint64_t* bvis, int32_t* q0, Ctl* ctl, hipStream_t s) {
    hipLaunchKernelGGL(k_bw_init, dim3(1), dim3(1), 0, s, bvis, q0, ctl);
    return hipGetLastError();
}

hipError_t launch_popcount(const uint64_t* bm, int64_t nwords, int64_t* out, int grid, hipStream_t s) {
    LAUNCH(k_popcount, grid, bm, nwords, out);
}

hipError_t launch_post_giant(int32_t V, int32_t* lab, uint64_t* act, const uint64_t* bvis, int32_t* fq, Ctl* ctl,
                             int grid, hipStream_t s) {
    LAUNCH(k_post_giant, grid, V, lab, act, bvis, fq, ctl);
}

hipError_t launch_trim1_sweep(const int32_t* fwb, const int32_t* fwa, const int32_t* bwb, const int32_t* bwa,
                              int32_t* lab, uint64_t* act, int32_t* const fq[2], Ctl* ctl, int sel, int grid,
                              hipStream_t s) {
    hipLaunchKernelGGL(k_prep_sweep, dim3(1), dim3(1), 0, s, ctl, sel);
    LAUNCH(k_trim1, grid, fwb, fwa, bwb, bwa, lab, act, fq[sel], fq[sel ^ 1], ctl, sel);
}

hipError_t launch_pattern_trim(const int32_t* fwb, const int32_t* fwa, const int32_t* bwb, const int32_t* bwa,
                               int32_t* lab, int32_t* stamp, const uint64_t* act, const int32_t* fq, Ctl* ctl,
                               int sel, int grid, hipStream_t s) {
    LAUNCH(k_pattern, grid, fwb, fwa, bwb, bwa, lab, stamp, act, fq, ctl, sel);
}

hipError_t launch_pattern_commit(int32_t* lab, uint64_t* act, const int32_t* stamp, const int32_t* fq, Ctl* ctl,
                                 int sel, int grid, hipStream_t s) {
    LAUNCH(k_pattern_commit, grid, lab, act, stamp, fq, ctl, sel);
}

hipError_t launch_wcc_uf(const int32_t* fwb, const int32_t* fwa, const uint64_t* act, const int32_t* fq,
                         int32_t* parent, Ctl* ctl, int sel, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_uf_init, dim3(grid), dim3(kBlock), 0, s, fq, parent, ctl, sel);
    hipLaunchKernelGGL(k_uf_hook, dim3(grid), dim3(kBlock), 0, s, fwb, fwa, act, fq, parent, ctl, sel);
    LAUNCH(k_uf_compress, grid, fq, parent, ctl, sel);
}

hipError_t launch_lp_round(const int32_t* fwb, const int32_t* fwa, const int32_t* bwb, const int32_t* bwa,
                           const uint64_t* act, const int32_t* fq, int32_t* color, Ctl* ctl, int sel, int grid,
                           hipStream_t s) {
    hipLaunchKernelGGL(k_clear_changed, dim3(1), dim3(1), 0, s, ctl);
    hipLaunchKernelGGL(k_lp_sweep, dim3(grid), dim3(kBlock), 0, s, fwb, fwa, bwb, bwa, act, fq, color, ctl, sel);
    LAUNCH(k_lp_jump, grid, fq, color, ctl, sel);
}

hipError_t launch_uf_init(const int32_t* fq, int32_t* parent, Ctl* ctl, int sel, int grid, hipStream_t s) {
    LAUNCH(k_uf_init, grid, fq, parent, ctl, sel);
}

hipError_t launch_mice_keys(const int32_t* fq, const int32_t* wcc, int32_t* stamp, uint64_t* keys, Ctl* ctl, int sel,
                            int grid, hipStream_t s) {
    LAUNCH(k_mice_keys, grid, fq, wcc, stamp, keys, ctl, sel);
}

hipError_t launch_mice_starts(const uint64_t* keys, int32_t n, int32_t* starts, Ctl* ctl, int grid, hipStream_t s) {
    LAUNCH(k_mice_starts, grid, keys, n, starts, ctl);
}

hipError_t launch_mice(const int32_t* fwb, const int32_t* fwa, const int32_t* bwb, const int32_t* bwa, int32_t* lab,
                       int32_t* stamp, const uint64_t* keys, int32_t n, const int32_t* starts, Ctl* ctl,
                       int32_t* qbuf, int32_t max_wcc, int grid, hipStream_t s) {
    LAUNCH(k_mice, grid, fwb, fwa, bwb, bwa, lab, stamp, keys, n, starts, ctl, qbuf, max_wcc);
}

hipError_t launch_count_unlabelled(const int32_t* lab, int32_t V, const uint64_t* sel, Ctl* ctl, int grid,
                                   hipStream_t s) {
    LAUNCH(k_count_unlabelled, grid, lab, V, sel, ctl);
}

hipError_t launch_check_csr(const int32_t* beg, const int32_t* adj, int32_t V, int64_t E, Ctl* ctl, int grid,
                            hipStream_t s) {
    LAUNCH(k_check_csr, grid, beg, adj, V, E, ctl);
}

hipError_t launch_stats(const int32_t* fwb, const int32_t* bwb, const int32_t* lab, const uint64_t* fvis,
                        const uint64_t* bvis, int32_t V, Ctl* ctl, int grid, hipStream_t s) {
    LAUNCH(k_stats, grid, fwb, bwb, lab, fvis, bvis, V, ctl);
}

hipError_t launch_scc_histogram(const int32_t* lab, int32_t V, int32_t* size, int64_t* out, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_i32, dim3(grid), dim3(kBlock), 0, s, size, 0, (int64_t)V);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(kBlock), 0, s, reinterpret_cast<int32_t*>(out), 0,
                       (int64_t)kHistSlots * 2);
    hipLaunchKernelGGL(k_scc_sizes, dim3(grid), dim3(kBlock), 0, s, lab, V, size);
    LAUNCH(k_scc_hist, grid, size, V, out);
}

}  // namespace ispan
