// Batched solves: many graphs in one pass of the pipeline over their
// disjoint union.  SCCs never cross graphs, so the union's SCCs are the
// graphs' SCCs; what changes is that every graph needs its own pivot (the
// reference's pivot_selection_from_fq, src/scc_common.h:92-150, per graph),
// its giant its own label, and the labels their graph's local ids.
//
// Input layout of the batch ABI (caller-owned, on the device): graph g's CSR pair with local
// ids, concatenated -- its V_g + 1 offsets start at beg[v_off[g] + g], its
// E_g adjacency entries at adj[e_off[g]].  A group of consecutive graphs
// whose union fits int32 offsets is rebuilt once into a union CSR (global
// ids, global offsets) that the single-graph kernels run on unchanged.
#include "common.h"
#include "device_util.h"

namespace ispan {
using namespace dev;

namespace {

constexpr int64_t kUnionChunk = 1 << 14;  // union building: vertices or edges per block

// Largest g with off[g] <= x, for off[0..n] ascending (x < off[n]).
__device__ __forceinline__ int32_t find_graph(const int64_t* off, int32_t n, int64_t x) {
    int32_t lo = 0, hi = n - 1;
    while (lo < hi) {
        const int32_t mid = (lo + hi + 1) >> 1;
        if (off[mid] <= x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Where graph g's arrays are: back to back (the batch ABI's layout).
struct CatSrc {
    const int32_t* beg;
    const int32_t* adj;
    const int64_t* v_off;
    const int64_t* e_off;
    __device__ const int32_t* beg_of(int32_t g) const { return beg + v_off[g] + g; }
    __device__ const int32_t* adj_of(int32_t g) const { return adj + e_off[g]; }
};

// Where graph i of a group is when the group takes graphs that are not
// consecutive in the caller's layout (a group of one size class): the
// caller's graph sg[i], whose offsets start at beg[sv[i] + sg[i]] and whose
// adjacency starts at adj[se[i]] (sv, se: the caller's v_off, e_off of it).
struct PermSrc {
    const int32_t* beg;
    const int32_t* adj;
    const int64_t* sv;
    const int64_t* se;
    const int32_t* sg;
    __device__ const int32_t* beg_of(int32_t i) const { return beg + sv[i] + sg[i]; }
    __device__ const int32_t* adj_of(int32_t i) const { return adj + se[i]; }
};

// Union offsets: ubeg[u] for the group's vertices u = v_off[g] - v_off[g0] + i
// is (e_off[g] - e_off[g0]) + beg_g[i].  Blocks take contiguous chunks, so a
// thread's graph index only moves forward.  bad |= 1: a graph's offsets do
// not start at 0 / end at E_g, or one lies outside [0, E_g] (an offset out
// of range is clamped, so no later kernel reads outside the adjacency).
template <class Src>
__global__ void k_union_beg(Src src, const int64_t* __restrict__ v_off, const int64_t* __restrict__ e_off, int32_t g0,
                            int32_t ng, int32_t* __restrict__ ubeg, int32_t* bad) {
    const int64_t vb = v_off[g0], Vu = v_off[g0 + ng] - vb, eb = e_off[g0];
    const int64_t c0 = (int64_t)blockIdx.x * kUnionChunk;
    if (c0 >= Vu) return;
    const int64_t c1 = min(c0 + kUnionChunk, Vu);
    int32_t g = find_graph(v_off + g0, ng, vb + c0 + threadIdx.x < vb + c1 ? vb + c0 + threadIdx.x : vb + c0) + g0;
    for (int64_t u = c0 + threadIdx.x; u < c1; u += blockDim.x) {
        while (vb + u >= v_off[g + 1]) ++g;
        const int64_t i = vb + u - v_off[g];
        const int64_t Vg = v_off[g + 1] - v_off[g], Eg = e_off[g + 1] - e_off[g];
        const int32_t* bg = src.beg_of(g);
        int64_t b = bg[i];
        if (b < 0 || b > Eg) {
            atomicOr(bad, 1);
            b = b < 0 ? 0 : Eg;
        }
        ubeg[u] = (int32_t)(e_off[g] - eb + b);
        if (i == 0 && (b != 0 || (int64_t)bg[Vg] != Eg)) atomicOr(bad, 1);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) ubeg[Vu] = (int32_t)(e_off[g0 + ng] - eb);
}

// Union adjacency: local ids shifted by the graph's vertex base.  bad |= 2: an
// id outside [0, V_g).
template <class Src>
__global__ void k_union_adj(Src src, const int64_t* __restrict__ v_off, const int64_t* __restrict__ e_off, int32_t g0,
                            int32_t ng, int32_t* __restrict__ uadj, int32_t* bad) {
    const int64_t vb = v_off[g0], eb = e_off[g0], Eu = e_off[g0 + ng] - eb;
    const int64_t c0 = (int64_t)blockIdx.x * kUnionChunk;
    if (c0 >= Eu) return;
    const int64_t c1 = min(c0 + kUnionChunk, Eu);
    const int64_t first = c0 + threadIdx.x < c1 ? c0 + threadIdx.x : c0;
    int32_t g = find_graph(e_off + g0, ng, eb + first) + g0;
    int any_bad = 0;
    for (int64_t k = c0 + threadIdx.x; k < c1; k += blockDim.x) {
        while (eb + k >= e_off[g + 1]) ++g;
        const int32_t w = src.adj_of(g)[eb + k - e_off[g]];
        const int64_t Vg = v_off[g + 1] - v_off[g];
        any_bad |= (w < 0 || w >= Vg);
        uadj[k] = (int32_t)(w + (v_off[g] - vb));
    }
    if (any_bad) atomicOr(bad, 2);
}

// Graph ranges of a batch: graph g owns vertices [vb[g], ve[g]) of the CSR
// being solved, split into kBatchChunk-vertex chunks; cstart[g] is the index of
// its first chunk (cstart[n] chunks in all).  One block per chunk, so the
// work follows the vertex count, not the number of graphs.
struct Ranges {
    const int64_t* vb;
    const int64_t* ve;
    const int64_t* cstart;
    int32_t n;
};

__device__ __forceinline__ void chunk_of(const Ranges& R, int64_t b, int32_t& g, int64_t& v0, int64_t& v1) {
    g = find_graph(R.cstart, R.n, b);
    v0 = R.vb[g] + (b - R.cstart[g]) * kBatchChunk;
    v1 = min(v0 + kBatchChunk, R.ve[g]);
}

// Selection bitmap of the batch's vertices (sel cleared beforehand).
__global__ void k_sel_ranges(Ranges R, uint64_t* sel) {
    int32_t g;
    int64_t v0, v1;
    chunk_of(R, blockIdx.x, g, v0, v1);
    for (int64_t w = (v0 >> 6) + threadIdx.x; w <= (v1 - 1) >> 6; w += blockDim.x) {
        uint64_t m = ~0ull;
        if (w == (v0 >> 6)) m &= ~0ull << (v0 & 63);
        if (w == ((v1 - 1) >> 6) && (v1 & 63)) m &= (1ull << (v1 & 63)) - 1;
        atomicOr(reinterpret_cast<unsigned long long*>(sel + w), (unsigned long long)m);  // edge words may be shared
    }
}

// Pivot candidates per chunk: the largest out*in among vertices with both
// degrees non-zero, the lowest id on ties (pivot_selection_from_fq,
// src/scc_common.h:92-150, restricted to one graph).
__global__ void k_chunk_pivot(const int32_t* __restrict__ fwb, const int32_t* __restrict__ bwb, Ranges R,
                              int32_t wrap32, int64_t* cp, int32_t* ci) {
    __shared__ int64_t sp[kBlock / kWave];
    __shared__ int32_t si[kBlock / kWave];
    int32_t g;
    int64_t v0, v1;
    chunk_of(R, blockIdx.x, g, v0, v1);
    int64_t bp = 0;
    int32_t bi = INT32_MAX;
    for (int64_t v = v0 + threadIdx.x; v < v1; v += blockDim.x) {
        const int32_t od = fwb[v + 1] - fwb[v], id = bwb[v + 1] - bwb[v];
        if (od > 0 && id > 0) {
            const int64_t p = wrap32 ? (int64_t)(int32_t)((uint32_t)od * (uint32_t)id) : (int64_t)od * id;
            if (p > 0 && pivot_better(p, (int32_t)v, bp, bi)) {
                bp = p;
                bi = (int32_t)v;
            }
        }
    }
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) {
        const int64_t op = __shfl_xor(bp, d, kWave);
        const int32_t oi = __shfl_xor(bi, d, kWave);
        if (pivot_better(op, oi, bp, bi)) {
            bp = op;
            bi = oi;
        }
    }
    if (lane_id() == 0) {
        sp[wave_id()] = bp;
        si[wave_id()] = bi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kBlock / kWave; ++w)
            if (pivot_better(sp[w], si[w], bp, bi)) {
                bp = sp[w];
                bi = si[w];
            }
        cp[blockIdx.x] = bp;
        ci[blockIdx.x] = bi;
    }
}

// Each graph's pivot from its chunks' candidates (-1: none); gmin reset.
__global__ void k_graph_pivot(Ranges R, const int64_t* __restrict__ cp, const int32_t* __restrict__ ci,
                              int32_t* roots, int32_t* gmin) {
    for (int32_t g = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x); g < R.n; g += gridDim.x * blockDim.x) {
        int64_t bp = 0;
        int32_t bi = INT32_MAX;
        for (int64_t c = R.cstart[g]; c < R.cstart[g + 1]; ++c)
            if (pivot_better(cp[c], ci[c], bp, bi)) {
                bp = cp[c];
                bi = ci[c];
            }
        roots[g] = (bp > 0 && bi != INT32_MAX) ? bi : -1;
        gmin[g] = INT32_MAX;
    }
}

// Smallest giant vertex of each graph (its canonical label): per chunk the
// first set bit of the BW-reach bitmap, folded into gmin[g].
__global__ void k_chunk_gmin(const uint64_t* __restrict__ bvis, Ranges R, const int32_t* __restrict__ roots,
                             int32_t* gmin) {
    __shared__ int32_t sm[kBlock / kWave];
    int32_t g;
    int64_t v0, v1;
    chunk_of(R, blockIdx.x, g, v0, v1);
    if (roots[g] < 0) return;
    int32_t m = INT32_MAX;
    for (int64_t w = (v0 >> 6) + threadIdx.x; w <= (v1 - 1) >> 6; w += blockDim.x) {
        uint64_t bits = bvis[w];
        if (w == (v0 >> 6)) bits &= ~0ull << (v0 & 63);
        if (w == ((v1 - 1) >> 6) && (v1 & 63)) bits &= (1ull << (v1 & 63)) - 1;
        if (bits) m = min(m, (int32_t)(w * 64 + __ffsll((unsigned long long)bits) - 1));
    }
    m = wave_min(m);
    if (lane_id() == 0) sm[wave_id()] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < kBlock / kWave; ++i) m = min(m, sm[i]);
        if (m != INT32_MAX) atomicMin(&gmin[g], m);
    }
}

// Final labels: giant vertices take their graph's smallest giant id, and
// every label becomes its graph's local id, written to that graph's output.
__global__ void k_chunk_finish(const int32_t* __restrict__ lab, const uint64_t* __restrict__ bvis, Ranges R,
                               const int32_t* __restrict__ gmin, int32_t* const* __restrict__ out) {
    int32_t g;
    int64_t v0, v1;
    chunk_of(R, blockIdx.x, g, v0, v1);
    const int64_t base = R.vb[g];
    const int32_t gm = gmin[g];
    int32_t* o = out[g];
    for (int64_t v = v0 + threadIdx.x; v < v1; v += blockDim.x) {
        const bool giant = gm != INT32_MAX && ((bvis[v >> 6] >> (v & 63)) & 1ull);
        o[v - base] = (int32_t)((giant ? gm : lab[v]) - base);
    }
}

// Appends one graph (local ids) to an arena CSR at vertex base vb and edge
// base eb: offsets shifted by eb (including the terminal one), ids by vb.
__global__ void k_arena_append(const int32_t* __restrict__ beg, const int32_t* __restrict__ adj, int32_t V, int64_t E,
                               int64_t vb, int64_t eb, int32_t* __restrict__ abeg, int32_t* __restrict__ aadj) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (int64_t i = t; i <= V; i += stride) abeg[vb + i] = (int32_t)(eb + beg[i]);
    for (int64_t k = t; k < E; k += stride) aadj[eb + k] = (int32_t)(adj[k] + vb);
}

int blocks_for(int64_t n) { return (int)std::max<int64_t>(1, (n + kUnionChunk - 1) / kUnionChunk); }

}  // namespace

template <class Src>
hipError_t union_launch(Src fw, Src bw, const int64_t* d_v_off, const int64_t* d_e_off, int32_t g0, int32_t ng,
                        int64_t Vu, int64_t Eu, int32_t* ufwb, int32_t* ufwa, int32_t* ubwb, int32_t* ubwa,
                        int32_t* bad, hipStream_t s) {
    hipLaunchKernelGGL(k_union_beg<Src>, dim3(blocks_for(Vu)), dim3(kBlock), 0, s, fw, d_v_off, d_e_off, g0, ng, ufwb,
                       bad);
    hipLaunchKernelGGL(k_union_beg<Src>, dim3(blocks_for(Vu)), dim3(kBlock), 0, s, bw, d_v_off, d_e_off, g0, ng, ubwb,
                       bad);
    if (Eu > 0) {
        hipLaunchKernelGGL(k_union_adj<Src>, dim3(blocks_for(Eu)), dim3(kBlock), 0, s, fw, d_v_off, d_e_off, g0, ng,
                           ufwa, bad);
        hipLaunchKernelGGL(k_union_adj<Src>, dim3(blocks_for(Eu)), dim3(kBlock), 0, s, bw, d_v_off, d_e_off, g0, ng,
                           ubwa, bad);
    }
    return hipGetLastError();
}

hipError_t launch_batch_union(const int32_t* fwb_cat, const int32_t* fwa_cat, const int32_t* bwb_cat,
                              const int32_t* bwa_cat, const int64_t* d_v_off, const int64_t* d_e_off, int32_t g0,
                              int32_t ng, int64_t Vu, int64_t Eu, int32_t* ufwb, int32_t* ufwa, int32_t* ubwb,
                              int32_t* ubwa, int32_t* bad, hipStream_t s) {
    return union_launch(CatSrc{fwb_cat, fwa_cat, d_v_off, d_e_off}, CatSrc{bwb_cat, bwa_cat, d_v_off, d_e_off}, d_v_off,
                        d_e_off, g0, ng, Vu, Eu, ufwb, ufwa, ubwb, ubwa, bad, s);
}

hipError_t launch_batch_union_perm(const int32_t* fwb_cat, const int32_t* fwa_cat, const int32_t* bwb_cat,
                                   const int32_t* bwa_cat, const int64_t* d_lv_off, const int64_t* d_le_off,
                                   const int64_t* d_sv, const int64_t* d_se, const int32_t* d_sg, int32_t ng,
                                   int64_t Vu, int64_t Eu, int32_t* ufwb, int32_t* ufwa, int32_t* ubwb, int32_t* ubwa,
                                   int32_t* bad, hipStream_t s) {
    return union_launch(PermSrc{fwb_cat, fwa_cat, d_sv, d_se, d_sg}, PermSrc{bwb_cat, bwa_cat, d_sv, d_se, d_sg},
                        d_lv_off, d_le_off, 0, ng, Vu, Eu, ufwb, ufwa, ubwb, ubwa, bad, s);
}

hipError_t launch_batch_roots(const int32_t* fwb, const int32_t* bwb, const BatchRanges& r, int32_t wrap32,
                              hipStream_t s) {
    const Ranges R{r.vb, r.ve, r.cstart, r.n};
    if (r.nchunks > 0)
        hipLaunchKernelGGL(k_chunk_pivot, dim3((unsigned)r.nchunks), dim3(kBlock), 0, s, fwb, bwb, R, wrap32, r.cp,
                           r.ci);
    hipLaunchKernelGGL(k_graph_pivot, dim3((r.n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, R, r.cp, r.ci, r.roots,
                       r.gmin);
    return hipGetLastError();
}

hipError_t launch_batch_finish(const int32_t* lab, const uint64_t* bvis, const BatchRanges& r, hipStream_t s) {
    const Ranges R{r.vb, r.ve, r.cstart, r.n};
    if (r.nchunks > 0) {
        hipLaunchKernelGGL(k_chunk_gmin, dim3((unsigned)r.nchunks), dim3(kBlock), 0, s, bvis, R, r.roots, r.gmin);
        hipLaunchKernelGGL(k_chunk_finish, dim3((unsigned)r.nchunks), dim3(kBlock), 0, s, lab, bvis, R, r.gmin, r.out);
    }
    return hipGetLastError();
}

hipError_t launch_batch_select(const BatchRanges& r, uint64_t* sel, int64_t nwords, hipStream_t s) {
    const Ranges R{r.vb, r.ve, r.cstart, r.n};
    hipError_t e = hipMemsetAsync(sel, 0, (size_t)std::max<int64_t>(nwords, 1) * 8, s);
    if (e != hipSuccess) return e;
    if (r.nchunks > 0) hipLaunchKernelGGL(k_sel_ranges, dim3((unsigned)r.nchunks), dim3(kBlock), 0, s, R, sel);
    return hipGetLastError();
}

hipError_t launch_arena_append(const int32_t* beg, const int32_t* adj, int32_t V, int64_t E, int64_t vb, int64_t eb,
                               int32_t* abeg, int32_t* aadj, hipStream_t s) {
    const int64_t n = std::max<int64_t>(E, (int64_t)V + 1);
    const int grid = (int)std::min<int64_t>(4096, (n + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_arena_append, dim3(std::max(grid, 1)), dim3(kBlock), 0, s, beg, adj, V, E, vb, eb, abeg, aadj);
    return hipGetLastError();
}

}  // namespace ispan
