// Batched solves: many graphs in one pass of the pipeline over their
// disjoint union.  SCCs never cross graphs, so the union's SCCs are the
// graphs' SCCs; what changes is that every graph needs its own pivot (the
// reference's pivot_selection_from_fq, src/scc_common.h:92-150, per graph),
// its giant its own label, and the labels their graph's local ids.
//
// Input layout (caller-owned, on the device): graph g's CSR pair with local
// ids, concatenated -- its V_g + 1 offsets start at beg[v_off[g] + g], its
// E_g adjacency entries at adj[e_off[g]].  A group of consecutive graphs
// whose union fits int32 offsets is rebuilt once into a union CSR (global
// ids, global offsets) that the single-graph kernels run on unchanged.
#include "common.h"
#include "device_util.h"

namespace ispan {
using namespace dev;

namespace {

constexpr int kChunk = 1 << 16;  // union-building work per block (vertices or edges)

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

// Where graph g's arrays are: back to back (the batch ABI's layout) or
// through a table of per-graph device pointers (the server's resident graphs).
struct CatSrc {
    const int32_t* beg;
    const int32_t* adj;
    const int64_t* v_off;
    const int64_t* e_off;
    __device__ const int32_t* beg_of(int32_t g) const { return beg + v_off[g] + g; }
    __device__ const int32_t* adj_of(int32_t g) const { return adj + e_off[g]; }
};

struct PtrSrc {
    const int32_t* const* beg;
    const int32_t* const* adj;
    __device__ const int32_t* beg_of(int32_t g) const { return beg[g]; }
    __device__ const int32_t* adj_of(int32_t g) const { return adj[g]; }
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
    const int64_t c0 = (int64_t)blockIdx.x * kChunk;
    if (c0 >= Vu) return;
    const int64_t c1 = min(c0 + kChunk, Vu);
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
    const int64_t c0 = (int64_t)blockIdx.x * kChunk;
    if (c0 >= Eu) return;
    const int64_t c1 = min(c0 + kChunk, Eu);
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

// Pivot of each graph: the largest out*in among vertices with both degrees
// non-zero, the lowest id on ties; -1 if there is none.  One block per graph.
__global__ void k_batch_roots(const int32_t* __restrict__ fwb, const int32_t* __restrict__ bwb,
                              const int64_t* __restrict__ goff, int32_t ng, int32_t wrap32, int32_t* roots) {
    __shared__ int64_t sp[kBlock / kWave];
    __shared__ int32_t si[kBlock / kWave];
    for (int32_t g = blockIdx.x; g < ng; g += gridDim.x) {
        int64_t bp = 0;
        int32_t bi = INT32_MAX;
        for (int64_t v = goff[g] + threadIdx.x; v < goff[g + 1]; v += blockDim.x) {
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
            roots[g] = (bp > 0 && bi != INT32_MAX) ? bi : -1;
        }
        __syncthreads();
    }
}

// Smallest vertex of each graph's giant (its canonical label): the first set
// bit of the BW-reach bitmap in the graph's range, scanned a block-width of
// words at a time from the start.  One block per graph.
__global__ void k_batch_gmin(const uint64_t* __restrict__ bvis, const int64_t* __restrict__ goff,
                             const int32_t* __restrict__ roots, int32_t ng, int32_t* gmin) {
    __shared__ int32_t sm[kBlock / kWave];
    for (int32_t g = blockIdx.x; g < ng; g += gridDim.x) {
        if (roots[g] < 0) {
            if (threadIdx.x == 0) gmin[g] = -1;
            continue;
        }
        const int64_t v0 = goff[g], v1 = goff[g + 1];
        int32_t found = INT32_MAX;
        for (int64_t wb = v0 >> 6; wb <= (v1 - 1) >> 6; wb += blockDim.x) {
            const int64_t w = wb + threadIdx.x;
            int32_t m = INT32_MAX;
            if (w <= (v1 - 1) >> 6) {
                uint64_t bits = bvis[w];
                if (w == (v0 >> 6)) bits &= ~0ull << (v0 & 63);
                if (w == ((v1 - 1) >> 6) && ((v1 & 63) != 0)) bits &= (1ull << (v1 & 63)) - 1;
                if (bits) m = (int32_t)(w * 64 + __ffsll((unsigned long long)bits) - 1);
            }
            m = wave_min(m);
            if (lane_id() == 0) sm[wave_id()] = m;
            __syncthreads();
            for (int i = 0; i < kBlock / kWave; ++i) found = min(found, sm[i]);
            __syncthreads();
            if (found != INT32_MAX) break;
        }
        if (threadIdx.x == 0) gmin[g] = found == INT32_MAX ? roots[g] : found;
    }
}

// Final labels of a group: giant vertices take their graph's smallest giant
// id, and every label becomes its graph's local id, written at the group's
// place in the batch's label array.
__global__ void k_batch_finish(const int32_t* __restrict__ lab, const uint64_t* __restrict__ bvis,
                               const int64_t* __restrict__ goff, const int32_t* __restrict__ gmin, int32_t ng,
                               int32_t* __restrict__ out) {
    const int64_t Vu = goff[ng];
    const int64_t c0 = (int64_t)blockIdx.x * kChunk;
    if (c0 >= Vu) return;
    const int64_t c1 = min(c0 + kChunk, Vu);
    int32_t g = find_graph(goff, ng, c0 + threadIdx.x < c1 ? c0 + threadIdx.x : c0);
    for (int64_t v = c0 + threadIdx.x; v < c1; v += blockDim.x) {
        while (v >= goff[g + 1]) ++g;
        const bool giant = (bvis[v >> 6] >> (v & 63)) & 1ull;
        const int32_t l = giant && gmin[g] >= 0 ? gmin[g] : lab[v];
        out[v] = (int32_t)(l - goff[g]);
    }
}

int blocks_for(int64_t n) { return (int)std::max<int64_t>(1, (n + kChunk - 1) / kChunk); }

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

hipError_t launch_batch_union_ptrs(const int32_t* const* fwb, const int32_t* const* fwa, const int32_t* const* bwb,
                                   const int32_t* const* bwa, const int64_t* d_v_off, const int64_t* d_e_off,
                                   int32_t ng, int64_t Vu, int64_t Eu, int32_t* ufwb, int32_t* ufwa, int32_t* ubwb,
                                   int32_t* ubwa, int32_t* bad, hipStream_t s) {
    return union_launch(PtrSrc{fwb, fwa}, PtrSrc{bwb, bwa}, d_v_off, d_e_off, 0, ng, Vu, Eu, ufwb, ufwa, ubwb, ubwa,
                        bad, s);
}

hipError_t launch_batch_roots(const int32_t* fwb, const int32_t* bwb, const int64_t* goff, int32_t ng, int32_t wrap32,
                              int32_t* roots, hipStream_t s) {
    hipLaunchKernelGGL(k_batch_roots, dim3(std::min(ng, 65536)), dim3(kBlock), 0, s, fwb, bwb, goff, ng, wrap32, roots);
    return hipGetLastError();
}

hipError_t launch_batch_finish(const int32_t* lab, const uint64_t* bvis, const int64_t* goff, const int32_t* roots,
                               int32_t ng, int64_t Vu, int32_t* gmin, int32_t* out, hipStream_t s) {
    hipLaunchKernelGGL(k_batch_gmin, dim3(std::min(ng, 65536)), dim3(kBlock), 0, s, bvis, goff, roots, ng, gmin);
    if (Vu > 0) hipLaunchKernelGGL(k_batch_finish, dim3(blocks_for(Vu)), dim3(kBlock), 0, s, lab, bvis, goff, gmin, ng, out);
    return hipGetLastError();
}

}  // namespace ispan
