// Device helpers shared by the kernel files (wave64 primitives, per-wave
// queue staging, traversal claims, early-exit adjacency scans).
#pragma once
#include "common.h"

namespace ispan {
namespace dev {

__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }
__device__ __forceinline__ int wave_id() { return threadIdx.x / kWave; }
__device__ __forceinline__ uint64_t lanemask_lt() { return (1ull << lane_id()) - 1ull; }

__device__ __forceinline__ bool bit_of(const uint64_t* bm, int32_t v) {
    return (bm[v >> 6] >> (v & 63)) & 1ull;
}

// Load that bypasses the non-coherent per-CU cache (reads what other CUs'
// atomics wrote).
template <class T>
__device__ __forceinline__ T load_agent(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Device-coherent store: written through to the device coherence point, so
// a block on another XCD that has invalidated its caches reads it.
template <class T>
__device__ __forceinline__ void st_agent(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// st_agent of four consecutive int32s (p 16-byte aligned) in one 16-byte
// store: the same write-through store (global_store_dwordx4 with sc1, as
// st_agent's global_store_dword), so a wave writes 1 KB per instruction.
// A VALU write to the data registers of a store wider than 8 bytes in the
// next cycle needs one wait state, which the compiler inserts for its own
// stores but not after inline asm: hence the s_nop.
__device__ __forceinline__ void st_agent4(int32_t* p, int32_t x, int32_t y, int32_t z, int32_t w) {
    typedef int32_t i32x4 __attribute__((ext_vector_type(4)));
    const i32x4 v = {x, y, z, w};
    asm volatile("global_store_dwordx4 %0, %1, off sc1\n\ts_nop 0" : : "v"(p), "v"(v) : "memory");
}

// v as held by lane src (any lane index per lane: ds_bpermute).
__device__ __forceinline__ uint64_t lane_from(uint64_t v, int src) {
    const uint32_t lo = (uint32_t)__shfl((int32_t)(uint32_t)v, src, kWave);
    const uint32_t hi = (uint32_t)__shfl((int32_t)(uint32_t)(v >> 32), src, kWave);
    return ((uint64_t)hi << 32) | lo;
}

// Lanes of one wave exchange data through memory: order and publish it.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// v as held by lane l, where l is the same in every lane: one v_readlane
// per 32 bits (__shfl would route through the LDS crossbar).
__device__ __forceinline__ int32_t lane_get(int32_t v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ uint64_t lane_get(uint64_t v, int l) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int32_t)(uint32_t)v, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int32_t)(uint32_t)(v >> 32), l);
    return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ int64_t lane_get(int64_t v, int l) { return (int64_t)lane_get((uint64_t)v, l); }

// x of lane (lane - N) in the same 16-lane row (DPP row_shr); lanes
// without one get `fill`.
template <int N>
__device__ __forceinline__ int32_t row_shr(int32_t x, int32_t fill) {
    return __builtin_amdgcn_update_dpp(fill, x, 0x110 + N, 0xF, 0xF, false);
}
template <int N>
__device__ __forceinline__ uint64_t row_shr(uint64_t x, uint64_t fill) {
    const uint32_t lo = (uint32_t)row_shr<N>((int32_t)(uint32_t)x, (int32_t)(uint32_t)fill);
    const uint32_t hi = (uint32_t)row_shr<N>((int32_t)(uint32_t)(x >> 32), (int32_t)(uint32_t)(fill >> 32));
    return ((uint64_t)hi << 32) | lo;
}
template <int N>
__device__ __forceinline__ int64_t row_shr(int64_t x, int64_t fill) {
    return (int64_t)row_shr<N>((uint64_t)x, (uint64_t)fill);
}

// Inclusive scan of op within each 16-lane row (lane 15 of a row holds the
// row's total); id is op's identity.
template <class T, class Op>
__device__ __forceinline__ T row_incl(T x, T id, Op op) {
    x = op(x, row_shr<1>(x, id));
    x = op(x, row_shr<2>(x, id));
    x = op(x, row_shr<4>(x, id));
    x = op(x, row_shr<8>(x, id));
    return x;
}

// Wave-wide inclusive sum: row scans, then the totals of earlier rows.
__device__ __forceinline__ int wave_incl_scan(int x) {
    x = row_incl(x, 0, [](int a, int b) { return a + b; });
    const int t0 = lane_get(x, 15), t1 = lane_get(x, 31), t2 = lane_get(x, 47);
    const int r = lane_id() >> 4;
    return x + (r >= 1 ? t0 : 0) + (r >= 2 ? t1 : 0) + (r >= 3 ? t2 : 0);
}

// Wave-wide sum and minimum; every lane gets the result.
__device__ __forceinline__ int wave_sum(int x) {
    x = row_incl(x, 0, [](int a, int b) { return a + b; });
    return lane_get(x, 15) + lane_get(x, 31) + lane_get(x, 47) + lane_get(x, 63);
}
__device__ __forceinline__ int64_t wave_sum(int64_t x) {
    x = row_incl(x, (int64_t)0, [](int64_t a, int64_t b) { return a + b; });
    return lane_get(x, 15) + lane_get(x, 31) + lane_get(x, 47) + lane_get(x, 63);
}
__device__ __forceinline__ int wave_min(int x) {
    x = row_incl(x, INT32_MAX, [](int a, int b) { return min(a, b); });
    return min(min(lane_get(x, 15), lane_get(x, 31)), min(lane_get(x, 47), lane_get(x, 63)));
}

// Block-wide exclusive scan (256 threads); every thread gets the total.
template <int NW = 4>
__device__ __forceinline__ void block_excl_scan(int x, int& excl, int& total, int* s_tmp /*[NW]*/) {
    static_assert(NW <= kWave, "one lane per wave total");
    const int incl = wave_incl_scan(x);
    if (lane_id() == kWave - 1) s_tmp[wave_id()] = incl;
    __syncthreads();
    // lane i of every wave scans the NW wave totals
    const int t = lane_id() < NW ? s_tmp[lane_id()] : 0;
    const int ti = wave_incl_scan(t);
    const int wv = wave_id();
    excl = lane_get(ti - t, wv) + incl - x;
    total = lane_get(ti, NW - 1);
    __syncthreads();
}

// Per-wave staging of queue pushes in LDS: one global atomic per flush
// instead of one per wave-step.  All 64 lanes must call push/flush together.
struct WaveQueue {
    int32_t* buf;  // kWaveBuf entries of LDS owned by this wave
    int n;         // wave-uniform fill
};

// COH: entries are written device-coherently (read by other blocks of the
// same launch after a grid barrier).
template <bool COH = false>
__device__ __forceinline__ void wq_flush(WaveQueue& wq, int32_t* gq, int32_t* gcount) {
    if (wq.n == 0) return;
    wave_sync();
    int base = 0;
    if (lane_id() == 0) base = atomicAdd(gcount, wq.n);
    base = lane_get(base, 0);
    for (int i = lane_id(); i < wq.n; i += kWave) {
        if (COH) st_agent(&gq[base + i], wq.buf[i]);
        else gq[base + i] = wq.buf[i];
    }
    wq.n = 0;
    wave_sync();
}

template <bool COH = false>
__device__ __forceinline__ void wq_push(WaveQueue& wq, bool pred, int32_t v, int32_t* gq, int32_t* gcount) {
    uint64_t m = __ballot(pred);
    if (m == 0) return;
    int c = __popcll(m);
    if (wq.n + c > kWaveBuf) wq_flush<COH>(wq, gq, gcount);
    if (pred) wq.buf[wq.n + __popcll(m & lanemask_lt())] = v;
    wq.n += c;
}

// Claim w for a traversal: it must be allowed (act for FW, fvis for BW) and
// not yet visited; the atomic makes every vertex claimed exactly once.
__device__ __forceinline__ bool try_claim(int32_t w, const uint64_t* __restrict__ allow, uint64_t* vis) {
    const int64_t wi = w >> 6;
    const uint64_t bit = 1ull << (w & 63);
    if (!(allow[wi] & bit)) return false;
    if (vis[wi] & bit) return false;
    unsigned long long old = atomicOr(reinterpret_cast<unsigned long long*>(vis + wi), (unsigned long long)bit);
    return !(old & bit);
}

// Wave-wide claim of targets w (lanes with want): only allowed, not yet
// visited vertices are claimed, each by exactly one lane.  Consecutive lanes
// whose targets share a bitmap word form a run; an OR-scan over the run gives
// one atomic per run instead of one per lane, so sorted lists that pack many
// targets into few words (a hub's list, a grid row) do not serialise on one
// address.  The exclusive part of the scan keeps a target repeated inside a
// run from being claimed twice.  Must be called by all 64 lanes.
__device__ __forceinline__ bool wave_try_claim(bool want, int32_t w, const uint64_t* __restrict__ allow,
                                               uint64_t* vis) {
    const int lane = lane_id();
    const int32_t wi = w >> 6;
    const uint64_t bit = 1ull << (w & 63);
    bool cand = false;
    if (want) cand = (allow[wi] & bit) && !(vis[wi] & bit);
    const int32_t key = cand ? wi : -1 - lane;  // lanes without a candidate are runs of their own
    const int32_t pkey = __shfl_up(key, 1, kWave);
    const uint64_t heads = __ballot(lane == 0 || pkey != key);
    const uint64_t le = lane == kWave - 1 ? ~0ull : ((2ull << lane) - 1);  // lanes <= this one
    const int start = 63 - __clzll(heads & le);
    const uint64_t after = heads & ~le;
    const int end = after ? __ffsll((unsigned long long)after) - 2 : kWave - 1;
    uint64_t incl = cand ? bit : 0ull;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const uint64_t y = __shfl_up(incl, d, kWave);
        if (lane - d >= start) incl |= y;
    }
    uint64_t excl = __shfl_up(incl, 1, kWave);
    if (lane == start) excl = 0;
    unsigned long long old = 0;
    if (lane == end && incl) old = atomicOr(reinterpret_cast<unsigned long long*>(vis + wi), (unsigned long long)incl);
    old = __shfl(old, end, kWave);
    return cand && !((old | excl) & bit);
}

// Does any edge in adj[p, end) satisfy pred(w, self)?  Each lane first walks
// its own list for up to kLaneEdges edges; the lists still open after that
// are finished by the whole wave, 64 edges per step, one list at a time.
// Must be called by all 64 lanes.
constexpr int kLaneEdges = 32;
template <class Pred>
__device__ __forceinline__ bool wave_any_edge(bool active, int32_t p, int32_t end, int32_t self,
                                              const int32_t* __restrict__ adj, Pred pred, int& scanned) {
    bool found = false;
    if (active) {
        const int32_t p0 = p;
        const int32_t lim = min(end, p + kLaneEdges);
        while (p + 4 <= lim && !found) {
            int32_t a0 = adj[p], a1 = adj[p + 1], a2 = adj[p + 2], a3 = adj[p + 3];
            found = pred(a0, self) || pred(a1, self) || pred(a2, self) || pred(a3, self);
            p += 4;
        }
        while (p < lim && !found) found = pred(adj[p++], self);
        scanned += p - p0;
    }
    // the rest of a long list: the whole wave, 4 x 64 entries per step with
    // their loads issued together (a hub's list would otherwise take one
    // dependent round trip per 64 entries)
    constexpr int kSteps = 4;
    uint64_t pend = __ballot(active && !found && p < end);
    while (pend) {
        const int l = __ffsll((unsigned long long)pend) - 1;
        const int32_t pp = lane_get(p, l), ee = lane_get(end, l), ss = lane_get(self, l);
        bool hit = false;
        int32_t qb = pp;
        for (; qb < ee; qb += kSteps * kWave) {
            int32_t a[kSteps];
#pragma unroll
            for (int k = 0; k < kSteps; ++k) {
                const int32_t q = qb + k * kWave + lane_id();
                a[k] = q < ee ? adj[q] : -1;
            }
            bool h = false;  // every predicate load issued, none waits on another
#pragma unroll
            for (int k = 0; k < kSteps; ++k) h |= pred(a[k] >= 0 ? a[k] : ss, ss) & (a[k] >= 0);
            if (__ballot(h)) {
                hit = true;
                break;
            }
        }
        if (lane_id() == l) {
            found = hit;
            scanned += min(qb + kSteps * kWave, ee) - pp;
        }
        pend &= pend - 1;
    }
    return found;
}

// Adds a per-thread count into a 64-bit global counter, one atomic per wave.
__device__ __forceinline__ void wave_add64(int64_t* dst, int64_t x) {
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) x += __shfl_xor(x, d, kWave);
    if (lane_id() == 0 && x) atomicAdd(reinterpret_cast<unsigned long long*>(dst), (unsigned long long)x);
}

__device__ __forceinline__ bool pivot_better(int64_t p, int32_t i, int64_t bp, int32_t bi) {
    return p > bp || (p == bp && i < bi);
}


// ---------------------------------------------------------- size-2/3 trims
// trim_2_from_fq (src/trim_2_3.h:146-283), trim_3_1_from_fq (:482-677) and
// trim_3_2_from_fq (:881-1080), checked against act as it stands when the
// pass starts: detections only write lab (and the pattern size into
// stamp); act is cleared afterwards.  With a fixed active set every
// detected group is an SCC of the active subgraph, so two detections of one
// vertex always agree on its label.

// Distinct active neighbours of x other than x, e1, e2: how many were seen
// (stopping at cap + 1) and the first two of them.
struct Nb {
    int n;
    int32_t g0, g1;
};

__device__ __forceinline__ Nb nb_scan(const int32_t* __restrict__ beg, const int32_t* __restrict__ adj,
                                      const uint64_t* __restrict__ act, int32_t x, int32_t e1, int32_t e2, int cap) {
    Nb r{0, -1, -1};
    const int32_t end = beg[x + 1];
    for (int32_t k = beg[x]; k < end; ++k) {
        const int32_t w = adj[k];
        if (w == x || w == e1 || w == e2 || !bit_of(act, w)) continue;
        if ((r.n >= 1 && w == r.g0) || (r.n >= 2 && w == r.g1)) continue;
        if (r.n == 0) r.g0 = w;
        else if (r.n == 1) r.g1 = w;
        if (++r.n > cap) break;
    }
    return r;
}

// nb_scan by a whole wave (every lane calls it with the same x, e1, e2,
// cap): 64 list entries per step, and within a step the lanes in list order,
// so the count and the first two neighbours are nb_scan's.
__device__ __forceinline__ Nb nb_scan_wave(const int32_t* __restrict__ beg, const int32_t* __restrict__ adj,
                                           const uint64_t* __restrict__ act, int32_t x, int32_t e1, int32_t e2,
                                           int cap) {
    Nb r{0, -1, -1};
    const int32_t end = beg[x + 1];
    for (int32_t k0 = beg[x]; k0 < end; k0 += kWave) {
        const int32_t k = k0 + (int32_t)lane_id();
        int32_t w = -1;
        bool ok = false;
        if (k < end) {
            w = adj[k];
            ok = w != x && w != e1 && w != e2 && bit_of(act, w);
        }
        while (true) {
            // the first entry (in list order) that is not one counted already
            const uint64_t m = __ballot(ok && !(r.n >= 1 && w == r.g0) && !(r.n >= 2 && w == r.g1));
            if (!m) break;
            const int32_t wl = lane_get(w, __ffsll((unsigned long long)m) - 1);
            if (r.n == 0) r.g0 = wl;
            else if (r.n == 1) r.g1 = wl;
            if (++r.n > cap) return r;
        }
    }
    return r;
}

__device__ __forceinline__ void pattern_mark(int32_t* lab, int32_t* stamp, int32_t a, int32_t b, int32_t c, int size) {
    int32_t m = min(a, b);
    if (size == 3) m = min(m, c);
    st_agent(&lab[a], m);
    st_agent(&lab[b], m);
    st_agent(&stamp[a], size);
    st_agent(&stamp[b], size);
    if (size == 3) {
        st_agent(&lab[c], m);
        st_agent(&stamp[c], size);
    }
}

// Checks vertex a against the size-2 pattern, then the size-3 cycle, then
// the size-3 hub, on out-edges and then in-edges; marks the first match.
// WAVE: the whole wave checks one vertex (a uniform, every lane calls), its
// lists scanned 64 entries at a time; otherwise each lane checks its own.
template <bool WAVE = false>
__device__ __forceinline__ void pattern_check(const int32_t* __restrict__ fwb, const int32_t* __restrict__ fwa,
                                              const int32_t* __restrict__ bwb, const int32_t* __restrict__ bwa,
                                              int32_t* lab, int32_t* stamp, const uint64_t* __restrict__ act,
                                              int32_t a) {
    auto nb_scan = [](const int32_t* beg, const int32_t* adj, const uint64_t* act, int32_t x, int32_t e1, int32_t e2,
                      int cap) {
        if constexpr (WAVE) return nb_scan_wave(beg, adj, act, x, e1, e2, cap);
        else return dev::nb_scan(beg, adj, act, x, e1, e2, cap);
    };
    auto pattern_mark = [](int32_t* lab, int32_t* stamp, int32_t a, int32_t b, int32_t c, int size) {
        if (!WAVE || lane_id() == 0) dev::pattern_mark(lab, stamp, a, b, c, size);
    };
    for (int dir = 0; dir < 2; ++dir) {
        const int32_t* beg = dir == 0 ? fwb : bwb;
        const int32_t* adj = dir == 0 ? fwa : bwa;
        const Nb na = nb_scan(beg, adj, act, a, -1, -1, 2);
        if (na.n == 1) {
            const int32_t b = na.g0;
            // size 2: b's only other one is a
            const Nb nb = nb_scan(beg, adj, act, b, -1, -1, 1);
            if (nb.n == 1 && nb.g0 == a) {
                pattern_mark(lab, stamp, a, b, -1, 2);
                return;
            }
            // size 3 cycle: a -> b -> c -> a
            const Nb nb2 = nb_scan(beg, adj, act, b, a, -1, 1);
            if (nb2.n == 1) {
                const int32_t c = nb2.g0;
                const Nb nc = nb_scan(beg, adj, act, c, b, -1, 1);
                if (nc.n == 1 && nc.g0 == a) {
                    pattern_mark(lab, stamp, a, b, c, 3);
                    return;
                }
            }
        } else if (na.n == 2) {
            // size 3 hub: a <-> b, a <-> c
            const int32_t b = na.g0, c = na.g1;
            const Nb nb = nb_scan(beg, adj, act, b, b, c, 1);
            if (nb.n == 1 && nb.g0 == a) {
                const Nb nc = nb_scan(beg, adj, act, c, b, c, 1);
                if (nc.n == 1 && nc.g0 == a) {
                    pattern_mark(lab, stamp, a, b, c, 3);
                    return;
                }
            }
        }
    }
}

}  // namespace dev
}  // namespace ispan
