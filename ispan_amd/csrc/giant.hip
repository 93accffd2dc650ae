// The giant-SCC phase as one cooperative, persistent kernel: trim-1-first and
// the pivot (src/trim_1_gfq.h:9-33, src/scc_common.h:92-150), the forward
// and backward direction-optimising traversals (src/fw_bw.h:396-698,
// 944-1248) and the relabelling of what they found, with grid-wide barriers
// between levels instead of host round trips.
//
// Every block runs the same control flow: after each barrier all blocks sum
// the same per-block partial counts and take the same top-down/bottom-up
// decision, so every block passes the same barriers.  Data written by other
// blocks before a barrier is read with plain loads after a barrier that
// invalidates the caches, and with device-coherent loads after one that does
// not; device-coherent loads are also kept for what other blocks change
// during the same phase.
#include <atomic>
#include "common.h"
#include "device_util.h"

namespace ispan {
using namespace dev;

namespace {

constexpr int kSpinCap = 1 << 24;  // barrier wait bound (~1 s); beyond it the kernel aborts
constexpr int kBarGroups = 32;     // arrival counters of the two-level barrier
// One 1024-thread block per CU: every block invalidates its caches after a
// barrier, so fewer, larger blocks make the barrier cheaper.
constexpr int kGBlock = 1024;
constexpr int kW = kGBlock / kWave;

// Partial-count slots: part[slot * grid + block].
enum Slot { kPivP, kPivI, kSurv, kPost, kEdgeTD, kEdgeBU, kFoundTD, kFoundBU, kGiantN };
static_assert(kGiantN < kGiantSlots, "slot count");

struct Smem {
    int32_t off[kGBlock];    // top-down tile: exclusive degree prefix, first edge
    int32_t start[kGBlock];
    int tmp[kW];
    int64_t red[kW];
    int64_t red2[kW];
    int64_t rres[4];  // reduction result of the last barrier
    int32_t bi[kW];
    int ok;
    int32_t nbar;   // barriers this block has arrived at in this launch
    int32_t gbase;  // barrier generation when this launch started
    int32_t wcnt[2][kW];    // block flushes: per-wave counts and bases of up to two counters
    int32_t wbase[2][kW];
    int32_t hv[kW][kWave];  // per-wave staged hub vertices: id, first edge, degree
    int32_t hs[kW][kWave];
    int32_t hd[kW][kWave];
};

// Per-wave staging, kPairCap 8-byte entries a wave: top-down queue entries,
// or kWaveBuf vertex ids in the phases that queue ids.  It is dynamic LDS,
// placed after Smem, so that Smem's arrays stay within the 64 KB an LDS
// instruction's immediate offset reaches.
constexpr int kPairCap = 512;
static_assert(kWaveBuf * 4 <= kPairCap * 8, "the id view fits the staging");
static_assert((kPairCap & (kPairCap - 1)) == 0, "the in-place ring indexes with a mask");
extern __shared__ uint64_t stage_lds[];

struct Ctx {
    const GiantArgs& a;
    int G, b, tid, lane, wave;
    Smem& sm;
    int64_t edges_td, edges_bu, found_td, found_bu;  // work counters (thread 0's are the block's)
    // this wave's staging, as 8-byte entries or as ids
    __device__ uint64_t* stage() const { return stage_lds + wave * kPairCap; }
    __device__ int32_t* stage32() const { return reinterpret_cast<int32_t*>(stage()); }
};

// Values a block contributes to the reduction a barrier carries: two sums
// and a max over non-negative keys (0 contributes nothing).
struct Red {
    int64_t s0, s1, s2, key;
};

// Waits until every vector-memory access this wave has issued (loads,
// stores, atomics) has completed: s_waitcnt vmcnt(0), with no cache
// maintenance.  An sc1 store or a device atomic has completed once it is
// visible at device scope, so after this wait the wave's device-coherent
// writes are visible to every other XCD.  (A release fence does not give
// this: at workgroup scope it emits no wait at all, since the block's waves
// share one L1; at agent scope it also writes back the whole L2.)
__device__ __forceinline__ void wait_vm() { __builtin_amdgcn_s_waitcnt(0x0F70); }  // vmcnt(0), others unchanged

// Two-level grid barrier: blocks arrive on one of 32 counters, the last of
// each group arrives on the top counter, the last overall bumps `gen`.
//
// Memory protocol: everything a block writes for other blocks is written
// device-coherently (atomics or st_agent), so arriving needs no L2
// writeback.  It does need every one of those writes to have completed
// before the arrival, and they are issued by all 16 waves while only thread
// 0 arrives: so every wave waits for its own accesses (wait_vm) before the
// block barrier, and thread 0 waits again before each of its arrival
// atomics (its reduction atomics and counter resets precede them).  The
// last block's reduction results and counter resets likewise complete
// before it bumps `gen`.  A block that sees `gen` move therefore sees every
// write made before the barrier, once it invalidates its L1 and its XCC's
// L2 (device-scope acquire) or reads the data with device-coherent loads.
//
// With RED, each block's `in` is folded in on the way: into its group's
// accumulator before arriving, the group total into the top accumulator by
// the group's last arriver, the grand total published by the last block
// before it bumps `gen`; every block gets the same `out`.
//
// ACQ: after the wait, invalidate this CU's L1 and this XCC's L2 (agent
// acquire), so that plain loads see what other blocks wrote before the
// barrier.  Without it only coherent loads and atomics may read such data
// (and loads that tolerate stale values).
//
// Returns false (and flags an error) if a wait exceeds kSpinCap.
//
// USE: which of s0, s1, s2, key (bits 0-3) the barrier carries (the same in
// every block); the others read back as 0.
template <bool RED = false, bool ACQ = true>
__device__ __forceinline__ bool gsync(GiantState* gs, int G, Smem& sm, int32_t* dbg, Red in = Red{0, 0, 0, 0},
                                      Red* out = nullptr, int use = 15) {
    using ull = unsigned long long;
    wait_vm();  // this wave's writes for other blocks have completed
    __syncthreads();
    if (threadIdx.x == 0) {
        int ok = 1;
        // gen is gbase when the launch starts and each barrier bumps it once,
        // so at this block's nb-th barrier it reads gbase + nb - 1: no load
        // needed
        const int32_t nb = ++sm.nbar;
        const int32_t g0 = sm.gbase + nb - 1;
        const int grp = blockIdx.x % kBarGroups;
        const int ngrp = min(kBarGroups, G);
        const int gsize = (G - grp + kBarGroups - 1) / kBarGroups;
        RLine& ga = gs->gacc[grp];
        if (RED) {
            if (in.s0) atomicAdd(reinterpret_cast<ull*>(&ga.sum[0]), (ull)in.s0);
            if (in.s1) atomicAdd(reinterpret_cast<ull*>(&ga.sum[1]), (ull)in.s1);
            if (in.s2) atomicAdd(reinterpret_cast<ull*>(&ga.sum[2]), (ull)in.s2);
            if (in.key) atomicMax(reinterpret_cast<ull*>(&ga.key), (ull)in.key);
        }
        wait_vm();  // the group accumulator holds this block's values before it arrives
        bool last = false;
        const int arrived = atomicAdd(&gs->sub[grp].v, 1);
        // diagnostics, stored while the arrival is in flight: barriers arrived
        // at, and passed (this block's previous one)
        dbg[blockIdx.x] = nb;
        dbg[G + blockIdx.x] = nb - 1;
        if (arrived == gsize - 1) {
            st_agent(&gs->sub[grp].v, 0);
            if (RED) {
                const ull a0 = (use & 1) ? atomicExch(reinterpret_cast<ull*>(&ga.sum[0]), 0ull) : 0ull;
                const ull a1 = (use & 2) ? atomicExch(reinterpret_cast<ull*>(&ga.sum[1]), 0ull) : 0ull;
                const ull a2 = (use & 4) ? atomicExch(reinterpret_cast<ull*>(&ga.sum[2]), 0ull) : 0ull;
                const ull ak = (use & 8) ? atomicExch(reinterpret_cast<ull*>(&ga.key), 0ull) : 0ull;
                if (a0) atomicAdd(reinterpret_cast<ull*>(&gs->tacc.sum[0]), a0);
                if (a1) atomicAdd(reinterpret_cast<ull*>(&gs->tacc.sum[1]), a1);
                if (a2) atomicAdd(reinterpret_cast<ull*>(&gs->tacc.sum[2]), a2);
                if (ak) atomicMax(reinterpret_cast<ull*>(&gs->tacc.key), ak);
            }
            wait_vm();  // the group counter's reset and the group's totals precede the top arrival
            if (atomicAdd(&gs->top.v, 1) == ngrp - 1) {
                st_agent(&gs->top.v, 0);
                if (RED) {
                    RLine& r = gs->res[(g0 + 1) & 1];
                    if (use & 1) st_agent(&r.sum[0], (int64_t)atomicExch(reinterpret_cast<ull*>(&gs->tacc.sum[0]), 0ull));
                    if (use & 2) st_agent(&r.sum[1], (int64_t)atomicExch(reinterpret_cast<ull*>(&gs->tacc.sum[1]), 0ull));
                    if (use & 4) st_agent(&r.sum[2], (int64_t)atomicExch(reinterpret_cast<ull*>(&gs->tacc.sum[2]), 0ull));
                    if (use & 8) st_agent(&r.key, (int64_t)atomicExch(reinterpret_cast<ull*>(&gs->tacc.key), 0ull));
                }
                last = true;
            }
        }
        if (last) {
            wait_vm();  // the top counter's reset and the results precede the new generation
            atomicAdd(&gs->gen.v, 1);
        } else {
            int spins = 0;
            while (load_agent(&gs->gen.v) == g0) {
                __builtin_amdgcn_s_sleep(2);
                if (++spins > kSpinCap) {
                    atomicOr(&gs->err.v, 1);
                    ok = 0;
                    break;
                }
            }
        }
        if (ACQ) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // this CU's L1 and this XCC's L2
        else __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        if (RED && ok) {
            const RLine& r = gs->res[(g0 + 1) & 1];
            sm.rres[0] = (use & 1) ? load_agent(&r.sum[0]) : 0;
            sm.rres[1] = (use & 2) ? load_agent(&r.sum[1]) : 0;
            sm.rres[2] = (use & 4) ? load_agent(&r.sum[2]) : 0;
            sm.rres[3] = (use & 8) ? load_agent(&r.key) : 0;
        }
        sm.ok = ok && load_agent(&gs->err.v) == 0;
    }
    __syncthreads();
    if (RED && out) *out = Red{sm.rres[0], sm.rres[1], sm.rres[2], sm.rres[3]};
    __syncthreads();
    return sm.ok;
}

// Sum of a per-thread value over the block (every thread gets it).
__device__ __forceinline__ int64_t block_sum(int64_t x, Smem& sm) {
    static_assert(kW == 16, "one DPP row holds the wave totals");
    x = wave_sum(x);
    if (lane_id() == 0) sm.red[wave_id()] = x;
    __syncthreads();
    x = row_incl(sm.red[lane_id() & 15], (int64_t)0, [](int64_t a, int64_t b) { return a + b; });
    x = lane_get(x, 15);
    __syncthreads();
    return x;
}

// Sum / minimum of one partial slot over all blocks (every block computes
// the same value).
__device__ __forceinline__ int64_t sum_slot(const int64_t* part, int slot, int G, Smem& sm) {
    int64_t x = 0;
    for (int i = threadIdx.x; i < G; i += kGBlock) x += load_agent(&part[(int64_t)slot * G + i]);
    return block_sum(x, sm);
}

// Per-wave staging of hub vertices in LDS; a flush turns them into
// kGiantChunk-edge work items with one atomic for the whole batch.
struct HubStage {
    int32_t* v;
    int32_t* s;
    int32_t* d;
    int n;  // wave-uniform fill
};

__device__ __forceinline__ void hub_flush(HubStage& hs, HubItem* hout, int32_t* hn_out) {
    if (hs.n == 0) return;
    wave_sync();
    const int lane = lane_id();
    int32_t v = 0, s0 = 0, nch = 0;
    if (lane < hs.n) {
        v = hs.v[lane];
        s0 = hs.s[lane];
        nch = (hs.d[lane] + kGiantChunk - 1) / kGiantChunk;
    }
    const int incl = wave_incl_scan(nch);
    int base = 0;
    if (lane == kWave - 1) base = atomicAdd(hn_out, incl);
    base = lane_get(base, kWave - 1) + incl - nch;
    for (int k = 0; k < nch; ++k) {
        st_agent(&hout[base + k].v, v);
        st_agent(&hout[base + k].start, s0 + k * kGiantChunk);
    }
    hs.n = 0;
    wave_sync();
}

__device__ __forceinline__ void hub_push(HubStage& hs, bool pred, int32_t v, int32_t s0, int32_t d, HubItem* hout,
                                         int32_t* hn_out) {
    const uint64_t m = __ballot(pred);
    if (m == 0) return;
    const int c = __popcll(m);
    if (hs.n + c > kWave) hub_flush(hs, hout, hn_out);
    if (pred) {
        const int i = hs.n + __popcll(m & lanemask_lt());
        hs.v[i] = v;
        hs.s[i] = s0;
        hs.d[i] = d;
    }
    hs.n += c;
}

// End-of-phase flushes of every wave's staged entries with one allocation
// per block: each wave's count goes through LDS, thread 0 reserves the
// block's total.  (One allocation per wave would put up to 4096 atomics on
// one counter at the end of every level.)  All threads call them.
// Room for every wave's entries in up to two global counters: lane 0 of
// each wave posts its counts, lanes 0-15 of wave 0 (one DPP row) scan them,
// one lane takes both counters' atomics back to back, and every wave reads
// its bases.  Two block barriers.  sm.wbase is rewritten only after the next
// call's first barrier, which every wave reaches after reading its bases.
__device__ __forceinline__ void block_reserve2(Ctx& c, int n0, int32_t* count0, int n1, int32_t* count1, int& base0,
                                               int& base1) {
    static_assert(kW == 16, "the wave counts fill one DPP row");
    Smem& sm = c.sm;
    if (c.lane == 0) {
        sm.wcnt[0][c.wave] = n0;
        sm.wcnt[1][c.wave] = n1;
    }
    __syncthreads();
    if (c.wave == 0) {
        const int x = c.lane < kW ? sm.wcnt[0][c.lane] : 0, y = c.lane < kW ? sm.wcnt[1][c.lane] : 0;
        const int xi = row_incl(x, 0, [](int p, int q) { return p + q; });
        const int yi = row_incl(y, 0, [](int p, int q) { return p + q; });
        const int xt = lane_get(xi, kW - 1), yt = lane_get(yi, kW - 1);
        int xb = 0, yb = 0;
        if (c.lane == 0) {
            if (xt) xb = atomicAdd(count0, xt);
            if (yt) yb = atomicAdd(count1, yt);
        }
        xb = lane_get(xb, 0);
        yb = lane_get(yb, 0);
        if (c.lane < kW) {
            sm.wbase[0][c.lane] = xb + xi - x;
            sm.wbase[1][c.lane] = yb + yi - y;
        }
    }
    __syncthreads();
    base0 = sm.wbase[0][c.wave];
    base1 = sm.wbase[1][c.wave];
}

__device__ __forceinline__ int block_reserve(Ctx& c, int wave_count, int32_t* gcount) {
    int base, unused;
    block_reserve2(c, wave_count, gcount, 0, gcount, base, unused);
    return base;
}

__device__ __forceinline__ void block_flush(Ctx& c, WaveQueue& wq, int32_t* gq, int32_t* gcount) {
    const int base = block_reserve(c, wq.n, gcount);
    for (int i = c.lane; i < wq.n; i += kWave) st_agent(&gq[base + i], wq.buf[i]);
    wq.n = 0;
    wave_sync();
}

// block_flush of the queue and of the staged hub vertices (as work items)
// with one reservation round.
__device__ __forceinline__ void block_flush_both(Ctx& c, WaveQueue& wq, int32_t* gq, int32_t* gcount, HubStage& hs,
                                                 HubItem* hout, int32_t* hn_out) {
    int32_t v = 0, s0 = 0, nch = 0;
    if (c.lane < hs.n) {
        v = hs.v[c.lane];
        s0 = hs.s[c.lane];
        nch = (hs.d[c.lane] + kGiantChunk - 1) / kGiantChunk;
    }
    const int incl = wave_incl_scan(nch);
    int qb, hb;
    block_reserve2(c, wq.n, gcount, lane_get(incl, kWave - 1), hn_out, qb, hb);
    for (int i = c.lane; i < wq.n; i += kWave) st_agent(&gq[qb + i], wq.buf[i]);
    hb += incl - nch;
    for (int k = 0; k < nch; ++k) {
        st_agent(&hout[hb + k].v, v);
        st_agent(&hout[hb + k].start, s0 + k * kGiantChunk);
    }
    wq.n = 0;
    hs.n = 0;
    wave_sync();
}

// A top-down queue entry: a vertex's first edge (high half) and degree (low
// half), which is all a top-down level needs to walk the vertex's list, so
// it loads no offsets; the claim that queues a vertex has just loaded them.
__device__ __forceinline__ uint64_t qentry(int32_t s0, int32_t d) {
    return ((uint64_t)(uint32_t)s0 << 32) | (uint32_t)d;
}
__device__ __forceinline__ int32_t qentry_start(uint64_t e) { return (int32_t)(uint32_t)(e >> 32); }
__device__ __forceinline__ int32_t qentry_degree(uint64_t e) { return (int32_t)(uint32_t)e; }

// A wave's staged top-down queue entries (in LDS) and its fill.
struct PairQueue {
    uint64_t* buf;
    int n;  // wave-uniform
};

__device__ __forceinline__ void pq_flush(PairQueue& pq, uint64_t* gq, int32_t* gcount) {
    if (pq.n == 0) return;
    wave_sync();
    int base = 0;
    if (lane_id() == 0) base = atomicAdd(gcount, pq.n);
    base = lane_get(base, 0);
    for (int i = lane_id(); i < pq.n; i += kWave) st_agent(&gq[base + i], pq.buf[i]);
    pq.n = 0;
    wave_sync();
}

__device__ __forceinline__ void pq_push(PairQueue& pq, bool pred, uint64_t e, uint64_t* gq, int32_t* gcount) {
    const uint64_t m = __ballot(pred);
    if (m == 0) return;
    const int c = __popcll(m);
    if (pq.n + c > kPairCap) pq_flush(pq, gq, gcount);
    if (pred) pq.buf[pq.n + __popcll(m & lanemask_lt())] = e;
    pq.n += c;
}

// block_flush_both for top-down queue entries.
__device__ __forceinline__ void block_flush_both(Ctx& c, PairQueue& pq, uint64_t* gq, int32_t* gcount, HubStage& hs,
                                                 HubItem* hout, int32_t* hn_out) {
    int32_t v = 0, s0 = 0, nch = 0;
    if (c.lane < hs.n) {
        v = hs.v[c.lane];
        s0 = hs.s[c.lane];
        nch = (hs.d[c.lane] + kGiantChunk - 1) / kGiantChunk;
    }
    const int incl = wave_incl_scan(nch);
    int qb, hb;
    block_reserve2(c, pq.n, gcount, lane_get(incl, kWave - 1), hn_out, qb, hb);
    for (int i = c.lane; i < pq.n; i += kWave) st_agent(&gq[qb + i], pq.buf[i]);
    hb += incl - nch;
    for (int k = 0; k < nch; ++k) {
        st_agent(&hout[hb + k].v, v);
        st_agent(&hout[hb + k].start, s0 + k * kGiantChunk);
    }
    pq.n = 0;
    hs.n = 0;
    wave_sync();
}

// Pushes a newly claimed vertex for expansion at the next level: vertices of
// degree >= kGiantHubDeg become kGiantChunk-edge work items, others go to
// the queue (degree-0 vertices have nothing to expand).  All lanes call it.
__device__ __forceinline__ void push_next(bool claim, int32_t w, const int32_t* __restrict__ beg, PairQueue& pq,
                                          uint64_t* qout, int32_t* qn_out, HubStage& hs, HubItem* hout,
                                          int32_t* hn_out, int64_t& pushed_edges) {
    int32_t d = 0, s0 = 0;
    if (claim) {
        s0 = beg[w];
        d = beg[w + 1] - s0;
        pushed_edges += d;
    }
    hub_push(hs, claim && d >= kGiantHubDeg, w, s0, d, hout, hn_out);
    pq_push(pq, claim && d > 0 && d < kGiantHubDeg, qentry(s0, d), qout, qn_out);
}

// The acquire gsync<.., true> performs, for a block that passed a barrier
// without it.
__device__ __forceinline__ void block_acquire() {
    if (threadIdx.x == 0) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    __syncthreads();
}

// Claims K targets per lane, each as wave_try_claim does, and loads the
// offsets of every candidate while the claim atomics are in flight: a
// claimed vertex's degree costs no round trip of its own.  All lanes call it.
template <int K>
__device__ __forceinline__ void claim_with_degree(const bool (&want)[K], const int32_t (&w)[K],
                                                  const uint64_t* __restrict__ allow, uint64_t* vis,
                                                  const int32_t* __restrict__ beg, bool (&claim)[K],
                                                  int32_t (&s0)[K], int32_t (&d)[K]) {
    const int lane = lane_id();
    uint64_t al[K], vi[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        al[k] = 0;
        vi[k] = 0;
        if (want[k]) {
            al[k] = allow[w[k] >> 6];
            vi[k] = vis[w[k] >> 6];
        }
    }
    bool cand[K];
#pragma unroll
    for (int k = 0; k < K; ++k) cand[k] = want[k] && ((al[k] & ~vi[k]) >> (w[k] & 63)) & 1ull;
    uint64_t incl[K], excl[K];
    int end[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int32_t key = cand[k] ? (w[k] >> 6) : -1 - lane;  // lanes without a candidate: runs of their own
        // key of lane - 1: DPP within rows, the three row starts by readlane
        int32_t pkey = row_shr<1>(key, key);
        {
            const int32_t k15 = lane_get(key, 15), k31 = lane_get(key, 31), k47 = lane_get(key, 47);
            pkey = lane == 16 ? k15 : lane == 32 ? k31 : lane == 48 ? k47 : pkey;
        }
        const uint64_t heads = __ballot(lane == 0 || pkey != key);
        const uint64_t le = lane == kWave - 1 ? ~0ull : ((2ull << lane) - 1);
        const int start = 63 - __clzll(heads & le);
        const uint64_t after = heads & ~le;
        end[k] = after ? __ffsll((unsigned long long)after) - 2 : kWave - 1;
        // OR over lanes start..lane: scans within rows, then the carries of
        // runs that cross a row boundary, row by row
        uint64_t in = cand[k] ? 1ull << (w[k] & 63) : 0ull;
        uint64_t y = row_shr<1>(in, 0ull);
        if (lane - 1 >= start) in |= y;
        y = row_shr<2>(in, 0ull);
        if (lane - 2 >= start) in |= y;
        y = row_shr<4>(in, 0ull);
        if (lane - 4 >= start) in |= y;
        y = row_shr<8>(in, 0ull);
        if (lane - 8 >= start) in |= y;
        const uint64_t c15 = lane_get(in, 15);
        if (lane >= 16 && lane < 32 && start <= 15) in |= c15;
        const uint64_t c31 = lane_get(in, 31);
        if (lane >= 32 && lane < 48 && start <= 31) in |= c31;
        const uint64_t c47 = lane_get(in, 47);
        if (lane >= 48 && start <= 47) in |= c47;
        // the run's OR before this lane
        uint64_t ex = row_shr<1>(in, 0ull);
        ex = lane == 16 ? c15 : lane == 32 ? c31 : lane == 48 ? c47 : ex;
        if (lane == start) ex = 0;
        incl[k] = in;
        excl[k] = ex;
    }
    unsigned long long old[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        old[k] = 0;
        if (lane == end[k] && incl[k])
            old[k] = atomicOr(reinterpret_cast<unsigned long long*>(vis + (w[k] >> 6)), (unsigned long long)incl[k]);
    }
    int32_t e1[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        s0[k] = 0;
        e1[k] = 0;
        if (cand[k]) {
            s0[k] = beg[w[k]];
            e1[k] = beg[w[k] + 1];
        }
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const uint64_t o = __shfl(old[k], end[k], kWave);
        claim[k] = cand[k] && !((o | excl[k]) & (1ull << (w[k] & 63)));
        d[k] = claim[k] ? e1[k] - s0[k] : 0;
    }
}

// push_next for K claimed targets whose offsets are known.
template <int K>
__device__ __forceinline__ void push_found_n(const bool (&claim)[K], const int32_t (&w)[K], const int32_t (&s0)[K],
                                             const int32_t (&d)[K], PairQueue& wq, uint64_t* qout, int32_t* qn_out,
                                             HubStage& hs, HubItem* hout, int32_t* hn_out, int64_t& pushed_edges) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
        pushed_edges += d[k];
        hub_push(hs, claim[k] && d[k] >= kGiantHubDeg, w[k], s0[k], d[k], hout, hn_out);
        pq_push(wq, claim[k] && d[k] > 0 && d[k] < kGiantHubDeg, qentry(s0[k], d[k]), qout, qn_out);
    }
}

// Top-down to exhaustion (mode 2) without level barriers in the thin parts:
// a wave expands the vertices it claimed itself, in the same level, instead
// of publishing them for the next level -- the reference's asynchronous
// top-down (src/fw_bw.h:506-544, 1068-1106), where each thread keeps
// expanding its own circular queue.  The wave's LDS staging buffer is that
// queue: up to kWave pending entries per step are flattened (wave-level
// degree scan) and walked one edge per lane; new claims of degree below
// kGiantHubDeg go back into the queue, hubs become next-level work items,
// and when the queue is full -- or after max_steps steps, so that a region
// one wave holds is shared again -- entries go to the next level's queue.
// Claims stay exact (the same wave-combined atomics), so the reach set and
// the found counts are those of a level-synchronous traversal; only the
// level at which a vertex is expanded changes.  All lanes call it.
template <bool BW>
__device__ __forceinline__ void expand_local(Ctx& c, PairQueue& wq, int32_t root, const int32_t* __restrict__ beg,
                                             const int32_t* __restrict__ adj, const uint64_t* __restrict__ allow,
                                             uint64_t* vis, uint64_t* qout, int32_t* qn_out, HubStage& hs,
                                             HubItem* hout, int32_t* hn_out, int64_t& my_found, int max_steps) {
    int32_t* off = c.sm.off + c.wave * kWave;  // this wave's slice of the tile arrays
    int32_t* stv = c.sm.start + c.wave * kWave;
    uint64_t* buf = wq.buf;
    constexpr int mask = kPairCap - 1;
    int head = 0, n = wq.n;  // pending: buf[head .. n) modulo kPairCap
    wave_sync();
    for (int step = 0; n > head && step < max_steps; ++step) {
        const int k = min(kWave, n - head);
        int32_t st = 0, dg = 0;
        if (c.lane < k) {
            const uint64_t q = buf[(head + c.lane) & mask];
            st = qentry_start(q);
            dg = qentry_degree(q);
        }
        head += k;
        const int incl = wave_incl_scan(dg);
        const int total = lane_get(incl, kWave - 1);
        off[c.lane] = incl - dg;
        stv[c.lane] = st;
        wave_sync();
        for (int eb = 0; eb < total; eb += kWave) {
            // the source of edge e: the last j < k with off[j] <= e
            const int e = eb + c.lane;
            int pos = 0;
#pragma unroll
            for (int sp = kWave / 2; sp > 0; sp >>= 1)
                if (pos + sp < k && off[pos + sp] <= e) pos += sp;
            bool want[1];
            int32_t w[1];
            w[0] = e < total ? adj[stv[pos] + (e - off[pos])] : 0;
            want[0] = e < total && w[0] != root;
            bool claim[1];
            int32_t s0[1], d[1];
            claim_with_degree<1>(want, w, allow, vis, beg, claim, s0, d);
            my_found += claim[0];
            hub_push(hs, claim[0] && d[0] >= kGiantHubDeg, w[0], s0[0], d[0], hout, hn_out);
            const bool keep = claim[0] && d[0] > 0 && d[0] < kGiantHubDeg;
            const uint64_t m = __ballot(keep);
            const int cnt = __popcll(m);
            if (cnt && n - head + cnt <= kPairCap) {
                if (keep) buf[(n + __popcll(m & lanemask_lt())) & mask] = qentry(s0[0], d[0]);
                n += cnt;
            } else if (cnt) {  // the queue is full: these go to the next level
                int base = 0;
                if (c.lane == 0) base = atomicAdd(qn_out, cnt);
                base = lane_get(base, 0);
                if (keep) st_agent(&qout[base + __popcll(m & lanemask_lt())], qentry(s0[0], d[0]));
            }
        }
        wave_sync();  // the queue entries and the slice are rewritten next step
    }
    const int left = n - head;  // the step budget ran out: the next level takes them
    if (left > 0) {
        int base = 0;
        if (c.lane == 0) base = atomicAdd(qn_out, left);
        base = lane_get(base, 0);
        for (int i = c.lane; i < left; i += kWave) st_agent(&qout[base + i], buf[(head + i) & mask]);
    }
    wq.n = 0;
    wave_sync();
}

// Level-synchronous top-down levels are expanded in place (as in mode 2)
// once their input gives every wave at least this many vertices (grid 4096:
// 2, 4, 8, 16 and 32 measured 4.30, 4.19, 4.01, 4.09 and 4.30 ms; R-MAT
// flat).
constexpr int kLocalPerWave = 8;

// Top-down: edges per lane per step.  One measured fastest: two let their
// load chains overlap, but the registers they hold push the kernel into
// spills on other paths.
constexpr int kTK = 1;

// Per-direction state of the joint traversal (the same in every block).
struct Dir {
    int mode;         // 0 top-down, 1 bottom-up, 2 top-down to exhaustion
    int64_t L;        // levels done
    int64_t visited;  // vertices claimed, the roots excluded
    bool live;        // still traversing
    bool gather;      // its last level was bottom-up and its next is top-down
};

// One level of one direction: top-down over the level's queue and hub items
// (then, where it applies, in place), or bottom-up over the candidate
// bitmap.  my_found (per thread) counts the vertices claimed; fedges, the
// out-degree sum of what was pushed for the next level (top-down) or the
// adjacency entries scanned (bottom-up).  Forward: from the pivot along
// out-edges, claiming active vertices into fvis.  Backward: along in-edges,
// claiming active vertices into bvis, not only those of the forward set: the
// giant SCC is fvis & bvis either way (a path from v to the pivot lies in
// the forward set when v does), and the two directions then need not wait
// for each other.
template <bool BW>
__device__ __forceinline__ void level_work(Ctx& c, const Dir& d, int32_t root, int64_t& my_found, int64_t& fedges) {
    const GiantArgs& a = c.a;
    Smem& sm = c.sm;
    const int32_t* beg = BW ? a.bwb : a.fwb;  // expansion (top-down) direction
    const int32_t* adj = BW ? a.bwa : a.fwa;
    const int32_t* rbeg = BW ? a.fwb : a.bwb;  // bottom-up scans the reverse direction
    const int32_t* radj = BW ? a.fwa : a.bwa;
    const int32_t* rfirst = BW ? a.fw_first : a.bw_first;
    const uint64_t* allow = a.act;
    uint64_t* vis = BW ? a.bvis : a.fvis;
    uint64_t* snap = BW ? a.snapb : a.snap;
    uint64_t* q0 = BW ? a.qb0 : a.q0;
    uint64_t* q1 = BW ? a.qb1 : a.q1;
    HubItem* h0 = BW ? a.hb0 : a.h0;
    HubItem* h1 = BW ? a.hb1 : a.h1;
    GiantState* gs = a.gs;
    const int dir = BW ? 1 : 0;
    PairQueue pq{c.stage(), 0};    // top-down: staged next-level entries
    WaveQueue wq{c.stage32(), 0};  // bottom-up: the same LDS, parked misses (ids)
    HubStage hs{sm.hv[c.wave], sm.hs[c.wave], sm.hd[c.wave], 0};
    const int mode = d.mode;
    const int64_t L = d.L;
    const int64_t total_waves = (int64_t)c.G * kW;
    const int64_t gwave = (int64_t)c.b * kW + c.wave;
    auto is_vis = [vis](int32_t u, int32_t) { return bit_of(vis, u); };
    const bool multi = a.n_roots > 0;
    my_found = 0;
    fedges = 0;
    const uint64_t* qin = (L & 1) ? q1 : q0;
    uint64_t* qout = (L & 1) ? q0 : q1;
    const HubItem* hin = (L & 1) ? h1 : h0;
    HubItem* hout = (L & 1) ? h0 : h1;
    int32_t* qn_out = &gs->qn[dir][(L + 1) % 3].v;
    int32_t* hn_out = &gs->hn[dir][(L + 1) % 3].v;
    const bool topdown = mode != 1;
    const bool tr = a.trace && L < kLevelTrace;
    const int64_t tl0 = tr ? (int64_t)wall_clock64() : 0;
    if (topdown) {
        // ---- top-down: tiles of the queue, then hub work items; level 0
        // expands the root's list in kGiantChunk-edge pieces.
        int64_t n = 0, nh = 0;
        int32_t rchunk = kGiantChunk;  // level 0: edges per piece of the root's list
        if (L == 0 && !multi) {
            if (c.b == 0 && c.tid == 0)
                atomicOr(reinterpret_cast<unsigned long long*>(vis + (root >> 6)), 1ull << (root & 63));
            // the root's list in about one piece per block
            const int32_t rdeg = beg[root + 1] - beg[root];
            rchunk = max(256, ((rdeg + c.G - 1) / c.G + 63) & ~63);
            rchunk = min(rchunk, kGiantChunk);
            nh = (rdeg + rchunk - 1) / rchunk;
        } else {
            n = load_agent(&gs->qn[dir][L % 3].v);
            nh = load_agent(&gs->hn[dir][L % 3].v);
            if (n < 0 || n > a.V || nh < 0 || nh > a.hub_cap) {  // cannot happen with exact claims
                if (c.tid == 0) atomicOr(&gs->err.v, 4);
                n = 0;
                nh = 0;
            }
        }
        static_assert((kGBlock & (kGBlock - 1)) == 0, "tile size");
        // entries per tile: a power of two from 64 to kGBlock, the
        // smallest that needs no more tiles than blocks, so a small
        // frontier is spread over many blocks
        int tile = kGBlock;
        while (tile > kWave && (int64_t)(tile / 2) * c.G >= n) tile /= 2;
        const int64_t ntiles = (n + tile - 1) / tile;
        // the backward direction's tiles start half the grid away, so that
        // when both directions' levels are small (fewer tiles than blocks)
        // they run on different blocks at the same time
        const int64_t b0 = BW ? ((int64_t)c.b + c.G / 2) % c.G : c.b;
        for (int64_t u = b0; u < ntiles + nh; u += c.G) {
            if (u < ntiles) {
                // a tile: queue entries on threads 0..tile-1
                const int64_t i = u * tile + c.tid;
                int32_t st = 0, dg = 0;
                if (c.tid < tile && i < n) {
                    const uint64_t q = load_agent(&qin[i]);
                    st = qentry_start(q);
                    dg = qentry_degree(q);
                }
                int excl, total;
                block_excl_scan<kW>(dg, excl, total, sm.tmp);
                sm.off[c.tid] = excl;
                sm.start[c.tid] = st;
                if (c.tid == 0) c.edges_td += total;
                __syncthreads();
                for (int eb = 0; eb < total; eb += kGBlock * kTK) {
                    // lane owns edges eb + k * kGBlock + tid; the source of
                    // edge e is the last j with off[j] <= e (off[0] = 0,
                    // non-decreasing), found by a fixed-depth search
                    int e[kTK], pos[kTK];
#pragma unroll
                    for (int k = 0; k < kTK; ++k) {
                        e[k] = eb + k * kGBlock + c.tid;
                        pos[k] = 0;
                    }
#pragma unroll
                    for (int sp = kGBlock / 2; sp > 0; sp >>= 1) {
#pragma unroll
                        for (int k = 0; k < kTK; ++k)
                            if (sm.off[pos[k] + sp] <= e[k]) pos[k] += sp;
                    }
                    bool want[kTK];
                    int32_t w[kTK];
#pragma unroll
                    for (int k = 0; k < kTK; ++k) {
                        w[k] = 0;
                        if (e[k] < total) w[k] = adj[sm.start[pos[k]] + (e[k] - sm.off[pos[k]])];
                    }
#pragma unroll
                    for (int k = 0; k < kTK; ++k) want[k] = e[k] < total && w[k] != root;
                    bool claim[kTK];
                    int32_t s0[kTK], d[kTK];
                    claim_with_degree<kTK>(want, w, allow, vis, beg, claim, s0, d);
#pragma unroll
                    for (int k = 0; k < kTK; ++k) {
                        my_found += claim[k];
                    }
                    push_found_n<kTK>(claim, w, s0, d, pq, qout, qn_out, hs, hout, hn_out, fedges);
                }
                __syncthreads();
            } else {
                int32_t v, s0;
                if (L == 0 && !multi) {
                    v = root;
                    s0 = beg[root] + (int32_t)(u - ntiles) * rchunk;
                } else {
                    v = load_agent(&hin[u - ntiles].v);
                    s0 = load_agent(&hin[u - ntiles].start);
                }
                const int32_t e1 = min(s0 + (L == 0 && !multi ? rchunk : kGiantChunk), beg[v + 1]);
                if (c.tid == 0) c.edges_td += e1 - s0;
                for (int32_t eb = s0; eb < e1; eb += kGBlock * kTK) {
                    bool want[kTK];
                    int32_t w[kTK];
#pragma unroll
                    for (int k = 0; k < kTK; ++k) {
                        const int32_t e = eb + k * kGBlock + c.tid;
                        w[k] = e < e1 ? adj[e] : 0;
                    }
#pragma unroll
                    for (int k = 0; k < kTK; ++k) want[k] = eb + k * kGBlock + c.tid < e1 && w[k] != root;
                    bool claim[kTK];
                    int32_t ws[kTK], wd[kTK];
                    claim_with_degree<kTK>(want, w, allow, vis, beg, claim, ws, wd);
#pragma unroll
                    for (int k = 0; k < kTK; ++k) {
                        my_found += claim[k];
                    }
                    push_found_n<kTK>(claim, w, ws, wd, pq, qout, qn_out, hs, hout, hn_out, fedges);
                }
            }
        }
        // in place: in top-down to exhaustion, and in any later top-down
        // level whose input gives every wave at least kLocalPerWave
        // vertices.  Where the frontier is small (the start of a
        // traversal on a high-diameter graph) a few waves would expand a
        // growing region alone: levels of up to 256-65536 entries
        // expanded so took the 4096^2 grid from 5.09 to 5.8-6.5 ms.
        if (a.td_local > 0 && (mode == 2 || (L > 0 && nh == 0 && n >= kLocalPerWave * total_waves)))
            expand_local<BW>(c, pq, root, beg, adj, allow, vis, qout, qn_out, hs, hout, hn_out, my_found,
                             a.td_local);
        block_flush_both(c, pq, qout, qn_out, hs, hout, hn_out);
    } else {
        // ---- bottom-up over the candidate bitmap, one wave per group of
        // kU consecutive 64-vertex words; only that wave writes those
        // words in this level.  Each stage issues its kU loads (first
        // in-neighbours, their visited bits) before any result is used.
        // Nothing is queued: each word's visited bits before the level go
        // to `snap`, and a switch back to top-down gathers vis & ~snap.
        constexpr int kU = 8;       // dense words per group
        constexpr int kSparse = 8;  // at most this many candidates: a lane's own work
        constexpr int kS = 4;       // sparse candidates per lane per step
        int scanned = 0;
        // candidates whose first probe misses are parked in the wave's LDS
        // buffer and scanned 64 at a time, one per lane, so their load
        // chains overlap instead of idling the other 63 lanes
        int nmiss = 0;  // wave-uniform
        auto scan_misses = [&]() {
            wave_sync();
            int32_t pv = -1;  // found in the previous batch: set behind this batch's loads
            for (int i0 = 0; i0 < nmiss; i0 += kWave) {
                const int i = i0 + c.lane;
                int32_t v = 0, p = 0, e = 0;
                if (i < nmiss) {
                    v = wq.buf[i];
                    p = rbeg[v] + 1;
                    e = rbeg[v + 1];
                }
                if (pv >= 0) atomicOr(reinterpret_cast<unsigned long long*>(vis + (pv >> 6)), 1ull << (pv & 63));
                pv = -1;
                const bool found = wave_any_edge(i < nmiss && p < e, p, e, v, radj, is_vis, scanned);
                if (found) {
                    pv = v;
                    my_found += 1;
                }
            }
            if (pv >= 0) atomicOr(reinterpret_cast<unsigned long long*>(vis + (pv >> 6)), 1ull << (pv & 63));
            nmiss = 0;
            wave_sync();
        };
        // misses with more list to scan go to the wave's buffer (all lanes call)
        auto park_misses = [&](bool miss, int32_t v) {
            const uint64_t mm = __ballot(miss);
            if (!mm) return;
            if (nmiss + __popcll(mm) > kWaveBuf) scan_misses();
            if (miss) wq.buf[nmiss + __popcll(mm & lanemask_lt())] = v;
            nmiss += __popcll(mm);
        };
        // A wave step covers kWave / kBUrun runs of kBUrun consecutive words
        // spread over the whole id range (run r of the step is run
        // cb + r * total_waves + gwave), so that the id ranges where the
        // work concentrates are shared by all blocks.
        constexpr int kBUrun = 8;
        const int64_t nruns = (a.nwords + kBUrun - 1) / kBUrun;
        // All waves sweep the id range in the same order (band by band:
        // vertices found early help later probes of the level hit), and
        // within a band wave g takes run (g * bu_mult) mod total_waves, a
        // permutation: with plain indices a block's runs would share
        // fixed bits of their vertex ids, and in graphs whose degrees
        // follow id bits (R-MAT) the blocks' work would differ severalfold.
        // (bu_mult < total_waves <= 65536: the product fits 32 bits)
        const int64_t pg = (int64_t)(((uint32_t)gwave * (uint32_t)a.bu_mult) % (uint32_t)total_waves);
        for (int64_t cb = 0; cb < nruns; cb += total_waves * (kWave / kBUrun)) {
            // lane j loads one word; then only words with candidates are
            // worked on, kU of them at a time
            const int64_t run = cb + (int64_t)(c.lane / kBUrun) * total_waves + pg;
            const int64_t wl = run * kBUrun + (c.lane % kBUrun);
            uint64_t lvw = 0, lcand = 0;
            if (wl < a.nwords) {
                lvw = vis[wl];
                lcand = allow[wl] & ~lvw;
            }
            // words with more than kSparse candidates: the wave works on
            // kU of them at a time, one lane per vertex
            const bool dense = __popcll(lcand) > kSparse;
            uint64_t busy = __ballot(dense);
            // Found bits of group member u wait in lane u and go out as
            // one atomic instruction behind the next group's loads: a
            // load's wait also waits for the atomics issued before it.
            int64_t pw = 0;
            uint64_t pm = 0;
            while (busy) {
                int js[kU];
#pragma unroll
                for (int u = 0; u < kU; ++u) {
                    js[u] = busy ? __ffsll((unsigned long long)busy) - 1 : -1;
                    busy &= busy ? busy - 1 : 0;
                }
                int32_t f[kU];
                int64_t ws[kU];  // word of group member u (wave-uniform)
#pragma unroll
                for (int u = 0; u < kU; ++u) {
                    f[u] = -1;
                    ws[u] = js[u] >= 0 ? lane_get(wl, js[u]) : 0;
                    if (js[u] >= 0 && ((lane_get(lcand, js[u]) >> c.lane) & 1ull))
                        f[u] = rfirst[ws[u] * 64 + c.lane];
                }
                if (pm) atomicOr(reinterpret_cast<unsigned long long*>(vis + pw), (unsigned long long)pm);
                pm = 0;
                uint32_t hits = 0;
#pragma unroll
                for (int u = 0; u < kU; ++u) hits |= (f[u] >= 0 && bit_of(vis, f[u])) ? 1u << u : 0u;
#pragma unroll
                for (int u = 0; u < kU; ++u) {
                    if (js[u] < 0) continue;
                    const uint64_t cm = lane_get(lcand, js[u]);
                    const int64_t w = ws[u];
                    const int32_t v = (int32_t)(w * 64 + c.lane);
                    const bool is_c = (cm >> c.lane) & 1ull;
                    const bool hit = is_c && ((hits >> u) & 1u);
                    scanned += is_c;
                    park_misses(is_c && !hit && f[u] >= 0, v);
                    // found bits are set within the level: later probes
                    // already see them
                    const uint64_t fm = __ballot(hit);
                    if (c.lane == u) {
                        pw = w;
                        pm = fm;
                    }
                    my_found += hit;
                }
            }
            if (tr && c.tid == 0 && cb == 0)
                atomicMax(reinterpret_cast<unsigned long long*>(&gs->lv_cp[dir][L][3]),
                          (unsigned long long)((int64_t)wall_clock64() - tl0));
            // sparse words: each lane works through its own word's
            // candidates, kS at a time, and sets its found bits at once
            uint64_t mine = dense ? 0ull : lcand, fm_lane = 0;
            while (__ballot(mine != 0)) {
                int32_t v[kS], f[kS];
#pragma unroll
                for (int k = 0; k < kS; ++k) {
                    v[k] = -1;
                    if (mine) {
                        v[k] = (int32_t)(wl * 64 + __ffsll((unsigned long long)mine) - 1);
                        mine &= mine - 1;
                    }
                }
#pragma unroll
                for (int k = 0; k < kS; ++k) f[k] = v[k] >= 0 ? rfirst[v[k]] : -1;
                bool hit[kS];
#pragma unroll
                for (int k = 0; k < kS; ++k) hit[k] = f[k] >= 0 && bit_of(vis, f[k]);
#pragma unroll
                for (int k = 0; k < kS; ++k) {
                    scanned += v[k] >= 0;
                    park_misses(v[k] >= 0 && !hit[k] && f[k] >= 0, v[k]);
                    if (hit[k]) {
                        fm_lane |= 1ull << (v[k] & 63);
                        my_found += 1;
                    }
                }
            }
            if (pm) atomicOr(reinterpret_cast<unsigned long long*>(vis + pw), (unsigned long long)pm);
            if (fm_lane) atomicOr(reinterpret_cast<unsigned long long*>(vis + wl), (unsigned long long)fm_lane);
            if (wl < a.nwords) st_agent(&snap[wl], lvw);  // the gather's "visited before this level"
        }
        if (tr && c.tid == 0)  // wave 0 of each block
            atomicMax(reinterpret_cast<unsigned long long*>(&gs->lv_cp[dir][L][1]),
                      (unsigned long long)((int64_t)wall_clock64() - tl0));
        scan_misses();
        if (tr && c.tid == 0)
            atomicMax(reinterpret_cast<unsigned long long*>(&gs->lv_cp[dir][L][2]),
                      (unsigned long long)((int64_t)wall_clock64() - tl0));
        fedges = scanned;  // bottom-up pushes nothing: the low half of the level sum counts the scan
    }
}

// The gather of a direction that switches from bottom-up back to top-down:
// the last level's finds (vis & ~snap) become the next level's queue and hub
// items.  Each wave loads 64 words at once, then expands the non-empty ones.
template <bool BW>
__device__ __forceinline__ void gather_finds(Ctx& c, const Dir& d) {
    const GiantArgs& a = c.a;
    Smem& sm = c.sm;
    GiantState* gs = a.gs;
    const int dir = BW ? 1 : 0;
    const int32_t* beg = BW ? a.bwb : a.fwb;
    const uint64_t* vis = BW ? a.bvis : a.fvis;
    const uint64_t* snap = BW ? a.snapb : a.snap;
    uint64_t* q_next = (d.L & 1) ? (BW ? a.qb1 : a.q1) : (BW ? a.qb0 : a.q0);
    HubItem* h_next = (d.L & 1) ? (BW ? a.hb1 : a.h1) : (BW ? a.hb0 : a.h0);
    int32_t* qn_next = &gs->qn[dir][d.L % 3].v;
    int32_t* hn_next = &gs->hn[dir][d.L % 3].v;
    PairQueue wq{c.stage(), 0};
    HubStage hs{sm.hv[c.wave], sm.hs[c.wave], sm.hd[c.wave], 0};
    const int64_t total_waves = (int64_t)c.G * kW;
    const int64_t gwave = (int64_t)c.b * kW + c.wave;
    int64_t unused = 0;
    for (int64_t wb = gwave * kWave; wb < a.nwords; wb += total_waves * kWave) {
        const int64_t w = wb + c.lane;
        uint64_t nw = 0;
        if (w < a.nwords) nw = vis[w] & ~snap[w];
        uint64_t busy = __ballot(nw != 0);
        while (busy) {
            const int l = __ffsll((unsigned long long)busy) - 1;
            busy &= busy - 1;
            const uint64_t m = lane_get(nw, l);
            const int32_t v = (int32_t)((wb + l) * 64 + c.lane);
            push_next((m >> c.lane) & 1ull, v, beg, wq, q_next, qn_next, hs, h_next, hn_next, unused);
        }
    }
    block_flush_both(c, wq, q_next, qn_next, hs, h_next, hn_next);
}

// The giant's traversals: forward and backward in the same levels, one grid
// barrier per level for both (fw_bfs_fq_queue, src/fw_bw.h:396-698, and
// bw_bfs_fq_queue, :944-1248, which the reference runs one after the
// other).  Each direction switches top-down / bottom-up / top-down to
// exhaustion on its own, by the reference's rules on alpha, beta and gamma
// (src/fw_bw.h:576-624) with set_size the active count for both; a
// direction that has finished idles while the other goes on.  Returns false
// after a failed barrier.  With a.n_roots > 0 every listed root seeds both
// directions (*seeded_out of them).
__device__ __forceinline__ bool traverse_joint(Ctx& c, int32_t root, int64_t set_size, Dir& f, Dir& b,
                                               int64_t* seeded_out) {
    const GiantArgs& a = c.a;
    Smem& sm = c.sm;
    GiantState* gs = a.gs;
    f = Dir{0, 0, 0, true, false};
    b = Dir{0, 0, 0, true, false};
    *seeded_out = 1;
    if (a.n_roots > 0) {
        // level-0 input of both directions: every graph's root, marked
        // visited, queued or split into hub items like any claimed vertex
        int64_t seeded = 0;
        for (int dir = 0; dir < 2; ++dir) {
            PairQueue wq{c.stage(), 0};
            HubStage hs{sm.hv[c.wave], sm.hs[c.wave], sm.hd[c.wave], 0};
            const int32_t* beg = dir ? a.bwb : a.fwb;
            uint64_t* vis = dir ? a.bvis : a.fvis;
            uint64_t* q0 = dir ? a.qb0 : a.q0;
            HubItem* h0 = dir ? a.hb0 : a.h0;
            int64_t unused = 0;
            for (int64_t ib = (int64_t)c.b * kGBlock + c.wave * kWave; ib < a.n_roots; ib += (int64_t)c.G * kGBlock) {
                const int64_t i = ib + c.lane;
                const int32_t r = i < a.n_roots ? a.roots[i] : -1;
                const bool ok = r >= 0 && r < a.V;
                if (ok) atomicOr(reinterpret_cast<unsigned long long*>(vis + (r >> 6)), 1ull << (r & 63));
                if (dir == 0) seeded += ok;
                push_next(ok, r, beg, wq, q0, &gs->qn[dir][0].v, hs, h0, &gs->hn[dir][0].v, unused);
            }
            block_flush_both(c, wq, q0, &gs->qn[dir][0].v, hs, h0, &gs->hn[dir][0].v);
        }
        Red r;
        if (!gsync<true>(gs, c.G, sm, a.dbg, Red{block_sum(seeded, sm), 0, 0, 0}, &r, 1)) return false;
        *seeded_out = r.s0;
    }
    while (f.live || b.live) {
        if (f.gather || b.gather) {
            // the visited bits were made exact after the level's barrier
            if (f.gather) gather_finds<false>(c, f);
            if (b.gather) gather_finds<true>(c, b);
            f.gather = b.gather = false;
            // the top-down level reads the queues coherently and visited bits
            // only as hints: no L2 invalidation
            if (!gsync<false, false>(gs, c.G, sm, a.dbg)) return false;
        }
        const bool tdf = f.mode != 1, tdb = b.mode != 1;
        const int64_t edges0 = c.edges_td + c.edges_bu;  // thread 0's are the block's
        const bool tr = a.trace != 0;
        const int64_t tl0 = tr ? (int64_t)wall_clock64() : 0;
        int64_t ff = 0, fe = 0, bf = 0, be = 0;  // per thread: found, pushed work / scanned entries
        if (f.live) level_work<false>(c, f, root, ff, fe);
        if (b.live) level_work<true>(c, b, root, bf, be);
        // ---- level totals: every block computes the same sums.  One sum per
        // direction carries the level's finds (high 32 bits) and the next
        // level's exact work or the scan (low 32 bits: at most E < 2^31, so no
        // carry), both reduced in one block round
        int64_t pf = wave_sum((ff << 32) + fe), pb = wave_sum((bf << 32) + be);
        if (c.lane == 0) {
            sm.red[c.wave] = pf;
            sm.red2[c.wave] = pb;
        }
        __syncthreads();
        pf = lane_get(row_incl(sm.red[c.lane & 15], (int64_t)0, [](int64_t x, int64_t y) { return x + y; }), 15);
        pb = lane_get(row_incl(sm.red2[c.lane & 15], (int64_t)0, [](int64_t x, int64_t y) { return x + y; }), 15);
        __syncthreads();
        auto block_counts = [&](const Dir& d, int64_t pk, bool td, int dir) {
            const int64_t blk = (int64_t)((uint64_t)pk >> 32);
            if (td) {
                c.found_td += blk;
            } else {
                const int64_t bscan = pk & 0xffffffffll;
                c.found_bu += blk;
                c.edges_bu += bscan;
                if (tr && d.L < kLevelTrace && c.tid == 0) {
                    atomicMax(reinterpret_cast<unsigned long long*>(&gs->lv_cp[dir][d.L][0]), (unsigned long long)bscan);
                    if (dir == 0 && c.b < 1024 && gs->bu1_scan[c.b] == 0) {
                        gs->bu1_scan[c.b] = bscan + 1;
                        gs->bu1_work[c.b] = (int64_t)wall_clock64() - tl0;
                    }
                }
            }
            if (tr && d.L < kLevelTrace && c.tid == 0) {
                const unsigned long long wt = (unsigned long long)((int64_t)wall_clock64() - tl0);
                atomicMax(reinterpret_cast<unsigned long long*>(&gs->lv_wmax[dir][d.L]), wt);
                atomicMax(reinterpret_cast<unsigned long long*>(&gs->lv_wmin[dir][d.L]), ~wt);
            }
            // counters of level L+2: last read at the start of level L-1,
            // first written in level L+1 (reset here, where the store overlaps
            // the barrier's own atomics)
            if (c.b == 0 && c.tid == 0) {
                __hip_atomic_store(&gs->qn[dir][(d.L + 2) % 3].v, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(&gs->hn[dir][(d.L + 2) % 3].v, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        };
        if (f.live) block_counts(f, pf, tdf, 0);
        if (b.live) block_counts(b, pb, tdb, 1);
        // Top-down levels read what other blocks wrote in the level before
        // through coherent loads (queue entries, hub items, counts), a
        // candidate bitmap that is fixed during the traversal, and visited
        // bits only as a hint for the claim atomics: their barriers skip the
        // L2 invalidation (the acquire below catches up where needed).
        Red r;
        if (!gsync<true, false>(gs, c.G, sm, a.dbg, Red{pf, pb, tr ? c.edges_td + c.edges_bu - edges0 : 0, 0}, &r,
                                3 | (tr ? 4 : 0)))
            return false;
        // per direction: the level's totals, the end test and the switches
        bool acquire = false, failed = false;
        auto decide = [&](Dir& d, int64_t pk, bool td, int dir) {
            const int64_t found = (int64_t)((uint64_t)pk >> 32);
            const int64_t next_edges = pk & 0xffffffffll;  // top-down levels: exact work of the next level
            if (c.b == 0 && c.tid == 0 && d.L < kLevelTrace) {
                gs->lv_ts[dir][d.L] = wall_clock64();
                gs->lv_found[dir][d.L] = found;
                gs->lv_edges[dir][d.L] = r.s2;
                gs->lv_mode[dir][d.L] = td ? 0 : 1;
            }
            d.visited += found;
            ++d.L;
            if (found == 0 ||
                (td && a.td_local > 0 &&  // the waves may have expanded everything they found
                 load_agent(&gs->qn[dir][d.L % 3].v) == 0 && load_agent(&gs->hn[dir][d.L % 3].v) == 0)) {
                d.live = false;
                acquire = true;  // what follows reads the visited bits exactly
                if (dir == 0 && c.b == 0 && c.tid == 0) gs->ts[2] = wall_clock64();  // the forward one's end
                return;
            }
            if (d.L > (int64_t)a.V + 2) {
                if (c.tid == 0) atomicOr(&gs->err.v, 2);
                failed = true;
                return;
            }
            if (d.mode == 0) {
                const double ef = a.switch_rule ? (double)next_edges : (double)found * a.avg_degree;
                const double er = (double)(set_size - d.visited) * a.avg_degree;
                if (er / a.alpha < ef) d.mode = 1;
            }
            // bottom-up levels and the gather read visited bits exactly
            if (!td || d.mode == 1) acquire = true;
            if (!td && ((set_size * 1.0 / a.beta) > found || d.L - 1 > a.gamma)) {  // this level was bottom-up
                d.mode = 2;
                d.gather = true;
            }
        };
        if (f.live) decide(f, r.s0, tdf, 0);
        if (b.live) decide(b, r.s1, tdb, 1);
        if (failed) return false;
        if (acquire) block_acquire();
    }
    return true;
}

// ---- phases after the giant (run when GiantArgs::rest != 0) --------------

// One trim-1 sweep (trim_1_from_fq_gfq, src/trim_1_gfq.h:365-458) over
// fin[0, n): act is updated in place, survivors are pushed to fout.
// Returns the number this block trimmed.
// With lo >= 0 the block sweeps only fin[lo, lo + n), entries it wrote
// itself, so no barrier has to make other blocks' entries visible first.
__device__ __forceinline__ int64_t trim_sweep(Ctx& c, const int32_t* fin, int64_t n, int32_t* fout, int32_t* fcount,
                                              int64_t lo = -1) {
    const GiantArgs& a = c.a;
    WaveQueue wq{c.stage32(), 0};
    const uint64_t* act = a.act;
    auto live = [act](int32_t w, int32_t self) { return w != self && bit_of(act, w); };
    int trimmed = 0, scanned = 0;
    const bool local = lo >= 0;
    const int64_t first = local ? c.wave * kWave : (int64_t)c.b * kGBlock + c.wave * kWave;
    const int64_t step = local ? kGBlock : (int64_t)c.G * kGBlock;
    if (local) fin += lo;
    for (int64_t ib = first; ib < n; ib += step) {
        const int64_t i = ib + c.lane;
        int32_t v = 0;
        bool alive = false;
        if (i < n) {
            v = local ? load_agent(&fin[i]) : fin[i];
            alive = bit_of(a.act, v);
        }
        const bool out = wave_any_edge(alive, alive ? a.fwb[v] : 0, alive ? a.fwb[v + 1] : 0, v, a.fwa, live, scanned);
        const bool in =
            wave_any_edge(alive && out, alive ? a.bwb[v] : 0, alive ? a.bwb[v + 1] : 0, v, a.bwa, live, scanned);
        const bool trim = alive && !(out && in);
        if (trim) {
            st_agent(&a.lab[v], v);
            atomicAnd(reinterpret_cast<unsigned long long*>(a.act + (v >> 6)), ~(1ull << (v & 63)));
            ++trimmed;
        }
        wq_push<true>(wq, alive && !trim, v, fout, fcount);
    }
    block_flush(c, wq, fout, fcount);
    return block_sum(trimmed, c.sm);
}

__device__ __forceinline__ int32_t uf_root_g(int32_t* parent, int32_t v) {
    int32_t nx = load_agent(&parent[v]);
    while (nx != v) {
        v = nx;
        nx = load_agent(&parent[v]);
    }
    return v;
}

// ---- small-SCC strategies after the trims (wcc_mode 1 and 2), run by
// k_tail: a second cooperative launch on the same stream, made only when the
// giant phase leaves active vertices.  A kernel of their own keeps k_giant's
// code (and its register allocation) as it is.
// col[v] = smallest id of v's WCC in the active subgraph, by label
// propagation over in- and out-edges with jumps along the colour chain, to a
// fixed point (coloring_wcc, src/color_propagation.h:468-593, which keeps
// the largest id).  The count of changed vertices rides on the barrier, so
// every block leaves after the same sweep.
__device__ __forceinline__ bool lp_wcc(Ctx& c, const int32_t* q, int64_t n, int64_t& rounds) {
    const GiantArgs& a = c.a;
    const int64_t nthreads = (int64_t)c.G * kGBlock;
    const int64_t gt = (int64_t)c.b * kGBlock + c.tid;
    int32_t* col = a.parent;
    for (int64_t i = gt; i < n; i += nthreads) st_agent(&col[q[i]], q[i]);
    if (!gsync(a.gs, c.G, c.sm, a.dbg)) return false;
    for (rounds = 1;; ++rounds) {
        int changed = 0;
        for (int64_t i = gt; i < n; i += nthreads) {
            const int32_t v = q[i];
            const int32_t c0 = load_agent(&col[v]);
            int32_t m = c0;
            for (int dir = 0; dir < 2; ++dir) {
                const int32_t* beg = dir ? a.bwb : a.fwb;
                const int32_t* adj = dir ? a.bwa : a.fwa;
                for (int32_t k = beg[v]; k < beg[v + 1]; ++k) {
                    const int32_t w = adj[k];
                    if (w != v && bit_of(a.act, w)) m = min(m, load_agent(&col[w]));
                }
            }
            // colours are ids of active members: follow the chain (at most
            // 100 hops, as the reference's path compression)
            for (int hop = 0; hop < 100; ++hop) {
                const int32_t nx = load_agent(&col[m]);
                if (nx >= m) break;
                m = nx;
            }
            if (m < c0) {
                atomicMin(&col[v], m);
                ++changed;
            }
        }
        Red r;
        if (!gsync<true>(a.gs, c.G, c.sm, a.dbg, Red{block_sum(changed, c.sm), 0, 0, 0}, &r, 1)) return false;
        if (r.s0 == 0) return true;
        if (rounds > (int64_t)a.V + 2) {
            if (c.tid == 0) atomicOr(&a.gs->err.v, 2);
            return false;
        }
    }
}

__device__ __forceinline__ int32_t key_colour(uint64_t k) { return (int32_t)(uint32_t)k; }

// The colouring SCC method (graph_color, src/color_propagation.h:9-467).  A
// vertex's priority is (active out-degree + 5) x active in-degree
// (degree_rank, :9-58; clipped to 31 bits here, which can only change ties:
// the priority chooses the work, never the labels).  Each round every active
// vertex takes the largest (priority, id) key among the vertices that reach
// it: propagation over in-edges with jumps along the colour chain
// (color_propagation, :64-200), to a fixed point.  A vertex whose key is its
// own is a root, and the vertices of its colour that reach it are its SCC
// (color_identify, :203-260; here a fixed point of backward sweeps for all
// roots at once instead of one search per root).  Those SCCs are labelled
// with their smallest id and leave the active set; the rest start the next
// round with their own keys again (color_init, :262-290).  Each round labels
// at least the SCC of the largest key of every weakly connected piece.
__device__ __forceinline__ bool colour_scc(Ctx& c, int& cur, int64_t n, int s, int64_t& rounds, int64_t& sccs) {
    const GiantArgs& a = c.a;
    GiantState* gs = a.gs;
    const int64_t nthreads = (int64_t)c.G * kGBlock;
    const int64_t gt = (int64_t)c.b * kGBlock + c.tid;
    auto buf = [&a](int k) { return k ? a.fq : a.fq_alt; };
    using ull = unsigned long long;
    {  // degree rank, once, into parent
        const int32_t* q = buf(cur);
        for (int64_t i = gt; i < n; i += nthreads) {
            const int32_t v = q[i];
            int64_t od = 0, id = 0;
            for (int32_t k = a.fwb[v]; k < a.fwb[v + 1]; ++k) od += bit_of(a.act, a.fwa[k]);
            for (int32_t k = a.bwb[v]; k < a.bwb[v + 1]; ++k) id += bit_of(a.act, a.bwa[k]);
            st_agent(&a.parent[v], (int32_t)min<int64_t>((od + 5) * id, INT32_MAX));
        }
    }
    sccs = 0;
    for (rounds = 1;; ++rounds) {
        const int32_t* q = buf(cur);
        int32_t* qn = buf(cur ^ 1);
        int32_t* cnt = &gs->fqn[s % 3].v;  // last read three rounds (or sweeps) ago
        if (c.b == 0 && c.tid == 0) __hip_atomic_store(cnt, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        for (int64_t i = gt; i < n; i += nthreads) {
            const int32_t v = q[i];
            st_agent(&a.ckey[v], ((uint64_t)(uint32_t)load_agent(&a.parent[v]) << 32) | (uint32_t)v);
            st_agent(&a.stamp[v], -1);
        }
        if (!gsync(gs, c.G, c.sm, a.dbg)) return false;
        for (int64_t it = 0;; ++it) {  // largest key among the vertices that reach v
            int changed = 0;
            for (int64_t i = gt; i < n; i += nthreads) {
                const int32_t v = q[i];
                const uint64_t k0 = load_agent(&a.ckey[v]);
                uint64_t k = k0;
                for (int32_t e = a.bwb[v]; e < a.bwb[v + 1]; ++e) {
                    const int32_t w = a.bwa[e];
                    if (w != v && bit_of(a.act, w)) k = max(k, load_agent(&a.ckey[w]));
                }
                for (int hop = 0; hop < 100; ++hop) {  // the colour's own key, and so on
                    const uint64_t kc = load_agent(&a.ckey[key_colour(k)]);
                    if (kc <= k) break;
                    k = kc;
                }
                if (k > k0) {
                    atomicMax(reinterpret_cast<ull*>(&a.ckey[v]), (ull)k);
                    ++changed;
                }
            }
            Red r;
            if (!gsync<true>(gs, c.G, c.sm, a.dbg, Red{block_sum(changed, c.sm), 0, 0, 0}, &r, 1)) return false;
            if (r.s0 == 0) break;
            if (it > (int64_t)a.V + 2) {
                if (c.tid == 0) atomicOr(&gs->err.v, 2);
                return false;
            }
        }
        for (int64_t i = gt; i < n; i += nthreads) {  // roots
            const int32_t v = q[i];
            if (key_colour(load_agent(&a.ckey[v])) == v) {
                st_agent(&a.stamp[v], v);
                st_agent(&a.aux[v], v);
            }
        }
        if (!gsync(gs, c.G, c.sm, a.dbg)) return false;
        for (int64_t it = 0;; ++it) {  // v joins its colour's SCC through an out-neighbour in it
            int changed = 0;
            for (int64_t i = gt; i < n; i += nthreads) {
                const int32_t v = q[i];
                if (load_agent(&a.stamp[v]) >= 0) continue;
                const int32_t r = key_colour(load_agent(&a.ckey[v]));
                for (int32_t e = a.fwb[v]; e < a.fwb[v + 1]; ++e) {
                    const int32_t u = a.fwa[e];
                    if (u != v && bit_of(a.act, u) && load_agent(&a.stamp[u]) == r) {
                        st_agent(&a.stamp[v], r);
                        ++changed;
                        break;
                    }
                }
            }
            Red r;
            if (!gsync<true>(gs, c.G, c.sm, a.dbg, Red{block_sum(changed, c.sm), 0, 0, 0}, &r, 1)) return false;
            if (r.s0 == 0) break;
            if (it > (int64_t)a.V + 2) {
                if (c.tid == 0) atomicOr(&gs->err.v, 2);
                return false;
            }
        }
        for (int64_t i = gt; i < n; i += nthreads) {  // smallest member per SCC
            const int32_t v = q[i];
            const int32_t r = load_agent(&a.stamp[v]);
            if (r >= 0 && r != v) atomicMin(&a.aux[r], v);
        }
        if (!gsync(gs, c.G, c.sm, a.dbg)) return false;
        // label the SCCs found, take them out of the active set, queue the rest
        WaveQueue wq{c.stage32(), 0};
        int found = 0;
        for (int64_t ib = (int64_t)c.b * kGBlock + c.wave * kWave; ib < n; ib += nthreads) {
            const int64_t i = ib + c.lane;
            int32_t v = 0, r = -1;
            if (i < n) {
                v = q[i];
                r = load_agent(&a.stamp[v]);
                if (r >= 0) {
                    st_agent(&a.lab[v], load_agent(&a.aux[r]));
                    atomicAnd(reinterpret_cast<ull*>(a.act + (v >> 6)), ~(1ull << (v & 63)));
                    found += r == v;
                }
            }
            wq_push<true>(wq, i < n && r < 0, v, qn, cnt);
        }
        block_flush(c, wq, qn, cnt);
        Red r;
        if (!gsync<true>(gs, c.G, c.sm, a.dbg, Red{block_sum(found, c.sm), 0, 0, 0}, &r, 1)) return false;
        sccs += r.s0;
        n = load_agent(cnt);
        cur ^= 1;
        ++s;
        if (n == 0) return true;
        if (rounds > (int64_t)a.V + 2) {
            if (c.tid == 0) atomicOr(&gs->err.v, 2);
            return false;
        }
    }
}

// Trims and WCC after the giant.  Returns false after a failed barrier.
// Pattern checks: a vertex with more entries than this in either list is
// checked by a whole wave instead of one lane.
constexpr int kPatternLong = 32;

// own_lo / own_n: the slice of the active queue this block wrote; the first
// trim sweep takes only it, so the queue needs no barrier before it (one is
// taken when no sweep comes first).
__device__ __forceinline__ bool rest_phase(Ctx& c, int64_t fq0, int64_t n, int64_t own_lo, int64_t own_n,
                                           int64_t& n_left) {
    const GiantArgs& a = c.a;
    GiantState* gs = a.gs;
    int cur = 1, s = 0;
    auto buf = [&a](int k) { return k ? a.fq : a.fq_alt; };
    int64_t prev = fq0, trim1 = 0, p2 = 0, p3 = 0;
    int it = 0;
    const int64_t nthreads = (int64_t)c.G * kGBlock;
    const int64_t gt = (int64_t)c.b * kGBlock + c.tid;
    // Sweep s pushes into fqn[s % 3]; all blocks read it after the barrier.
    // It clears fqn[(s + 1) % 3] for the next sweep: that counter was last read
    // after sweep s - 2, i.e. before the barrier that ended sweep s - 1.
    auto sweep = [&]() -> bool {
        if (c.b == 0 && c.tid == 0)
            __hip_atomic_store(&gs->fqn[(s + 1) % 3].v, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const int64_t t = s == 0 ? trim_sweep(c, buf(cur), own_n, buf(cur ^ 1), &gs->fqn[s % 3].v, own_lo)
                                 : trim_sweep(c, buf(cur), n, buf(cur ^ 1), &gs->fqn[s % 3].v);
        Red r;
        if (!gsync<true>(gs, c.G, c.sm, a.dbg, Red{t, 0, 0, 0}, &r, 1)) return false;
        trim1 += r.s0;
        n = load_agent(&gs->fqn[s % 3].v);
        cur ^= 1;
        ++s;
        return true;
    };
    // P7: up to trim_times sweeps with the theta stop rule (src/scc_core.cpp:562-602);
    // the first comparison is against the pre-giant queue, as there.
    while (prev > 0 && it < a.trim_times && n > 0) {
        if (!sweep()) return false;
        if (prev - n < prev * a.theta) break;
        prev = n;
        ++it;
    }
    bool wcc = false;
    if (n > 0 && s == 0 && !gsync(gs, c.G, c.sm, a.dbg)) return false;  // no sweep: the queue's barrier
    if (n > 0) {
        // P8-P10: pattern detection against the current act, then commit.
        // Over the active queue (trim_*_from_fq, src/trim_2_3.h:146-283,
        // 482-677, 881-1080) or, with trim_scope 1, over every active vertex
        // of the graph (trim_*_from_graph, src/trim_2_3.h:9-144, 287-480,
        // 682-879): the same detections, a different walk.
        const int32_t* f = buf(cur);
        auto each_active = [&](auto&& fn) {
            if (a.trim_scope) {
                for (int64_t w = gt; w < a.nwords; w += nthreads) {
                    uint64_t m = a.act[w];
                    while (m) {
                        fn((int32_t)(w * 64 + __ffsll((unsigned long long)m) - 1));
                        m &= m - 1;
                    }
                }
            } else {
                for (int64_t i = gt; i < n; i += nthreads) {
                    const int32_t v = f[i];
                    if (bit_of(a.act, v)) fn(v);
                }
            }
        };
        // A lane checks its own vertex, except one with a long list (a lane
        // would walk it entry by entry): those are queued in the wave's LDS
        // buffer and checked afterwards by the whole wave, 64 entries a step.
        {
            int32_t* lq = c.stage32();
            int32_t& nlq = c.sm.wcnt[0][c.wave];
            if (c.lane == 0) nlq = 0;
            wave_sync();
            each_active([&](int32_t v) {
                const int32_t dv = max(a.fwb[v + 1] - a.fwb[v], a.bwb[v + 1] - a.bwb[v]);
                if (dv > kPatternLong) {
                    const int slot = atomicAdd(&nlq, 1);
                    if (slot < kWaveBuf) {
                        lq[slot] = v;
                        return;
                    }
                }
                pattern_check(a.fwb, a.fwa, a.bwb, a.bwa, a.lab, a.stamp, a.act, v);
            });
            wave_sync();
            const int nl = min(nlq, kWaveBuf);
            for (int j = 0; j < nl; ++j)
                pattern_check<true>(a.fwb, a.fwa, a.bwb, a.bwa, a.lab, a.stamp, a.act, lq[j]);
        }
        if (!gsync(gs, c.G, c.sm, a.dbg)) return false;
        int c2 = 0, c3 = 0;
        each_active([&](int32_t v) {
            if (a.lab[v] >= 0) {
                atomicAnd(reinterpret_cast<unsigned long long*>(a.act + (v >> 6)), ~(1ull << (v & 63)));
                if (a.stamp[v] == 2) ++c2;
                else ++c3;
            }
        });
        const int64_t b2 = block_sum(c2, c.sm), b3 = block_sum(c3, c.sm);
        Red r;
        if (!gsync<true>(gs, c.G, c.sm, a.dbg, Red{b2, b3, 0, 0}, &r, 3)) return false;
        p2 = r.s0;
        p3 = r.s1;
        // P11: one more trim-1 sweep
        if (!sweep()) return false;
        // (wcc_mode 1 and 2 run in k_tail, launched when something is left)
        if (n > 0 && a.wcc_mode == 0) {
            // P12: WCC by union-find; parent = smallest id of the WCC
            const int32_t* g = buf(cur);
            for (int64_t i = gt; i < n; i += nthreads) {
                const int32_t v = g[i];
                st_agent(&a.parent[v], v);
            }
            if (!gsync(gs, c.G, c.sm, a.dbg)) return false;
            for (int64_t i = gt; i < n; i += nthreads) {
                const int32_t v = g[i];
                for (int32_t k = a.fwb[v]; k < a.fwb[v + 1]; ++k) {
                    const int32_t w = a.fwa[k];
                    if (w == v || !bit_of(a.act, w)) continue;
                    int32_t x = v, y = w;
                    while (true) {
                        x = uf_root_g(a.parent, x);
                        y = uf_root_g(a.parent, y);
                        if (x == y) break;
                        if (x < y) {
                            const int32_t t = x;
                            x = y;
                            y = t;
                        }
                        if (atomicCAS(&a.parent[x], x, y) == x) break;
                    }
                }
            }
            if (!gsync(gs, c.G, c.sm, a.dbg)) return false;
            for (int64_t i = gt; i < n; i += nthreads) {
                const int32_t v = g[i];
                st_agent(&a.parent[v], uf_root_g(a.parent, v));
            }
            wcc = true;
        }
    }
    if (c.b == 0 && c.tid == 0) {
        gs->n_final = n;
        gs->final_sel = cur;
        gs->trim1 = trim1;
        gs->pat2 = p2;
        gs->pat3 = p3;
        gs->sweeps = s;
        gs->wcc_done = wcc;
        a.fq_n_ctl[cur] = (int32_t)n;
    }
    n_left = n;
    return true;
}

}  // namespace

// Start of a k_giant launch: block 0 clears the state and the control block
// (device-coherent stores, completed before the epoch is published); the
// other blocks wait, bounded, for the launch's epoch and then acquire.
// Returns false after a timed-out wait.
__device__ __forceinline__ bool launch_start(const GiantArgs& a, Smem& sm) {
    GiantState* gs = a.gs;
    if (blockIdx.x == 0) {
        int64_t* p = reinterpret_cast<int64_t*>(gs);
        for (int i = threadIdx.x; i < a.state_words; i += kGBlock) st_agent(&p[i], (int64_t)0);
        int64_t* q = reinterpret_cast<int64_t*>(a.ctl);
        for (int i = threadIdx.x; i < a.ctl_words; i += kGBlock) st_agent(&q[i], (int64_t)0);
        wait_vm();
        __syncthreads();
        if (threadIdx.x == 0) st_agent(&gs->epoch, a.epoch);
        return true;
    }
    if (threadIdx.x == 0) {
        int ok = 1, spins = 0;
        while (load_agent(&gs->epoch) != a.epoch) {
            __builtin_amdgcn_s_sleep(2);
            if (++spins > kSpinCap) {
                atomicOr(&gs->err.v, 1);
                ok = 0;
                break;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        sm.ok = ok;
    }
    __syncthreads();
    return sm.ok;
}

__global__ __launch_bounds__(kGBlock) void k_giant(const GiantArgs a) {
    __shared__ Smem sm;
    if (threadIdx.x == 0) {
        a.dbg[blockIdx.x] = 0;
        a.dbg[gridDim.x + blockIdx.x] = 0;
        sm.nbar = 0;
        sm.gbase = 0;  // block 0 clears the state (gen included) before any block starts
    }
    if (!launch_start(a, sm)) return;
    Ctx c{a, (int)gridDim.x, (int)blockIdx.x, (int)threadIdx.x, lane_id(), wave_id(), sm, 0, 0, 0, 0};
    GiantState* gs = a.gs;
    const int64_t total_waves = (int64_t)c.G * kW;
    const int64_t gwave = (int64_t)c.b * kW + c.wave;
    const bool stamp = c.b == 0 && c.tid == 0;  // phase timestamps (constant-rate clock)
    if (stamp) gs->ts[0] = wall_clock64();
    if (c.tid == 0) {  // launch ramp: when each block started
        const unsigned long long t = wall_clock64();
        atomicMax(reinterpret_cast<unsigned long long*>(&gs->t_start[0]), ~t);  // min as max of ~t (state starts at 0)
        atomicMax(reinterpret_cast<unsigned long long*>(&gs->t_start[1]), t);
    }

    // ---- phase A: trim-1-first (raw degree 0 either way), act bitmap,
    // cleared traversal bitmaps, pivot = largest out*in (lowest id on ties).
    int64_t bestp = 0, surv = 0;
    int32_t besti = INT32_MAX;
    // a wave takes kUA consecutive words at a time and issues all their offset
    // loads before it uses any of them; lane i loads beg[v] once and takes
    // beg[v + 1] from lane i + 1 (lane 63 from the next word's lane 0)
    constexpr int kUA = 16;
    // lane u < kUA holds the previous step's word w0 + u until the next
    // step's loads are out: a load's wait also waits for earlier stores
    int64_t pword = -1;
    uint64_t pmine = 0;
    for (int64_t w0 = gwave * kUA; w0 < a.nwords; w0 += total_waves * kUA) {
        int32_t fo[kUA], bo[kUA];
#pragma unroll
        for (int u = 0; u < kUA; ++u) {
            const int64_t v = (w0 + u) * 64 + c.lane;
            fo[u] = v <= a.V ? a.fwb[v] : 0;
            bo[u] = v <= a.V ? a.bwb[v] : 0;
        }
        int32_t fl = 0, bl = 0;  // offsets of the vertex after the last word
        const int64_t vl = (w0 + kUA) * 64;
        if (c.lane == kWave - 1 && vl <= a.V) {
            fl = a.fwb[vl];
            bl = a.bwb[vl];
        }
        // vertices outside the selection never become active: lane u < kUA
        // loads word w0 + u's selection
        const uint64_t lsel = (c.lane < kUA && w0 + c.lane < a.nwords) ? (a.sel ? a.sel[w0 + c.lane] : ~0ull) : 0ull;
        if (pword >= 0) {
            st_agent(&a.act[pword], pmine);
            st_agent(&a.fvis[pword], (uint64_t)0);
            st_agent(&a.bvis[pword], (uint64_t)0);
            pword = -1;
        }
        uint64_t mine = 0;  // lane u < kUA keeps word w0 + u's survivors
#pragma unroll
        for (int u = 0; u < kUA; ++u) {
            const int64_t w = w0 + u, v = w * 64 + c.lane;
            const int32_t f0 = u + 1 < kUA ? lane_get(fo[u + 1 < kUA ? u + 1 : u], 0) : fl;
            const int32_t b0 = u + 1 < kUA ? lane_get(bo[u + 1 < kUA ? u + 1 : u], 0) : bl;
            int32_t fn = __shfl_down(fo[u], 1, kWave), bn = __shfl_down(bo[u], 1, kWave);
            if (c.lane == kWave - 1) {
                fn = f0;
                bn = b0;
            }
            const uint64_t sw = lane_get(lsel, u);
            const bool on = v < a.V && ((sw >> c.lane) & 1ull);
            const int32_t od = fn - fo[u], id = bn - bo[u];
            const bool s = on && od > 0 && id > 0;
            if (s) {
                const int64_t p = a.wrap32 ? (int64_t)(int32_t)((uint32_t)od * (uint32_t)id) : (int64_t)od * id;
                if (p > 0 && pivot_better(p, (int32_t)v, bestp, besti)) {
                    bestp = p;
                    besti = (int32_t)v;
                }
            }
            const uint64_t m = __ballot(s);
            if (c.lane == u) mine = m;
        }
        if (c.lane < kUA && w0 + c.lane < a.nwords) {
            pword = w0 + c.lane;
            pmine = mine;
            surv += __popcll(mine);
        }
    }
    if (pword >= 0) {
        st_agent(&a.act[pword], pmine);
        st_agent(&a.fvis[pword], (uint64_t)0);
        st_agent(&a.bvis[pword], (uint64_t)0);
    }
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) {
        const int64_t op = __shfl_xor(bestp, d, kWave);
        const int32_t oi = __shfl_xor(besti, d, kWave);
        if (pivot_better(op, oi, bestp, besti)) {
            bestp = op;
            besti = oi;
        }
    }
    if (c.lane == 0) {
        sm.red2[c.wave] = bestp;
        sm.bi[c.wave] = besti;
    }
    const int64_t surv_blk = block_sum(surv, sm);
    if (c.tid == 0) {
        for (int w = 1; w < kW; ++w)
            if (pivot_better(sm.red2[w], sm.bi[w], bestp, besti)) {
                bestp = sm.red2[w];
                besti = sm.bi[w];
            }
        st_agent(&a.part[(int64_t)kPivP * c.G + c.b], (int64_t)(bestp));
        st_agent(&a.part[(int64_t)kPivI * c.G + c.b], (int64_t)(besti));
        st_agent(&a.part[(int64_t)kSurv * c.G + c.b], (int64_t)(surv_blk));
        const unsigned long long t = wall_clock64();  // trim-1-first done in this block
        atomicMax(reinterpret_cast<unsigned long long*>(&gs->t_arrive[0]), ~t);
        atomicMax(reinterpret_cast<unsigned long long*>(&gs->t_arrive[1]), t);
    }
    if (!gsync(gs, c.G, sm, a.dbg)) return;
    // every block reduces the pivot candidates and the survivor count
    bestp = 0;
    besti = INT32_MAX;
    for (int i = c.tid; i < c.G; i += kGBlock) {
        const int64_t p = load_agent(&a.part[(int64_t)kPivP * c.G + i]);
        const int32_t id = (int32_t)load_agent(&a.part[(int64_t)kPivI * c.G + i]);
        if (pivot_better(p, id, bestp, besti)) {
            bestp = p;
            besti = id;
        }
    }
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) {
        const int64_t op = __shfl_xor(bestp, d, kWave);
        const int32_t oi = __shfl_xor(besti, d, kWave);
        if (pivot_better(op, oi, bestp, besti)) {
            bestp = op;
            besti = oi;
        }
    }
    if (c.lane == 0) {
        sm.red2[c.wave] = bestp;
        sm.bi[c.wave] = besti;
    }
    __syncthreads();
    for (int w = 0; w < kW; ++w)
        if (pivot_better(sm.red2[w], sm.bi[w], bestp, besti)) {
            bestp = sm.red2[w];
            besti = sm.bi[w];
        }
    __syncthreads();
    const int64_t fq0 = sum_slot(a.part, kSurv, c.G, sm);
    const int32_t root = (a.n_roots == 0 && bestp > 0 && besti < a.V) ? besti : -1;
    if (stamp) gs->ts[1] = wall_clock64();

    int64_t fw_count = 0, bw_reach = 0, fw_levels = 0, bw_levels = 0;
    int32_t gmin = root;
    const bool multi = a.n_roots > 0;
    const bool traversed = multi || root >= 0;
    if (traversed) {
        int64_t seeded = 0;
        Dir f, b;
        if (!traverse_joint(c, root, fq0, f, b, &seeded)) return;
        fw_count = f.visited + seeded;
        bw_reach = b.visited + seeded;
        fw_levels = f.L;
        bw_levels = b.L;
    }
    if (stamp) {
        if (!traversed) gs->ts[2] = wall_clock64();
        gs->ts[3] = wall_clock64();  // both directions done
    }

    // ---- post: giant labels, act minus giant, every label written once;
    // each block owns a contiguous word range so the rebuilt active queue
    // comes out ascending without atomics.  First the block's words are
    // narrowed and counted, and the giant's smallest id (its label: no larger
    // than the pivot, which the giant contains) is found among words up to
    // the pivot's; the barrier after the counts reduces it as its key.
    const int64_t per = (a.nwords + c.G - 1) / c.G;
    const int64_t wb = min((int64_t)c.b * per, a.nwords), we = min(wb + per, a.nwords);
    const bool want_min = traversed && !multi;
    int64_t cnt = 0, gcnt = 0;
    int32_t mn = INT32_MAX;
    for (int64_t w = wb + c.tid; w < we; w += kGBlock) {
        const uint64_t a0 = a.act[w];
        uint64_t g = 0;
        if (traversed) {  // the giant: reached both ways; bvis keeps only it
            const uint64_t bw = a.bvis[w];
            g = a.fvis[w] & bw;
            if (g != bw) st_agent(&a.bvis[w], g);
        }
        if (g) st_agent(&a.act[w], a0 & ~g);
        // the vertices left at -1 below: only they can still be unlabelled at
        // the end (the traversal no longer reads this bitmap)
        st_agent(&a.snap[w], a0 & ~g);
        cnt += __popcll(a0 & ~g);
        gcnt += __popcll(g);
        if (want_min && g && w <= (root >> 6)) mn = min(mn, (int32_t)(w * 64 + __ffsll((unsigned long long)g) - 1));
    }
    const int64_t cnt_blk = block_sum(cnt, sm), giant_blk = block_sum(gcnt, sm);
    if (want_min) {
        mn = wave_min(mn);
        if (c.lane == 0) sm.tmp[c.wave] = mn;
        __syncthreads();
        mn = lane_get(row_incl(sm.tmp[c.lane & 15], INT32_MAX, [](int32_t x, int32_t y) { return min(x, y); }), 15);
        __syncthreads();
    }
    if (a.trace && c.tid == 0) atomicMax((unsigned long long*)&gs->post_cp[0], (unsigned long long)wall_clock64());
    if (c.tid == 0) {
        st_agent(&a.part[(int64_t)kPost * c.G + c.b], (int64_t)(cnt_blk));
        st_agent(&a.part[(int64_t)kGiantN * c.G + c.b], (int64_t)(giant_blk));
        st_agent(&a.part[(int64_t)kEdgeTD * c.G + c.b], (int64_t)(c.edges_td));
        st_agent(&a.part[(int64_t)kEdgeBU * c.G + c.b], (int64_t)(c.edges_bu));
        st_agent(&a.part[(int64_t)kFoundTD * c.G + c.b], (int64_t)(c.found_td));
        st_agent(&a.part[(int64_t)kFoundBU * c.G + c.b], (int64_t)(c.found_bu));
    }
    {
        Red r;
        if (!gsync<true>(gs, c.G, sm, a.dbg, Red{0, 0, 0, mn < INT32_MAX ? (int64_t)INT32_MAX - mn : 0}, &r, 8))
            return;
        if (want_min) gmin = (int32_t)((int64_t)INT32_MAX - r.key);
    }
    if (a.trace && c.tid == 0) atomicMax((unsigned long long*)&gs->post_cp[1], (unsigned long long)wall_clock64());
    // the labels: one word per lane, the wave loads 64 words at once (its own
    // block's, narrowed above), then writes them four words at a time, four
    // consecutive labels a lane (coalesced 1 KB stores, no loads in the loop)
    for (int64_t base = wb + (int64_t)c.wave * kWave; base < we; base += (int64_t)kW * kWave) {
        const int64_t w = base + c.lane;
        uint64_t ac = 0, g = 0, sw = 0;
        if (w < we) {
            ac = a.act[w];
            g = traversed ? a.bvis[w] : 0ull;
            sw = a.sel ? a.sel[w] : ~0ull;  // unselected vertices keep whatever label they had
        }
        const int nw = (int)min<int64_t>(kWave, we - base);
        const int q = (c.lane & 15) * 4;  // this lane's first bit in its word
        for (int j = 0; j < nw; j += 4) {
            const int src = j + (c.lane >> 4);  // this lane's word in the group of four
            const uint64_t aj = lane_from(ac, src), gj = lane_from(g, src), sj = lane_from(sw, src);
            if (src >= nw) continue;
            const int64_t v0 = (base + src) * 64 + q;
            int32_t l[4];
            bool all = v0 + 3 < a.V;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int64_t v = v0 + k;
                const bool gi = (gj >> (q + k)) & 1ull, act_v = (aj >> (q + k)) & 1ull;
                l[k] = gi ? (multi ? (int32_t)v : gmin) : (act_v ? -1 : (int32_t)v);
                all = all && ((sj >> (q + k)) & 1ull);
            }
            if (all) {
                st_agent4(&a.lab[v0], l[0], l[1], l[2], l[3]);
            } else {
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    if (v0 + k < a.V && ((sj >> (q + k)) & 1ull)) st_agent(&a.lab[v0 + k], l[k]);
            }
        }
    }
    // exclusive prefix of the block counts, then an ordered block-local scan
    int64_t off, tot;
    {
        int64_t x0 = 0, xt = 0;
        for (int i = c.tid; i < c.G; i += kGBlock) {
            const int64_t x = load_agent(&a.part[(int64_t)kPost * c.G + i]);
            xt += x;
            if (i < c.b) x0 += x;
        }
        off = block_sum(x0, sm);
        tot = block_sum(xt, sm);
    }
    // this block's slice of the queue, the active vertices of its own words.
    // The first trim sweep takes it alone (no barrier before it): a sweep
    // clears act bits only of the vertices it sweeps, so no block's sweep
    // touches the words another block is still listing.
    const int64_t own_lo = off;
    for (int64_t base = wb; base < we; base += kGBlock) {
        const int64_t w = base + c.tid;
        const uint64_t m = w < we ? a.act[w] : 0ull;
        int excl, total;
        block_excl_scan<kW>(__popcll(m), excl, total, sm.tmp);
        uint64_t mm = m;
        int64_t o = off + excl;
        while (mm) {
            const int bit = __ffsll((unsigned long long)mm) - 1;
            st_agent(&a.fq[o++], (int32_t)(w * 64 + bit));
            mm &= mm - 1;
        }
        off += total;
    }
    if (a.trace && c.tid == 0) atomicMax((unsigned long long*)&gs->post_cp[2], (unsigned long long)wall_clock64());
    if (c.b == 0) {
        const int64_t w0 = sum_slot(a.part, kEdgeTD, c.G, sm), w1 = sum_slot(a.part, kEdgeBU, c.G, sm);
        const int64_t w2 = sum_slot(a.part, kFoundTD, c.G, sm), w3 = sum_slot(a.part, kFoundBU, c.G, sm);
        const int64_t giant = sum_slot(a.part, kGiantN, c.G, sm);
        if (c.tid == 0) {
            gs->root = root;
            gs->fq0 = fq0;
            gs->fw_count = fw_count;
            gs->giant = giant;
            gs->bw_reach = bw_reach;
            gs->gmin = gmin;
            gs->fq_n = tot;
            gs->fw_levels = fw_levels;
            gs->bw_levels = bw_levels;
            gs->work[0] = w0;
            gs->work[1] = w1;
            gs->work[2] = w2;
            gs->work[3] = w3;
            *a.fq_n_out = (int32_t)tot;
            gs->ts[4] = wall_clock64();
        }
    }
    if (a.rest) {
        __syncthreads();  // the block's queue slice is written
        int64_t n_left = 0;
        if (!rest_phase(c, fq0, tot, own_lo, off - own_lo, n_left)) return;
        if (n_left == 0) {
            // Every label is final (the rest phase's last barrier acquired
            // them).  The post phase gave every selected vertex outside the
            // post-time active set (`snap`) a label >= 0, so only that set
            // is checked for -1, each block over its own word range.
            int64_t left = 0;
            for (int64_t w = wb + c.tid; w < we; w += kGBlock) {
                uint64_t m = a.snap[w];
                while (m) {
                    const int64_t v = w * 64 + __ffsll((unsigned long long)m) - 1;
                    left += a.lab[v] < 0;
                    m &= m - 1;
                }
            }
            left = block_sum(left, sm);
            if (c.tid == 0) {
                if (left) atomicAdd(reinterpret_cast<unsigned long long*>(&gs->unlabelled), (unsigned long long)left);
                if (c.b == 0) gs->unl_checked = 1;
            }
        }
        if (stamp) gs->ts[5] = wall_clock64();
    }
}

__global__ __launch_bounds__(kGBlock) void k_tail(const GiantArgs a) {
    __shared__ Smem sm;
    GiantState* gs = a.gs;
    if (threadIdx.x == 0) {
        a.dbg[blockIdx.x] = 0;
        a.dbg[gridDim.x + blockIdx.x] = 0;
        sm.nbar = 0;
        sm.gbase = gs->gen.v;  // barrier generations continue from k_giant's
        sm.ok = gs->err.v == 0;
    }
    __syncthreads();
    if (!sm.ok) return;  // k_giant failed; the host reports its error
    const int64_t n = gs->n_final;
    int cur = (int)gs->final_sel;
    if (n == 0) return;
    Ctx c{a, (int)gridDim.x, (int)blockIdx.x, (int)threadIdx.x, lane_id(), wave_id(), sm, 0, 0, 0, 0};
    const bool lead = c.b == 0 && c.tid == 0;
    int64_t rounds = 0, sccs = 0;
    if (a.wcc_mode == 1) {  // P12 by label propagation: parent = smallest id of the WCC
        if (!lp_wcc(c, cur ? a.fq : a.fq_alt, n, rounds)) return;
        if (lead) {
            gs->wcc_done = 1;
            gs->tail_rounds = rounds;
        }
    } else if (a.wcc_mode == 2) {  // P12-P14 replaced: nothing is left for the mice phase
        if (!colour_scc(c, cur, n, (int)gs->sweeps, rounds, sccs)) return;
        if (lead) {
            gs->n_final = 0;
            gs->final_sel = cur;
            gs->tail_rounds = rounds;
            gs->tail_sccs = sccs;
            a.fq_n_ctl[cur] = 0;
        }
    }
    if (lead) gs->ts[5] = wall_clock64();
}

__global__ void k_zero_i64(int64_t* p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

// first[v] = the first entry of v's list (-1 if empty): the bottom-up sweep's
// first probe reads 4 dense bytes instead of a cache line of the list.
__global__ void k_first(const int32_t* __restrict__ beg, const int32_t* __restrict__ adj, int32_t V, int32_t* first) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (int64_t)gridDim.x * blockDim.x)
        first[v] = beg[v + 1] > beg[v] ? adj[beg[v]] : -1;
}

constexpr size_t kStageBytes = (size_t)kW * kPairCap * 8;  // dynamic LDS of k_giant and k_tail

int giant_max_blocks_per_cu() {
    for (const void* k : {reinterpret_cast<const void*>(k_giant), reinterpret_cast<const void*>(k_tail)})
        if (hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kStageBytes) != hipSuccess)
            return 0;
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_giant, kGBlock, kStageBytes) != hipSuccess) return 0;
    return nb;
}

int64_t giant_bu_mult(int grid) {
    const int64_t t = (int64_t)grid * kW;
    int64_t m = 2654435761ll % t;  // Knuth's multiplicative hash constant, then the next value coprime to t
    for (;; ++m) {
        int64_t x = m, y = t;
        while (y) {
            const int64_t r = x % y;
            x = y;
            y = r;
        }
        if (x == 1) return m;
    }
}

hipError_t launch_giant(const GiantArgs& a, Ctl* ctl, int grid, hipStream_t s, hipEvent_t before) {
    static_assert(kGiantStateHead % 8 == 0 && sizeof(Ctl) % 8 == 0 && offsetof(GiantState, epoch) % 8 == 0,
                  "state sizes");
    static std::atomic<int64_t> epochs{0};
    hipError_t e;
    if (before && (e = hipEventRecord(before, s)) != hipSuccess) return e;
    GiantArgs copy = a;
    copy.ctl = ctl;
    copy.epoch = ++epochs;
    copy.state_words = (int32_t)((a.trace ? offsetof(GiantState, epoch) : kGiantStateHead) / 8);
    copy.ctl_words = (int32_t)(sizeof(Ctl) / 8);
    void* args[] = {&copy};
    return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_giant), dim3(grid), dim3(kGBlock), args,
                                      kStageBytes, s);
}

hipError_t launch_tail(const GiantArgs& a, int grid, hipStream_t s) {
    GiantArgs copy = a;
    void* args[] = {&copy};
    return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_tail), dim3(grid), dim3(kGBlock), args,
                                      kStageBytes, s);
}

hipError_t launch_first(const int32_t* beg, const int32_t* adj, int32_t V, int32_t* first, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_first, dim3(grid), dim3(256), 0, s, beg, adj, V, first);
    return hipGetLastError();
}

}  // namespace ispan
