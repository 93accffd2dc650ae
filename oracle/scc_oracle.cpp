// ============================================================================
// TEST INFRASTRUCTURE -- THE ORACLE.  Not part of the product.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
// this library, and only to check results against (or as the timed CPU
// baseline).  The product (libispan_hip.so) never links or calls it.
//
// Contents (all exported with a C ABI for ctypes):
//   * orc_rmat_edges / orc_grid_edges : seeded counter-hash graph generators
//   * orc_build_csr                   : edge list -> deduplicated, self-loop-free
//                                       int32 CSR in both directions (iSpan layout)
//   * orc_tarjan                      : serial iterative Tarjan -> canonical labels
//   * orc_canon_ref_raw               : reference raw scc_id encoding -> canonical
//   * orc_ispan                       : CPU (OpenMP) restatement of the iSpan
//                                       shared-memory phase sequence
//
// "Canonical labels": label[v] = smallest vertex id in v's SCC.
//
// The iSpan restatement follows the phase order of src/scc_core.cpp:116-1070
// and the phase algorithms cited at each function below.  It deliberately
// differs from the reference where the reference is wrong or racy:
//   H0/H1  mice labels live in their own array (-1 = unassigned), so pivots 0
//          and 1 cannot collide with "unassigned" or with the giant label
//          (reference: src/fw_bw.h:1658 vs :974).
//   H2     the mice visit stamps are reset by member id, not by queue position
//          (reference: src/fw_bw.h:1560-1563).
//   H3     pivot degree product in int64 by default; `pivot_wrap32` restores
//          the reference's int32 product (src/scc_common.h:117).
//   H5     an empty frontier queue skips the giant phase instead of forcing
//          vertex 0 into it (src/fw_bw.h:974).
//   H6     per-thread queues grow instead of overflowing (src/fw_bw.h:472,537).
//   H7     size-2/3 pattern trims detect against the labels as they stood at
//          the start of the pattern phase and commit after a barrier, so a
//          half-written triple can never be mistaken for a pair.
// Every one of these changes only which code path finds an SCC, never the
// partition, so canonical labels are unchanged wherever the reference is right.
// ============================================================================
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <sys/time.h>
#include <omp.h>
#include <algorithm>
#include <vector>
#include "gen.h"

namespace {

inline double now_s() {
    struct timeval tv;
    gettimeofday(&tv, nullptr);
    return tv.tv_sec + tv.tv_usec * 1e-6;
}

inline int32_t ld(const int32_t* p) { return __atomic_load_n(p, __ATOMIC_RELAXED); }
inline void st(int32_t* p, int32_t v) { __atomic_store_n(p, v, __ATOMIC_RELAXED); }
inline bool cas(int32_t* p, int32_t expect, int32_t want) {
    return __atomic_compare_exchange_n(p, &expect, want, false, __ATOMIC_RELAXED, __ATOMIC_RELAXED);
}

// Static block split used throughout the reference (src/scc_core.cpp:148-151).
inline void split(int64_t n, int T, int t, int64_t* b, int64_t* e) {
    int64_t step = n / T;
    *b = (int64_t)t * step;
    *e = (t == T - 1) ? n : *b + step;
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------- generators

// R-MAT: m draws at `scale` bit levels; draws with an endpoint >= V are
// dropped (V need not be a power of two).  Returns the number of edges kept.
int64_t orc_rmat_edges(int32_t V, int scale, int64_t m, double a, double b, double c,
                       uint64_t seed, int32_t* src, int32_t* dst) {
    orc_rmat_thr t = orc_rmat_thresholds(a, b, c);
    uint64_t base = orc_stream(seed, 1);
    if (((int64_t)1 << scale) == V) {
#pragma omp parallel for schedule(static)
        for (int64_t e = 0; e < m; ++e) {
            uint32_t s, d;
            orc_rmat_edge(base, (uint64_t)e, scale, t, &s, &d);
            src[e] = (int32_t)s;
            dst[e] = (int32_t)d;
        }
        return m;
    }
    int64_t k = 0;
    for (int64_t e = 0; e < m; ++e) {
        uint32_t s, d;
        orc_rmat_edge(base, (uint64_t)e, scale, t, &s, &d);
        if (s < (uint32_t)V && d < (uint32_t)V) {
            src[k] = (int32_t)s;
            dst[k] = (int32_t)d;
            ++k;
        }
    }
    return k;
}

// n x n lattice: each of the L = 2n(n-1) lattice edges is bidirectional when
// its 32-bit draw is below floor(p_bidir * 2^32), otherwise one direction
// chosen by the draw's high bit; then L/100 uniformly random directed edges.
// Capacity needed: 2L + L/100.  Returns the number of edges written.
int64_t orc_grid_edges(int32_t n, double p_bidir, uint64_t seed, int32_t* src, int32_t* dst) {
    const int64_t H = (int64_t)n * (n - 1);
    const int64_t L = 2 * H;
    const int64_t R = L / 100;
    const uint64_t b1 = orc_stream(seed, 2), b2 = orc_stream(seed, 3);
    const uint32_t tb = (uint32_t)(p_bidir * 4294967296.0);
    const uint64_t V = (uint64_t)n * (uint64_t)n;
    int64_t k = 0;
    for (int64_t i = 0; i < L; ++i) {
        int64_t u, v;
        if (i < H) {
            int64_t r = i / (n - 1), cc = i % (n - 1);
            u = r * n + cc;
            v = u + 1;
        } else {
            int64_t j = i - H;
            int64_t r = j / n, cc = j % n;
            u = r * n + cc;
            v = u + n;
        }
        uint64_t h = orc_mix64(b1 ^ (uint64_t)i);
        if ((uint32_t)h < tb) {
            src[k] = (int32_t)u; dst[k] = (int32_t)v; ++k;
            src[k] = (int32_t)v; dst[k] = (int32_t)u; ++k;
        } else if ((h >> 63) == 0) {
            src[k] = (int32_t)u; dst[k] = (int32_t)v; ++k;
        } else {
            src[k] = (int32_t)v; dst[k] = (int32_t)u; ++k;
        }
    }
    for (int64_t i = 0; i < R; ++i) {
        uint64_t h = orc_mix64(b2 ^ (uint64_t)i);
        src[k] = (int32_t)(((h & 0xffffffffull) * V) >> 32);
        dst[k] = (int32_t)(((h >> 32) * V) >> 32);
        ++k;
    }
    return k;
}

// Edge list -> CSR pair.  Self-loops dropped, duplicates removed, every row
// sorted ascending.  fw_adj/bw_adj need capacity m.  Returns E (or -1 on an
// endpoint outside [0, V)).
int64_t orc_build_csr(int32_t V, int64_t m, const int32_t* src, const int32_t* dst,
                      int32_t* fw_beg, int32_t* fw_adj, int32_t* bw_beg, int32_t* bw_adj) {
    for (int64_t i = 0; i < m; ++i)
        if (src[i] < 0 || src[i] >= V || dst[i] < 0 || dst[i] >= V) return -1;
    std::vector<int64_t> cnt((size_t)V + 1, 0);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < m; ++i)
        if (src[i] != dst[i]) __atomic_fetch_add(&cnt[src[i]], 1, __ATOMIC_RELAXED);
    std::vector<int64_t> off((size_t)V + 1, 0);
    for (int32_t v = 0; v < V; ++v) off[v + 1] = off[v] + cnt[v];
    std::vector<int32_t> tmp((size_t)off[V]);
    std::vector<int64_t> pos(off.begin(), off.end() - 1);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < m; ++i) {
        if (src[i] == dst[i]) continue;
        int64_t p = __atomic_fetch_add(&pos[src[i]], 1, __ATOMIC_RELAXED);
        tmp[p] = dst[i];
    }
    std::vector<int64_t> ucnt((size_t)V, 0);
#pragma omp parallel for schedule(dynamic, 1024)
    for (int32_t v = 0; v < V; ++v) {
        int32_t* b = tmp.data() + off[v];
        int32_t* e = tmp.data() + off[v + 1];
        std::sort(b, e);
        ucnt[v] = std::unique(b, e) - b;
    }
    int64_t E = 0;
    for (int32_t v = 0; v < V; ++v) {
        fw_beg[v] = (int32_t)E;
        E += ucnt[v];
    }
    fw_beg[V] = (int32_t)E;
#pragma omp parallel for schedule(dynamic, 1024)
    for (int32_t v = 0; v < V; ++v)
        memcpy(fw_adj + fw_beg[v], tmp.data() + off[v], sizeof(int32_t) * ucnt[v]);
    // transpose
    std::fill(cnt.begin(), cnt.end(), 0);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < E; ++i) __atomic_fetch_add(&cnt[fw_adj[i]], 1, __ATOMIC_RELAXED);
    int64_t acc = 0;
    for (int32_t v = 0; v < V; ++v) {
        bw_beg[v] = (int32_t)acc;
        pos[v] = acc;
        acc += cnt[v];
    }
    bw_beg[V] = (int32_t)acc;
#pragma omp parallel for schedule(dynamic, 1024)
    for (int32_t u = 0; u < V; ++u)
        for (int32_t k = fw_beg[u]; k < fw_beg[u + 1]; ++k) {
            int64_t p = __atomic_fetch_add(&pos[fw_adj[k]], 1, __ATOMIC_RELAXED);
            bw_adj[p] = u;
        }
#pragma omp parallel for schedule(dynamic, 1024)
    for (int32_t v = 0; v < V; ++v) std::sort(bw_adj + bw_beg[v], bw_adj + bw_beg[v + 1]);
    return E;
}

// ------------------------------------------------------------------- Tarjan

// Serial iterative Tarjan (Tarjan 1972) over a forward CSR; writes the
// canonical labelling.  `mask` (may be NULL) restricts to vertices with
// mask[v] != 0; other vertices get label -1.  Returns the number of SCCs.
int64_t orc_tarjan_masked(int32_t V, const int32_t* beg, const int32_t* adj,
                          const uint8_t* mask, int32_t* label) {
    std::vector<int32_t> index((size_t)V, -1), low((size_t)V, 0);
    std::vector<uint8_t> on((size_t)V, 0);
    std::vector<int32_t> stk;
    std::vector<std::pair<int32_t, int32_t>> call;  // (vertex, next edge)
    int32_t next = 0;
    int64_t nscc = 0;
    for (int32_t v = 0; v < V; ++v) label[v] = -1;
    for (int32_t r = 0; r < V; ++r) {
        if (index[r] != -1 || (mask && !mask[r])) continue;
        call.push_back({r, beg[r]});
        index[r] = low[r] = next++;
        stk.push_back(r);
        on[r] = 1;
        while (!call.empty()) {
            int32_t v = call.back().first;
            int32_t& k = call.back().second;
            if (k < beg[v + 1]) {
                int32_t w = adj[k++];
                if (mask && !mask[w]) continue;
                if (index[w] == -1) {
                    index[w] = low[w] = next++;
                    stk.push_back(w);
                    on[w] = 1;
                    call.push_back({w, beg[w]});
                } else if (on[w]) {
                    low[v] = std::min(low[v], index[w]);
                }
                continue;
            }
            call.pop_back();
            if (!call.empty()) {
                int32_t p = call.back().first;
                low[p] = std::min(low[p], low[v]);
            }
            if (low[v] == index[v]) {
                size_t top = stk.size();
                int32_t mn = v;
                size_t i = top;
                do {
                    --i;
                    mn = std::min(mn, stk[i]);
                } while (stk[i] != v);
                for (size_t j = i; j < top; ++j) {
                    label[stk[j]] = mn;
                    on[stk[j]] = 0;
                }
                stk.resize(i);
                ++nscc;
            }
        }
    }
    return nscc;
}

int64_t orc_tarjan(int32_t V, const int32_t* beg, const int32_t* adj, int32_t* label) {
    return orc_tarjan_masked(V, beg, adj, nullptr, label);
}

// Reference raw labels (src/scc_core.cpp scc_id; encoding documented at
// src/fw_bw.h:974,1658, src/trim_1_gfq.h:21, src/trim_2_3.h:216,579,979) ->
// canonical.  1 and each v >= 2 are groups; -1 is a singleton; -2/-3/-4
// groups are recovered by an SCC run on the subgraph they induce.
// Returns 0, or -1 if a raw label is 0 (unassigned) or unknown.
int orc_canon_ref_raw(int32_t V, const int32_t* beg, const int32_t* adj,
                      const int32_t* raw, int32_t* label) {
    int32_t giant_min = INT32_MAX;
    std::vector<int32_t> gmin((size_t)V + 1, INT32_MAX);
    std::vector<uint8_t> mask((size_t)V, 0);
    bool any_small = false;
    for (int32_t v = 0; v < V; ++v) {
        int32_t r = raw[v];
        if (r == 1) giant_min = std::min(giant_min, v);
        else if (r >= 2 && r < V) gmin[r] = std::min(gmin[r], v);
        else if (r >= 2) return -1;
        else if (r == -2 || r == -3 || r == -4) { mask[v] = 1; any_small = true; }
        else if (r != -1) return -1;
    }
    std::vector<int32_t> sub;
    if (any_small) {
        sub.resize((size_t)V);
        orc_tarjan_masked(V, beg, adj, mask.data(), sub.data());
    }
    for (int32_t v = 0; v < V; ++v) {
        int32_t r = raw[v];
        if (r == 1) label[v] = giant_min;
        else if (r >= 2) label[v] = gmin[r];
        else if (r == -1) label[v] = v;
        else label[v] = sub[v];
    }
    return 0;
}

// ------------------------------------------------------- iSpan restatement

typedef struct {
    int32_t alpha, beta, gamma;  // DO-BFS switch thresholds (src/fw_bw.h:581,599)
    double theta;                // trim-1 loop cutoff (src/scc_core.cpp:597)
    int32_t trim_times;          // TRIM_TIMES (src/util.h:34)
    int32_t threads;
    int32_t pivot_wrap32;        // 1: reference int32 degree product
} orc_params;

// times[]: same slot map as the reference's avg_time (src/load_print.cpp:67-68,
// src/scc_core.cpp:1038-1052), in ms:
//   0 trim total, 1 elephant (pivot+fw+bw), 2 mice (wcc+mice fw-bw), 3 total,
//   4 trim size_1, 5 trim size_2, 6 pivot, 7 fw, 8 bw, 9 wcc, 10 mice fw-bw,
//   13 trim size_3.
// counters[]: 0 fq after trim-1-first, 1 pivot, 2 |FW|, 3 |giant|,
//   4 trim-2 vertices, 5 trim-3 vertices, 6 WCC count, 7 mice vertices,
//   8 fw levels, 9 bw levels.
// Returns 0, or -2 if a vertex is left unlabelled (never expected).
int orc_ispan(int32_t V, const int32_t* fwb, const int32_t* fwa, const int32_t* bwb,
              const int32_t* bwa, const orc_params* P, int32_t* lab, double* times,
              int64_t* counters) {
    const int T = P->threads > 0 ? P->threads : omp_get_max_threads();
    const int64_t E = V > 0 ? fwb[V] : 0;
    const double avg_degree = V > 0 ? (double)E / V : 0.0;
    const int32_t FREE = -1, GIANT = -2;

    std::vector<int32_t> fw_lvl((size_t)V), bw_lvl((size_t)V), color((size_t)V), pend((size_t)V);
    std::vector<int32_t> qa((size_t)V + 1), qb((size_t)V + 1), fwset((size_t)V + 1), cur((size_t)V + 1);
    std::vector<int64_t> cnt(T * 8), pre(T * 8);
    std::vector<int32_t> wcc;
    for (int32_t i = 0; i < V; ++i) {  // untimed init, as src/scc_core.cpp:116-125
        color[i] = i;
        fw_lvl[i] = bw_lvl[i] = -1;
        lab[i] = FREE;
        pend[i] = -1;
    }
    double tm[16] = {0};
    int64_t ctr[16] = {0};
    int32_t* fq = qa.data();
    int32_t* fq_alt = qb.data();
    int64_t fq_n = 0, fwset_n = 0, cur_n = 0;
    int32_t root = -1;
    bool any_flag = false;

    double t_start = now_s();
#pragma omp parallel num_threads(T)
    {
        const int tid = omp_get_thread_num();
        int64_t vb, ve;
        std::vector<int32_t> myq;
        myq.reserve(1024);
        double t0 = now_s();

        // P1 trim_1_first (src/trim_1_gfq.h:9-33): raw out- or in-degree 0.
        split(V, T, tid, &vb, &ve);
        for (int64_t v = vb; v < ve; ++v)
            if (fwb[v + 1] == fwb[v] || bwb[v + 1] == bwb[v]) lab[v] = (int32_t)v;
#pragma omp barrier
        double t_first = now_s() - t0;

        // P2 generate_frontier_queue (src/trim_1_gfq.h:486-529): ascending compaction.
        t0 = now_s();
        {
            int64_t c = 0;
            for (int64_t v = vb; v < ve; ++v) c += lab[v] == FREE;
            cnt[tid * 8] = c;
#pragma omp barrier
            int64_t p = 0;
            for (int i = 0; i < tid; ++i) p += cnt[i * 8];
            for (int64_t v = vb; v < ve; ++v)
                if (lab[v] == FREE) fq[p++] = (int32_t)v;
#pragma omp barrier
#pragma omp single
            {
                fq_n = 0;
                for (int i = 0; i < T; ++i) fq_n += cnt[i * 8];
                ctr[0] = fq_n;
            }
        }
        double t_size1 = now_s() - t0;

        // P3 pivot_selection_from_fq (src/scc_common.h:92-150): first maximum of
        // out*in over the ascending queue, i.e. the lowest id on ties.
        t0 = now_s();
        {
            int64_t qb0, qe0;
            split(fq_n, T, tid, &qb0, &qe0);
            int64_t best = 0;
            int32_t bestv = -1;
            for (int64_t i = qb0; i < qe0; ++i) {
                int32_t v = fq[i];
                int64_t od = fwb[v + 1] - fwb[v], id = bwb[v + 1] - bwb[v];
                int64_t prod = P->pivot_wrap32 ? (int64_t)(int32_t)((uint32_t)od * (uint32_t)id) : od * id;
                if (prod > best) { best = prod; bestv = v; }
            }
            cnt[tid * 8] = best;
            pre[tid * 8] = bestv;
#pragma omp barrier
#pragma omp single
            {
                int64_t b = 0;
                root = -1;
                for (int i = 0; i < T; ++i)
                    if (cnt[i * 8] > b) { b = cnt[i * 8]; root = (int32_t)pre[i * 8]; }
                if (root < 0 && fq_n > 0 && P->pivot_wrap32) root = 0;  // reference default
                if (root >= 0 && lab[root] != FREE) root = -1;
                ctr[1] = root;
            }
        }
        double t_pivot = now_s() - t0;

        // ---- generic DO-BFS step shared by FW (src/fw_bw.h:396-698) and BW
        // (src/fw_bw.h:944-1248).  Modes: 0 sync top-down, 1 Rsync bottom-up,
        // 2 async top-down to exhaustion.
        auto concat = [&](int32_t* dstq, int64_t* n_out) {
            cnt[tid * 8] = (int64_t)myq.size();
#pragma omp barrier
            int64_t p = 0;
            for (int i = 0; i < tid; ++i) p += cnt[i * 8];
            if (!myq.empty()) memcpy(dstq + p, myq.data(), myq.size() * sizeof(int32_t));
#pragma omp barrier
#pragma omp single
            {
                int64_t s = 0;
                for (int i = 0; i < T; ++i) s += cnt[i * 8];
                *n_out = s;
            }
        };

        double t_fw = 0, t_bw = 0;
        if (root >= 0) {
            // P4 forward reachability from root among FREE vertices.
            t0 = now_s();
            int32_t level = 0;
            int mode = 0;
            int64_t visited = 0;
#pragma omp single
            {
                fw_lvl[root] = 0;
                cur[0] = root;
                cur_n = 1;
            }
            int64_t fb, fe;
            split(fq_n, T, tid, &fb, &fe);
            while (true) {
                myq.clear();
                int64_t local = 0;
                if (mode == 0 || mode == 2) {
                    int64_t b, e;
                    split(cur_n, T, tid, &b, &e);
                    if (mode == 0) {
                        for (int64_t i = b; i < e; ++i) {
                            int32_t v = cur[i];
                            for (int32_t k = fwb[v]; k < fwb[v + 1]; ++k) {
                                int32_t w = fwa[k];
                                if (ld(&lab[w]) == FREE && ld(&fw_lvl[w]) == -1 && cas(&fw_lvl[w], -1, level + 1))
                                    myq.push_back(w);
                            }
                        }
                        local = (int64_t)myq.size();
                    } else {
                        for (int64_t i = b; i < e; ++i) myq.push_back(cur[i]);
                        for (size_t h = 0; h < myq.size(); ++h) {
                            int32_t v = myq[h];
                            for (int32_t k = fwb[v]; k < fwb[v + 1]; ++k) {
                                int32_t w = fwa[k];
                                if (ld(&lab[w]) == FREE && ld(&fw_lvl[w]) == -1 && cas(&fw_lvl[w], -1, level + 1))
                                    myq.push_back(w);
                            }
                        }
                        local = 0;  // async mode ends the loop (src/fw_bw.h:506-544)
                    }
                } else {
                    for (int64_t i = fb; i < fe; ++i) {
                        int32_t v = fq[i];
                        if (ld(&lab[v]) != FREE || ld(&fw_lvl[v]) != -1) continue;
                        for (int32_t k = bwb[v]; k < bwb[v + 1]; ++k) {
                            int32_t u = bwa[k];
                            if (ld(&lab[u]) == FREE && ld(&fw_lvl[u]) != -1) {
                                st(&fw_lvl[v], level + 1);
                                ++local;
                                break;
                            }
                        }
                    }
                }
                cnt[tid * 8] = local;
#pragma omp barrier
                int64_t frontier = 0;
                for (int i = 0; i < T; ++i) frontier += cnt[i * 8];
                visited += frontier;
#pragma omp barrier
                if (frontier == 0) break;
                bool gather = false;
                if (mode == 0) {
                    double ef = (double)frontier * avg_degree;
                    double er = (double)(fq_n - visited) * avg_degree;
                    if (er / P->alpha < ef) mode = 1;
                } else if (mode == 1 && ((fq_n * 1.0 / P->beta) > frontier || level > P->gamma)) {
                    myq.clear();
                    for (int64_t i = fb; i < fe; ++i) {
                        int32_t v = fq[i];
                        if (lab[v] == FREE && fw_lvl[v] == level + 1) myq.push_back(v);
                    }
                    mode = 2;
                    gather = true;
                }
                if (mode == 0 || gather) concat(cur.data(), &cur_n);
                ++level;
            }
#pragma omp single
            ctr[8] = level;
            t_fw = now_s() - t0;

            // P5 bw_gfq_from_fw (src/trim_1_gfq.h:583-633): FW set, ascending.
            t0 = now_s();
            {
                int64_t c = 0;
                for (int64_t i = fb; i < fe; ++i) c += fw_lvl[fq[i]] != -1;
                cnt[tid * 8] = c;
#pragma omp barrier
                int64_t p = 0;
                for (int i = 0; i < tid; ++i) p += cnt[i * 8];
                for (int64_t i = fb; i < fe; ++i)
                    if (fw_lvl[fq[i]] != -1) fwset[p++] = fq[i];
#pragma omp barrier
#pragma omp single
                {
                    fwset_n = 0;
                    for (int i = 0; i < T; ++i) fwset_n += cnt[i * 8];
                    ctr[2] = fwset_n;
                }
            }

            // P6 backward reachability inside the FW set -> giant SCC.
            level = 0;
            mode = 0;
            visited = 0;
#pragma omp single
            {
                bw_lvl[root] = 0;
                lab[root] = GIANT;
                cur[0] = root;
                cur_n = 1;
            }
            int64_t sb, se;
            split(fwset_n, T, tid, &sb, &se);
            while (true) {
                myq.clear();
                int64_t local = 0;
                if (mode == 0 || mode == 2) {
                    int64_t b, e;
                    split(cur_n, T, tid, &b, &e);
                    if (mode == 2)
                        for (int64_t i = b; i < e; ++i) myq.push_back(cur[i]);
                    size_t h = 0;
                    int64_t i = b;
                    while (true) {
                        int32_t v;
                        if (mode == 0) {
                            if (i >= e) break;
                            v = cur[i++];
                        } else {
                            if (h >= myq.size()) break;
                            v = myq[h++];
                        }
                        for (int32_t k = bwb[v]; k < bwb[v + 1]; ++k) {
                            int32_t w = bwa[k];
                            if (ld(&lab[w]) == FREE && ld(&bw_lvl[w]) == -1 && ld(&fw_lvl[w]) != -1 &&
                                cas(&bw_lvl[w], -1, level + 1)) {
                                st(&lab[w], GIANT);
                                myq.push_back(w);
                            }
                        }
                    }
                    local = mode == 0 ? (int64_t)myq.size() : 0;
                } else {
                    for (int64_t i = sb; i < se; ++i) {
                        int32_t v = fwset[i];
                        if (ld(&lab[v]) != FREE) continue;
                        for (int32_t k = fwb[v]; k < fwb[v + 1]; ++k) {
                            if (ld(&lab[fwa[k]]) == GIANT) {
                                st(&bw_lvl[v], level + 1);
                                st(&lab[v], GIANT);
                                ++local;
                                break;
                            }
                        }
                    }
                }
                cnt[tid * 8] = local;
#pragma omp barrier
                int64_t frontier = 0;
                for (int i = 0; i < T; ++i) frontier += cnt[i * 8];
                visited += frontier;
#pragma omp barrier
                if (frontier == 0) break;
                bool gather = false;
                if (mode == 0) {
                    double ef = (double)frontier * avg_degree;
                    double er = (double)(fwset_n - visited) * avg_degree;
                    if (er / P->alpha < ef) mode = 1;
                } else if (mode == 1 && ((fwset_n * 1.0 / P->beta) > frontier || level > P->gamma)) {
                    myq.clear();
                    for (int64_t i = sb; i < se; ++i) {
                        int32_t v = fwset[i];
                        if (lab[v] == GIANT && bw_lvl[v] == level + 1) myq.push_back(v);
                    }
                    mode = 2;
                    gather = true;
                }
                if (mode == 0 || gather) concat(cur.data(), &cur_n);
                ++level;
            }
#pragma omp single
            ctr[9] = level;
            t_bw = now_s() - t0;
        }
#pragma omp barrier

        // trim_1_from_fq_gfq (src/trim_1_gfq.h:365-458): a FREE vertex with no
        // FREE non-self out- or in-neighbour is a singleton SCC.  Updates are
        // in place (a vertex trimmed now is seen by later checks in the same
        // sweep); the queue is compacted ascending.  Returns the new size.
        auto trim1 = [&]() -> int64_t {
            int64_t b, e;
            split(fq_n, T, tid, &b, &e);
            int64_t c = 0;
            for (int64_t i = b; i < e; ++i) {
                int32_t v = fq[i];
                if (ld(&lab[v]) != FREE) continue;
                bool out = false, in = false;
                for (int32_t k = fwb[v]; k < fwb[v + 1] && !out; ++k) {
                    int32_t w = fwa[k];
                    out = w != v && ld(&lab[w]) == FREE;
                }
                if (out)
                    for (int32_t k = bwb[v]; k < bwb[v + 1] && !in; ++k) {
                        int32_t w = bwa[k];
                        in = w != v && ld(&lab[w]) == FREE;
                    }
                if (!out || !in) st(&lab[v], v);
                else ++c;
            }
            cnt[tid * 8] = c;
#pragma omp barrier
            int64_t p = 0;
            for (int i = 0; i < tid; ++i) p += cnt[i * 8];
            for (int64_t i = b; i < e; ++i)
                if (lab[fq[i]] == FREE) fq_alt[p++] = fq[i];
#pragma omp barrier
#pragma omp single
            {
                std::swap(fq, fq_alt);
                fq_n = 0;
                for (int i = 0; i < T; ++i) fq_n += cnt[i * 8];
            }
            return fq_n;
        };

        // P7: up to trim_times sweeps, stop when fewer than theta*prev trimmed
        // (src/scc_core.cpp:562-602).
        t0 = now_s();
        {
            int64_t prev = fq_n;
            int it = 0;
            while (prev > 0 && it < P->trim_times) {
                int64_t n = trim1();
                if (prev - n < prev * P->theta) break;
                prev = n;
                ++it;
            }
        }
        t_size1 += now_s() - t0;

        double t_size2 = 0, t_size3 = 0, t_wcc = 0, t_mice = 0;
        if (fq_n > 0) {
            // ---- pattern trims (src/trim_2_3.h).  nb(): the distinct FREE
            // neighbours of x outside {x1, x2}; stops after `cap` + 1.
            auto nb = [&](const int32_t* beg, const int32_t* adj, int32_t x, int32_t x1, int32_t x2,
                          int cap, int32_t* got) -> int {
                int n = 0;
                for (int32_t k = beg[x]; k < beg[x + 1]; ++k) {
                    int32_t w = adj[k];
                    if (w == x || w == x1 || w == x2 || lab[w] != FREE) continue;
                    bool seen = false;
                    for (int j = 0; j < n && j < cap; ++j) seen |= got[j] == w;
                    if (seen) continue;
                    if (n < cap) got[n] = w;
                    if (++n > cap) break;
                }
                return n;
            };
            auto mark = [&](int32_t a, int32_t b2, int32_t c2) {
                int32_t m = std::min(a, b2);
                if (c2 >= 0) m = std::min(m, c2);
                st(&pend[a], m);
                st(&pend[b2], m);
                if (c2 >= 0) st(&pend[c2], m);
            };
            auto commit = [&](int64_t* nlab) {
#pragma omp barrier
                int64_t b, e, c = 0;
                split(fq_n, T, tid, &b, &e);
                for (int64_t i = b; i < e; ++i) {
                    int32_t v = fq[i];
                    if (pend[v] != -1) {
                        lab[v] = pend[v];
                        pend[v] = -1;
                        ++c;
                    }
                }
                cnt[tid * 8] = c;
#pragma omp barrier
#pragma omp single
                for (int i = 0; i < T; ++i) *nlab += cnt[i * 8];
            };
            int64_t qb1, qe1;
            split(fq_n, T, tid, &qb1, &qe1);

            // P8 trim_2_from_fq (src/trim_2_3.h:146-283): {v,u} with v's only
            // FREE out-neighbour u and u's only one v; same on in-edges.
            t0 = now_s();
            for (int64_t i = qb1; i < qe1; ++i) {
                int32_t v = fq[i], g[2], h[2];
                if (lab[v] != FREE) continue;
                for (int dir = 0; dir < 2; ++dir) {
                    const int32_t* beg = dir == 0 ? fwb : bwb;
                    const int32_t* adj = dir == 0 ? fwa : bwa;
                    if (nb(beg, adj, v, -1, -1, 1, g) == 1 && nb(beg, adj, g[0], -1, -1, 1, h) == 1 && h[0] == v) {
                        mark(v, g[0], -1);
                        break;
                    }
                }
            }
            commit(&ctr[4]);
            t_size2 = now_s() - t0;

            // P9 trim_3_1_from_fq (src/trim_2_3.h:482-677): cycle A->B->C->A
            // closed under out-edges (or in-edges).
            t0 = now_s();
            for (int64_t i = qb1; i < qe1; ++i) {
                int32_t a = fq[i], g[2], h[2], k2[2];
                if (lab[a] != FREE) continue;
                for (int dir = 0; dir < 2; ++dir) {
                    const int32_t* beg = dir == 0 ? fwb : bwb;
                    const int32_t* adj = dir == 0 ? fwa : bwa;
                    if (nb(beg, adj, a, -1, -1, 1, g) != 1) continue;
                    int32_t b2 = g[0];
                    if (nb(beg, adj, b2, a, -1, 1, h) != 1) continue;
                    int32_t c2 = h[0];
                    if (nb(beg, adj, c2, b2, -1, 1, k2) == 1 && k2[0] == a) {
                        mark(a, b2, c2);
                        break;
                    }
                }
            }
            commit(&ctr[5]);

            // P10 trim_3_2_from_fq (src/trim_2_3.h:881-1080): hub A with
            // exactly two FREE out-neighbours B, C whose only other one is A.
            for (int64_t i = qb1; i < qe1; ++i) {
                int32_t a = fq[i], g[3], h[2], k2[2];
                if (lab[a] != FREE) continue;
                for (int dir = 0; dir < 2; ++dir) {
                    const int32_t* beg = dir == 0 ? fwb : bwb;
                    const int32_t* adj = dir == 0 ? fwa : bwa;
                    if (nb(beg, adj, a, -1, -1, 2, g) != 2) continue;
                    int32_t b2 = g[0], c2 = g[1];
                    if (nb(beg, adj, b2, b2, c2, 1, h) == 1 && h[0] == a && nb(beg, adj, c2, b2, c2, 1, k2) == 1 &&
                        k2[0] == a) {
                        mark(a, b2, c2);
                        break;
                    }
                }
            }
            commit(&ctr[5]);
            t_size3 = now_s() - t0;

            // P11 one more trim-1 sweep (src/scc_core.cpp:757-795).
            t0 = now_s();
            trim1();
            t_size1 += now_s() - t0;

            // P12 coloring_wcc (src/color_propagation.h:468-593): max-id label
            // propagation over FREE in/out neighbours + pointer jumping, to a
            // fixed point.  Result: color = largest id of the WCC.
            t0 = now_s();
            {
                int64_t b, e;
                split(fq_n, T, tid, &b, &e);
                while (true) {
#pragma omp barrier
                    bool changed = false;
                    for (int64_t i = b; i < e; ++i) {
                        int32_t v = fq[i];
                        int32_t cv = ld(&color[v]);
                        for (int dir = 0; dir < 2; ++dir) {
                            const int32_t* beg = dir == 0 ? bwb : fwb;
                            const int32_t* adj = dir == 0 ? bwa : fwa;
                            for (int32_t k = beg[v]; k < beg[v + 1]; ++k) {
                                int32_t w = adj[k];
                                if (w == v || lab[w] != FREE) continue;
                                int32_t cw = ld(&color[w]);
                                if (cv < cw) {
                                    cv = cw;
                                    changed = true;
                                }
                            }
                        }
                        if (cv != ld(&color[v])) st(&color[v], cv);
                    }
#pragma omp barrier
                    if (changed) {
                        for (int64_t i = b; i < e; ++i) {
                            int32_t v = fq[i];
                            int32_t r = ld(&color[v]);
                            if (r == v) continue;
                            for (int hop = 0; hop < 100; ++hop) {
                                int32_t nr = ld(&color[r]);
                                if (nr == r) break;
                                r = nr;
                            }
                            // compress v's chain towards r, never lowering a
                            // colour.  The walk stops at a root (colour ==
                            // id) and where another thread has already raised
                            // the chain past r: colours only grow, so climbing
                            // on from there could end at a root above r and
                            // spin on it (the reference writes root
                            // unconditionally instead, :560-565).
                            int32_t x = v;
                            while (x != r) {
                                int32_t nx = ld(&color[x]);
                                if (nx == r) break;
                                if (nx < r) st(&color[x], r);
                                if (nx == x || nx > r) break;
                                x = nx;
                            }
                        }
                    }
                    cnt[tid * 8] = changed;
#pragma omp barrier
                    bool any = false;
                    for (int i = 0; i < T; ++i) any |= cnt[i * 8] != 0;
                    if (!any) break;
                }
            }
            t_wcc = now_s() - t0;

            // P13 init_fw_sa (src/fw_bw.h:1548-1583), with the stamp reset done
            // by member id: the sorted list of distinct WCC colours.
            t0 = now_s();
            {
                int64_t b, e;
                split(fq_n, T, tid, &b, &e);
                for (int64_t i = b; i < e; ++i) fw_lvl[fq[i]] = -1;
            }
#pragma omp single
            {
                wcc.resize((size_t)fq_n);
                for (int64_t i = 0; i < fq_n; ++i) wcc[i] = color[fq[i]];
                std::sort(wcc.begin(), wcc.end());
                wcc.erase(std::unique(wcc.begin(), wcc.end()), wcc.end());
                ctr[6] = (int64_t)wcc.size();
                ctr[7] = fq_n;
            }
            // P14 mice_fw_bw (src/fw_bw.h:1585-1691): each thread owns a
            // contiguous chunk of the sorted colours; in ascending queue order,
            // an unlabelled vertex v of an owned WCC is the pivot of a FW search
            // (stamp v) and a BW search inside it; its SCC gets label v, the
            // smallest id of that SCC since every smaller FREE id of the WCC is
            // labelled already.
            {
                const int64_t W = (int64_t)wcc.size();
                const int64_t step = W / T;
                std::vector<int32_t> q;
                for (int64_t i = 0; i < fq_n; ++i) {
                    int32_t v = fq[i];
                    int64_t rank = std::lower_bound(wcc.begin(), wcc.end(), color[v]) - wcc.begin();
                    int64_t owner = step > 0 ? std::min<int64_t>(rank / step, T - 1) : T - 1;
                    if (owner != tid || lab[v] != FREE) continue;  // only the owner writes this WCC
                    q.clear();
                    fw_lvl[v] = v;
                    q.push_back(v);
                    for (size_t h = 0; h < q.size(); ++h) {
                        int32_t x = q[h];
                        for (int32_t k = fwb[x]; k < fwb[x + 1]; ++k) {
                            int32_t w = fwa[k];
                            if (lab[w] == FREE && fw_lvl[w] != v) {
                                fw_lvl[w] = v;
                                q.push_back(w);
                            }
                        }
                    }
                    q.clear();
                    lab[v] = v;
                    q.push_back(v);
                    for (size_t h = 0; h < q.size(); ++h) {
                        int32_t x = q[h];
                        for (int32_t k = bwb[x]; k < bwb[x + 1]; ++k) {
                            int32_t w = bwa[k];
                            if (lab[w] == FREE && fw_lvl[w] == v) {
                                lab[w] = v;
                                q.push_back(w);
                            }
                        }
                    }
                }
            }
#pragma omp barrier
            t_mice = now_s() - t0;
        }
        if (tid == 0) {
            tm[0] = (t_first + t_size1 + t_size2 + t_size3) * 1e3;
            tm[1] = (t_pivot + t_fw + t_bw) * 1e3;
            tm[2] = (t_wcc + t_mice) * 1e3;
            tm[4] = (t_first + t_size1) * 1e3;
            tm[5] = t_size2 * 1e3;
            tm[6] = t_pivot * 1e3;
            tm[7] = t_fw * 1e3;
            tm[8] = t_bw * 1e3;
            tm[9] = t_wcc * 1e3;
            tm[10] = t_mice * 1e3;
            tm[13] = t_size3 * 1e3;
        }
    }
    tm[3] = (now_s() - t_start) * 1e3;

    // giant label = smallest member id (canonical form), then sanity check.
    int32_t gmin = INT32_MAX;
    int64_t gsize = 0, unl = 0;
#pragma omp parallel for reduction(min : gmin) reduction(+ : gsize, unl) num_threads(T)
    for (int32_t v = 0; v < V; ++v) {
        if (lab[v] == GIANT) {
            gmin = std::min(gmin, v);
            ++gsize;
        } else if (lab[v] < 0) {
            ++unl;
        }
    }
    if (gsize) {
#pragma omp parallel for num_threads(T)
        for (int32_t v = 0; v < V; ++v)
            if (lab[v] == GIANT) lab[v] = gmin;
    }
    ctr[3] = gsize;
    if (times) memcpy(times, tm, sizeof(tm));
    if (counters) memcpy(counters, ctr, sizeof(ctr));
    return unl ? -2 : 0;
}

int orc_num_procs(void) { return omp_get_num_procs(); }

}  // extern "C"
