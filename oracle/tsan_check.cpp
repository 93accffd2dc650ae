// ThreadSanitizer driver for the oracle's OpenMP restatement (test
// infrastructure only; tests/test_oracle_tsan.py builds it with clang's
// -fsanitize=thread against LLVM's OpenMP runtime and runs it with the Archer
// tool, which tells ThreadSanitizer about OpenMP's barriers and tasks).
//
// The restatement shares its flags and labels between threads through
// __atomic builtins on purpose, so a report here is an unintended race.  The
// graphs cover every phase: R-MAT (giant, trims), a lattice (long
// traversals), and a chain of small cycles in one weakly connected piece
// (size-2/3 patterns, the WCC's label propagation and path compression, the
// mice phase).  Each run's labels are compared with Tarjan's.  Exit status:
// the number of runs whose labels differ (ThreadSanitizer sets its own on a
// report).
#include <cstdint>
#include <cstdio>
#include <vector>

extern "C" {
int64_t orc_rmat_edges(int32_t V, int scale, int64_t m, double a, double b, double c, uint64_t seed,
                       int32_t* src, int32_t* dst);
int64_t orc_grid_edges(int32_t n, double p_bidir, uint64_t seed, int32_t* src, int32_t* dst);
int64_t orc_build_csr(int32_t V, int64_t m, const int32_t* src, const int32_t* dst, int32_t* fb, int32_t* fa,
                      int32_t* bb, int32_t* ba);
int64_t orc_tarjan(int32_t V, const int32_t* beg, const int32_t* adj, int32_t* label);
typedef struct {
    int32_t alpha, beta, gamma;
    double theta;
    int32_t trim_times;
    int32_t threads;
    int32_t pivot_wrap32;
} orc_params;
int orc_ispan(int32_t V, const int32_t* fwb, const int32_t* fwa, const int32_t* bwb, const int32_t* bwa,
              const orc_params* P, int32_t* lab, double* times, int64_t* counters);
}

namespace {

struct Edges {
    int32_t V = 0;
    std::vector<int32_t> s, d;
};

Edges rmat(int scale, int ef, uint64_t seed) {
    Edges g;
    g.V = 1 << scale;
    const int64_t m = (int64_t)g.V * ef;
    g.s.resize(m);
    g.d.resize(m);
    const int64_t k = orc_rmat_edges(g.V, scale, m, 0.57, 0.19, 0.19, seed, g.s.data(), g.d.data());
    g.s.resize(k);
    g.d.resize(k);
    return g;
}

Edges lattice(int n, uint64_t seed) {
    Edges g;
    g.V = n * n;
    const int64_t L = 2LL * n * (n - 1);
    g.s.resize(2 * L + L / 100);
    g.d.resize(2 * L + L / 100);
    const int64_t k = orc_grid_edges(n, 0.6, seed, g.s.data(), g.d.data());
    g.s.resize(k);
    g.d.resize(k);
    return g;
}

// Groups of 1-6 vertices on a cycle each, group g pointing forward to groups
// g+1..g+3: every group is an SCC, and the sizes cover the trims (1), the
// patterns (2, 3) and the mice phase (4-6).
Edges cycle_chain(int groups, uint64_t seed) {
    Edges g;
    std::vector<int32_t> first;
    uint64_t h = seed * 0x9E3779B97F4A7C15ull + 1;
    auto next = [&h]() {
        h ^= h >> 33;
        h *= 0xff51afd7ed558ccdull;
        h ^= h >> 33;
        return h;
    };
    for (int i = 0; i < groups; ++i) {
        const int k = 1 + (int)(next() % 6);
        first.push_back(g.V);
        for (int j = 0; j < k; ++j) {
            g.s.push_back(g.V + j);
            g.d.push_back(g.V + (j + 1) % k);
        }
        g.V += k;
    }
    first.push_back(g.V);
    for (int i = 0; i < groups; ++i)
        for (int t = 1; t <= 3 && i + t < groups; ++t) {
            g.s.push_back(first[i] + (int32_t)(next() % (uint64_t)(first[i + 1] - first[i])));
            g.d.push_back(first[i + t]);
        }
    return g;
}

int check(const char* name, const Edges& e, const orc_params& p) {
    const int32_t V = e.V;
    const int64_t m = (int64_t)e.s.size();
    std::vector<int32_t> fb(V + 1), bb(V + 1), fa(m > 0 ? m : 1), ba(m > 0 ? m : 1);
    orc_build_csr(V, m, e.s.data(), e.d.data(), fb.data(), fa.data(), bb.data(), ba.data());
    std::vector<int32_t> tj(V), lab(V);
    orc_tarjan(V, fb.data(), fa.data(), tj.data());
    double t[16] = {0};
    int64_t c[16] = {0};
    const int rc = orc_ispan(V, fb.data(), fa.data(), bb.data(), ba.data(), &p, lab.data(), t, c);
    int64_t bad = 0;
    for (int32_t v = 0; v < V; ++v) bad += lab[v] != tj[v];
    printf("%-12s V=%d E=%lld threads=%d alpha=%d beta=%d: rc %d, %lld labels differ\n", name, V,
           (long long)fb[V], p.threads, p.alpha, p.beta, rc, (long long)bad);
    return rc != 0 || bad != 0;
}

}  // namespace

int main() {
    setvbuf(stdout, nullptr, _IONBF, 0);  // progress lines as they happen
    const Edges graphs[] = {rmat(14, 8, 7), lattice(96, 3), cycle_chain(3000, 5)};
    const char* names[] = {"rmat-14-8", "lattice-96", "cycle-chain"};
    int failed = 0;
    for (int i = 0; i < 3; ++i)
        for (int T : {2, 4, 8}) {
            failed += check(names[i], graphs[i], orc_params{30, 200, 10, 0.1, 3, T, 0});
            failed += check(names[i], graphs[i], orc_params{1, 1, 0, 0.9, 1, T, 0});
        }
    return failed;
}
