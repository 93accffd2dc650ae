"""CPU tests of the oracle: the iSpan restatement, Tarjan, the generators and
the reference canonicaliser, pinned against the reference's own fixture,
hand graphs, brute force, scipy and (where it was built) the compiled
reference."""
import numpy as np
import pytest

import oracle as O
from tests import graphs as G

scipy_csgraph = pytest.importorskip("scipy.sparse.csgraph")
from scipy.sparse import csr_matrix  # noqa: E402


def scipy_labels(g):
    A = csr_matrix((np.ones(g.E), g.fw_adj, g.fw_beg), shape=(g.V, g.V))
    n, lab = scipy_csgraph.connected_components(A, directed=True, connection="strong")
    mn = np.full(n, g.V, np.int64)
    np.minimum.at(mn, lab, np.arange(g.V))
    return mn[lab].astype(np.int32)


def is_canonical(lab):
    v = np.arange(lab.shape[0])
    return bool(np.all(lab <= v) and np.all(lab[lab] == lab))


# ---------------------------------------------------------------- toy fixture

def test_toy_csr_matches_shipped_files():
    g, d = G.toy_graph()
    src = np.repeat(np.arange(g.V, dtype=np.int32), np.diff(g.fw_beg))
    rebuilt = O.build_csr(g.V, src, g.fw_adj)
    assert np.array_equal(rebuilt.fw_beg, d["fw_begin"]) and np.array_equal(rebuilt.fw_adj, d["fw_adjacent"])


def test_toy_labels():
    g, d = G.toy_graph()
    assert d["canonical"] == [0, 0, 0, 0, 0, 5, 5]
    assert O.tarjan(g).tolist() == d["canonical"]
    lab, _t, c = O.ispan(g, threads=2)
    assert lab.tolist() == d["canonical"]
    assert c["giant_size"] == 5 and c["trim2"] == 2
    # the reference's raw encoding (1 giant, -2 size-2) canonicalises to the same partition
    assert d["ref_raw_scc_id"] == [1, 1, 1, 1, 1, -2, -2]
    assert O.canon_ref_raw(g, np.asarray(d["ref_raw_scc_id"], np.int32)).tolist() == d["canonical"]


@pytest.mark.skipif(not O.have_ref(), reason="compiled reference not present")
def test_toy_compiled_reference():
    g, d = G.toy_graph()
    raw, out = O.run_ref(g, threads=2)
    assert raw.tolist() == d["ref_raw_scc_id"]
    assert "largest, 5" in out


# ------------------------------------------------------------ hand graphs

@pytest.mark.parametrize("name,d", G.pattern_cases(), ids=[n for n, _ in G.pattern_cases()])
def test_pattern_graphs(name, d):
    g = G.pattern_graph(d)
    tj = O.tarjan(g)
    assert tj.tolist() == d["canonical"]
    assert np.array_equal(G.brute_force_labels(g), tj)
    for threads in (1, 3):
        lab, _t, _c = O.ispan(g, threads=threads)
        assert np.array_equal(lab, tj), (name, threads)


def test_pattern_rules_fire():
    """Each size-2/3 hand graph is peeled by the pattern phase, not by mice."""
    for name in ("pair_sink", "pair_source"):
        _l, _t, c = O.ispan(G.pattern_graph(dict(G.pattern_cases())[name]), threads=1)
        assert c["trim2"] == 2, name
    for name in ("cycle3_sink", "cycle3_source", "hub3_sink", "hub3_source"):
        _l, _t, c = O.ispan(G.pattern_graph(dict(G.pattern_cases())[name]), threads=1)
        assert c["trim3"] == 3, name


def test_reference_hazards_recorded():
    """The fixtures record that the compiled reference livelocks on H0 and
    returns a wrong partition on H1/H2; the restatement returns Tarjan's."""
    cases = dict(G.pattern_cases())
    assert cases["H0_pivot0"]["ref"].startswith("timeout")
    assert cases["H1_pivot1"]["ref"] == "wrong partition"
    assert cases["H2_stale_stamp"]["ref"] == "wrong partition"
    for name in ("H0_pivot0", "H1_pivot1", "H2_stale_stamp"):
        g = G.pattern_graph(cases[name])
        lab, _t, _c = O.ispan(g, threads=2)
        assert np.array_equal(lab, O.tarjan(g)), name


# --------------------------------------------------------- random graphs

@pytest.mark.parametrize("seed", range(60))
def test_random_graphs_three_ways(seed):
    g = G.random_graph(seed, vmax=400, density=[0.5, 1.5, 4.0][seed % 3])
    tj = O.tarjan(g)
    assert np.array_equal(tj, scipy_labels(g))
    assert is_canonical(tj)
    lab, _t, _c = O.ispan(g, threads=1 + seed % 4, pivot_wrap32=bool(seed % 2))
    assert np.array_equal(lab, tj)


@pytest.mark.parametrize("seed", range(20))
def test_tiny_graphs_brute_force(seed):
    g = G.random_graph(1000 + seed, vmax=12, density=2.0)
    assert np.array_equal(O.tarjan(g), G.brute_force_labels(g))


@pytest.mark.parametrize("seed", range(12))
def test_cycle_forests(seed):
    g = G.cycles_graph(seed, n_cycles=30 + 10 * seed)
    tj = O.tarjan(g)
    assert np.array_equal(tj, scipy_labels(g))
    for threads in (1, 4):
        lab, _t, _c = O.ispan(g, threads=threads, theta=[0.1, 0.9][seed % 2], trim_times=seed % 4)
        assert np.array_equal(lab, tj)


@pytest.mark.parametrize("alpha,beta,gamma", [(1, 1, 0), (30, 200, 10), (1000000, 2, 1), (2, 1000000, 100)])
def test_switch_thresholds_do_not_change_labels(alpha, beta, gamma):
    g = O.rmat(14, 8, 7)
    lab, _t, _c = O.ispan(g, alpha=alpha, beta=beta, gamma=gamma, threads=4)
    assert np.array_equal(lab, O.tarjan(g))


# ------------------------------------------------------- seeded generators

@pytest.mark.parametrize("key,d", G.hash_cases(max_v=1 << 20), ids=[k for k, _ in G.hash_cases(max_v=1 << 20)])
def test_hashed_graphs(key, d):
    g = G.build_hashed(d)
    assert (g.V, g.E) == (d["V"], d["E"])
    assert G.sha(g.fw_beg, g.fw_adj, g.bw_beg, g.bw_adj) == d["csr_sha256"]
    tj = O.tarjan(g)
    assert G.sha(tj) == d["labels_sha256"]
    lab, _t, _c = O.ispan(g, threads=4)
    assert G.sha(lab) == d["labels_sha256"]
    assert int(np.count_nonzero(lab == np.arange(g.V))) == d["n_scc"]
    assert int(np.bincount(lab).max()) == d["giant"]
    assert d["ref"] == "agrees"


def test_generator_properties():
    g = O.rmat(14, 16, 3)
    for beg, adj in ((g.fw_beg, g.fw_adj), (g.bw_beg, g.bw_adj)):
        assert beg[0] == 0 and beg[-1] == g.E and np.all(np.diff(beg) >= 0)
        src = np.repeat(np.arange(g.V), np.diff(beg))
        assert not np.any(src == adj)  # no self-loops
        keys = src.astype(np.int64) * g.V + adj
        assert np.all(np.diff(keys) > 0)  # rows sorted, no duplicates
    # transpose consistency
    fk = np.sort(np.repeat(np.arange(g.V), np.diff(g.fw_beg)).astype(np.int64) * g.V + g.fw_adj)
    bk = np.sort(g.bw_adj.astype(np.int64) * g.V + np.repeat(np.arange(g.V), np.diff(g.bw_beg)))
    assert np.array_equal(fk, bk)
    # R-MAT skew: vertex 0 has the largest out-degree
    assert int(np.argmax(np.diff(g.fw_beg))) == 0
    # lattice: L = 2n(n-1) lattice edges, 60% both ways, plus L/100 random
    n = 64
    V, s, _d = O.grid_edges(n, 0.6, 1)
    L = 2 * n * (n - 1)
    assert L + L // 100 <= s.shape[0] <= 2 * L + L // 100
    assert abs((s.shape[0] - L - L // 100) / L - 0.6) < 0.05


def test_non_power_of_two_rmat():
    V, s, d = O.rmat_edges(12, 8, 5, V=3000)
    assert s.max() < 3000 and d.max() < 3000
    g = O.build_csr(V, s, d)
    assert np.array_equal(O.ispan(g)[0], O.tarjan(g))


# ------------------------------------------------- the checks catch mistakes

def test_planted_mistakes_are_caught():
    """A wrong labelling of a real graph must fail the comparisons used by
    the parity tests (label-equality, canonical form, scipy cross-check)."""
    g = O.rmat(12, 8, 11)
    tj = O.tarjan(g)
    sizes = np.bincount(tj)
    pair = np.nonzero(sizes == 2)[0]
    big = int(np.argmax(sizes))
    bad = tj.copy()
    if pair.size:  # split a 2-SCC
        members = np.nonzero(tj == pair[0])[0]
        bad[members[1]] = members[1]
    else:  # split the giant
        members = np.nonzero(tj == big)[0]
        bad[members[-1]] = members[-1]
    assert not np.array_equal(bad, scipy_labels(g))
    merged = tj.copy()  # merge two SCCs
    singles = np.nonzero(sizes == 1)[0]
    merged[singles[1]] = singles[0]
    assert not np.array_equal(merged, tj)
    assert not is_canonical(np.where(tj == big, big + 1, tj).clip(max=g.V - 1))
    # an unassigned raw reference label (0) is rejected by the canonicaliser
    raw = np.full(g.V, -1, np.int32)
    raw[5] = 0
    with pytest.raises(ValueError):
        O.canon_ref_raw(g, raw)


def test_pivot_int32_product_changes_work_not_labels():
    """H3: at R-MAT scale 21 vertex 0's out*in exceeds 2^31; the reference's
    int32 product picks another pivot, the labels stay the same."""
    g = O.rmat(21, 16, 1)
    od, idg = np.diff(g.fw_beg).astype(np.int64), np.diff(g.bw_beg).astype(np.int64)
    assert od[0] * idg[0] > 2 ** 31
    lab64, _t, c64 = O.ispan(g, threads=8)
    lab32, _t, c32 = O.ispan(g, threads=8, pivot_wrap32=True)
    assert c64["pivot"] == 0 and c32["pivot"] != 0
    assert np.array_equal(lab64, lab32)
    assert np.array_equal(lab64, O.tarjan(g))


@pytest.mark.skipif(not O.have_ref(), reason="compiled reference not present")
def test_compiled_reference_c1():
    """C1 (R-MAT s16, seed 1): the compiled reference, the restatement and
    Tarjan give one partition."""
    d = dict(G.hash_cases())["rmat-16-16-seed1"]
    g = G.build_hashed(d)
    raw, out = O.run_ref(g, threads=4, run_times=2)
    assert G.sha(O.canon_ref_raw(g, raw)) == d["labels_sha256"]
    avg = O.parse_ref_average(out)
    assert set(avg) >= {"Trim", "Elephant SCC", "Mice SCC", "Total time", "FW BFS", "BW BFS"}


@pytest.mark.timeout(300, method="thread")  # a spin inside the C call ends the run instead of hanging it
def test_wcc_path_compression_terminates():
    """Path compression in the restatement's WCC stops at a root and where
    another thread has raised the chain past its target: with the walk left
    to climb from there, long chains of small SCCs could spin forever at a
    root above the target under contention."""
    for groups in (500, 600, 800, 3000):
        g = G.cycle_chain(groups, groups)
        tj = O.tarjan(g)
        for threads in (2, 4, 8):
            for _ in range(3):
                lab, _t, c = O.ispan(g, threads=threads)
                assert np.array_equal(lab, tj), (groups, threads)
                assert c["n_wcc"] >= 1
