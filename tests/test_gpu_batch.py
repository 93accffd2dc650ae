"""GPU tests of batched solves: many graphs in one pass over their disjoint
union.  Every graph's labels must equal Tarjan's on that graph alone, and the
labels of a single solve of the same graph."""
import numpy as np
import pytest

import ispan_amd as I
import oracle as O
from tests import graphs as G

pytestmark = pytest.mark.gpu


def host(g):
    return I.HostGraph(g.fw_beg, g.fw_adj, g.bw_beg, g.bw_adj)


def mixed_graphs(seed: int, n: int):
    """Uneven sizes: random graphs (with self-loops and duplicates), cycle
    forests, small R-MATs, the pattern fixtures, empty and edge-less graphs."""
    rng = np.random.default_rng(seed)
    out = []
    pats = [G.pattern_graph(d) for _, d in G.pattern_cases()]
    for i in range(n):
        k = int(rng.integers(0, 6))
        if k == 0:
            out.append(G.random_graph(seed * 1000 + i, vmax=400))
        elif k == 1:
            out.append(G.cycles_graph(seed * 1000 + i, n_cycles=int(rng.integers(2, 40))))
        elif k == 2:
            out.append(O.rmat(int(rng.integers(4, 12)), float(rng.integers(2, 16)), seed * 1000 + i))
        elif k == 3:
            out.append(pats[i % len(pats)])
        elif k == 4:
            V = int(rng.integers(0, 5))
            out.append(O.build_csr(V, np.zeros(0, np.int32), np.zeros(0, np.int32)))
        else:
            out.append(O.grid(int(rng.integers(2, 24)), 0.6, seed * 1000 + i))
    return out


def solve_batch(solver, graphs, max_group_edges=0, **kw):
    b = I.DeviceBatch.from_host([host(g) for g in graphs])
    try:
        solver.batch_bind(b, max_group_edges)
    finally:
        b.free()  # the bound batch holds its own union copies
    st = solver.batch_solve(I.default_params(**kw))
    return solver.batch_labels(), st


@pytest.mark.parametrize("seed", range(4))
def test_batch_matches_tarjan(solver, seed):
    graphs = mixed_graphs(seed, 40)
    labs, st = solve_batch(solver, graphs)
    assert solver.batch_groups() == 1
    for g, lab in zip(graphs, labs):
        assert np.array_equal(lab, O.tarjan(g))
    assert st.engine_used == 0


def test_batch_split_into_groups(solver):
    graphs = mixed_graphs(7, 60)
    ref = [O.tarjan(g) for g in graphs]
    labs, _ = solve_batch(solver, graphs, max_group_edges=2000)
    assert solver.batch_groups() > 3
    for lab, r in zip(labs, ref):
        assert np.array_equal(lab, r)


def test_batch_equals_single_solves(solver):
    graphs = [O.rmat(s, ef, 10 + s) for s, ef in ((10, 4), (12, 16), (9, 8), (13, 6), (11, 12))]
    graphs += [G.pattern_graph(d) for _, d in G.pattern_cases()]
    graphs += [G.cycles_graph(3, n_cycles=50)]
    labs, st = solve_batch(solver, graphs)
    giants = 0
    for g, lab in zip(graphs, labs):
        solver.upload(host(g))
        s1 = solver.solve(I.default_params())
        assert np.array_equal(lab, solver.labels())
        giants += s1.giant_size
    assert st.giant_size == giants


def test_batch_resolve_is_stable(solver):
    graphs = mixed_graphs(11, 30)
    labs, _ = solve_batch(solver, graphs)
    st2 = solver.batch_solve(I.default_params(timing=1))
    labs2 = solver.batch_labels()
    for a, b in zip(labs, labs2):
        assert np.array_equal(a, b)
    assert st2.ms[3] > 0


def test_batch_from_device_graphs(solver):
    dgs = [solver.gen_rmat(V, ef, seed) for V, ef, seed in ((1000, 4.0, 1), (70000, 16.0, 2), (4096, 8.0, 3),
                                                            (123457, 11.0, 4), (1 << 15, 16.0, 5))]
    try:
        b = I.DeviceBatch.from_device(dgs)
        try:
            solver.batch_bind(b)
        finally:
            b.free()
        solver.batch_solve(I.default_params())
        labs = solver.batch_labels()
        for dg, lab in zip(dgs, labs):
            solver.bind(dg)
            solver.solve(I.default_params())
            assert np.array_equal(lab, solver.labels())
            v = np.arange(dg.V)
            assert np.all(lab <= v) and np.all(lab[lab] == lab)
    finally:
        for dg in dgs:
            dg.free()


def test_batch_rejects_bad_graphs(solver):
    good = O.rmat(8, 8, 1)
    bad = O.rmat(8, 8, 2)
    adj = bad.fw_adj.copy()
    adj[3] = len(bad.fw_beg) - 1  # one past the graph's last id: inside the union, outside the graph
    with pytest.raises(I.IspanError) as e:
        b = I.DeviceBatch.from_host([I.HostGraph(bad.fw_beg, adj, bad.bw_beg, bad.bw_adj), host(good)])
        try:
            solver.batch_bind(b)
        finally:
            b.free()
    assert e.value.code == -2
    beg = good.fw_beg.copy()
    beg[-1] += 1  # offsets end past E_g
    with pytest.raises(I.IspanError) as e:
        b = I.DeviceBatch.from_host([I.HostGraph(beg, good.fw_adj, good.bw_beg, good.bw_adj), host(bad)])
        try:
            solver.batch_bind(b)
        finally:
            b.free()
    assert e.value.code == -2
    with pytest.raises(I.IspanError):
        solver.batch_solve()  # nothing bound after a failed bind
    solver.upload(host(good))  # the context still solves single graphs
    solver.solve()
    assert np.array_equal(solver.labels(), O.tarjan(good))


@pytest.mark.parametrize("max_group_edges", [0, 60_000_000])
def test_batch_c4_distribution(solver, max_group_edges):
    """256 graphs drawn as the batched workload draws them: V log-uniform in
    [1e3, 1e6], edge factor uniform in [4, 16], (a,b,c,d) = (0.57,0.19,
    0.19,0.05), generated on the device; one group, then split into groups of
    at most 60M edges.  Every graph's labels equal Tarjan's on that graph."""
    rng = np.random.default_rng(2024)
    n = 256
    Vs = np.exp(rng.uniform(np.log(1e3), np.log(1e6), n)).astype(np.int64)
    efs = rng.uniform(4.0, 16.0, n)
    dgs = []
    try:
        for i in range(n):
            dgs.append(solver.gen_rmat(int(Vs[i]), float(efs[i]), 77_000_003 + i))
        b = I.DeviceBatch.from_device(dgs)
        try:
            solver.batch_bind(b, max_group_edges)
        finally:
            b.free()
        if max_group_edges:
            assert solver.batch_groups() >= 3
        st = solver.batch_solve(I.default_params())
        labs = solver.batch_labels()
        giants = 0
        for dg, lab in zip(dgs, labs):
            hg = dg.download()
            tj = O.tarjan(O.Graph(hg.fw_beg, hg.fw_adj, hg.bw_beg, hg.bw_adj))
            assert np.array_equal(lab, tj), f"V={dg.V} E={dg.E}"
        assert st.engine_used == 0 and st.giant_size > 0
    finally:
        for dg in dgs:
            dg.free()


@pytest.mark.parametrize("cap", [0, 3000])
def test_batch_by_class_and_tuned_solve(solver, cap):
    """Binding by size class packs graphs in class order (scattered graphs
    gathered into one union); per-class tuning fills the classes of the
    groups' largest graphs; the tuned solve and the by-class solve give every
    graph Tarjan's labels."""
    graphs = mixed_graphs(21, 60) + [O.rmat(s, 8, 900 + s) for s in (12, 14, 9, 14, 12)]
    ref = [O.tarjan(g) for g in graphs]
    b = I.DeviceBatch.from_host([host(g) for g in graphs])
    try:
        solver.batch_bind(b, cap, I.default_params(batch_by_class=1))
    finally:
        b.free()
    classes = {int(np.floor(np.log2(max(len(g.fw_beg) - 1, 1)))) for g in graphs}
    assert (solver.batch_groups() == 1) == (cap == 0)
    solver.batch_solve()
    for lab, r in zip(solver.batch_labels(), ref):
        assert np.array_equal(lab, r)
    cands = [I.default_params(), I.default_params(alpha=2, beta=5000), I.default_params(wcc_mode=2),
             I.default_params(wcc_mode=1, trim_scope=1)]
    table = solver.autotune_classes(cands, reps=1)
    tuned = set(table.classes())
    assert tuned <= classes and max(classes) in tuned and (len(tuned) == 1) == (cap == 0)
    st = solver.batch_solve_tuned(table)
    assert st.ms[3] > 0
    for lab, r in zip(solver.batch_labels(), ref):
        assert np.array_equal(lab, r)


def test_autotune_classes_single_graph(solver):
    g = O.rmat(13, 8, 3)
    solver.upload(host(g))
    prior = I.TuningTable()
    prior.tuned[5] = 1
    prior.params[5] = I.default_params(alpha=7)
    t = solver.autotune_classes([I.default_params(), I.default_params(wcc_mode=2, alpha=4)], reps=2, table=prior)
    assert set(t.classes()) == {5, 13} and t.params[5].alpha == 7 and t.ms[13] > 0
    solver.solve(t.lookup(g.V))
    assert np.array_equal(solver.labels(), O.tarjan(g))
