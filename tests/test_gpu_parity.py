"""GPU parity: canonical labels from the HIP pipeline (through the C ABI) must
equal the oracle's -- Tarjan's labelling, the golden fixtures, and the
sha256 of Tarjan's labels for the large seeded graphs.  Integer work: the
bar is bit-exact equality."""
import ctypes as C
import functools
import os
import subprocess

import numpy as np
import pytest

import ispan_amd as I
import oracle as O
from tests import graphs as G
from ispan_amd import hiprt as H

pytestmark = pytest.mark.gpu


def solve(solver, g, **kw):
    solver.upload(I.HostGraph(g.fw_beg, g.fw_adj, g.bw_beg, g.bw_adj))
    st = solver.solve(I.default_params(**kw))
    return solver.labels(), st


def check(solver, g, **kw):
    lab, st = solve(solver, g, **kw)
    tj = O.tarjan(g)
    bad = np.nonzero(lab != tj)[0]
    assert bad.size == 0, f"{bad.size} mismatches, first {bad[:8]}: gpu {lab[bad[:8]]} tarjan {tj[bad[:8]]}"
    return st


# ------------------------------------------------------------ fixtures

def test_toy(solver):
    g, d = G.toy_graph()
    lab, st = solve(solver, g)
    assert lab.tolist() == d["canonical"]
    assert st.giant_size == 5 and st.trim2 == 2 and st.pivot == 1


@pytest.mark.parametrize("name,d", G.pattern_cases(), ids=[n for n, _ in G.pattern_cases()])
@pytest.mark.parametrize("wcc_mode,trim_scope", [(0, 0), (1, 0), (2, 0), (0, 1), (2, 1)])
def test_pattern_graphs(solver, name, d, wcc_mode, trim_scope):
    g = G.pattern_graph(d)
    lab, st = solve(solver, g, wcc_mode=wcc_mode, trim_scope=trim_scope)
    assert lab.tolist() == d["canonical"]
    if name.startswith("pair"):
        assert st.trim2 == 2
    if name.startswith(("cycle3", "hub3")):
        assert st.trim3 == 3


def long_list_patterns(seed: int):
    """A dense strongly connected core (the giant) and size-2/3 pattern groups
    whose vertices have long lists once the core is removed: 40 out-edges into
    the core, or 40 in-edges from sources that trim-1-first removes, plus
    duplicate edges and self-loops.  Every pattern vertex keeps out*in below
    the core's, so the pivot is in the core.  Returns the graph and the
    numbers of pair and triple vertices."""
    rng = np.random.default_rng(seed)
    n_core = 2000
    edges = [(i, (i + 1) % n_core) for i in range(n_core)]
    edges += [(int(u), int(v)) for u, v in rng.integers(0, n_core, size=(n_core * 20, 2)) if u != v]
    nxt = n_core
    pairs = triples = 0

    def tails(x, long_in):
        nonlocal nxt
        if long_in:  # 40 sources into x, one edge from x into the core
            for _ in range(40):
                edges.append((nxt, x))
                nxt += 1
            edges.append((x, int(rng.integers(0, n_core))))
        else:  # 40 edges from x into the core, some repeated
            for w in rng.integers(0, n_core, size=40):
                edges.append((x, int(w)))
            edges.append((x, int(rng.integers(0, n_core))))
        edges.append((x, x))

    for k in range(30):
        long_in = k % 2 == 1
        kind = k % 3
        if kind == 0:  # pair a <-> b
            a_, b_ = nxt, nxt + 1
            nxt += 2
            edges += [(a_, b_), (b_, a_), (a_, b_)]
            for x in (a_, b_):
                tails(x, long_in)
            pairs += 2
        elif kind == 1:  # cycle a -> b -> c -> a
            a_, b_, c_ = nxt, nxt + 1, nxt + 2
            nxt += 3
            edges += [(a_, b_), (b_, c_), (c_, a_)]
            for x in (a_, b_, c_):
                tails(x, long_in)
            triples += 3
        else:  # hub a <-> b, a <-> c
            a_, b_, c_ = nxt, nxt + 1, nxt + 2
            nxt += 3
            edges += [(a_, b_), (b_, a_), (a_, c_), (c_, a_)]
            for x in (a_, b_, c_):
                tails(x, long_in)
            triples += 3
    return O.graph_from_edges(nxt, edges), pairs, triples


@pytest.mark.parametrize("wcc_mode,trim_scope", [(0, 0), (2, 0), (0, 1)])
@pytest.mark.parametrize("seed", [1, 2])
def test_patterns_with_long_lists(solver, seed, wcc_mode, trim_scope):
    """Pattern vertices with more than 32 entries in a list are checked by a
    whole wave; the detections and labels are those of the per-lane check."""
    g, pairs, triples = long_list_patterns(seed)
    st = check(solver, g, wcc_mode=wcc_mode, trim_scope=trim_scope)
    assert st.giant_size == 2000
    assert st.trim2 == pairs and st.trim3 == triples


def core_and_chain(groups: int, seed: int):
    """A dense strongly connected core (the giant) and, apart from it, a
    chain of small cycles (tests.graphs.cycle_chain) whose last vertex points
    into the core: after the giant the rest is one WCC of small SCCs."""
    rng = np.random.default_rng(seed)
    n_core = 3000
    core = np.concatenate([np.stack([np.arange(n_core), (np.arange(n_core) + 1) % n_core], 1),
                           rng.integers(0, n_core, size=(n_core * 20, 2))])
    ch = G.cycle_chain(groups, seed)
    src = np.repeat(np.arange(ch.V, dtype=np.int64), np.diff(ch.fw_beg)) + n_core
    dst = ch.fw_adj.astype(np.int64) + n_core
    s_all = np.concatenate([core[:, 0], src, [n_core + ch.V - 1]]).astype(np.int32)
    d_all = np.concatenate([core[:, 1], dst, [0]]).astype(np.int32)
    return O.build_csr(n_core + ch.V, s_all, d_all)


@pytest.mark.parametrize("wcc_mode", [0, 1])
def test_long_wcc_goes_to_colouring(solver, wcc_mode):
    """A WCC of more than ISPAN_MICE_MAX_WCC vertices left after the trims is
    labelled by the colouring method instead of FW-BW per WCC (whose searches
    would each run down the rest of the chain); labels are Tarjan's."""
    g = core_and_chain(2000, 3)
    st = check(solver, g, wcc_mode=wcc_mode)
    assert st.giant_size == 3000
    assert st.max_wcc > 1024 and st.coloured > 0 and st.mice == 0


def test_small_wccs_stay_in_the_mice_phase(solver):
    g = G.cycles_graph(5, n_cycles=400)
    st = check(solver, g)
    if st.n_wcc > 0:
        assert st.max_wcc <= 1024 and st.coloured == 0 and st.mice > 0


# --------------------------------------------------------- random graphs

PARAMS = [dict(), dict(wcc_mode=1), dict(pivot_wrap32=1), dict(alpha=1, beta=1, gamma=0),
          dict(alpha=1000000, beta=2, gamma=1), dict(alpha=2, beta=1000000, gamma=100, async_batch=1),
          dict(theta=0.9, trim_times=0), dict(theta=0.0, trim_times=5), dict(wcc_mode=2), dict(trim_scope=1),
          dict(wcc_mode=2, trim_scope=1, theta=0.9, trim_times=0), dict(wcc_mode=1, engine=1), dict(switch_rule=0),
          dict(wcc_mode=1, trim_times=0), dict(alpha=1, beta=1, gamma=0, td_local=4),
          dict(alpha=1, beta=1, gamma=0, td_local=1000), dict(td_local=16)]


@pytest.mark.parametrize("seed", range(96))
def test_random_graphs(solver, seed):
    g = G.random_graph(seed, vmax=[50, 400, 3000][seed % 3], density=[0.6, 1.5, 4.0][(seed // 3) % 3])
    check(solver, g, **PARAMS[seed % len(PARAMS)])


@pytest.mark.parametrize("seed", range(24))
def test_cycle_forests(solver, seed):
    g = G.cycles_graph(seed, n_cycles=20 + 40 * seed)
    st = check(solver, g, **PARAMS[seed % len(PARAMS)])
    assert st.n_wcc >= 0


@pytest.mark.parametrize("scale,ef,seed", [(8, 16, 1), (12, 4, 2), (14, 2, 3), (16, 16, 4), (18, 8, 5)])
def test_rmat_vs_oracle_restatement(solver, scale, ef, seed):
    g = O.rmat(scale, ef, seed)
    lab, st = solve(solver, g)
    olab, _t, oc = O.ispan(g, threads=4)
    assert np.array_equal(lab, olab)
    assert st.fq0 == oc["fq0"] and st.pivot == oc["pivot"] and st.giant_size == oc["giant_size"]
    assert st.fw_size == oc["fw_size"]


@pytest.mark.parametrize("n,p,seed", [(16, 0.6, 1), (64, 0.3, 2), (256, 0.6, 3), (256, 0.1, 4), (700, 0.5, 5)])
def test_grids(solver, n, p, seed):
    check(solver, O.grid(n, p, seed), wcc_mode=seed % 2)


# ----------------------------------------------------------- edge cases

def test_empty_and_edgeless(solver):
    empty = O.Graph(np.zeros(1, np.int32), np.zeros(0, np.int32), np.zeros(1, np.int32), np.zeros(0, np.int32))
    lab, st = solve(solver, empty)
    assert lab.shape == (0,)
    g = O.build_csr(9, np.zeros(0, np.int32), np.zeros(0, np.int32))
    lab, st = solve(solver, g)
    assert lab.tolist() == list(range(9)) and st.pivot == -1


def test_duplicates_self_loops_unsorted_rows(solver):
    """The reference converter keeps duplicates, self-loops and input order
    (graph_converter/txt_to_bin_fw_int.cpp:146-181); the pipeline must give
    the same labels as on the cleaned graph."""
    rng = np.random.default_rng(3)
    V, m = 500, 2500
    e = rng.integers(0, V, size=(m, 2)).astype(np.int32)
    e = np.concatenate([e, e[:400], np.stack([np.arange(50), np.arange(50)], 1).astype(np.int32)])
    order = np.argsort(e[:, 0], kind="stable")
    src, dst = e[order, 0], e[order, 1]
    fw_beg = np.zeros(V + 1, np.int32)
    np.add.at(fw_beg, src + 1, 1)
    fw_beg = np.cumsum(fw_beg).astype(np.int32)
    o2 = np.argsort(dst, kind="stable")
    bw_beg = np.zeros(V + 1, np.int32)
    np.add.at(bw_beg, dst[o2] + 1, 1)
    bw_beg = np.cumsum(bw_beg).astype(np.int32)
    raw = O.Graph(fw_beg, dst.astype(np.int32), bw_beg, src[o2].astype(np.int32))
    clean = O.build_csr(V, e[:, 0].copy(), e[:, 1].copy())
    lab, _st = solve(solver, raw)
    assert np.array_equal(lab, O.tarjan(clean))


def test_repeat_and_regrow(solver):
    big, small = O.rmat(16, 16, 9), O.rmat(10, 8, 9)
    first, _ = solve(solver, big)
    for _ in range(3):
        st = solver.solve(I.default_params())
        assert np.array_equal(solver.labels(), first)
    lab_s, _ = solve(solver, small)
    assert np.array_equal(lab_s, O.tarjan(small))
    lab_b, _ = solve(solver, big)
    assert np.array_equal(lab_b, first)


def test_error_codes():
    with I.Solver(0) as s:
        with pytest.raises(I.IspanError) as e:
            s.solve()
        assert e.value.code == -8  # no graph bound
        g, _d = G.toy_graph()
        bad = g.fw_adj.copy()
        bad[0] = 99
        with pytest.raises(I.IspanError) as e:
            s.upload(I.HostGraph(g.fw_beg, bad, g.bw_beg, g.bw_adj))
        assert e.value.code == -2
        s.upload(I.HostGraph(g.fw_beg, g.fw_adj, g.bw_beg, g.bw_adj))
        with pytest.raises(I.IspanError) as e:
            s.solve(I.default_params(alpha=0))
        assert e.value.code == -1
        assert s.solve().giant_size == 5


def test_one_shot_entry_point():
    g = O.rmat(12, 16, 4)
    lab = I.scc_detection(I.HostGraph(g.fw_beg, g.fw_adj, g.bw_beg, g.bw_adj), 30, 200, 10, 0.1, 56)
    assert np.array_equal(lab, O.tarjan(g))


def test_labels_device_copy(solver):
    g = O.rmat(12, 8, 6)
    solver.upload(I.HostGraph(g.fw_beg, g.fw_adj, g.bw_beg, g.bw_adj))
    out = H.DeviceArray(np.full(g.V, -7, np.int32))
    try:
        solver.solve(I.default_params(), d_labels=out.ptr)
        solver.synchronize()
        assert np.array_equal(out.download(), O.tarjan(g))
    finally:
        out.free()
    assert solver.labels_device_ptr() != 0


def test_external_stream(solver):
    g = O.rmat(13, 8, 8)
    st = H.Stream()
    solver.set_stream(st.handle)
    try:
        lab, _ = solve(solver, g)
        st.synchronize()
    finally:
        solver.set_stream(None)
        st.destroy()
    assert np.array_equal(lab, O.tarjan(g))


def test_bind_caller_device_csr(solver):
    """Solve on a CSR that lives in caller-allocated device memory."""
    g = O.rmat(14, 8, 9)
    bufs = [H.DeviceArray(np.asarray(a, np.int32)) for a in (g.fw_beg, g.fw_adj, g.bw_beg, g.bw_adj)]
    try:
        solver.bind_device(g.V, g.E, *[b.ptr for b in bufs])
        solver.solve(I.default_params())
        assert np.array_equal(solver.labels(), O.tarjan(g))
    finally:
        for b in bufs:
            b.free()
    solver.upload(I.HostGraph(g.fw_beg, g.fw_adj, g.bw_beg, g.bw_adj))  # drop the freed pointers


def test_stats_consistency(solver):
    g = O.rmat(16, 16, 1)
    lab, st = solve(solver, g, timing=2)
    solver.solve_stats(st)
    sizes = np.bincount(lab)
    assert st.giant_size == sizes.max()
    assert st.n_scc == np.count_nonzero(sizes)
    assert st.fw_size >= st.giant_size
    fw_edges_lb = int(np.diff(g.fw_beg)[lab == lab[st.pivot]].sum())
    assert st.fw_edges >= fw_edges_lb and st.bw_edges > 0
    kp = solver.kernel_profile()
    assert kp.launches[3] == 1 and kp.found[3] == st.fq0
    # every vertex each direction reaches is claimed once, but the two roots;
    # the backward direction reaches the giant and what reaches it
    assert st.bw_reach >= st.giant_size
    assert sum(kp.found[:3]) == st.fw_size + st.bw_reach - 2


@functools.lru_cache(maxsize=1)
def uniform_digraph(V, E, seed):
    rng = np.random.default_rng(seed)
    return O.graph_from_edges(V, np.stack([rng.integers(0, V, E), rng.integers(0, V, E)], axis=1))


@pytest.mark.parametrize("engine", [0, 1])
def test_large_top_down_frontiers(solver, engine):
    # uniform random digraph, 10^6 vertices, out-degree 4: with alpha = 1 the
    # forward frontier grows 1 -> 11 -> ... -> 42 K -> 148 K -> 355 K -> 333 K
    # top-down before the switch, so its levels stage queue tiles of 64, 256
    # and 1024 entries (td_local = 0: every level synchronous)
    g = uniform_digraph(1_000_000, 4_000_000, 11)
    st = check(solver, g, engine=engine, alpha=1, td_local=0)
    assert st.fw_levels >= 10 and st.giant_size > 900_000
    # with in-place expansion (engine 0; engine 1 ignores td_local) the large
    # levels are walked by the waves that found them: fewer levels, the same
    # labels.  A bottom-up level's finds serve as parents within that level,
    # so level counts of two solves can differ by one: engine 1 compares none.
    st2 = check(solver, g, engine=engine, alpha=1)
    assert st2.giant_size == st.giant_size
    if engine == 0:
        assert st2.fw_levels < st.fw_levels


@pytest.mark.parametrize("engine", [0, 1])
@pytest.mark.parametrize("params", [dict(alpha=1), dict()], ids=["topdown", "default"])
def test_hub_split_at_a_later_level(solver, engine, params):
    # the pivot 0 reaches the hub 1 (and 1000 others) at level 0; the hub's
    # list (20 000 vertices that point back to 0, then 100 sinks) is split
    # into several work items, the last one partial, and expanded at level 1;
    # alpha = 1 keeps the forward traversal top-down
    A, B, S = range(2, 1002), range(1002, 21002), range(21002, 21102)
    edges = [(0, 1)] + [(0, a) for a in A] + [(a, 0) for a in A]
    edges += [(1, b) for b in B] + [(b, 0) for b in B] + [(1, s) for s in S]
    g = O.graph_from_edges(21102, edges)
    st = check(solver, g, engine=engine, timing=2, **params)
    assert st.pivot == 0 and st.giant_size == 21002 and st.fw_size == 21102 - 100
    if params:
        solver.solve_stats(st)
        kp = solver.kernel_profile()
        # forward alone claims the hub and A at level 0, then B from the hub's
        # items, all top-down (the backward traversal may switch to bottom-up)
        assert kp.found[0] + kp.found[1] >= 1001 + 20000
        assert sum(kp.found[:3]) == st.fw_size + st.bw_reach - 2


# ------------------------------------------------ device graph producers

@pytest.mark.parametrize("key,d", G.hash_cases(), ids=[k for k, _ in G.hash_cases()])
def test_hashed_graphs_on_device(solver, key, d):
    """Device generator == oracle generator (CSR sha256) and device labels ==
    Tarjan's (label sha256), up to the full-size configurations."""
    dg = solver.gen_rmat(1 << d["size"], float(d["param"]), d["seed"]) if d["kind"] == "rmat" else \
        solver.gen_grid(d["size"], float(d["param"]), d["seed"])
    try:
        assert (dg.V, dg.E) == (d["V"], d["E"])
        hg = dg.download()
        assert G.sha(hg.fw_beg, hg.fw_adj, hg.bw_beg, hg.bw_adj) == d["csr_sha256"]
        solver.bind(dg)
        st = solver.solve(I.default_params())
        lab = solver.labels()
        assert G.sha(lab) == d["labels_sha256"]
        assert st.giant_size == d["giant"]
        v = np.arange(dg.V)
        assert np.all(lab <= v) and np.all(lab[lab] == lab)  # canonical form
        assert int(np.count_nonzero(lab == v)) == d["n_scc"]
    finally:
        dg.free()


def test_build_csr_device_matches_oracle(solver):
    rng = np.random.default_rng(12)
    V, m = 5000, 40000
    e = rng.integers(0, V, size=(m, 2)).astype(np.int32)
    src = H.DeviceArray(e[:, 0].copy())
    dst = H.DeviceArray(e[:, 1].copy())
    dg = solver.build_csr(V, m, src.ptr, dst.ptr)
    try:
        hg = dg.download()
        og = O.build_csr(V, e[:, 0].copy(), e[:, 1].copy())
        for a, b in zip((hg.fw_beg, hg.fw_adj, hg.bw_beg, hg.bw_adj), (og.fw_beg, og.fw_adj, og.bw_beg, og.bw_adj)):
            assert np.array_equal(a, b)
    finally:
        dg.free()
    e[7, 1] = V  # out of range
    dst.upload(e[:, 1].copy())
    try:
        with pytest.raises(I.IspanError) as ex:
            solver.build_csr(V, m, src.ptr, dst.ptr)
        assert ex.value.code == -2
    finally:
        src.free()
        dst.free()


# ----------------------------------------------------------- drop-in CLI

def test_cli_drop_in(tmp_path):
    """The `ispan` executable takes the reference argv and prints its
    grammar (src/scc_core.cpp:1066, src/load_print.cpp:58,69-90)."""
    d = dict(G.hash_cases())["rmat-16-16-seed1"]
    g = G.build_hashed(d)
    paths = O.write_ispan_files(g, str(tmp_path))
    out = tmp_path / "labels.bin"
    r = subprocess.run([I.CLI_PATH, *paths, "56", "30", "200", "10", "0.1", "3"], capture_output=True, text=True,
                       env=dict(os.environ, ISPAN_LABELS_OUT=str(out)), timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.startswith("args = 11\nThread = 56, alpha = 30, beta = 200, gamma = 10, theta = 0.1, run_times = 3")
    assert r.stdout.count("\nRuntime: ") == 3 and r.stdout.count("\nResult:\n") == 3
    assert f"largest, {d['giant']}" in r.stdout and f"total, {d['n_scc']}" in r.stdout
    avg = O.parse_ref_average(r.stdout)
    assert set(avg) >= {"Trim", "Elephant SCC", "Mice SCC", "Total time", "FW BFS", "BW BFS", "Wcc", "Mice fw-bw"}
    assert G.sha(np.fromfile(out, np.int32)) == d["labels_sha256"]


# ------------------------------------------------ text edge-list converter

def test_convert_toy_matches_reference_files(solver, tmp_path):
    """Faithful conversion of graph_converter/toy_edge_list.txt reproduces the
    reference converter's shipped fw_begin.bin, fw_adjacent.bin, head.bin and
    out_degree.bin byte for byte."""
    g, d = G.toy_graph()
    txt = tmp_path / "toy_edge_list.txt"
    txt.write_text(d["edge_list_text"])
    V, E = solver.convert_edge_list(str(txt), str(tmp_path), clean=False)
    assert (V, E) == (7, 16)
    rd = lambda n: np.fromfile(tmp_path / n, np.int32).tolist()
    assert rd("fw_begin.bin") == d["fw_begin"] and rd("fw_adjacent.bin") == d["fw_adjacent"]
    assert rd("head.bin") == d["head"] and rd("out_degree.bin") == d["out_degree"]
    assert rd("bw_begin.bin") == d["bw_begin"] and rd("bw_adjacent.bin") == d["bw_adjacent"]
    assert rd("in_degree.bin") == np.diff(d["bw_begin"]).tolist()
    hg = I.graph_load(*[str(tmp_path / n) for n in ("fw_begin.bin", "fw_adjacent.bin", "bw_begin.bin",
                                                      "bw_adjacent.bin")])
    solver.upload(hg)
    solver.solve()
    assert solver.labels().tolist() == d["canonical"]


def test_convert_order_duplicates_and_clean(solver, tmp_path):
    rng = np.random.default_rng(5)
    V0, m = 300, 2000
    e = rng.integers(0, V0, size=(m, 2)) + 17  # ids shifted by the smallest id
    e = np.concatenate([e, e[:100], np.array([[20, 20], [21, 21]])])
    txt = tmp_path / "edges.txt"
    txt.write_text("% header\n" + "\n".join(f"{a} {b}" for a, b in e) + "\n")
    lo = int(e.min())
    src, dst = (e[:, 0] - lo).astype(np.int32), (e[:, 1] - lo).astype(np.int32)
    V = int(e.max()) - lo + 1
    raw_dir, clean_dir = tmp_path / "raw", tmp_path / "clean"
    raw_dir.mkdir()
    clean_dir.mkdir()
    Vr, Er = solver.convert_edge_list(str(txt), str(raw_dir), clean=False)
    assert (Vr, Er) == (V, e.shape[0])
    fb = np.fromfile(raw_dir / "fw_begin.bin", np.int32)
    fa = np.fromfile(raw_dir / "fw_adjacent.bin", np.int32)
    order = np.argsort(src, kind="stable")  # rows in input order
    assert np.array_equal(fa, dst[order])
    assert np.array_equal(np.diff(fb), np.bincount(src, minlength=V))
    bb = np.fromfile(raw_dir / "bw_begin.bin", np.int32)
    ba = np.fromfile(raw_dir / "bw_adjacent.bin", np.int32)
    assert np.array_equal(ba, src[np.argsort(dst, kind="stable")])
    Vc, Ec = solver.convert_edge_list(str(txt), str(clean_dir), clean=True)
    og = O.build_csr(V, src, dst)
    assert Ec == og.E
    for name, ref in (("fw_begin.bin", og.fw_beg), ("fw_adjacent.bin", og.fw_adj), ("bw_begin.bin", og.bw_beg),
                      ("bw_adjacent.bin", og.bw_adj)):
        assert np.array_equal(np.fromfile(clean_dir / name, np.int32), ref), name
    # the raw (duplicated, self-looped, unsorted) files solve to the same labels
    hg = I.graph_load(*[str(raw_dir / n) for n in ("fw_begin.bin", "fw_adjacent.bin", "bw_begin.bin",
                                                     "bw_adjacent.bin")])
    solver.upload(hg)
    solver.solve()
    assert np.array_equal(solver.labels(), O.tarjan(og))
    bad = tmp_path / "bad.txt"
    bad.write_text("1 2\n3\n")
    with pytest.raises(I.IspanError) as ex:
        solver.convert_edge_list(str(bad), str(tmp_path))
    assert ex.value.code == -7


# ---------------------------------------------- both giant-phase engines

@pytest.mark.parametrize("engine", [0, 1])
@pytest.mark.parametrize("seed", range(16))
def test_engines_random(solver, engine, seed):
    g = G.random_graph(500 + seed, vmax=[60, 800, 5000][seed % 3], density=[0.7, 2.0, 5.0][seed % 3])
    check(solver, g, **dict(PARAMS[seed % len(PARAMS)], engine=engine))


@pytest.mark.parametrize("engine", [0, 1])
def test_engines_structured(solver, engine):
    for g in (O.rmat(17, 16, 2), O.rmat(15, 3, 4), O.grid(300, 0.5, 6), G.cycles_graph(3, n_cycles=400)):
        check(solver, g, engine=engine)
    for name, d in G.pattern_cases():
        lab, _st = solve(solver, G.pattern_graph(d), engine=engine)
        assert lab.tolist() == d["canonical"], name


def test_autotune_picks_a_candidate_and_labels_do_not_change(solver):
    g = O.grid(64, 0.6, 3)
    solver.upload(I.HostGraph(g.fw_beg, g.fw_adj, g.bw_beg, g.bw_adj))
    cands = I.tuning_grid(alphas=(2, 30), betas=(50, 200))
    best, ms = solver.autotune(cands, reps=2)
    assert best in cands and len(ms) == 4 and min(ms) > 0
    ref = O.tarjan(g)
    for p in cands:
        solver.solve(p)
        assert np.array_equal(solver.labels(), ref)


# ------------------------------------------------------ SCC size histogram

def size_bins(lab: np.ndarray, n_bins: int = 64) -> np.ndarray:
    """Log2 bins of the SCC sizes of canonical labels (numpy, from the labels)."""
    sizes = np.bincount(lab)
    sizes = sizes[sizes > 0]
    k = np.floor(np.log2(sizes)).astype(np.int64)
    return np.bincount(np.minimum(k, n_bins - 1), minlength=n_bins)


@pytest.mark.parametrize("name", ["toy", "rmat-16-16-seed1", "grid-512-0.6-seed1", "cycles", "pair_sink"])
def test_scc_histogram(solver, name):
    """Device histogram == bins of Tarjan's SCC sizes; its summary ==
    ispan_summarize of the same labels."""
    if name == "toy":
        g = G.toy_graph()[0]
    elif name == "cycles":
        g = G.cycles_graph(3, n_cycles=300)
    elif name == "pair_sink":
        g = G.pattern_graph(dict(G.pattern_cases())[name])
    else:
        g = G.build_hashed(dict(G.hash_cases())[name])
    lab, _ = solve(solver, g)
    tj = O.tarjan(g)
    assert np.array_equal(lab, tj)
    bins, summ = solver.scc_histogram(64)
    assert np.array_equal(bins, size_bins(tj))
    assert summ == I.get_scc_result(tj)
    assert int(bins.sum()) == summ["total"] == int(np.count_nonzero(tj == np.arange(g.V)))
    b8, _ = solver.scc_histogram(8)  # the last bin takes every larger size
    assert np.array_equal(b8, size_bins(tj, 8))


def test_cli_histogram(tmp_path):
    """`--histogram` adds one line per non-empty size range after each
    result block; the result block itself is unchanged."""
    d = dict(G.hash_cases())["rmat-16-16-seed1"]
    g = G.build_hashed(d)
    paths = O.write_ispan_files(g, str(tmp_path))
    r = subprocess.run([I.CLI_PATH, *paths, "8", "30", "200", "10", "0.1", "2", "--histogram"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.count("\nSCC size histogram:\n") == 2
    assert f"largest, {d['giant']}" in r.stdout and f"total, {d['n_scc']}" in r.stdout
    block = r.stdout.split("\nSCC size histogram:\n")[1].split("\n\n")[0]
    got = {}
    for line in block.strip().splitlines():
        key, n = line.split(", ")
        lo = int(key.split()[1].split("-")[0])
        got[int(np.log2(lo))] = int(n)
    want = size_bins(O.tarjan(g))
    assert got == {k: int(n) for k, n in enumerate(want) if n}
    bad = subprocess.run([I.CLI_PATH, *paths, "8", "30", "200", "10", "0.1", "2", "--bogus"], capture_output=True,
                         text=True, timeout=60)
    assert bad.returncode == 255 and "Usage" in bad.stdout


# ------------------------------------- switch rule and pivot product width

def closed_form_pivot(g, wrap32: bool) -> int:
    """Lowest id maximising out*in over vertices with both degrees > 0
    (src/scc_common.h:92-150); wrap32 takes the product modulo 2^32 as a
    signed int32, as the reference's index_t product does (:117)."""
    od, idg = np.diff(g.fw_beg).astype(np.int64), np.diff(g.bw_beg).astype(np.int64)
    p = od * idg
    if wrap32:
        p = ((p + 2 ** 31) % 2 ** 32) - 2 ** 31
    p = np.where((od > 0) & (idg > 0), p, 0)
    return int(np.argmax(p)) if p.max() > 0 else -1


@pytest.mark.parametrize("wrap32", [0, 1])
@pytest.mark.parametrize("engine", [0, 1])
def test_pivot_wrap32_rmat21(solver, wrap32, engine):
    """H3 at R-MAT s21: vertex 0's out*in passes 2^31, so the int32 product
    picks another pivot.  The device pivot equals the closed form (and the
    oracle restatement's) either way; the labels equal Tarjan's either way."""
    g = O.rmat(21, 16, 1)
    lab, st = solve(solver, g, pivot_wrap32=wrap32, engine=engine)
    want = closed_form_pivot(g, bool(wrap32))
    assert (want == 0) == (not wrap32)
    assert st.pivot == want
    assert O.ispan(g, threads=8, pivot_wrap32=bool(wrap32))[2]["pivot"] == want
    assert np.array_equal(lab, O.tarjan(g))


@pytest.mark.parametrize("key", ["rmat-20-16-seed1", "grid-1024-0.6-seed1", "rmat-18-8-seed3"])
def test_reference_switch_estimate(solver, key):
    """switch_rule = 0 (the reference's frontier estimate, src/fw_bw.h:578-581)
    on the cooperative engine: same labels as the golden hash."""
    d = dict(G.hash_cases())[key]
    g = G.build_hashed(d)
    lab, st = solve(solver, g, switch_rule=0)
    assert G.sha(lab) == d["labels_sha256"]
    assert st.engine_used == 0 and st.giant_size == d["giant"]



# ----------------------------------------------- small-SCC strategy arms

@pytest.mark.parametrize("key", ["rmat-20-16-seed1", "rmat-18-8-seed3", "grid-1024-0.6-seed1", "grid-4096-0.6-seed1"])
@pytest.mark.parametrize("wcc_mode,trim_scope", [(1, 0), (2, 0), (2, 1), (0, 1)])
def test_small_scc_arms_hashed(solver, key, wcc_mode, trim_scope):
    """Every small-SCC strategy on the cooperative engine gives the golden
    labels: label-propagation WCC then mice (1), the colouring SCC method (2),
    and the full-graph pattern walk (trim_scope 1).  The colouring method
    leaves nothing to the host-driven mice phase: at most one more launch
    and host wait after the giant phase's."""
    d = dict(G.hash_cases())[key]
    dg = solver.gen_rmat(1 << d["size"], float(d["param"]), d["seed"]) if d["kind"] == "rmat" else \
        solver.gen_grid(d["size"], float(d["param"]), d["seed"])
    try:
        solver.bind(dg)
        st = solver.solve(I.default_params(wcc_mode=wcc_mode, trim_scope=trim_scope))
        assert G.sha(solver.labels()) == d["labels_sha256"]
        assert st.engine_used == 0
        if wcc_mode == 2:
            assert st.host_syncs <= 2 and st.mice == 0
        elif st.mice > 0:
            assert st.n_wcc > 0
    finally:
        dg.free()


@pytest.mark.parametrize("key", ["rmat-20-16-seed1", "rmat-24-16-seed1", "grid-1024-0.6-seed1", "grid-4096-0.6-seed1"])
@pytest.mark.parametrize("td_local", [1, 16, 64, 1000])
def test_local_expansion_hashed(solver, key, td_local):
    """Top-down to exhaustion with waves expanding their own finds in place
    (td_local steps) reaches the same sets: golden labels, and the same FW
    reach and giant sizes as the level-synchronous traversal."""
    d = dict(G.hash_cases())[key]
    dg = solver.gen_rmat(1 << d["size"], float(d["param"]), d["seed"]) if d["kind"] == "rmat" else \
        solver.gen_grid(d["size"], float(d["param"]), d["seed"])
    try:
        solver.bind(dg)
        s0 = solver.solve(I.default_params(td_local=0))  # every level synchronous
        st = solver.solve(I.default_params(td_local=td_local))
        assert G.sha(solver.labels()) == d["labels_sha256"]
        assert (st.fw_size, st.giant_size) == (s0.fw_size, s0.giant_size)
        assert st.fw_levels <= s0.fw_levels and st.bw_levels <= s0.bw_levels
    finally:
        dg.free()
