"""GPU tests of the continuous-batching server: every request's labels are the
same solved alone and in a mix of concurrent requests, and equal Tarjan's or
a single solve's; batching happens; the HBM budget gates registration."""
import threading

import numpy as np
import pytest

import ispan_amd as I
import oracle as O
from tests import graphs as G

pytestmark = pytest.mark.gpu


def host(g):
    return I.HostGraph(g.fw_beg, g.fw_adj, g.bw_beg, g.bw_adj)


@pytest.fixture(scope="module")
def pool():
    graphs = [O.rmat(s, 16, 40 + s) for s in (6, 9, 11, 12, 13)]
    graphs += [G.cycles_graph(5, n_cycles=60), G.random_graph(3, vmax=800), O.grid(40, 0.6, 2)]
    graphs += [G.pattern_graph(d) for _, d in G.pattern_cases()[:3]]
    graphs += [O.build_csr(3, np.zeros(0, np.int32), np.zeros(0, np.int32))]
    return graphs, [O.tarjan(g) for g in graphs]


def test_alone_and_in_mix(pool):
    graphs, ref = pool
    with I.Server(0) as srv:
        ids = [srv.add_graph(host(g)) for g in graphs]
        alone = []
        for gid in ids:  # one at a time: every batch holds one request
            lab, info = srv.wait(srv.submit(gid))
            assert info.status == 0 and info.batch_graphs == 1
            alone.append(lab)
        for lab, r in zip(alone, ref):
            assert np.array_equal(lab, r)
        # four client threads, each submitting the whole pool three times
        results, errors = [], []

        def client(seed):
            try:
                order = np.random.default_rng(seed).permutation(np.repeat(np.arange(len(ids)), 3))
                rids = [(int(k), srv.submit(ids[int(k)])) for k in order]
                for k, rid in rids:
                    lab, info = srv.wait(rid)
                    results.append((k, lab, info.batch_graphs))
            except Exception as e:  # noqa: BLE001 - reported below
                errors.append(e)

        ts = [threading.Thread(target=client, args=(s,)) for s in range(4)]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        assert not errors
        assert len(results) == 4 * 3 * len(ids)
        for k, lab, _ in results:
            assert np.array_equal(lab, alone[k])
        assert max(b for _, _, b in results) > 1  # requests were batched together
        st = srv.stats()
        assert st.requests_done == len(ids) + len(results)
        assert st.batches < st.requests_done
        assert st.p99_ms >= st.p50_ms > 0


def test_device_graphs_match_single_solves(solver):
    dgs = [solver.gen_rmat(V, ef, seed) for V, ef, seed in ((5000, 8.0, 1), (1 << 16, 16.0, 2), (30000, 4.0, 3))]
    try:
        with I.Server(0) as srv:
            ids = [srv.add_graph(dg) for dg in dgs]
            rids = [srv.submit(gid) for gid in ids for _ in range(2)]
            labs = [srv.wait(r)[0] for r in rids]
        for i, dg in enumerate(dgs):
            solver.bind(dg)
            solver.solve(I.default_params())
            single = solver.labels()
            assert np.array_equal(labs[2 * i], single) and np.array_equal(labs[2 * i + 1], single)
    finally:
        for dg in dgs:
            dg.free()


def test_budget_admission_and_removal():
    small = O.rmat(8, 8, 1)
    big = O.rmat(14, 16, 1)
    with I.Server(0, hbm_budget=1 << 20) as srv:
        gid = srv.add_graph(host(small))
        with pytest.raises(I.IspanError) as e:
            srv.add_graph(host(big))
        assert e.value.code == -9
        lab, _ = srv.wait(srv.submit(gid))
        assert np.array_equal(lab, O.tarjan(small))
        before = srv.stats().bytes_resident
        srv.remove_graph(gid)
        assert srv.stats().bytes_resident < before
        with pytest.raises(I.IspanError) as e:
            srv.submit(gid)
        assert e.value.code == -11


def test_remove_with_requests_in_flight(pool):
    graphs, ref = pool
    with I.Server(0) as srv:
        gid = srv.add_graph(host(graphs[3]))
        rids = [srv.submit(gid) for _ in range(5)]
        srv.remove_graph(gid)  # memory goes once these finish
        for rid in rids:
            lab, info = srv.wait(rid)
            assert np.array_equal(lab, ref[3])
        assert srv.stats().graphs_resident == 0


def test_bad_inputs():
    with I.Server(0) as srv:
        g = O.rmat(7, 4, 2)
        adj = g.fw_adj.copy()
        adj[0] = len(g.fw_beg) + 5
        with pytest.raises(I.IspanError) as e:
            srv.add_graph(I.HostGraph(g.fw_beg, adj, g.bw_beg, g.bw_adj))
        assert e.value.code == -2
        with pytest.raises(I.IspanError) as e:
            srv.submit(12345)
        assert e.value.code == -11
        with pytest.raises(I.IspanError) as e:
            srv.wait(777)
        assert e.value.code == -11


def test_labels_in_device_memory(pool):
    from ispan_amd import hiprt
    graphs, ref = pool
    with I.Server(0) as srv:
        ids = [srv.add_graph(host(g)) for g in graphs[:5]]
        rids = [srv.submit(gid) for gid in ids]
        for k, rid in enumerate(rids):
            ptr, info = srv.wait_device(rid)
            V = len(graphs[k].fw_beg) - 1
            out = np.empty(V, np.int32)
            hiprt.rt().hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2)
            srv.release(rid)
            assert np.array_equal(out, ref[k])
        with pytest.raises(I.IspanError) as e:
            srv.release(rids[0])
        assert e.value.code == -11


def test_arena_grows_while_serving(pool):
    graphs, ref = pool
    with I.Server(0) as srv:
        first = srv.add_graph(host(graphs[0]))
        stop, errors, done = [False], [], []

        def client():
            try:
                while not stop[0]:
                    lab, _ = srv.wait(srv.submit(first))
                    done.append(np.array_equal(lab, ref[0]))
            except Exception as e:  # noqa: BLE001 - reported below
                errors.append(e)

        t = threading.Thread(target=client)
        t.start()
        try:
            ids = [srv.add_graph(host(g)) for g in graphs[1:]]  # each one may grow the arena
        finally:
            stop[0] = True
            t.join()
        assert not errors and done and all(done)
        for gid, r in zip(ids, ref[1:]):
            lab, _ = srv.wait(srv.submit(gid))
            assert np.array_equal(lab, r)


def test_arena_space_reused_after_removal(pool):
    graphs, ref = pool
    with I.Server(0) as srv:
        a = srv.add_graph(host(graphs[2]))
        b = srv.add_graph(host(graphs[4]))
        used = srv.stats().bytes_device
        srv.remove_graph(b)
        c = srv.add_graph(host(graphs[3]))  # smaller than b: fits where b was
        assert srv.stats().bytes_device == used
        for gid, r in ((a, ref[2]), (c, ref[3])):
            lab, _ = srv.wait(srv.submit(gid))
            assert np.array_equal(lab, r)


def test_submit_many_shares_one_batch(pool):
    graphs, ref = pool
    with I.Server(0) as srv:
        ids = [srv.add_graph(host(g)) for g in graphs]
        rids = srv.submit_many(ids + ids[:3])  # the last three repeat graphs of the batch
        for k, rid in enumerate(rids):
            lab, info = srv.wait(rid)
            assert info.batch_graphs == len(rids)
            assert np.array_equal(lab, ref[k % len(ids)])
        with pytest.raises(I.IspanError) as e:
            srv.submit_many([ids[0], 999])
        assert e.value.code == -11
        assert srv.stats().requests_done == len(rids)


def test_served_scales_interleaved(solver):
    """R-MAT scales 12..16 (EF16) resident in one server; two client threads
    submit interleaved requests (single submits and submit_many bursts);
    every result equals Tarjan's labels of its graph."""
    dgs = [solver.gen_rmat(1 << s, 16.0, 500 + 10 * s + k) for s in range(12, 17) for k in range(2)]
    refs = []
    try:
        for dg in dgs:
            hg = dg.download()
            refs.append(O.tarjan(O.Graph(hg.fw_beg, hg.fw_adj, hg.bw_beg, hg.bw_adj)))
        with I.Server(0) as srv:
            ids = [srv.add_graph(dg) for dg in dgs]
            errors, checked = [], []

            def client(seed):
                try:
                    rng = np.random.default_rng(seed)
                    for _ in range(6):
                        picks = rng.integers(0, len(ids), int(rng.integers(1, 6)))
                        if rng.random() < 0.5:
                            rids = srv.submit_many([ids[int(k)] for k in picks])
                        else:
                            rids = [srv.submit(ids[int(k)]) for k in picks]
                        for k, rid in zip(picks, rids):
                            lab, info = srv.wait(rid)
                            assert info.status == 0
                            checked.append(bool(np.array_equal(lab, refs[int(k)])))
                except Exception as e:  # noqa: BLE001 - reported below
                    errors.append(e)

            ts = [threading.Thread(target=client, args=(s,)) for s in (1, 2)]
            for t in ts:
                t.start()
            for t in ts:
                t.join()
            assert not errors, errors
            assert checked and all(checked)
    finally:
        for dg in dgs:
            dg.free()


def test_free_memory_caps_admission(solver):
    """A configured budget far above the device's memory admits only what the
    device can still hold: with all but ~1.5 GB of HBM taken by another
    allocation, a graph whose arena and solve state need ~2 GB is rejected
    (ISPAN_ERR_BUDGET) although the 16 TiB budget would take it; once that
    memory is released, the same graph is admitted and solved."""
    from ispan_amd import hiprt
    dg = solver.gen_rmat(1 << 23, 16.0, 9)
    hog = []
    try:
        _, total = hiprt.mem_info()
        with I.Server(0, hbm_budget=1 << 44) as srv:
            assert srv.stats().bytes_budget <= total
            while True:  # take memory in pieces until ~1.5 GB is left
                free, _ = hiprt.mem_info()
                if free <= (1536 << 20):
                    break
                hog.append(hiprt.malloc(min(free - (1536 << 20), 8 << 30)))
            with pytest.raises(I.IspanError) as e:
                srv.add_graph(dg)
            assert e.value.code == -9
            while hog:
                hiprt.free(hog.pop())
            gid = srv.add_graph(dg)
            lab, info = srv.wait(srv.submit(gid))
            assert info.status == 0
        solver.bind(dg)
        solver.solve(I.default_params())
        assert np.array_equal(lab, solver.labels())
    finally:
        while hog:
            hiprt.free(hog.pop())
        dg.free()


def test_server_with_tuning_table(pool):
    """Batches solved by the table's parameters for their largest graph's
    size class give the same labels."""
    graphs, ref = pool
    t = I.TuningTable()
    for k in range(I.SIZE_CLASSES):
        t.tuned[k] = 1
        t.params[k] = I.default_params(alpha=2 + k % 3, beta=[50, 5000][k % 2], wcc_mode=k % 3, trim_scope=k % 2)
    with I.Server(0) as srv:
        srv.set_tuning(t)
        ids = [srv.add_graph(host(g)) for g in graphs]
        rids = srv.submit_many(ids) + [srv.submit(gid) for gid in ids]
        for k, rid in enumerate(rids):
            lab, info = srv.wait(rid)
            assert info.status == 0 and np.array_equal(lab, ref[k % len(ids)])
        srv.set_tuning(None)
        lab, _ = srv.wait(srv.submit(ids[3]))
        assert np.array_equal(lab, ref[3])
