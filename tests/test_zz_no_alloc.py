"""The allocation-free entry point ispan_scc_detection_noalloc against the
allocating one-shot ispan_scc_detection on the same graphs: labels must be
bit-exact equal (and equal to Tarjan's), call after call on the same
device array, the second call on a smaller graph than the first."""
import ctypes as C

import numpy as np
import pytest

import ispan_amd as I
import oracle as O
from tests import graphs as G
from ispan_amd import hiprt as H

pytestmark = pytest.mark.gpu

SCRATCH_ERR = -12


def one_shot(g, **kw):
    """The existing entry point: host CSR in, host labels out."""
    lab = np.empty(g.V, np.int32)
    st = I.Stats()
    rc = I.lib().ispan_scc_detection(g.V, g.E, g.fw_beg.ctypes.data, g.fw_adj.ctypes.data, g.bw_beg.ctypes.data,
                                     g.bw_adj.ctypes.data, C.byref(I.default_params(**kw)), lab.ctypes.data,
                                     C.byref(st))
    assert rc == 0, I.ERRORS.get(rc, rc)
    return lab, st


def no_alloc(g, stream=None, **kw):
    """The allocation-free entry point: device CSR in, device labels out."""
    arrs = [H.DeviceArray(np.ascontiguousarray(a, np.int32)) for a in (g.fw_beg, g.fw_adj, g.bw_beg, g.bw_adj)]
    out = H.DeviceArray(np.full(max(g.V, 1), -7, np.int32))
    try:
        st = I.Stats()
        rc = I.lib().ispan_scc_detection_noalloc(g.V, g.E, *(a.ptr for a in arrs), C.byref(I.default_params(**kw)),
                                                 out.ptr, stream, C.byref(st))
        lab = out.download()[:g.V]
    finally:
        for a in arrs + [out]:
            a.free()
    return rc, lab, st


def case(name):
    print(f"[no_alloc] {name}", flush=True)


def test_two_calls_in_a_row_second_smaller():
    big = O.rmat(16, 16, 1)
    small = G.cycles_graph(5, n_cycles=400)  # trims, patterns, WCC and the mice tail
    assert small.V < big.V
    for name, g in (("R-MAT s16", big), ("cycle forest", small)):
        case(name)
        rc, lab, st = no_alloc(g)
        assert rc == 0, I.ERRORS.get(rc, rc)
        ref, rst = one_shot(g)
        assert np.array_equal(lab, ref), f"{name}: {np.count_nonzero(lab != ref)} labels differ"
        assert np.array_equal(lab, O.tarjan(g)), name
        assert (st.pivot, st.fw_size, st.giant_size, st.mice) == (rst.pivot, rst.fw_size, rst.giant_size, rst.mice)


def test_stale_scratch_and_caller_stream():
    """Graphs of decreasing, then increasing size on one device array, on a
    caller-owned stream: nothing carries over from the previous call."""
    s = H.Stream()
    try:
        for seed, vmax in ((11, 3000), (12, 400), (13, 50), (14, 3000)):
            g = G.random_graph(seed, vmax=vmax, density=1.5)
            case(f"random seed {seed} V={g.V}")
            rc, lab, _ = no_alloc(g, stream=s.handle, wcc_mode=seed % 2)
            assert rc == 0, I.ERRORS.get(rc, rc)
            assert np.array_equal(lab, one_shot(g)[0])
    finally:
        s.destroy()


@pytest.mark.parametrize("scale,V", [(19, None), (20, 750_000)])
def test_largest_documented_size_fits(scale, V):
    """The sizes the header names: R-MAT scale 19 and 750,000 vertices at
    edge factor 16."""
    g = O.rmat(scale, 16, 2, V=V)
    case(f"R-MAT s{scale} V={g.V} E={g.E}")
    rc, lab, _ = no_alloc(g)
    assert rc == 0, I.ERRORS.get(rc, rc)
    assert np.array_equal(lab, one_shot(g)[0])


def test_rejects_graph_beyond_scratch():
    """2^21 vertices need more than the fixed array: rejected before any
    device work (the graph itself is valid: no edges)."""
    V = 1 << 21
    g = O.Graph(np.zeros(V + 1, np.int32), np.zeros(0, np.int32), np.zeros(V + 1, np.int32), np.zeros(0, np.int32))
    case(f"edgeless V={V}")
    rc, _, _ = no_alloc(g)
    assert rc == SCRATCH_ERR
    assert b"scratch" in I.lib().ispan_strerror(rc)
