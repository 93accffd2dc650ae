"""CPU tests of the C ABI: the library loads, exports every entry point the
header declares, and its host-only functions (file loading, CSR checks,
summaries, error codes) behave; no kernel is launched."""
import os
import re
import subprocess

import numpy as np
import pytest

import ispan_amd as I
import oracle as O
from tests import graphs as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ispan_hip.h")

pytestmark = pytest.mark.skipif(not I.library_available(), reason="libispan_hip.so not built")


def header_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ispan_[a-z0-9_]+)\s*\(", txt)))


def test_exports_match_header():
    declared = header_functions()
    assert len(declared) >= 20
    out = subprocess.run(["nm", "-D", "--defined-only", I.LIB_PATH], capture_output=True, text=True, check=True)
    exported = {line.split()[-1] for line in out.stdout.splitlines()}
    missing = [f for f in declared if f not in exported]
    assert not missing, missing
    assert sorted(I.EXPORTS) == declared
    L = I.lib()
    for f in declared:
        assert getattr(L, f) is not None


def test_library_is_gfx950():
    """The device code object embedded in the library targets gfx950."""
    blob = open(I.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob


def test_default_params_and_strerror():
    p = I.default_params()
    assert (p.alpha, p.beta, p.gamma, p.trim_times, p.validate) == (30, 200, 10, 3, 1)
    assert (p.switch_rule, p.td_local, p.wcc_mode, p.trim_scope, p.engine) == (1, 64, 0, 0, 0)
    assert abs(p.theta - 0.1) < 1e-12
    L = I.lib()
    for code, name in I.ERRORS.items():
        assert L.ispan_strerror(code)
    assert L.ispan_strerror(0) == b"ok"
    assert b"gfx950" in L.ispan_version()


def test_csr_check_host():
    g, _d = G.toy_graph()
    hg = I.HostGraph(g.fw_beg, g.fw_adj, g.bw_beg, g.bw_adj)
    assert I.csr_check(hg)
    bad = I.HostGraph(g.fw_beg, g.fw_adj.copy(), g.bw_beg, g.bw_adj)
    bad.fw_adj[3] = 7  # id == V
    assert not I.csr_check(bad)
    bad2 = I.HostGraph(g.fw_beg.copy(), g.fw_adj, g.bw_beg, g.bw_adj)
    bad2.fw_beg[2] = bad2.fw_beg[3] + 1  # non-monotone offsets
    assert not I.csr_check(bad2)


def test_summarize_matches_sizes():
    g = O.rmat(12, 8, 2)
    lab = O.tarjan(g)
    s = I.get_scc_result(lab)
    sizes = np.bincount(lab)
    sizes = sizes[sizes > 0]
    assert s["total"] == sizes.size
    assert s["largest"] == sizes.max()
    assert s["size_1"] == np.count_nonzero(sizes == 1)
    assert s["size_2"] == np.count_nonzero(sizes == 2)
    assert s["size_3"] == np.count_nonzero(sizes == 3)
    assert s["size_1"] + s["size_2"] + s["size_3"] + s["others"] + 1 == s["total"]
    with pytest.raises(I.IspanError):
        I.get_scc_result(np.array([0, 2, 1], np.int32))  # not canonical


def test_graph_load_roundtrip_and_errors(tmp_path):
    g, _d = G.toy_graph()
    paths = O.write_ispan_files(g, str(tmp_path))
    hg = I.graph_load(*paths)
    assert hg.V == 7 and hg.E == 16
    for a, b in zip((hg.fw_beg, hg.fw_adj, hg.bw_beg, hg.bw_adj), (g.fw_beg, g.fw_adj, g.bw_beg, g.bw_adj)):
        assert np.array_equal(a, b)
    with pytest.raises(I.IspanError) as e:
        I.graph_load(str(tmp_path / "missing.bin"), *paths[1:])
    assert e.value.code == -7
    short = tmp_path / "short.bin"
    np.asarray(g.bw_adj[:-1], np.int32).tofile(short)
    with pytest.raises(I.IspanError) as e:
        I.graph_load(paths[0], paths[1], paths[2], str(short))
    assert e.value.code == -7


def test_context_without_gpu_fails_cleanly():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    with pytest.raises(I.IspanError) as e:
        I.Solver(0)
    assert e.value.code == -1


def test_cli_usage_without_gpu():
    if not os.path.exists(I.CLI_PATH):
        pytest.skip("ispan CLI not built")
    r = subprocess.run([I.CLI_PATH, "a", "b"], capture_output=True, text=True)
    assert r.returncode == 255
    assert r.stdout.startswith("args = 3")
    assert "Usage:" in r.stdout


def test_tuning_table_file_roundtrip(tmp_path):
    """A per-size-class table survives save/load; lookup falls back to the
    defaults for untuned classes; a malformed file is rejected."""
    t = I.TuningTable()
    for k, (a, b, w) in {9: (4, 5000, 2), 14: (2, 50, 1), 20: (30, 200, 0)}.items():
        t.tuned[k] = 1
        t.params[k] = I.default_params(alpha=a, beta=b, wcc_mode=w, trim_scope=k % 2, theta=0.05)
        t.ms[k] = k / 3
    path = str(tmp_path / "table.txt")
    t.save(path)
    u = I.TuningTable.load(path)
    assert sorted(u.classes()) == [9, 14, 20]
    for k in (9, 14, 20):
        p, q = t.params[k], u.params[k]
        assert (p.alpha, p.beta, p.wcc_mode, p.trim_scope, p.theta) == (q.alpha, q.beta, q.wcc_mode, q.trim_scope,
                                                                      q.theta)
        assert abs(u.ms[k] - k / 3) < 1e-5
    assert u.lookup(700).alpha == 4 and u.lookup(1023).alpha == 4 and u.lookup(1024).alpha == 30
    assert u.lookup(1 << 14).beta == 50 and u.lookup(5).alpha == 30  # class 2: untuned, defaults
    bad = tmp_path / "bad.txt"
    bad.write_text("9 4 5000\n")
    with pytest.raises(I.IspanError) as e:
        I.TuningTable.load(str(bad))
    assert e.value.code == -7
