"""TEST INFRASTRUCTURE -- the oracle.  Not part of the product.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only to check results against (or as
the timed CPU baseline).  The product library never loads it.

ctypes wrappers over ``oracle/liboracle.so`` (built by ``oracle/Makefile``):

* ``rmat_edges`` / ``grid_edges`` -- seeded counter-hash generators
* ``build_csr``                   -- edge list -> iSpan int32 CSR pair
* ``tarjan``                      -- serial Tarjan, canonical labels
* ``ispan``                       -- OpenMP restatement of the iSpan phase
                                     sequence (src/scc_core.cpp:116-1070)
* ``run_ref`` / ``canon_ref_raw`` -- the compiled reference in ``oracle/_ref``
                                     (present only where it was built) and the
                                     canonicaliser for its raw ``scc_id``.

Canonical labels: ``label[v]`` = smallest vertex id in v's SCC.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import tempfile
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_BIN = os.path.join(HERE, "_ref", "ispan_ref")

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")


class _Params(C.Structure):
    _fields_ = [("alpha", C.c_int32), ("beta", C.c_int32), ("gamma", C.c_int32),
                ("theta", C.c_double), ("trim_times", C.c_int32), ("threads", C.c_int32),
                ("pivot_wrap32", C.c_int32)]


_lib = None


def build(force: bool = False) -> None:
    """Compile liboracle.so (and oracle/_ref when the reference sources exist)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE, "-B" if force else "all"], check=True)
    if os.path.isdir("/root/reference/src"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_rmat_edges.restype = C.c_int64
        L.orc_rmat_edges.argtypes = [C.c_int32, C.c_int, C.c_int64, C.c_double, C.c_double, C.c_double,
                                     C.c_uint64, _i32p, _i32p]
        L.orc_grid_edges.restype = C.c_int64
        L.orc_grid_edges.argtypes = [C.c_int32, C.c_double, C.c_uint64, _i32p, _i32p]
        L.orc_build_csr.restype = C.c_int64
        L.orc_build_csr.argtypes = [C.c_int32, C.c_int64, _i32p, _i32p, _i32p, _i32p, _i32p, _i32p]
        L.orc_tarjan.restype = C.c_int64
        L.orc_tarjan.argtypes = [C.c_int32, _i32p, _i32p, _i32p]
        L.orc_canon_ref_raw.restype = C.c_int
        L.orc_canon_ref_raw.argtypes = [C.c_int32, _i32p, _i32p, _i32p, _i32p]
        L.orc_ispan.restype = C.c_int
        L.orc_ispan.argtypes = [C.c_int32, _i32p, _i32p, _i32p, _i32p, C.POINTER(_Params), _i32p, _f64p, _i64p]
        L.orc_num_procs.restype = C.c_int
        _lib = L
    return _lib


@dataclass
class Graph:
    """iSpan CSR pair (src/graph.h:9-31): fw_beg[V+1], fw_adj[E], bw_beg[V+1], bw_adj[E]."""
    fw_beg: np.ndarray
    fw_adj: np.ndarray
    bw_beg: np.ndarray
    bw_adj: np.ndarray

    @property
    def V(self) -> int:
        return int(self.fw_beg.shape[0] - 1)

    @property
    def E(self) -> int:
        return int(self.fw_adj.shape[0])


RMAT_ABC = (0.57, 0.19, 0.19)


def rmat_edges(scale: int, edge_factor: float, seed: int, V: int | None = None, abc=RMAT_ABC):
    """R-MAT draws; V defaults to 2**scale, otherwise endpoints >= V are dropped."""
    if V is None:
        V = 1 << scale
    m = int(round(edge_factor * V))
    src = np.empty(m, np.int32)
    dst = np.empty(m, np.int32)
    k = lib().orc_rmat_edges(V, scale, m, abc[0], abc[1], abc[2], seed, src, dst)
    return V, src[:k], dst[:k]


def grid_edges(n: int, p_bidir: float, seed: int):
    L = 2 * n * (n - 1)
    cap = 2 * L + L // 100
    src = np.empty(cap, np.int32)
    dst = np.empty(cap, np.int32)
    k = lib().orc_grid_edges(n, p_bidir, seed, src, dst)
    return n * n, src[:k], dst[:k]


def build_csr(V: int, src: np.ndarray, dst: np.ndarray) -> Graph:
    src = np.ascontiguousarray(src, np.int32)
    dst = np.ascontiguousarray(dst, np.int32)
    m = src.shape[0]
    fb = np.empty(V + 1, np.int32)
    bb = np.empty(V + 1, np.int32)
    fa = np.empty(max(m, 1), np.int32)
    ba = np.empty(max(m, 1), np.int32)
    E = lib().orc_build_csr(V, m, src, dst, fb, fa, bb, ba)
    if E < 0:
        raise ValueError("edge endpoint outside [0, V)")
    return Graph(fb, fa[:E].copy(), bb, ba[:E].copy())


def graph_from_edges(V: int, edges) -> Graph:
    e = np.asarray(list(edges), np.int32).reshape(-1, 2)
    return build_csr(V, e[:, 0].copy(), e[:, 1].copy())


def rmat(scale: int, edge_factor: float = 16, seed: int = 1, V: int | None = None) -> Graph:
    V, s, d = rmat_edges(scale, edge_factor, seed, V)
    return build_csr(V, s, d)


def grid(n: int, p_bidir: float = 0.6, seed: int = 1) -> Graph:
    V, s, d = grid_edges(n, p_bidir, seed)
    return build_csr(V, s, d)


def tarjan(g: Graph) -> np.ndarray:
    lab = np.empty(g.V, np.int32)
    lib().orc_tarjan(g.V, g.fw_beg, _nz(g.fw_adj), lab)
    return lab


def _nz(a: np.ndarray) -> np.ndarray:
    return a if a.size else np.zeros(1, np.int32)


TIME_SLOTS = {0: "trim", 1: "elephant", 2: "mice", 3: "total", 4: "trim_size_1", 5: "trim_size_2",
              6: "pivot", 7: "fw_bfs", 8: "bw_bfs", 9: "wcc", 10: "mice_fw_bw", 13: "trim_size_3"}
COUNTER_SLOTS = ["fq0", "pivot", "fw_size", "giant_size", "trim2", "trim3", "n_wcc", "mice", "fw_levels",
                 "bw_levels"]


def ispan(g: Graph, alpha=30, beta=200, gamma=10, theta=0.1, trim_times=3, threads=0, pivot_wrap32=False):
    """CPU restatement of scc_detection (src/scc_core.cpp:27-1094) -> (labels, times, counters)."""
    p = _Params(int(alpha), int(beta), int(gamma), float(theta), int(trim_times), int(threads),
                int(bool(pivot_wrap32)))
    lab = np.empty(max(g.V, 1), np.int32)
    t = np.zeros(16, np.float64)
    c = np.zeros(16, np.int64)
    rc = lib().orc_ispan(g.V, g.fw_beg, _nz(g.fw_adj), g.bw_beg, _nz(g.bw_adj), C.byref(p), lab, t, c)
    if rc != 0:
        raise RuntimeError(f"orc_ispan failed rc={rc}")
    times = {name: float(t[i]) for i, name in TIME_SLOTS.items()}
    counters = {name: int(c[i]) for i, name in enumerate(COUNTER_SLOTS)}
    return lab[:g.V], times, counters


def canon_ref_raw(g: Graph, raw: np.ndarray) -> np.ndarray:
    lab = np.empty(g.V, np.int32)
    rc = lib().orc_canon_ref_raw(g.V, g.fw_beg, _nz(g.fw_adj), np.ascontiguousarray(raw, np.int32), lab)
    if rc != 0:
        raise ValueError("reference raw labels contain unassigned/unknown entries")
    return lab


def write_ispan_files(g: Graph, d: str) -> list[str]:
    """Write the four iSpan binaries (headerless int32; src/graph.cpp:13-14)."""
    names = ["fw_begin.bin", "fw_adjacent.bin", "bw_begin.bin", "bw_adjacent.bin"]
    paths = [os.path.join(d, n) for n in names]
    for p, a in zip(paths, [g.fw_beg, g.fw_adj, g.bw_beg, g.bw_adj]):
        np.ascontiguousarray(a, np.int32).tofile(p)
    return paths


def read_ispan_files(paths) -> Graph:
    arrs = [np.fromfile(p, dtype=np.int32) for p in paths]
    return Graph(*arrs)


def have_ref() -> bool:
    return os.path.exists(REF_BIN)


def run_ref(g: Graph, threads: int = 4, alpha=30, beta=200, gamma=10, theta=0.1, run_times=1,
            timeout: float = 120.0, workdir: str | None = None):
    """Run the compiled reference on g; returns (raw scc_id[V], stdout)."""
    if not have_ref():
        raise FileNotFoundError(REF_BIN)
    with tempfile.TemporaryDirectory(dir=workdir) as d:
        paths = write_ispan_files(g, d)
        out = os.path.join(d, "labels.bin")
        env = dict(os.environ, ISPAN_REF_LABELS=out, OMP_NUM_THREADS=str(threads))
        r = subprocess.run([REF_BIN, *paths, str(threads), str(alpha), str(beta), str(gamma), str(theta),
                            str(run_times)], env=env, capture_output=True, text=True, timeout=timeout)
        if r.returncode != 0:
            raise RuntimeError(f"reference exited {r.returncode}: {r.stdout[-2000:]} {r.stderr[-2000:]}")
        raw = np.fromfile(out, dtype=np.int32)
    return raw, r.stdout


def parse_ref_average(stdout: str) -> dict:
    """Parse the 'Average Time Consumption' block (src/load_print.cpp:69-90)."""
    res = {}
    seen = False
    for line in stdout.splitlines():
        if line.startswith("Average Time Consumption"):
            seen = True
            continue
        if seen and "," in line:
            k, v = line.rsplit(",", 1)
            try:
                res[k.strip()] = float(v)
            except ValueError:
                pass
    return res
