"""ispan_amd: MI355X (gfx950) strongly-connected-components pipeline of iSpan.

Python binding over the C ABI of ``ispan_amd/lib/libispan_hip.so``
(declared in ``include/ispan_hip.h``).  The GPU library is the only compute
path: importing works without it, but every call that would compute raises
``NativeLibraryMissing`` when it has not been built.

Mirrors the reference's C++ entry points (src/scc_common.h:7-33):

* ``graph_load(fw_beg, fw_adj, bw_beg, bw_adj)``  -> ``HostGraph``
* ``scc_detection(g, alpha, beta, gamma, theta, thread_count)`` -> labels
* ``get_scc_result(labels)`` -> SCC size summary
plus a ``Solver`` context for device-resident graphs and repeated solves.
Labels are canonical: ``label[v]`` = smallest vertex id in v's SCC.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libispan_hip.so")
CLI_PATH = os.path.join(PKG, "bin", "ispan")

ISPAN_OK = 0
ERRORS = {-1: "ISPAN_ERR_ARG", -2: "ISPAN_ERR_CSR", -3: "ISPAN_ERR_NOMEM", -4: "ISPAN_ERR_HIP",
          -5: "ISPAN_ERR_ITER", -6: "ISPAN_ERR_UNLABELLED", -7: "ISPAN_ERR_IO", -8: "ISPAN_ERR_NOGRAPH",
          -9: "ISPAN_ERR_BUDGET", -10: "ISPAN_ERR_TIMEOUT", -11: "ISPAN_ERR_NOTFOUND", -12: "ISPAN_ERR_SCRATCH"}

# Every symbol include/ispan_hip.h declares.
EXPORTS = [
    "ispan_default_params", "ispan_strerror", "ispan_ctx_create", "ispan_ctx_destroy", "ispan_ctx_set_stream",
    "ispan_ctx_synchronize", "ispan_ctx_last_error", "ispan_graph_upload", "ispan_graph_bind_device", "ispan_solve", "ispan_labels_download",
    "ispan_labels_device", "ispan_solve_stats", "ispan_scc_detection", "ispan_graph_load", "ispan_graph_free",
    "ispan_csr_check", "ispan_summarize", "ispan_gen_rmat_device", "ispan_gen_grid_device",
    "ispan_build_csr_device", "ispan_device_free", "ispan_version", "ispan_kernel_profile_get",
    "ispan_convert_edge_list", "ispan_autotune", "ispan_batch_bind_device", "ispan_batch_groups", "ispan_batch_solve",
    "ispan_batch_labels_download", "ispan_batch_labels_device", "ispan_server_create", "ispan_server_destroy",
    "ispan_server_add_graph", "ispan_server_add_graph_device", "ispan_server_remove_graph", "ispan_server_submit", "ispan_server_submit_many",
    "ispan_server_wait", "ispan_server_wait_device", "ispan_server_release", "ispan_server_stats_get",
    "ispan_server_reset_stats", "ispan_server_last_error", "ispan_scc_detection_noalloc",
    "ispan_scc_histogram", "ispan_autotune_classes", "ispan_batch_solve_tuned", "ispan_tuning_lookup",
    "ispan_tuning_save", "ispan_tuning_load", "ispan_server_set_tuning",
]


class NativeLibraryMissing(RuntimeError):
    pass


class IspanError(RuntimeError):
    def __init__(self, code: int, msg: str = ""):
        self.code = code
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")


class Params(C.Structure):
    """ispan_params (include/ispan_hip.h)."""
    _fields_ = [("alpha", C.c_int32), ("beta", C.c_int32), ("gamma", C.c_int32), ("theta", C.c_double),
                ("trim_times", C.c_int32), ("thread_count", C.c_int32), ("pivot_wrap32", C.c_int32),
                ("timing", C.c_int32), ("validate", C.c_int32), ("wcc_mode", C.c_int32),
                ("async_batch", C.c_int32), ("engine", C.c_int32), ("switch_rule", C.c_int32),
                ("trim_scope", C.c_int32), ("batch_by_class", C.c_int32), ("td_local", C.c_int32)]


SIZE_CLASSES = 32


class TuningTable(C.Structure):
    """ispan_tuning_table: tuned parameters per size class floor(log2 V)."""
    _fields_ = [("tuned", C.c_int32 * SIZE_CLASSES), ("params", Params * SIZE_CLASSES),
                ("ms", C.c_double * SIZE_CLASSES)]

    def classes(self) -> dict:
        return {k: self.params[k] for k in range(SIZE_CLASSES) if self.tuned[k]}

    def lookup(self, V: int) -> Params:
        p = Params()
        lib().ispan_tuning_lookup(C.byref(self), V, C.byref(p))
        return p

    def save(self, path: str):
        _check(lib().ispan_tuning_save(C.byref(self), path.encode()))

    @classmethod
    def load(cls, path: str) -> "TuningTable":
        t = cls()
        _check(lib().ispan_tuning_load(path.encode(), C.byref(t)))
        return t


class Stats(C.Structure):
    """ispan_stats (include/ispan_hip.h)."""
    _fields_ = [("ms", C.c_double * 16), ("fq0", C.c_int64), ("pivot", C.c_int64), ("fw_size", C.c_int64),
                ("giant_size", C.c_int64), ("fw_edges", C.c_int64), ("bw_edges", C.c_int64),
                ("trim1", C.c_int64), ("trim2", C.c_int64), ("trim3", C.c_int64), ("n_wcc", C.c_int64),
                ("mice", C.c_int64), ("fw_levels", C.c_int64), ("bw_levels", C.c_int64), ("n_scc", C.c_int64),
                ("host_syncs", C.c_int64), ("engine_used", C.c_int64), ("bw_reach", C.c_int64),
                ("max_wcc", C.c_int64), ("coloured", C.c_int64), ("reserved", C.c_int64 * 5)]

    def as_dict(self) -> dict:
        d = {k: getattr(self, k) for k, _ in self._fields_ if k not in ("ms", "reserved")}
        names = {0: "trim", 1: "elephant", 2: "mice_total", 3: "total", 4: "trim_size_1", 5: "trim_size_2",
                 6: "pivot", 7: "fw_bfs", 8: "bw_bfs", 9: "wcc", 10: "mice_fw_bw", 11: "trim_1_first",
                 13: "trim_size_3", 14: "trim_1_sweeps"}
        d["ms"] = {n: self.ms[i] for i, n in names.items()}
        return d


class ServerConfig(C.Structure):
    """ispan_server_config (include/ispan_hip.h)."""
    _fields_ = [("hbm_budget", C.c_int64), ("max_batch_graphs", C.c_int32), ("reserved0", C.c_int32),
                ("max_batch_vertices", C.c_int64), ("reserved", C.c_int64 * 5)]


class RequestInfo(C.Structure):
    _fields_ = [("submit_ms", C.c_double), ("start_ms", C.c_double), ("done_ms", C.c_double),
                ("latency_ms", C.c_double), ("batch", C.c_int64), ("batch_graphs", C.c_int32),
                ("status", C.c_int32)]


class ServerStats(C.Structure):
    _fields_ = [("requests_done", C.c_int64), ("batches", C.c_int64), ("graphs_resident", C.c_int64),
                ("bytes_resident", C.c_int64), ("bytes_budget", C.c_int64), ("bytes_device", C.c_int64),
                ("p50_ms", C.c_double), ("p99_ms", C.c_double), ("mean_ms", C.c_double), ("max_ms", C.c_double),
                ("mean_batch_graphs", C.c_double), ("reserved", C.c_int64 * 5)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class KernelProfile(C.Structure):
    """ispan_kernel_profile: classes 0 top-down, 1 hub items, 2 bottom-up, 3 trim-1-first."""
    _fields_ = [("launches", C.c_int64 * 4), ("ms", C.c_double * 4), ("edges", C.c_int64 * 4),
                ("found", C.c_int64 * 4)]
    CLASSES = ("k_td", "k_td_hub", "k_bu", "k_trim1_first")
    CLASSES_COOP = ("giant_topdown", "-", "giant_bottomup", "k_giant")

    def as_dict(self) -> dict:
        return {n: {"launches": self.launches[i], "ms": self.ms[i], "edges": self.edges[i], "found": self.found[i]}
                for i, n in enumerate(self.CLASSES)}


class HostGraphC(C.Structure):
    _fields_ = [("V", C.c_int32), ("E", C.c_int64), ("fw_beg", C.POINTER(C.c_int32)),
                ("fw_adj", C.POINTER(C.c_int32)), ("bw_beg", C.POINTER(C.c_int32)),
                ("bw_adj", C.POINTER(C.c_int32))]


class Summary(C.Structure):
    _fields_ = [("largest", C.c_int64), ("size_1", C.c_int64), ("size_2", C.c_int64), ("size_3", C.c_int64),
                ("others", C.c_int64), ("total", C.c_int64)]


_lib = None
_vp = C.c_void_p
_i32pp = C.POINTER(C.POINTER(C.c_int32))


def library_available() -> bool:
    return os.path.exists(LIB_PATH)


def lib():
    """Load libispan_hip.so (raises NativeLibraryMissing if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise NativeLibraryMissing(f"{LIB_PATH} not built; run `python -m ispan_amd.build`")
    L = C.CDLL(LIB_PATH)
    sig = {
        "ispan_default_params": (None, [C.POINTER(Params)]),
        "ispan_strerror": (C.c_char_p, [C.c_int]),
        "ispan_version": (C.c_char_p, []),
        "ispan_ctx_create": (C.c_int, [C.POINTER(_vp), C.c_int]),
        "ispan_ctx_destroy": (C.c_int, [_vp]),
        "ispan_ctx_set_stream": (C.c_int, [_vp, _vp]),
        "ispan_ctx_synchronize": (C.c_int, [_vp]),
        "ispan_ctx_last_error": (C.c_char_p, [_vp]),
        "ispan_graph_upload": (C.c_int, [_vp, C.c_int32, C.c_int64, _vp, _vp, _vp, _vp, C.POINTER(Params)]),
        "ispan_graph_bind_device": (C.c_int, [_vp, C.c_int32, C.c_int64, _vp, _vp, _vp, _vp, C.POINTER(Params)]),
        "ispan_solve": (C.c_int, [_vp, C.POINTER(Params), _vp, C.POINTER(Stats)]),
        "ispan_labels_download": (C.c_int, [_vp, _vp]),
        "ispan_labels_device": (_vp, [_vp]),
        "ispan_solve_stats": (C.c_int, [_vp, C.POINTER(Stats)]),
        "ispan_scc_detection": (C.c_int, [C.c_int32, C.c_int64, _vp, _vp, _vp, _vp, C.POINTER(Params), _vp,
                                          C.POINTER(Stats)]),
        "ispan_scc_histogram": (C.c_int, [_vp, _vp, C.c_int32, C.POINTER(Summary)]),
        "ispan_autotune_classes": (C.c_int, [_vp, C.POINTER(Params), C.c_int32, C.c_int32, C.POINTER(TuningTable)]),
        "ispan_batch_solve_tuned": (C.c_int, [_vp, C.POINTER(TuningTable), C.POINTER(Stats)]),
        "ispan_tuning_lookup": (None, [C.POINTER(TuningTable), C.c_int64, C.POINTER(Params)]),
        "ispan_tuning_save": (C.c_int, [C.POINTER(TuningTable), C.c_char_p]),
        "ispan_tuning_load": (C.c_int, [C.c_char_p, C.POINTER(TuningTable)]),
        "ispan_server_set_tuning": (C.c_int, [_vp, C.POINTER(TuningTable)]),
        "ispan_scc_detection_noalloc": (C.c_int, [C.c_int32, C.c_int64, _vp, _vp, _vp, _vp, C.POINTER(Params), _vp,
                                                  _vp, C.POINTER(Stats)]),
        "ispan_graph_load": (C.c_int, [C.c_char_p] * 4 + [C.POINTER(HostGraphC)]),
        "ispan_graph_free": (None, [C.POINTER(HostGraphC)]),
        "ispan_csr_check": (C.c_int, [C.c_int32, C.c_int64, _vp, _vp, _vp, _vp]),
        "ispan_summarize": (C.c_int, [C.c_int32, _vp, C.POINTER(Summary)]),
        "ispan_gen_rmat_device": (C.c_int, [_vp, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_double,
                                            C.c_uint64, _i32pp, _i32pp, _i32pp, _i32pp, C.POINTER(C.c_int64)]),
        "ispan_gen_grid_device": (C.c_int, [_vp, C.c_int32, C.c_double, C.c_uint64, _i32pp, _i32pp, _i32pp,
                                            _i32pp, C.POINTER(C.c_int64)]),
        "ispan_build_csr_device": (C.c_int, [_vp, C.c_int32, C.c_int64, _vp, _vp, _i32pp, _i32pp, _i32pp, _i32pp,
                                             C.POINTER(C.c_int64)]),
        "ispan_device_free": (None, [_vp]),
        "ispan_kernel_profile_get": (C.c_int, [_vp, C.POINTER(KernelProfile)]),
        "ispan_server_create": (C.c_int, [C.POINTER(_vp), C.c_int, C.POINTER(ServerConfig), C.POINTER(Params)]),
        "ispan_server_destroy": (C.c_int, [_vp]),
        "ispan_server_add_graph": (C.c_int, [_vp, C.c_int32, C.c_int64, _vp, _vp, _vp, _vp, C.POINTER(C.c_int64)]),
        "ispan_server_add_graph_device": (C.c_int, [_vp, C.c_int32, C.c_int64, _vp, _vp, _vp, _vp,
                                                    C.POINTER(C.c_int64)]),
        "ispan_server_remove_graph": (C.c_int, [_vp, C.c_int64]),
        "ispan_server_submit": (C.c_int, [_vp, C.c_int64, C.POINTER(C.c_int64)]),
        "ispan_server_wait": (C.c_int, [_vp, C.c_int64, _vp, C.POINTER(RequestInfo), C.c_int64]),
        "ispan_server_submit_many": (C.c_int, [_vp, C.c_int32, _vp, _vp]),
        "ispan_server_wait_device": (C.c_int, [_vp, C.c_int64, C.POINTER(_vp), C.POINTER(RequestInfo), C.c_int64]),
        "ispan_server_release": (C.c_int, [_vp, C.c_int64]),
        "ispan_server_stats_get": (C.c_int, [_vp, C.POINTER(ServerStats)]),
        "ispan_server_reset_stats": (C.c_int, [_vp]),
        "ispan_server_last_error": (C.c_char_p, [_vp]),
        "ispan_batch_bind_device": (C.c_int, [_vp, C.c_int32, _vp, _vp, _vp, _vp, _vp, _vp, C.c_int64,
                                              C.POINTER(Params)]),
        "ispan_batch_groups": (C.c_int, [_vp]),
        "ispan_autotune": (C.c_int, [_vp, _vp, C.c_int32, C.c_int32, C.POINTER(C.c_int32), _vp]),
        "ispan_batch_solve": (C.c_int, [_vp, C.POINTER(Params), C.POINTER(Stats)]),
        "ispan_batch_labels_download": (C.c_int, [_vp, _vp]),
        "ispan_batch_labels_device": (_vp, [_vp]),
        "ispan_convert_edge_list": (C.c_int, [_vp, C.c_char_p, C.c_char_p, C.c_int32, C.POINTER(C.c_int32),
                                              C.POINTER(C.c_int64)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def default_params(**kw) -> Params:
    p = Params()
    lib().ispan_default_params(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise TypeError(f"unknown parameter {k}")
        setattr(p, k, v)
    return p


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data if a.size else 0


def _check(rc: int, ctx=None):
    if rc != ISPAN_OK:
        msg = lib().ispan_ctx_last_error(ctx).decode() if ctx else lib().ispan_strerror(rc).decode()
        raise IspanError(rc, msg)


@dataclass
class HostGraph:
    """iSpan CSR pair in host memory (src/graph.h:9-31)."""
    fw_beg: np.ndarray
    fw_adj: np.ndarray
    bw_beg: np.ndarray
    bw_adj: np.ndarray

    @property
    def V(self) -> int:
        return int(self.fw_beg.shape[0]) - 1

    @property
    def E(self) -> int:
        return int(self.fw_adj.shape[0])


def graph_load(fw_beg_file: str, fw_adj_file: str, bw_beg_file: str, bw_adj_file: str) -> HostGraph:
    """Read the four iSpan binaries (graph_load, src/load_print.cpp:6-22)."""
    g = HostGraphC()
    _check(lib().ispan_graph_load(fw_beg_file.encode(), fw_adj_file.encode(), bw_beg_file.encode(),
                                  bw_adj_file.encode(), C.byref(g)))
    try:
        V, E = g.V, g.E
        out = HostGraph(np.ctypeslib.as_array(g.fw_beg, (V + 1,)).copy(),
                        np.ctypeslib.as_array(g.fw_adj, (max(E, 1),))[:E].copy(),
                        np.ctypeslib.as_array(g.bw_beg, (V + 1,)).copy(),
                        np.ctypeslib.as_array(g.bw_adj, (max(E, 1),))[:E].copy())
    finally:
        lib().ispan_graph_free(C.byref(g))
    return out


@dataclass
class DeviceGraph:
    """CSR pair produced on the device by the library; free() releases it."""
    V: int
    E: int
    fw_beg: int
    fw_adj: int
    bw_beg: int
    bw_adj: int

    def free(self):
        for p in (self.fw_beg, self.fw_adj, self.bw_beg, self.bw_adj):
            if p:
                lib().ispan_device_free(p)
        self.fw_beg = self.fw_adj = self.bw_beg = self.bw_adj = 0

    def download(self) -> HostGraph:
        out = []
        for p, n in ((self.fw_beg, self.V + 1), (self.fw_adj, self.E), (self.bw_beg, self.V + 1),
                     (self.bw_adj, self.E)):
            a = np.empty(n, np.int32)
            if n:
                _check(_hiprt().hipMemcpy(a.ctypes.data, p, a.nbytes, 2) and -4)  # hipMemcpyDeviceToHost
            out.append(a)
        return HostGraph(*out)


_hip = None


def _hiprt():
    """The HIP runtime, for plain device<->host copies of library-owned arrays."""
    global _hip
    if _hip is None:
        for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _hip = C.CDLL(name)
                break
            except OSError:
                continue
        if _hip is None:
            raise NativeLibraryMissing("libamdhip64.so not found")
        _hip.hipMemcpy.restype = C.c_int
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return _hip


class Solver:
    """One ispan_ctx: a device, a stream, device-resident graph and labels."""

    def __init__(self, device: int = 0):
        self._ctx = C.c_void_p()
        _check(lib().ispan_ctx_create(C.byref(self._ctx), device))
        self.V = -1
        self.E = 0
        self._keep = None

    def close(self):
        if self._ctx:
            lib().ispan_ctx_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def handle(self):
        return self._ctx

    def set_stream(self, stream_handle: int | None):
        _check(lib().ispan_ctx_set_stream(self._ctx, stream_handle or None), self._ctx)

    def synchronize(self):
        """Wait for all work queued by this context."""
        _check(lib().ispan_ctx_synchronize(self._ctx), self._ctx)

    def upload(self, g: HostGraph, params: Params | None = None):
        arrs = [np.ascontiguousarray(a, np.int32) for a in (g.fw_beg, g.fw_adj, g.bw_beg, g.bw_adj)]
        V, E = int(arrs[0].shape[0]) - 1, int(arrs[1].shape[0])
        _check(lib().ispan_graph_upload(self._ctx, V, E, *[_ptr(a) for a in arrs],
                                        C.byref(params) if params else None), self._ctx)
        self.V, self.E = V, E
        self._keep = None

    def bind_device(self, V: int, E: int, fw_beg: int, fw_adj: int, bw_beg: int, bw_adj: int,
                    params: Params | None = None, keep=None):
        _check(lib().ispan_graph_bind_device(self._ctx, V, E, fw_beg, fw_adj, bw_beg, bw_adj,
                                             C.byref(params) if params else None), self._ctx)
        self.V, self.E = V, E
        self._keep = keep

    def bind(self, dg: DeviceGraph, params: Params | None = None):
        self.bind_device(dg.V, dg.E, dg.fw_beg, dg.fw_adj, dg.bw_beg, dg.bw_adj, params, keep=dg)

    def solve(self, params: Params | None = None, d_labels: int = 0) -> Stats:
        st = Stats()
        _check(lib().ispan_solve(self._ctx, C.byref(params) if params else None, d_labels or None, C.byref(st)),
               self._ctx)
        return st

    def solve_stats(self, st: Stats) -> Stats:
        _check(lib().ispan_solve_stats(self._ctx, C.byref(st)), self._ctx)
        return st

    def kernel_profile(self) -> KernelProfile:
        kp = KernelProfile()
        _check(lib().ispan_kernel_profile_get(self._ctx, C.byref(kp)), self._ctx)
        return kp

    def labels(self) -> np.ndarray:
        out = np.empty(max(self.V, 0), np.int32)
        _check(lib().ispan_labels_download(self._ctx, _ptr(out) if out.size else None), self._ctx)
        return out

    def labels_device_ptr(self) -> int:
        return lib().ispan_labels_device(self._ctx) or 0

    def scc_histogram(self, n_bins: int = 64):
        """SCC-size histogram of the last solve (computed on the device):
        bins[k] counts SCCs of size in [2^k, 2^(k+1)), plus the summary dict
        of get_scc_result."""
        bins = np.zeros(n_bins, np.int64)
        sm = Summary()
        _check(lib().ispan_scc_histogram(self._ctx, _ptr(bins) if n_bins else None, n_bins, C.byref(sm)), self._ctx)
        return bins, {k: int(getattr(sm, k)) for k, _ in Summary._fields_}

    # ---- batched solves (many graphs, one pass over their disjoint union)
    def batch_bind(self, batch: "DeviceBatch", max_group_edges: int = 0, params: Params | None = None):
        v = np.ascontiguousarray(batch.v_off, np.int64)
        e = np.ascontiguousarray(batch.e_off, np.int64)
        _check(lib().ispan_batch_bind_device(self._ctx, len(v) - 1, _ptr(v), _ptr(e), batch.fw_beg, batch.fw_adj,
                                             batch.bw_beg, batch.bw_adj, max_group_edges,
                                             C.byref(params) if params else None), self._ctx)
        self._batch_v_off = v
        self.V = -1

    def batch_groups(self) -> int:
        return lib().ispan_batch_groups(self._ctx)

    def batch_solve(self, params: Params | None = None) -> Stats:
        st = Stats()
        _check(lib().ispan_batch_solve(self._ctx, C.byref(params) if params else None, C.byref(st)), self._ctx)
        return st

    def batch_labels(self) -> list[np.ndarray]:
        """Each graph's labels (smallest local id of the vertex's SCC)."""
        v = self._batch_v_off
        out = np.empty(max(int(v[-1]), 1), np.int32)
        _check(lib().ispan_batch_labels_download(self._ctx, _ptr(out)), self._ctx)
        return [out[v[g]:v[g + 1]] for g in range(len(v) - 1)]

    def autotune(self, candidates: list[Params], reps: int = 3) -> tuple[Params, list[float]]:
        """Time each candidate on the bound graph or batch; -> (fastest, median ms of each)."""
        arr = (Params * len(candidates))(*candidates)
        ms = np.zeros(len(candidates), np.float64)
        best = C.c_int32()
        _check(lib().ispan_autotune(self._ctx, arr, len(candidates), reps, C.byref(best), _ptr(ms)), self._ctx)
        return candidates[best.value], ms.tolist()

    def autotune_classes(self, candidates: list[Params], reps: int = 3,
                         table: "TuningTable | None" = None) -> "TuningTable":
        """Fastest candidate per size class of the bound graph or batch, into
        (a copy of) table."""
        t = TuningTable()
        if table is not None:
            C.memmove(C.byref(t), C.byref(table), C.sizeof(TuningTable))
        arr = (Params * len(candidates))(*candidates)
        _check(lib().ispan_autotune_classes(self._ctx, arr, len(candidates), reps, C.byref(t)), self._ctx)
        return t

    def batch_solve_tuned(self, table: "TuningTable") -> Stats:
        st = Stats()
        _check(lib().ispan_batch_solve_tuned(self._ctx, C.byref(table), C.byref(st)), self._ctx)
        return st

    def gen_rmat(self, V: int, edge_factor: float = 16.0, seed: int = 1, a=0.57, b=0.19, c=0.19) -> DeviceGraph:
        ptrs = [C.POINTER(C.c_int32)() for _ in range(4)]
        E = C.c_int64()
        _check(lib().ispan_gen_rmat_device(self._ctx, V, edge_factor, a, b, c, seed, *[C.byref(p) for p in ptrs],
                                           C.byref(E)), self._ctx)
        return DeviceGraph(V, E.value, *[C.cast(p, C.c_void_p).value or 0 for p in ptrs])

    def gen_grid(self, n: int, p_bidir: float = 0.6, seed: int = 1) -> DeviceGraph:
        ptrs = [C.POINTER(C.c_int32)() for _ in range(4)]
        E = C.c_int64()
        _check(lib().ispan_gen_grid_device(self._ctx, n, p_bidir, seed, *[C.byref(p) for p in ptrs], C.byref(E)),
               self._ctx)
        return DeviceGraph(n * n, E.value, *[C.cast(p, C.c_void_p).value or 0 for p in ptrs])

    def convert_edge_list(self, txt_path: str, out_dir: str, clean: bool = False) -> tuple[int, int]:
        """Text edge list -> the reference's binary files (graph_converter/)."""
        V, E = C.c_int32(), C.c_int64()
        _check(lib().ispan_convert_edge_list(self._ctx, txt_path.encode(), out_dir.encode(), int(bool(clean)),
                                             C.byref(V), C.byref(E)), self._ctx)
        return V.value, E.value

    def build_csr(self, V: int, m: int, d_src: int, d_dst: int) -> DeviceGraph:
        ptrs = [C.POINTER(C.c_int32)() for _ in range(4)]
        E = C.c_int64()
        _check(lib().ispan_build_csr_device(self._ctx, V, m, d_src, d_dst, *[C.byref(p) for p in ptrs],
                                            C.byref(E)), self._ctx)
        return DeviceGraph(V, E.value, *[C.cast(p, C.c_void_p).value or 0 for p in ptrs])


class Server:
    """Continuous-batching SCC server (ispan_server_*): graphs resident on the
    device, requests batched as they arrive."""

    def __init__(self, device: int = 0, hbm_budget: int = 0, max_batch_graphs: int = 0, max_batch_vertices: int = 0,
                 params: Params | None = None):
        cfg = ServerConfig(hbm_budget=hbm_budget, max_batch_graphs=max_batch_graphs,
                           max_batch_vertices=max_batch_vertices)
        h = C.c_void_p()
        _check(lib().ispan_server_create(C.byref(h), device, C.byref(cfg), C.byref(params) if params else None))
        self._srv = h.value
        self._V = {}      # graph id -> V
        self._rid_V = {}  # request id -> V of its graph

    def _err(self, rc: int):
        if rc != 0:
            raise IspanError(rc, (lib().ispan_server_last_error(self._srv) or b"").decode())

    def add_graph(self, g: "HostGraph | DeviceGraph") -> int:
        gid = C.c_int64()
        if isinstance(g, DeviceGraph):
            rc = lib().ispan_server_add_graph_device(self._srv, g.V, g.E, g.fw_beg, g.fw_adj, g.bw_beg, g.bw_adj,
                                                     C.byref(gid))
            V = g.V
        else:
            arrs = [np.ascontiguousarray(a, np.int32) for a in (g.fw_beg, g.fw_adj, g.bw_beg, g.bw_adj)]
            V, E = len(arrs[0]) - 1, len(arrs[1])
            rc = lib().ispan_server_add_graph(self._srv, V, E, *[_ptr(a) for a in arrs], C.byref(gid))
        self._err(rc)
        self._V[gid.value] = V
        return gid.value

    def remove_graph(self, gid: int):
        self._err(lib().ispan_server_remove_graph(self._srv, gid))

    def submit(self, gid: int) -> int:
        rid = C.c_int64()
        self._err(lib().ispan_server_submit(self._srv, gid, C.byref(rid)))
        self._rid_V[rid.value] = self._V[gid]
        return rid.value

    def submit_many(self, gids) -> list[int]:
        g = np.ascontiguousarray(gids, np.int64)
        r = np.empty(len(g), np.int64)
        self._err(lib().ispan_server_submit_many(self._srv, len(g), _ptr(g), _ptr(r)))
        for rid, gid in zip(r.tolist(), g.tolist()):
            self._rid_V[rid] = self._V[gid]
        return r.tolist()

    def wait(self, rid: int, timeout_ms: int = -1, want_labels: bool = True):
        """-> (labels or None, RequestInfo)."""
        V = self._rid_V.pop(rid, 0)
        out = np.empty(max(V, 1), np.int32) if want_labels else None
        info = RequestInfo()
        rc = lib().ispan_server_wait(self._srv, rid, _ptr(out) if out is not None else None, C.byref(info),
                                     timeout_ms)
        if rc == -10:
            self._rid_V[rid] = V
        self._err(rc)
        return (out[:V] if out is not None else None), info

    def wait_device(self, rid: int, timeout_ms: int = -1):
        """-> (device pointer of the labels, RequestInfo); release(rid) frees them."""
        p = C.c_void_p()
        info = RequestInfo()
        self._err(lib().ispan_server_wait_device(self._srv, rid, C.byref(p), C.byref(info), timeout_ms))
        return p.value or 0, info

    def release(self, rid: int):
        self._rid_V.pop(rid, None)
        self._err(lib().ispan_server_release(self._srv, rid))

    def stats(self) -> ServerStats:
        st = ServerStats()
        self._err(lib().ispan_server_stats_get(self._srv, C.byref(st)))
        return st

    def reset_stats(self):
        self._err(lib().ispan_server_reset_stats(self._srv))

    def set_tuning(self, table: "TuningTable | None"):
        """Solve each batch with the table's parameters for its largest graph's size class."""
        self._err(lib().ispan_server_set_tuning(self._srv, C.byref(table) if table is not None else None))

    def close(self):
        if self._srv:
            lib().ispan_server_destroy(self._srv)
            self._srv = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class DeviceBatch:
    """Graphs back to back in device memory, in the layout
    ispan_batch_bind_device reads: graph g's V_g + 1 offsets at
    fw_beg[v_off[g] + g], its adjacency at fw_adj[e_off[g]] (local ids)."""

    def __init__(self, v_off, e_off, ptrs):
        self.v_off = np.asarray(v_off, np.int64)
        self.e_off = np.asarray(e_off, np.int64)
        self.fw_beg, self.fw_adj, self.bw_beg, self.bw_adj = ptrs

    @property
    def n(self) -> int:
        return len(self.v_off) - 1

    @staticmethod
    def _layout(sizes):
        v_off = np.zeros(len(sizes) + 1, np.int64)
        e_off = np.zeros(len(sizes) + 1, np.int64)
        for i, (V, E) in enumerate(sizes):
            v_off[i + 1] = v_off[i] + V
            e_off[i + 1] = e_off[i] + E
        return v_off, e_off

    @classmethod
    def from_host(cls, graphs: list[HostGraph]) -> "DeviceBatch":
        from . import hiprt
        sizes = [(len(g.fw_beg) - 1, len(g.fw_adj)) for g in graphs]
        v_off, e_off = cls._layout(sizes)
        n = len(graphs)
        nb, na = int(v_off[-1] + n) * 4, int(e_off[-1]) * 4
        ptrs = [hiprt.malloc(nb), hiprt.malloc(na), hiprt.malloc(nb), hiprt.malloc(na)]
        for k, fld in enumerate(("fw_beg", "fw_adj", "bw_beg", "bw_adj")):
            beg = fld.endswith("beg")
            cat = np.concatenate([np.asarray(getattr(g, fld), np.int32) for g in graphs]) if n else \
                np.empty(0, np.int32)
            if cat.size:
                hiprt.upload(ptrs[k], cat)
            assert cat.nbytes == (nb if beg else na)
        return cls(v_off, e_off, ptrs)

    @classmethod
    def from_device(cls, graphs: list[DeviceGraph]) -> "DeviceBatch":
        from . import hiprt
        sizes = [(g.V, g.E) for g in graphs]
        v_off, e_off = cls._layout(sizes)
        n = len(graphs)
        nb, na = int(v_off[-1] + n) * 4, int(e_off[-1]) * 4
        ptrs = [hiprt.malloc(nb), hiprt.malloc(na), hiprt.malloc(nb), hiprt.malloc(na)]
        for i, g in enumerate(graphs):
            bo, eo = int(v_off[i] + i) * 4, int(e_off[i]) * 4
            hiprt.device_copy(ptrs[0] + bo, g.fw_beg, (g.V + 1) * 4)
            hiprt.device_copy(ptrs[2] + bo, g.bw_beg, (g.V + 1) * 4)
            if g.E:
                hiprt.device_copy(ptrs[1] + eo, g.fw_adj, g.E * 4)
                hiprt.device_copy(ptrs[3] + eo, g.bw_adj, g.E * 4)
        return cls(v_off, e_off, ptrs)

    def free(self):
        from . import hiprt
        for p in (self.fw_beg, self.fw_adj, self.bw_beg, self.bw_adj):
            if p:
                hiprt.free(p)
        self.fw_beg = self.fw_adj = self.bw_beg = self.bw_adj = 0


def tuning_grid(alphas=(4, 10, 30, 100), betas=(50, 200, 800), **kw) -> list[Params]:
    """Candidate parameter sets for Solver.autotune: every (alpha, beta) pair."""
    return [default_params(alpha=a, beta=b, **kw) for a in alphas for b in betas]


def class_tuning_grid(alphas=(2, 10, 30), betas=(50, 200, 5000)) -> list[Params]:
    """Candidates for Solver.autotune_classes: every (alpha, beta) pair, then
    the defaults with one other knob changed at a time -- the reference's
    frontier estimate, each small-SCC strategy, the full-graph pattern walk,
    a longer trim-1 phase, another bottom-up level cap, and in-place
    expansion off or with a larger step budget."""
    c = tuning_grid(alphas, betas)
    c += [default_params(switch_rule=0), default_params(wcc_mode=1), default_params(wcc_mode=2),
          default_params(trim_scope=1), default_params(theta=0.01, trim_times=10), default_params(gamma=30),
          default_params(td_local=0), default_params(td_local=1024)]
    return c


def scc_detection(g: HostGraph, alpha: int = 30, beta: int = 200, gamma: int = 10, theta: float = 0.1,
                  thread_count: int = 1, **kw) -> np.ndarray:
    """scc_detection (src/scc_core.cpp:27-1094) on the GPU -> canonical labels."""
    p = default_params(alpha=int(alpha), beta=int(beta), gamma=int(gamma), theta=float(theta),
                       thread_count=int(thread_count), **kw)
    arrs = [np.ascontiguousarray(a, np.int32) for a in (g.fw_beg, g.fw_adj, g.bw_beg, g.bw_adj)]
    V, E = int(arrs[0].shape[0]) - 1, int(arrs[1].shape[0])
    lab = np.empty(max(V, 1), np.int32)
    st = Stats()
    _check(lib().ispan_scc_detection(V, E, *[_ptr(a) for a in arrs], C.byref(p), _ptr(lab), C.byref(st)))
    return lab[:V]


def get_scc_result(labels: np.ndarray) -> dict:
    """SCC size summary (get_scc_result, src/load_print.cpp:24-60, without its phantom entry)."""
    lab = np.ascontiguousarray(labels, np.int32)
    s = Summary()
    _check(lib().ispan_summarize(int(lab.shape[0]), _ptr(lab), C.byref(s)))
    return {k: getattr(s, k) for k, _ in Summary._fields_}


def csr_check(g: HostGraph) -> bool:
    arrs = [np.ascontiguousarray(a, np.int32) for a in (g.fw_beg, g.fw_adj, g.bw_beg, g.bw_adj)]
    return lib().ispan_csr_check(g.V, g.E, *[_ptr(a) for a in arrs]) == ISPAN_OK
