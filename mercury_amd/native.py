"""ctypes binding of ``libmgpm.so`` (the C ABI in ``include/mgpm.h``).

This is a thin host-side mirror for tests, benchmarks and Python callers;
the counting itself always runs in the native library's HIP kernels.  If the
library has not been built, importing raises instead of falling back.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path
from typing import Optional

import numpy as np

_LIB_PATH = Path(__file__).resolve().parent / "libmgpm.so"

TC, P1, P2, P3, P4, P5, P6, P7, P8 = range(9)
PATTERN_NAMES = {"TC": TC, **{f"P{i}": i for i in range(1, 9)}}


class MgpmError(RuntimeError):
    def __init__(self, code: int, where: str, detail: str):
        self.code = code
        super().__init__(f"{where}: {strerror(code)} ({code}): {detail}")


class Tune(C.Structure):
    _fields_ = [("small_max_deg", C.c_int32), ("wave_max_deg", C.c_int32),
                ("block_max_deg", C.c_int32), ("chunk_work", C.c_int32),
                ("grid_blocks", C.c_int32), ("pattern_waves", C.c_int32),
                ("group_work", C.c_int32), ("tc_items", C.c_int32),
                ("tc_cursor_min", C.c_int32), ("pattern_split", C.c_int32),
                ("pattern_census", C.c_int32), ("tc_layout", C.c_int32),
                ("tc_merge_max", C.c_int32), ("tc_half_waves", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("kernel_ms", C.c_double), ("e2e_ms", C.c_double),
                ("prep_ms", C.c_double), ("n", C.c_uint64),
                ("m", C.c_uint64), ("m_plus", C.c_uint64),
                ("s_plus", C.c_uint64), ("probes", C.c_uint64),
                ("bytes_alg", C.c_uint64), ("path", C.c_int32),
                ("reserved", C.c_int32 * 3)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_
                if name != "reserved"}


class Query(C.Structure):
    _fields_ = [("graph", C.c_void_p), ("pattern", C.c_int32),
                ("status", C.c_int32), ("count", C.c_uint64)]


class BatchOpts(C.Structure):
    _fields_ = [("lanes", C.c_int32), ("use_graph", C.c_int32),
                ("autotune", C.c_int32), ("pattern_census", C.c_int32),
                ("reserved", C.c_int32 * 4)]


class ServerConfig(C.Structure):
    _fields_ = [("device", C.c_int32), ("slots", C.c_int32),
                ("max_batch", C.c_int32), ("lanes", C.c_int32),
                ("hbm_budget", C.c_uint64), ("pattern_census", C.c_int32),
                ("reserved32", C.c_int32), ("reserved", C.c_uint64 * 1)]


class Request(C.Structure):
    _fields_ = [("graph", C.c_int32), ("pattern", C.c_int32),
                ("arrival_us", C.c_uint64), ("count", C.c_uint64),
                ("status", C.c_int32), ("batch_size", C.c_int32),
                ("latency_ms", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64),
                ("failed", C.c_uint64), ("launches", C.c_uint64),
                ("resident_bytes", C.c_uint64),
                ("workspace_bytes", C.c_uint64),
                ("budget_bytes", C.c_uint64), ("busy_ms", C.c_double),
                ("deferred", C.c_uint64), ("deferred_failed", C.c_uint64)]


class GraphInfo(C.Structure):
    _fields_ = [("n", C.c_int32), ("max_degree", C.c_int32),
                ("nnz", C.c_int64), ("meta_max_degree", C.c_int32),
                ("reserved", C.c_int32)]


def _load() -> C.CDLL:
    if not _LIB_PATH.exists():
        raise ImportError(
            f"{_LIB_PATH} is not built; run `python -m mercury_amd.build` "
            "(there is no non-native fallback)")
    lib = C.CDLL(str(_LIB_PATH))
    P = C.c_void_p
    i32, i64, u64 = C.c_int32, C.c_int64, C.c_uint64
    pp = C.POINTER(C.c_void_p)
    sig = {
        "mgpm_strerror": (C.c_char_p, [C.c_int]),
        "mgpm_last_error": (C.c_char_p, []),
        "mgpm_abi_version": (C.c_int, []),
        "mgpm_graph_load": (C.c_int, [C.c_char_p, pp]),
        "mgpm_graph_load_ex": (C.c_int, [C.c_char_p, i32, pp]),
        "mgpm_graph_labels": (C.c_int, [P, pp, C.POINTER(i32), pp,
                                        C.POINTER(i32)]),
        "mgpm_graph_set_labels": (C.c_int, [P, i32, P, i32, P]),
        "mgpm_graph_label_frequency": (C.c_int, [P, P]),
        "mgpm_graph_nlf": (C.c_int, [P, P, P, P]),
        "mgpm_graph_from_csr": (C.c_int, [i32, i64, P, P, pp]),
        "mgpm_graph_read_edges": (C.c_int, [C.c_char_p, i32, i32, pp]),
        "mgpm_graph_write": (C.c_int, [P, C.c_char_p]),
        "mgpm_graph_info_get": (C.c_int, [P, C.POINTER(GraphInfo)]),
        "mgpm_graph_csr": (C.c_int, [P, pp, pp]),
        "mgpm_graph_free": (C.c_int, [P]),
        "mgpm_graph_generate_rmat": (C.c_int, [i32, i32, C.c_double,
                                               C.c_double, C.c_double, u64,
                                               pp]),
        "mgpm_graph_generate_er": (C.c_int, [i32, i64, u64, pp]),
        "mgpm_graph_orientation": (C.c_int, [P, i32, P, P, P, P, P, P, P]),
        "mgpm_device_count": (C.c_int, [C.POINTER(C.c_int)]),
        "mgpm_device_synchronize": (C.c_int, [C.c_int]),
        "mgpm_dgraph_upload": (C.c_int, [P, C.c_int, P, pp]),
        "mgpm_dgraph_upload_ex": (C.c_int, [P, C.c_int, P, i32, pp]),
        "mgpm_dgraph_footprint": (C.c_int, [P, C.POINTER(u64)]),
        "mgpm_dgraph_free": (C.c_int, [P]),
        "mgpm_tc_count": (C.c_int, [P, C.POINTER(Tune), P, C.POINTER(u64),
                                    C.POINTER(Stats)]),
        "mgpm_pattern_count": (C.c_int, [P, C.c_int, C.POINTER(Tune), P,
                                         C.POINTER(u64), C.POINTER(Stats)]),
        "mgpm_pattern_count_noalloc": (C.c_int, [P, C.c_int, P,
                                                 C.POINTER(u64)]),
        "mgpm_dgraph_orientation": (C.c_int, [P, i32, P, P, P, P, P, P, P, P,
                                              C.POINTER(Stats)]),
        "mgpm_autotune": (C.c_int, [P, C.c_int, i32, C.POINTER(Tune),
                                    C.POINTER(C.c_double)]),
        "mgpm_batch_create": (C.c_int, [C.c_int, C.POINTER(BatchOpts), pp]),
        "mgpm_batch_run": (C.c_int, [P, C.POINTER(Query), C.c_int, P,
                                     C.POINTER(Stats)]),
        "mgpm_batch_footprint": (C.c_int, [P, C.POINTER(u64)]),
        "mgpm_batch_free": (C.c_int, [P]),
        "mgpm_server_create": (C.c_int, [C.POINTER(ServerConfig), pp]),
        "mgpm_server_add_graph": (C.c_int, [P, P, C.POINTER(i32)]),
        "mgpm_server_submit": (C.c_int, [P, i32, i32, C.POINTER(u64)]),
        "mgpm_server_wait": (C.c_int, [P, u64, i32, C.POINTER(Request)]),
        "mgpm_server_replay": (C.c_int, [P, C.POINTER(Request), C.c_int]),
        "mgpm_server_stats_get": (C.c_int, [P, C.POINTER(ServerStats)]),
        "mgpm_server_destroy": (C.c_int, [P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()

# Every symbol include/mgpm.h declares (checked by the CPU test suite).
EXPORTED = [
    "mgpm_strerror", "mgpm_abi_version", "mgpm_last_error",
    "mgpm_graph_load", "mgpm_graph_from_csr", "mgpm_graph_read_edges",
    "mgpm_graph_write", "mgpm_graph_load_ex", "mgpm_graph_labels",
    "mgpm_graph_set_labels", "mgpm_graph_label_frequency", "mgpm_graph_nlf",
    "mgpm_graph_info_get", "mgpm_graph_csr", "mgpm_graph_free",
    "mgpm_graph_generate_rmat", "mgpm_graph_generate_er",
    "mgpm_graph_orientation", "mgpm_device_count", "mgpm_device_synchronize",
    "mgpm_dgraph_upload", "mgpm_dgraph_upload_ex", "mgpm_dgraph_footprint",
    "mgpm_dgraph_free",
    "mgpm_tc_count", "mgpm_pattern_count", "mgpm_autotune",
    "mgpm_pattern_count_noalloc",
    "mgpm_dgraph_orientation",
    "mgpm_batch_create",
    "mgpm_batch_run", "mgpm_batch_footprint", "mgpm_batch_free",
    "mgpm_server_create", "mgpm_server_add_graph", "mgpm_server_submit",
    "mgpm_server_wait", "mgpm_server_replay", "mgpm_server_stats_get",
    "mgpm_server_destroy",
]


def strerror(code: int) -> str:
    return lib.mgpm_strerror(code).decode()


def check(code: int, where: str) -> None:
    if code != 0:
        raise MgpmError(code, where, lib.mgpm_last_error().decode())


def _ptr(a: np.ndarray) -> C.c_void_p:
    return C.c_void_p(a.ctypes.data)


def device_count() -> int:
    c = C.c_int(0)
    check(lib.mgpm_device_count(C.byref(c)), "mgpm_device_count")
    return c.value


def make_tune(**kw) -> Tune:
    t = Tune()
    for k, v in kw.items():
        setattr(t, k, int(v))
    return t


class Graph:
    """Host CSR graph (``mgpm_graph``)."""

    def __init__(self, handle: C.c_void_p):
        self._h = handle

    # -- constructors ------------------------------------------------------
    @classmethod
    def load(cls, prefix: str, vlabels: bool = False,
             elabels: bool = False) -> "Graph":
        """Reference format (graph.cc:6-92); with vlabels / elabels also
        <prefix>.vlabel.bin / <prefix>.elabel.bin (graph.cc:37-82)."""
        h = C.c_void_p()
        flags = (1 if vlabels else 0) | (2 if elabels else 0)
        check(lib.mgpm_graph_load_ex(str(prefix).encode(), flags,
                                     C.byref(h)),
              f"mgpm_graph_load({prefix})")
        return cls(h)

    @classmethod
    def read_edges(cls, path: str, fmt: int = 0, base: int = -1) -> "Graph":
        """MatrixMarket (fmt 1) or edge-list (fmt 2; 0 = by extension) file;
        ids start at `base` (-1 = 1)."""
        h = C.c_void_p()
        check(lib.mgpm_graph_read_edges(str(path).encode(), fmt, base,
                                        C.byref(h)),
              f"mgpm_graph_read_edges({path})")
        return cls(h)

    @classmethod
    def from_csr(cls, rowptr, colidx) -> "Graph":
        rp = np.ascontiguousarray(rowptr, dtype=np.int64)
        ci = np.ascontiguousarray(colidx, dtype=np.int32)
        h = C.c_void_p()
        check(lib.mgpm_graph_from_csr(len(rp) - 1, len(ci), _ptr(rp),
                                      _ptr(ci), C.byref(h)),
              "mgpm_graph_from_csr")
        return cls(h)

    @classmethod
    def rmat(cls, scale: int, edge_factor: int = 16, a: float = 0.57,
             b: float = 0.19, c: float = 0.19, seed: int = 1) -> "Graph":
        h = C.c_void_p()
        check(lib.mgpm_graph_generate_rmat(scale, edge_factor, a, b, c, seed,
                                           C.byref(h)),
              "mgpm_graph_generate_rmat")
        return cls(h)

    @classmethod
    def er(cls, n: int, draws: int, seed: int) -> "Graph":
        h = C.c_void_p()
        check(lib.mgpm_graph_generate_er(n, draws, seed, C.byref(h)),
              "mgpm_graph_generate_er")
        return cls(h)

    # -- accessors -----------------------------------------------------------
    def info(self) -> GraphInfo:
        gi = GraphInfo()
        check(lib.mgpm_graph_info_get(self._h, C.byref(gi)), "graph_info")
        return gi

    @property
    def n(self) -> int:
        return self.info().n

    @property
    def nnz(self) -> int:
        return self.info().nnz

    def csr(self) -> tuple[np.ndarray, np.ndarray]:
        """Copies of (rowptr int64[n+1], colidx int32[nnz])."""
        gi = self.info()
        rp, ci = C.c_void_p(), C.c_void_p()
        check(lib.mgpm_graph_csr(self._h, C.byref(rp), C.byref(ci)),
              "mgpm_graph_csr")
        rowptr = np.ctypeslib.as_array(
            C.cast(rp, C.POINTER(C.c_int64)), shape=(gi.n + 1,)).copy()
        if gi.nnz:
            colidx = np.ctypeslib.as_array(
                C.cast(ci, C.POINTER(C.c_int32)), shape=(gi.nnz,)).copy()
        else:
            colidx = np.zeros(0, dtype=np.int32)
        return rowptr, colidx

    def labels(self) -> tuple:
        """(vlabels uint8[n] or None, num_vertex_classes, elabels
        uint16[nnz] or None, num_edge_classes), copies."""
        vp, ep = C.c_void_p(), C.c_void_p()
        nv, ne = C.c_int32(), C.c_int32()
        check(lib.mgpm_graph_labels(self._h, C.byref(vp), C.byref(nv),
                                    C.byref(ep), C.byref(ne)),
              "mgpm_graph_labels")
        gi = self.info()
        v = e = None
        if nv.value:
            v = (np.ctypeslib.as_array(C.cast(vp, C.POINTER(C.c_uint8)),
                                       shape=(gi.n,)).copy()
                 if gi.n else np.zeros(0, np.uint8))
        if ne.value:
            e = (np.ctypeslib.as_array(C.cast(ep, C.POINTER(C.c_uint16)),
                                       shape=(gi.nnz,)).copy()
                 if gi.nnz else np.zeros(0, np.uint16))
        return v, nv.value, e, ne.value

    def set_labels(self, vlabels=None, num_vertex_classes: int = 0,
                   elabels=None, num_edge_classes: int = 0) -> None:
        v = None if vlabels is None else np.ascontiguousarray(vlabels,
                                                              np.uint8)
        e = None if elabels is None else np.ascontiguousarray(elabels,
                                                              np.uint16)
        if v is not None and len(v) != self.n:
            raise ValueError("vlabels: one per vertex")
        if e is not None and len(e) != self.nnz:
            raise ValueError("elabels: one per CSR entry")
        check(lib.mgpm_graph_set_labels(
            self._h, num_vertex_classes if v is not None else 0,
            None if v is None else _ptr(v),
            num_edge_classes if e is not None else 0,
            None if e is None else _ptr(e)), "mgpm_graph_set_labels")

    def label_frequency(self) -> np.ndarray:
        """Vertices per label 0..num_vertex_classes (graph.cc:440-457)."""
        nv = self.labels()[1]
        f = np.zeros(max(nv, 0) + 1, dtype=np.int64)
        check(lib.mgpm_graph_label_frequency(self._h, _ptr(f)),
              "mgpm_graph_label_frequency")
        return f

    def nlf(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """Neighbourhood label frequency of every vertex (graph.cc:483-495)
        as (offsets int64[n+1], labels int32, counts int32)."""
        off = np.zeros(self.n + 1, dtype=np.int64)
        check(lib.mgpm_graph_nlf(self._h, _ptr(off), None, None),
              "mgpm_graph_nlf")
        lab = np.zeros(max(int(off[-1]), 1), dtype=np.int32)
        cnt = np.zeros(max(int(off[-1]), 1), dtype=np.int32)
        check(lib.mgpm_graph_nlf(self._h, _ptr(off), _ptr(lab), _ptr(cnt)),
              "mgpm_graph_nlf")
        return off, lab[:off[-1]], cnt[:off[-1]]

    def write(self, prefix: str) -> None:
        check(lib.mgpm_graph_write(self._h, str(prefix).encode()),
              f"mgpm_graph_write({prefix})")

    def orientation(self, deg_thd: int = 128) -> dict:
        """DAG + small/large division (graph.cc:196-321 semantics)."""
        n = self.n
        sizes = np.zeros(3, dtype=np.int64)
        check(lib.mgpm_graph_orientation(self._h, deg_thd, _ptr(sizes), None,
                                         None, None, None, None, None),
              "mgpm_graph_orientation")
        out = {k: np.zeros(n + 1, dtype=np.int64)
               for k in ("dag_rowptr", "small_rowptr", "large_rowptr")}
        out["dag_colidx"] = np.zeros(max(sizes[0], 1), dtype=np.int32)
        out["small_colidx"] = np.zeros(max(sizes[1], 1), dtype=np.int32)
        out["large_colidx"] = np.zeros(max(sizes[2], 1), dtype=np.int32)
        check(lib.mgpm_graph_orientation(
            self._h, deg_thd, _ptr(sizes), _ptr(out["dag_rowptr"]),
            _ptr(out["dag_colidx"]), _ptr(out["small_rowptr"]),
            _ptr(out["small_colidx"]), _ptr(out["large_rowptr"]),
            _ptr(out["large_colidx"])), "mgpm_graph_orientation")
        for k, s in zip(("dag_colidx", "small_colidx", "large_colidx"),
                        sizes):
            out[k] = out[k][:s]
        return out

    def close(self) -> None:
        if self._h:
            lib.mgpm_graph_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceGraph:
    """Device-resident graph (``mgpm_dgraph``) on one GPU."""

    UPLOAD_DEVICE_PLAN = 1

    def __init__(self, graph: Graph, device: int = 0,
                 stream: Optional[int] = None, device_plan: bool = False):
        """device_plan: build the DAG and the triangle plan on the device
        (``MGPM_UPLOAD_DEVICE_PLAN``) instead of on the host."""
        self._h = C.c_void_p()
        flags = self.UPLOAD_DEVICE_PLAN if device_plan else 0
        check(lib.mgpm_dgraph_upload_ex(graph._h, device,
                                        C.c_void_p(stream or 0), flags,
                                        C.byref(self._h)),
              "mgpm_dgraph_upload_ex")
        self.device = device

    def footprint(self) -> int:
        b = C.c_uint64()
        check(lib.mgpm_dgraph_footprint(self._h, C.byref(b)), "footprint")
        return b.value

    def tc_count(self, tune: Optional[Tune] = None,
                 stream: Optional[int] = None) -> tuple[int, Stats]:
        cnt, st = C.c_uint64(), Stats()
        check(lib.mgpm_tc_count(self._h, C.byref(tune) if tune else None,
                                C.c_void_p(stream or 0), C.byref(cnt),
                                C.byref(st)), "mgpm_tc_count")
        return cnt.value, st

    def pattern_count(self, pattern: int, tune: Optional[Tune] = None,
                      stream: Optional[int] = None) -> tuple[int, Stats]:
        cnt, st = C.c_uint64(), Stats()
        check(lib.mgpm_pattern_count(self._h, int(pattern),
                                     C.byref(tune) if tune else None,
                                     C.c_void_p(stream or 0), C.byref(cnt),
                                     C.byref(st)), "mgpm_pattern_count")
        return cnt.value, st

    def pattern_count_noalloc(self, pattern: int,
                              stream: Optional[int] = None) -> int:
        """mgpm_pattern_count_noalloc: the count with every scratch buffer
        in the library's fixed per-GPU array (no device allocation)."""
        cnt = C.c_uint64()
        check(lib.mgpm_pattern_count_noalloc(self._h, int(pattern),
                                             C.c_void_p(stream or 0),
                                             C.byref(cnt)),
              "mgpm_pattern_count_noalloc")
        return cnt.value

    def orientation(self, deg_thd: int = 128,
                    stream: Optional[int] = None) -> tuple[dict, Stats]:
        """DAG and small/large division built on the device (the arrays of
        Graph.orientation), copied back; with the device build's Stats."""
        import numpy as np
        sizes = (C.c_int64 * 3)()
        st = Stats()
        check(lib.mgpm_dgraph_orientation(self._h, deg_thd,
                                          C.c_void_p(stream or 0), sizes,
                                          None, None, None, None, None, None,
                                          C.byref(st)),
              "mgpm_dgraph_orientation")
        n1 = st.n + 1
        out = {}
        for k, name in enumerate(("dag", "small", "large")):
            out[name + "_rowptr"] = np.empty(n1, dtype=np.int64)
            out[name + "_colidx"] = np.empty(max(sizes[k], 1), dtype=np.int32)
        ptr = lambda a: a.ctypes.data_as(C.c_void_p)
        check(lib.mgpm_dgraph_orientation(
            self._h, deg_thd, C.c_void_p(stream or 0), sizes,
            ptr(out["dag_rowptr"]), ptr(out["dag_colidx"]),
            ptr(out["small_rowptr"]), ptr(out["small_colidx"]),
            ptr(out["large_rowptr"]), ptr(out["large_colidx"]), C.byref(st)),
            "mgpm_dgraph_orientation")
        for k, name in enumerate(("dag", "small", "large")):
            out[name + "_colidx"] = out[name + "_colidx"][:sizes[k]]
        return out, st

    def autotune(self, pattern: int = 0, reps: int = 0) -> tuple[Tune, float]:
        """Times candidate knobs on the device, keeps the fastest resident
        and returns it with its kernel time in ms."""
        t, ms = Tune(), C.c_double()
        check(lib.mgpm_autotune(self._h, int(pattern), reps, C.byref(t),
                                C.byref(ms)), "mgpm_autotune")
        return t, ms.value

    def close(self) -> None:
        if self._h:
            lib.mgpm_dgraph_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many (graph, pattern) queries per call, overlapped over `lanes`
    streams, optionally replayed from a recorded HIP graph
    (``mgpm_batch``)."""

    def __init__(self, device: int = 0, lanes: int = 0,
                 use_graph: bool = False, autotune: bool = False,
                 census: bool = True):
        self._h = C.c_void_p()
        opts = BatchOpts(lanes=lanes, use_graph=int(bool(use_graph)),
                         autotune=int(bool(autotune)),
                         pattern_census=0 if census else -1)
        check(lib.mgpm_batch_create(device, C.byref(opts), C.byref(self._h)),
              "mgpm_batch_create")

    def run(self, queries, stream: Optional[int] = None):
        """queries: list of (DeviceGraph, pattern).  Returns (counts, Stats)."""
        arr = (Query * len(queries))()
        for i, (dg, p) in enumerate(queries):
            arr[i].graph = dg._h
            arr[i].pattern = int(p)
        st = Stats()
        check(lib.mgpm_batch_run(self._h, arr, len(queries),
                                 C.c_void_p(stream or 0), C.byref(st)),
              "mgpm_batch_run")
        return [arr[i].count for i in range(len(queries))], st

    def footprint(self) -> int:
        b = C.c_uint64()
        check(lib.mgpm_batch_footprint(self._h, C.byref(b)), "footprint")
        return b.value

    def close(self) -> None:
        if self._h:
            lib.mgpm_batch_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching count server (``mgpm_server``)."""

    def __init__(self, device: int = 0, slots: int = 0, max_batch: int = 0,
                 hbm_budget: int = 0, lanes: int = 0, census: bool = True):
        cfg = ServerConfig(device=device, slots=slots, max_batch=max_batch,
                           lanes=lanes, hbm_budget=hbm_budget,
                           pattern_census=0 if census else -1)
        self._h = C.c_void_p()
        check(lib.mgpm_server_create(C.byref(cfg), C.byref(self._h)),
              "mgpm_server_create")

    def add_graph(self, graph: Graph) -> int:
        gid = C.c_int32()
        check(lib.mgpm_server_add_graph(self._h, graph._h, C.byref(gid)),
              "mgpm_server_add_graph")
        return gid.value

    def submit(self, graph_id: int, pattern: int) -> int:
        t = C.c_uint64()
        check(lib.mgpm_server_submit(self._h, graph_id, pattern, C.byref(t)),
              "mgpm_server_submit")
        return t.value

    def wait(self, ticket: int, timeout_ms: int = -1) -> Request:
        r = Request()
        check(lib.mgpm_server_wait(self._h, ticket, timeout_ms, C.byref(r)),
              "mgpm_server_wait")
        return r

    def replay(self, schedule):
        """schedule: list of (graph_id, pattern, arrival_us).  Returns the
        filled Request array."""
        arr = (Request * len(schedule))()
        for i, (g, p, t) in enumerate(schedule):
            arr[i].graph, arr[i].pattern, arr[i].arrival_us = g, p, int(t)
        rc = lib.mgpm_server_replay(self._h, arr, len(schedule))
        if rc not in (0, -13):
            check(rc, "mgpm_server_replay")
        return arr

    def stats(self) -> ServerStats:
        st = ServerStats()
        check(lib.mgpm_server_stats_get(self._h, C.byref(st)), "stats")
        return st

    def close(self) -> None:
        if self._h:
            lib.mgpm_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
