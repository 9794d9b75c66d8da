"""ctypes wrapper of the CPU oracle (oracle/build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the parity reference / CPU baseline.  The product
never imports it.  Build with ``make -C oracle`` (or __graft_entry__.build()).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "build" / "liboracle.so"
REF_PATH = HERE / "_ref" / "libmercury_ref.so"

BRUTE_KEYS = ["triangle", "claw", "path4", "paw", "k14", "fork", "bull",
              "diamond_pendant", "k113", "k4_pendant", "k4", "k5"]

_lib = None
_ref = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} missing; run `make -C oracle`")
        L = C.CDLL(str(LIB_PATH))
        P, i32 = C.c_void_p, C.c_int32
        L.orc_orientation.restype = i32
        L.orc_orientation.argtypes = [i32, P, P, i32, P, P, P, P, P, P, P]
        L.orc_tc.restype = C.c_uint64
        L.orc_tc.argtypes = [i32, P, P, i32]
        L.orc_tc_dag.restype = C.c_uint64
        L.orc_tc_dag.argtypes = [i32, P, P, i32]
        L.orc_tc_dag_range.restype = C.c_uint64
        L.orc_tc_dag_range.argtypes = [i32, P, P, i32, i32, i32]
        L.orc_pattern.restype = C.c_uint64
        L.orc_pattern.argtypes = [i32, P, P, i32, i32, P]
        L.orc_pattern_range.restype = C.c_uint64
        L.orc_pattern_range.argtypes = [i32, P, P, i32, i32, i32, i32, P]
        L.orc_pattern_work.restype = C.c_uint64
        L.orc_pattern_work.argtypes = [i32, P, P, i32, i32, P, P]
        L.orc_brute.restype = i32
        L.orc_brute.argtypes = [i32, P, P, P]
        _lib = L
    return _lib


def ref_available() -> bool:
    return REF_PATH.exists()


def ref() -> C.CDLL:
    """The reference's own host sources compiled into oracle/_ref."""
    global _ref
    if _ref is None:
        if not REF_PATH.exists():
            raise ImportError(f"{REF_PATH} missing; run `make -C oracle ref`")
        R = C.CDLL(str(REF_PATH))
        P = C.c_void_p
        R.ref_orientation.restype = C.c_int
        R.ref_orientation.argtypes = [C.c_char_p, C.c_int, P, P, P, P, P, P,
                                      P]
        R.ref_edgelist.restype = C.c_int64
        R.ref_edgelist.argtypes = [C.c_char_p, P, P]
        _ref = R
    return _ref


def _p(a: np.ndarray):
    return C.c_void_p(a.ctypes.data)


def _csr(rowptr, colidx):
    rp = np.ascontiguousarray(rowptr, dtype=np.int64)
    ci = np.ascontiguousarray(colidx, dtype=np.int32)
    if ci.size == 0:
        ci = np.zeros(1, dtype=np.int32)
    return rp, ci


def _orient_with(fn, first_args, n, deg_thd):
    sizes = np.zeros(3, dtype=np.int64)
    fn(*first_args, deg_thd, _p(sizes), None, None, None, None, None, None)
    out = {k: np.zeros(n + 1, dtype=np.int64)
           for k in ("dag_rowptr", "small_rowptr", "large_rowptr")}
    for k, s in zip(("dag_colidx", "small_colidx", "large_colidx"), sizes):
        out[k] = np.zeros(max(int(s), 1), dtype=np.int32)
    out["max_out"] = fn(*first_args, deg_thd, _p(sizes),
                        _p(out["dag_rowptr"]), _p(out["dag_colidx"]),
                        _p(out["small_rowptr"]), _p(out["small_colidx"]),
                        _p(out["large_rowptr"]), _p(out["large_colidx"]))
    for k, s in zip(("dag_colidx", "small_colidx", "large_colidx"), sizes):
        out[k] = out[k][:int(s)]
    return out


def orientation(rowptr, colidx, deg_thd: int = 128) -> dict:
    rp, ci = _csr(rowptr, colidx)
    n = len(rp) - 1
    return _orient_with(lib().orc_orientation, (n, _p(rp), _p(ci)), n,
                        deg_thd)


def ref_orientation(prefix: str, n: int, deg_thd: int = 128) -> dict:
    return _orient_with(ref().ref_orientation, (str(prefix).encode(),), n,
                        deg_thd)


def ref_edgelist(prefix: str) -> tuple[np.ndarray, np.ndarray]:
    R = ref()
    nnz = R.ref_edgelist(str(prefix).encode(), None, None)
    src = np.zeros(max(nnz, 1), dtype=np.int32)
    dst = np.zeros(max(nnz, 1), dtype=np.int32)
    R.ref_edgelist(str(prefix).encode(), _p(src), _p(dst))
    return src[:nnz], dst[:nnz]


def tc(rowptr, colidx, threads: int = 0) -> int:
    rp, ci = _csr(rowptr, colidx)
    return int(lib().orc_tc(len(rp) - 1, _p(rp), _p(ci), threads))


def tc_dag(dag_rowptr, dag_colidx, threads: int = 0, v_lo: int = 0,
           v_hi: int | None = None) -> int:
    rp, ci = _csr(dag_rowptr, dag_colidx)
    n = len(rp) - 1
    return int(lib().orc_tc_dag_range(n, _p(rp), _p(ci), v_lo,
                                      n if v_hi is None else v_hi, threads))


def pattern(rowptr, colidx, p: int, threads: int = 0, v_lo: int = 0,
            v_hi: int | None = None) -> tuple[int, int, int]:
    """Returns (combined count, counters[1], counters[2])."""
    rp, ci = _csr(rowptr, colidx)
    n = len(rp) - 1
    c = np.zeros(2, dtype=np.uint64)
    r = lib().orc_pattern_range(n, _p(rp), _p(ci), p, v_lo,
                                n if v_hi is None else v_hi, threads, _p(c))
    return int(r), int(c[0]), int(c[1])


def pattern_work(rowptr, colidx, p: int, threads: int = 0) -> tuple[int, int]:
    """Returns (combined count, set-op work W = sum of |a| + |b| over the
    set operations of the reference's algebra)."""
    rp, ci = _csr(rowptr, colidx)
    c = np.zeros(2, dtype=np.uint64)
    w = np.zeros(1, dtype=np.uint64)
    r = lib().orc_pattern_work(len(rp) - 1, _p(rp), _p(ci), p, threads,
                               _p(c), _p(w))
    return int(r), int(w[0])


def brute(rowptr, colidx) -> dict:
    rp, ci = _csr(rowptr, colidx)
    out = np.zeros(12, dtype=np.uint64)
    if lib().orc_brute(len(rp) - 1, _p(rp), _p(ci), _p(out)) != 0:
        raise ValueError("brute force needs n <= 64")
    return {k: int(v) for k, v in zip(BRUTE_KEYS, out)}


def expected_patterns(census: dict) -> dict:
    """Pattern counts implied by an induced-subgraph census: the
    multiplicities with which the reference's set algebra sees each induced
    subgraph (both directions of every task edge, set differences that keep
    v1 / v2 in the frontier)."""
    return {
        0: census["triangle"],
        1: census["claw"],
        2: census["fork"],
        3: 4 * census["k14"] + 3 * census["claw"],
        4: census["path4"],
        5: 2 * census["bull"],
        6: 2 * census["diamond_pendant"] + 2 * census["paw"],
        7: 2 * census["k113"],
        8: 3 * census["k4_pendant"],
    }
