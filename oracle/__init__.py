"""ctypes bindings of the CPU oracle (oracle/oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker or the timed CPU baseline. The
product (forkmatch_amd) never imports it.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "liboracle.so"
_lib = None

u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is not built; run `python -m forkmatch_amd.build`")
        h = C.CDLL(str(LIB_PATH))
        h.oracle_graph_new.restype = C.c_void_p
        h.oracle_graph_new.argtypes = [C.c_uint32, C.c_uint64, u32p, u32p, u32p]
        h.oracle_graph_free.restype = None
        h.oracle_graph_free.argtypes = [C.c_void_p]
        h.oracle_graph_symmetric.restype = C.c_int
        h.oracle_graph_symmetric.argtypes = [C.c_void_p]
        h.oracle_plan.restype = C.c_int
        h.oracle_plan.argtypes = [C.c_uint32, C.c_uint32, u32p, u32p, u32p, u32p, u32p, u32p, u32p, u32p, u32p,
                                  u32p]
        h.oracle_match.restype = C.c_int
        h.oracle_match.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, u32p, u32p, u32p, u32p, C.c_int, C.c_uint64,
                                   C.c_uint64, C.c_uint64, C.c_int, C.c_uint64, u32p, C.c_int, u64p, u64p, u64p, u64p,
                                   C.POINTER(C.c_double)]
        h.oracle_match_budget.restype = C.c_int
        h.oracle_match_budget.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, u32p, u32p, u32p, u32p, C.c_int,
                                          C.c_uint64, C.c_uint64, C.c_uint64, C.c_int, C.c_uint64, u32p, C.c_int,
                                          C.c_double, u64p, u64p, u64p, u64p, C.POINTER(C.c_double),
                                          C.POINTER(C.c_int)]
        _lib = h
    return _lib


def _u32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint32))


def _p(a):
    return a.ctypes.data_as(u32p)


class OracleGraph:
    def __init__(self, vertices: int, src, dst, label):
        self.V = int(vertices)
        self._s, self._d, self._l = _u32(src), _u32(dst), _u32(label)
        self._h = lib().oracle_graph_new(self.V, len(self._s), _p(self._s), _p(self._d), _p(self._l))

    @property
    def symmetric(self) -> bool:
        return bool(lib().oracle_graph_symmetric(self._h))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().oracle_graph_free(self._h)
            self._h = None


@dataclass
class OracleResult:
    count: int
    table: np.ndarray | None
    row_words: int
    partials: list
    seconds: float


def _order(order, n):
    if order is None:
        return None, None
    o = _u32(order)
    if len(o) != n:
        raise ValueError("order must name every query vertex once")
    return o, _p(o)


def plan(n: int, edges, order=None):
    """The reference's plan, or the requirements of a given matching order."""
    keep, op = _order(order, n)
    e = np.asarray(edges, dtype=np.uint32).reshape(-1, 3)
    s, d, l = _u32(e[:, 0]), _u32(e[:, 1]), _u32(e[:, 2])
    cap = max(1, 4 * len(e) + 4)
    rb = np.zeros(n + 1, np.uint32)
    an, lb, dr = (np.zeros(cap, np.uint32) for _ in range(3))
    nr = C.c_uint32()
    out = np.zeros(max(n, 1), np.uint32)
    rc = lib().oracle_plan(n, len(e), _p(s), _p(d), _p(l), op, _p(out), _p(rb), _p(an), _p(lb), _p(dr),
                           C.byref(nr))
    if rc:
        raise ValueError(f"oracle planner error {rc}")
    k = nr.value
    return out[:n].tolist(), rb.tolist(), an[:k].tolist(), lb[:k].tolist(), dr[:k].tolist()


def match(graph: OracleGraph, n: int, edges, threads: int = 1, table: bool = True, max_rows: int | None = None,
          instrument: bool = False, roots: tuple[int, int] | None = None, stride: int = 1,
          order=None) -> OracleResult:
    """Full enumeration; rows sorted lexicographically, columns in query vertex order.
    `order` replaces the reference's matching order (same embeddings, its
    own partials per depth)."""
    keep, op = _order(order, n)
    e = np.asarray(edges, dtype=np.uint32).reshape(-1, 3)
    s, d, l = _u32(e[:, 0]), _u32(e[:, 1]), _u32(e[:, 2])
    rb, re_ = (0, 2 ** 64 - 1) if roots is None else roots
    cnt, rows, words = C.c_uint64(), C.c_uint64(), C.c_uint64()
    partials = np.zeros(max(n, 1), np.uint64)
    secs = C.c_double()
    if table:
        if max_rows is None:
            # a count-only pass sizes the table exactly
            first = match(graph, n, edges, threads, table=False, roots=roots, stride=stride, order=order)
            max_rows = first.count
        buf = np.zeros((max(max_rows, 1), n), np.uint32)
    else:
        max_rows = 0
        buf = np.zeros((1, max(n, 1)), np.uint32)
    rc = lib().oracle_match(graph._h, n, len(e), _p(s), _p(d), _p(l), op, threads, rb, re_, stride, 1 if table else 0,
                            max_rows, _p(buf), 1 if instrument else 0, C.byref(cnt), C.byref(rows),
                            C.byref(words), partials.ctypes.data_as(u64p), C.byref(secs))
    if rc:
        raise ValueError(f"oracle error {rc}")
    tab = None
    if table:
        tab = buf[: rows.value]
        if len(tab):
            tab = tab[np.lexsort(tab.T[::-1])]
    return OracleResult(int(cnt.value), tab, int(words.value), [int(x) for x in partials[:n]], float(secs.value))


def timed_count(graph: OracleGraph, n: int, edges, threads: int, max_seconds: float,
                roots: tuple[int, int] | None = None, stride: int = 1) -> tuple[int, float, bool]:
    """A throughput sample for the timed CPU baseline: count-only
    enumeration that stops after max_seconds. Returns (embeddings counted,
    seconds, stopped early); a stopped run's count is partial."""
    e = np.asarray(edges, dtype=np.uint32).reshape(-1, 3)
    s, d, l = _u32(e[:, 0]), _u32(e[:, 1]), _u32(e[:, 2])
    rb, re_ = (0, 2 ** 64 - 1) if roots is None else roots
    cnt, rows, words = C.c_uint64(), C.c_uint64(), C.c_uint64()
    partials = np.zeros(max(n, 1), np.uint64)
    secs, stopped = C.c_double(), C.c_int()
    buf = np.zeros((1, max(n, 1)), np.uint32)
    rc = lib().oracle_match_budget(graph._h, n, len(e), _p(s), _p(d), _p(l), None, threads, rb, re_, stride, 0, 0,
                                   _p(buf), 0, max_seconds, C.byref(cnt), C.byref(rows), C.byref(words),
                                   partials.ctypes.data_as(u64p), C.byref(secs), C.byref(stopped))
    if rc:
        raise ValueError(f"oracle error {rc}")
    return int(cnt.value), float(secs.value), bool(stopped.value)
