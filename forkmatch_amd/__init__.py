"""ForkMatch for MI355X: host-side Python mirror of the C ABI.

The reference's surface (src/smsm.cu:1182-1393) is ``match(query, data)``
run by the ``ForkMatch <query.g> <data.g>`` CLI; it loads the data graph
(``parseGraph``), the query (``fileParse`` + ``preProcessQuery``) and
launches the matcher. This module exposes the same steps as objects:

    g = Graph.load_g("data.g")          # parseGraph, uploaded to HBM
    q = Query.load_g("query.g")         # fileParse + preProcessQuery
    with Engine() as e:                 # prealloc + populatePool
        r = e.match(g, q)               # initPosting + isoLoop + isoWrite
        r.count(0), r.table(0)

Errors raise :class:`ForkMatchError` carrying the reference's exit code
where one exists (233 malformed query, 76/77/78 malformed data graph, ...).
Every call goes to libforkmatch.so; nothing runs on the CPU in its place.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L

FM_MODE_COUNT = 0
FM_MODE_TABLE = 1
FM_RESULT_TRUNCATED = 1
FM_RESULT_ABORTED = 2
PLAN_COST = 0       # order from the data graph's label statistics
PLAN_REFERENCE = 1  # the reference's query-degree order
PREALLOC_UNLIMITED = (1 << 64) - 1  # Engine(prealloc=...): no device-memory budget
NO_DEADLINE = 0xFFFFFFFF  # match(timeout_ms=...): the launch has no deadline

STATUS_NAMES = {
    0: "FM_OK", 1: "FM_ERR_GENERIC", 12: "FM_ERR_DATA_MISSING_FILE", 76: "FM_ERR_DATA_BAD_START",
    77: "FM_ERR_DATA_BAD_END", 78: "FM_ERR_DATA_BAD_HEADER", 79: "FM_ERR_DATA_BAD_BODY", 99: "FM_ERR_ARGV",
    232: "FM_ERR_QUERY_MISSING_FILE", 233: "FM_ERR_QUERY_INVALID", 234: "FM_ERR_QUERY_MISSING_DATA",
    300: "FM_ERR_INVALID_ARG", 301: "FM_ERR_DEVICE", 302: "FM_ERR_OUT_OF_MEMORY",
    303: "FM_ERR_POOL_EXHAUSTED", 304: "FM_ERR_QUEUE_FULL", 305: "FM_ERR_TIMEOUT",
    306: "FM_ERR_COUNT_LIMIT", 307: "FM_ERR_QUERY_TOO_LARGE", 308: "FM_ERR_STALE_RESULT",
    309: "FM_ERR_NOT_SUPPORTED", 310: "FM_ERR_INTERNAL", 311: "FM_ERR_BUSY",
}


class ForkMatchError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(code, code)} ({code}): {message}")
        self.code = code
        self.message = message


def _check(rc: int) -> None:
    if rc != 0:
        raise ForkMatchError(rc, L.lib().fm_last_error().decode(errors="replace"))


def _u32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint32))


def _p32(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def device_count() -> int:
    return int(L.lib().fm_device_count())


class Graph:
    """Label-partitioned CSR data graph (host copy, plus HBM once uploaded)."""

    def __init__(self, handle: C.c_void_p):
        self._h = handle

    # -- constructors -------------------------------------------------------
    @classmethod
    def parse_g(cls, path: str) -> "Graph":
        h = C.c_void_p()
        _check(L.lib().fm_graph_parse_g(str(path).encode(), C.byref(h)))
        return cls(h)

    @classmethod
    def load_g(cls, path: str) -> "Graph":
        h = C.c_void_p()
        _check(L.lib().fm_graph_load_g(str(path).encode(), C.byref(h)))
        return cls(h)

    @classmethod
    def from_edges(cls, vertices: int, src, dst, label) -> "Graph":
        s, d, l = _u32(src), _u32(dst), _u32(label)
        if not (len(s) == len(d) == len(l)):
            raise ValueError("src, dst and label must have the same length")
        h = C.c_void_p()
        _check(L.lib().fm_graph_from_edges(vertices, len(s), _p32(s), _p32(d), _p32(l), C.byref(h)))
        return cls(h)

    @classmethod
    def load_bin(cls, path: str) -> "Graph":
        h = C.c_void_p()
        _check(L.lib().fm_graph_load_bin(str(path).encode(), C.byref(h)))
        return cls(h)

    @classmethod
    def erdos_renyi(cls, vertices: int, undirected_edges: int, nlabels: int, zipf_s: float = 0.0,
                    seed: int = 0, vlabels: int = 1) -> "Graph":
        h = C.c_void_p()
        dist = 1 if zipf_s > 0 else 0
        _check(L.lib().fm_gen_erdos_renyi(vertices, undirected_edges, nlabels, dist, zipf_s, vlabels, seed,
                                          C.byref(h)))
        return cls(h)

    @classmethod
    def rmat(cls, scale: int, undirected_edges: int, nlabels: int, zipf_s: float = 0.0, seed: int = 0,
             a: float = 0.57, b: float = 0.19, c: float = 0.19, vlabels: int = 1) -> "Graph":
        h = C.c_void_p()
        dist = 1 if zipf_s > 0 else 0
        _check(L.lib().fm_gen_rmat(scale, undirected_edges, a, b, c, nlabels, dist, zipf_s, vlabels, seed,
                                   C.byref(h)))
        return cls(h)

    # -- queries about the graph --------------------------------------------
    def info(self) -> L.GraphInfo:
        gi = L.GraphInfo()
        _check(L.lib().fm_graph_info_get(self._h, C.byref(gi)))
        return gi

    @property
    def vertices(self) -> int:
        return int(self.info().vertices)

    @property
    def edges(self) -> int:
        return int(self.info().edges)

    def csr(self, direction: int = 0, device: bool = False):
        """(ptr, col, label_values) of the out-rows (0) or in-rows (1)."""
        gi = self.info()
        ptr = np.zeros(gi.labels * gi.vertices + 1, dtype=np.uint32)
        col = np.zeros(max(gi.edges, 1), dtype=np.uint32)
        labels = np.zeros(max(gi.labels, 1), dtype=np.uint32)
        if device:
            _check(L.lib().fm_graph_copy_device_csr(self._h, direction, _p32(ptr), _p32(col)))
            _check(L.lib().fm_graph_copy_csr(self._h, direction, None, None, _p32(labels)))
        else:
            _check(L.lib().fm_graph_copy_csr(self._h, direction, _p32(ptr), _p32(col), _p32(labels)))
        return ptr, col[: gi.edges], labels[: gi.labels]

    def edge_list(self):
        """Directed (src, dst, label) arrays, one entry per stored edge."""
        ptr, col, labels = self.csr(0)
        V = self.vertices
        counts = np.diff(ptr.astype(np.int64))
        rows = np.repeat(np.arange(len(counts), dtype=np.int64), counts)
        src = (rows % V).astype(np.uint32)
        lab = labels[(rows // V)].astype(np.uint32) if len(labels) else np.zeros(0, np.uint32)
        return src, col.copy(), lab

    def upload(self) -> "Graph":
        _check(L.lib().fm_graph_upload(self._h))
        return self

    def save_bin(self, path: str) -> None:
        _check(L.lib().fm_graph_save_bin(self._h, str(path).encode()))

    def write_g(self, path: str) -> None:
        _check(L.lib().fm_graph_write_g(self._h, str(path).encode()))

    def close(self) -> None:
        if self._h:
            L.lib().fm_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class Plan:
    """Matching order and per-depth requirements (src/ccsr.cuh:229-326)."""
    order: list
    req_begin: list
    anchor: list
    label: list
    direction: list

    def requirements(self, depth: int):
        b, e = self.req_begin[depth], self.req_begin[depth + 1]
        return [(self.anchor[i], self.label[i], self.direction[i]) for i in range(b, e)]


class Query:
    """A parsed query graph and its matching plan."""

    def __init__(self, handle: C.c_void_p):
        self._h = handle
        self.count_only = False

    @classmethod
    def load_g(cls, path: str) -> "Query":
        h = C.c_void_p()
        _check(L.lib().fm_query_load_g(str(path).encode(), C.byref(h)))
        return cls(h)

    @classmethod
    def from_edges(cls, vertices: int, edges) -> "Query":
        e = np.asarray(edges, dtype=np.uint32).reshape(-1, 3)
        s, d, l = _u32(e[:, 0]), _u32(e[:, 1]), _u32(e[:, 2])
        h = C.c_void_p()
        _check(L.lib().fm_query_from_edges(vertices, len(s), _p32(s), _p32(d), _p32(l), C.byref(h)))
        return cls(h)

    @classmethod
    def random_walk(cls, graph: Graph, vertices: int, shape: int = 0, min_cycles: int = 0,
                    seed: int = 0) -> "Query":
        h = C.c_void_p()
        _check(L.lib().fm_gen_query_walk(graph._h, vertices, shape, min_cycles, seed, C.byref(h)))
        return cls(h)

    @property
    def vertices(self) -> int:
        n = C.c_uint32()
        _check(L.lib().fm_query_info(self._h, C.byref(n), None))
        return int(n.value)

    def edges(self) -> np.ndarray:
        n, m = C.c_uint32(), C.c_uint32()
        _check(L.lib().fm_query_info(self._h, C.byref(n), C.byref(m)))
        s = np.zeros(max(m.value, 1), np.uint32)
        d = np.zeros_like(s)
        l = np.zeros_like(s)
        _check(L.lib().fm_query_edges(self._h, _p32(s), _p32(d), _p32(l)))
        return np.stack([s, d, l], axis=1)[: m.value]

    def plan(self, graph: "Graph | None" = None, plan: int = None) -> Plan:
        """The reference's plan, or with `graph` the plan matching uses on
        that graph (`plan` = PLAN_COST, the default there, or PLAN_REFERENCE)."""
        info = L.PlanInfo()
        if graph is None:
            def call(*a):
                return L.lib().fm_query_plan(self._h, *a)
        else:
            mode = PLAN_COST if plan is None else plan

            def call(*a):
                return L.lib().fm_query_plan_on(graph._h, self._h, mode, *a)
        _check(call(C.byref(info), None, None, None, None, None))
        n, r = info.vertices, max(info.nreqs, 1)
        order = np.zeros(n, np.uint32)
        rb = np.zeros(n + 1, np.uint32)
        an, lb, dr = (np.zeros(r, np.uint32) for _ in range(3))
        _check(call(C.byref(info), _p32(order), _p32(rb), _p32(an), _p32(lb), _p32(dr)))
        k = info.nreqs
        return Plan(order.tolist(), rb.tolist(), an[:k].tolist(), lb[:k].tolist(), dr[:k].tolist())

    def symmetry(self, max_auts: int = 48):
        """(automorphisms, constraints): auts[k][u] = p_k(u) (identity
        first; empty when there are more than max_auts) and pairs (a, b)
        meaning m(a) < m(b) (symmetry breaking)."""
        na, nl = C.c_uint32(), C.c_uint32()
        _check(L.lib().fm_query_symmetry(self._h, max_auts, C.byref(na), None, 0, C.byref(nl), None, 0))
        n = self.vertices
        auts = np.zeros(max(na.value * n, 1), np.uint32)
        less = np.zeros(max(2 * nl.value, 1), np.uint32)
        _check(L.lib().fm_query_symmetry(self._h, max_auts, C.byref(na), _p32(auts), len(auts), C.byref(nl),
                                         _p32(less), len(less)))
        return (auts[: na.value * n].reshape(na.value, n).tolist(),
                [tuple(x) for x in less[: 2 * nl.value].reshape(nl.value, 2).tolist()])

    def write_g(self, path: str) -> None:
        _check(L.lib().fm_query_write_g(self._h, str(path).encode()))

    def set_count_only(self, count_only: bool = True) -> "Query":
        """In a table-mode launch, count this query without storing its
        rows (its table stays empty)."""
        _check(L.lib().fm_query_set_count_only(self._h, 1 if count_only else 0))
        self.count_only = bool(count_only)
        return self

    def close(self) -> None:
        if self._h:
            L.lib().fm_query_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class QueryOutcome:
    count: int
    rows: int
    status: int
    flags: int
    vertices: int
    alg_row_bytes: int
    row_lookups: int
    partials: list = field(default_factory=list)
    candidates: int = 0
    probes: int = 0

    @property
    def truncated(self) -> bool:
        return bool(self.flags & FM_RESULT_TRUNCATED)


class Result:
    """Counts, tables and launch statistics of one match launch."""

    def __init__(self, handle: C.c_void_p):
        self._h = handle

    def __len__(self) -> int:
        return int(L.lib().fm_result_queries(self._h))

    def query(self, i: int) -> QueryOutcome:
        qr = L.QueryResult()
        _check(L.lib().fm_result_query(self._h, i, C.byref(qr)))
        return QueryOutcome(int(qr.count), int(qr.rows), int(qr.status), int(qr.flags), int(qr.vertices),
                            int(qr.alg_bytes), int(qr.row_lookups), [int(x) for x in qr.partials[: qr.vertices]],
                            int(qr.candidates), int(qr.probes))

    def count(self, i: int = 0) -> int:
        return self.query(i).count

    def stats(self) -> L.LaunchStats:
        st = L.LaunchStats()
        _check(L.lib().fm_result_stats(self._h, C.byref(st)))
        return st

    def table(self, i: int = 0, sorted: bool = True, max_rows: int | None = None) -> np.ndarray:
        q = self.query(i)
        rows = q.rows if max_rows is None else min(q.rows, max_rows)
        out = np.zeros((max(rows, 1), q.vertices), dtype=np.uint32)
        copied = C.c_uint64()
        _check(L.lib().fm_result_copy_table(self._h, i, _p32(out), rows, 1 if sorted else 0, C.byref(copied)))
        return out[: copied.value]

    def close(self) -> None:
        if self._h:
            L.lib().fm_result_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Engine:
    """Device context: bank arena, output arena, work queues, stream."""

    def __init__(self, soln_prealloc: int = 0, output_prealloc: int = 0, bank_size: int = 0,
                 output_page_bytes: int = 0, queue_capacity: int = 0, max_queries: int = 0, device: int = 0,
                 prealloc: int = 0):
        """prealloc: device bytes the arenas and the matched graph may take
        together (PREALLOC; 0 = the build default, PREALLOC_UNLIMITED = no
        limit). Beyond it the engine and the launches fail with status 1."""
        cfg = L.EngineConfig()
        cfg.prealloc = prealloc
        cfg.soln_prealloc = soln_prealloc
        cfg.output_prealloc = output_prealloc
        cfg.bank_size = bank_size
        cfg.output_page_bytes = output_page_bytes
        cfg.queue_capacity = queue_capacity
        cfg.max_queries = max_queries
        cfg.device = device
        h = C.c_void_p()
        _check(L.lib().fm_engine_create(C.byref(cfg), C.byref(h)))
        self._h = h

    def info(self) -> L.EngineInfo:
        ei = L.EngineInfo()
        _check(L.lib().fm_engine_info_get(self._h, C.byref(ei)))
        return ei

    @staticmethod
    def _opts(mode, instrument, max_rows, max_count, timeout_ms, grid_waves, task_rows, root_chunk,
              max_partials=0, plan=0, symmetry=False, waves_per_query=0, tile_parents=0):
        o = L.MatchOpts()
        o.waves_per_query = waves_per_query
        o.tile_parents = tile_parents
        o.max_partials = max_partials
        o.plan = plan
        o.symmetry = 1 if symmetry else 0
        o.mode = mode
        o.instrument = 1 if instrument else 0
        o.max_rows = max_rows
        o.max_count = max_count
        o.timeout_ms = timeout_ms
        o.grid_waves = grid_waves
        o.task_rows = task_rows
        o.root_chunk = root_chunk
        return o

    def match(self, graph: Graph, query: Query, table: bool = True, **kw) -> Result:
        return self.match_batch(graph, [query], table=table, **kw)

    def match_batch(self, graph: Graph, queries, table: bool = True, instrument: bool = False,
                    max_rows: int = 0, max_count: int = 0, timeout_ms: int = 0, grid_waves: int = 0,
                    task_rows: int = 0, root_chunk: int = 0, max_partials: int = 0,
                    allow_errors: bool = False, plan: int = 0, symmetry: bool = False, waves_per_query: int = 0,
                    tile_parents: int = 0) -> Result:
        """Matches the queries in one persistent launch. Launch knobs (0 =
        the library's default): grid_waves, task_rows (parents per queued
        task: the fork threshold), root_chunk (roots per claim),
        waves_per_query (a query's roots cut into this many claims, in place
        of root_chunk) and tile_parents (parents staged in LDS per step)."""
        o = self._opts(FM_MODE_TABLE if table else FM_MODE_COUNT, instrument, max_rows, max_count, timeout_ms,
                       grid_waves, task_rows, root_chunk, max_partials, plan, symmetry, waves_per_query,
                       tile_parents)
        arr = (C.c_void_p * len(queries))(*[q._h for q in queries])
        h = C.c_void_p()
        rc = L.lib().fm_match_batch(self._h, graph._h, arr, len(queries), C.byref(o), C.byref(h))
        if rc != 0 and not (allow_errors and h.value):
            msg = L.lib().fm_last_error().decode(errors="replace")
            if h.value:
                L.lib().fm_result_destroy(h)
            raise ForkMatchError(rc, msg)
        return Result(h)

    def close(self) -> None:
        if self._h:
            L.lib().fm_engine_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


FM_ERR_BUSY = 311


class Server:
    """Continuous-batching server: the graph and the bank pool stay on the
    device; submitted queries join the running persistent launch.

        with Server(engine, graph) as srv:
            t = srv.submit(q)          # None when refused (slots / bank budget)
            for c in srv.poll(): ...   # completions: ticket, count, rows, status
            rows = srv.table(t); srv.release(t)
    """

    def __init__(self, engine: Engine, graph: Graph, slots: int = 256, slot_pages: int = 256, bank_budget: int = 0,
                 idle_ms: int = 20, root_chunk: int = 0, task_rows: int = 0, table: bool = True,
                 max_launch_ms: int = 0, plan: int = 0, slice_ms: int = 0, symmetry: bool = False):
        cfg = L.ServerConfig()
        cfg.slots, cfg.slot_pages, cfg.bank_budget, cfg.idle_ms = slots, slot_pages, bank_budget, idle_ms
        cfg.root_chunk, cfg.task_rows, cfg.table = root_chunk, task_rows, 1 if table else 0
        cfg.max_launch_ms, cfg.plan, cfg.slice_ms = max_launch_ms, plan, slice_ms
        cfg.symmetry = 1 if symmetry else 0
        h = C.c_void_p()
        _check(L.lib().fm_server_create(engine._h, graph._h, C.byref(cfg), C.byref(h)))
        self._h = h
        self._keep = (engine, graph)

    def submit(self, query: Query, max_rows: int = 0):
        t = C.c_uint32()
        rc = L.lib().fm_server_submit(self._h, query._h, max_rows, C.byref(t))
        if rc == FM_ERR_BUSY:
            return None
        _check(rc)
        return int(t.value)

    def poll(self, max_items: int = 1024):
        arr = (L.ServerCompletion * max_items)()
        n = C.c_uint32()
        _check(L.lib().fm_server_poll(self._h, arr, max_items, C.byref(n)))
        return [arr[i] for i in range(n.value)]

    def table(self, ticket: int, rows: int, vertices: int, sorted: bool = True) -> np.ndarray:
        out = np.zeros((max(rows, 1), vertices), dtype=np.uint32)
        copied = C.c_uint64()
        _check(L.lib().fm_server_copy_table(self._h, ticket, _p32(out), rows, 1 if sorted else 0, C.byref(copied)))
        return out[: copied.value]

    def release(self, ticket: int) -> None:
        _check(L.lib().fm_server_release(self._h, ticket))

    def stats(self) -> L.ServerStats:
        st = L.ServerStats()
        _check(L.lib().fm_server_stats_get(self._h, C.byref(st)))
        return st

    def close(self) -> None:
        if self._h:
            L.lib().fm_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceBuffer:
    """A hipMalloc'd buffer owned from Python (for the intersection kernel)."""

    def __init__(self, nbytes: int):
        self.ptr = C.c_void_p()
        self.nbytes = int(nbytes)
        _check(L.lib().fm_device_alloc(self.nbytes, C.byref(self.ptr)))

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(max(a.nbytes, 4))
        if a.nbytes:
            _check(L.lib().fm_copy_to_device(b.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes))
        return b

    def to_array(self, dtype, count: int) -> np.ndarray:
        out = np.zeros(max(count, 1), dtype=dtype)
        nbytes = count * out.itemsize
        if nbytes:
            _check(L.lib().fm_copy_to_host(out.ctypes.data_as(C.c_void_p), self.ptr, nbytes))
        return out[:count]

    def free(self) -> None:
        if self.ptr:
            L.lib().fm_device_free(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def intersect_pairs(engine: Engine, data, a_off, a_len, b_off, b_len, out_off=None, out_words: int = 0,
                    buffers: dict | None = None, noalloc: bool = False) -> dict:
    """|A_i n B_i| (and the intersections when out_off is given) on the device.

    Lists are slices of `data` (u32, each slice sorted ascending). Returns
    {"counts", "out", "kernel_ms"}. `buffers` may hold device copies from an
    earlier call to skip the uploads (keys data/a_off/a_len/b_off/b_len).
    noalloc: run fm_intersect_pairs_noalloc (the library's fixed scratch)."""
    npairs = len(a_len)
    bufs = buffers if buffers is not None else {}
    if "data" not in bufs:
        bufs["data"] = DeviceBuffer.from_array(np.asarray(data, np.uint32))
        bufs["a_off"] = DeviceBuffer.from_array(np.asarray(a_off, np.uint64))
        bufs["a_len"] = DeviceBuffer.from_array(np.asarray(a_len, np.uint32))
        bufs["b_off"] = DeviceBuffer.from_array(np.asarray(b_off, np.uint64))
        bufs["b_len"] = DeviceBuffer.from_array(np.asarray(b_len, np.uint32))
    counts = DeviceBuffer(4 * max(npairs, 1))
    oo = DeviceBuffer.from_array(np.asarray(out_off, np.uint64)) if out_off is not None else None
    out = DeviceBuffer(4 * max(out_words, 1)) if out_off is not None else None
    ms = C.c_double()
    fn = L.lib().fm_intersect_pairs_noalloc if noalloc else L.lib().fm_intersect_pairs
    _check(fn(engine._h, bufs["data"].ptr, bufs["a_off"].ptr, bufs["a_len"].ptr, bufs["b_off"].ptr, bufs["b_len"].ptr,
              npairs, counts.ptr, oo.ptr if oo else None, out.ptr if out else None, C.byref(ms)))
    res = {"counts": counts.to_array(np.uint32, npairs), "kernel_ms": ms.value, "buffers": bufs}
    res["out"] = out.to_array(np.uint32, out_words) if out is not None else None
    return res


def match(query_path: str, data_path: str, **engine_kw) -> tuple[int, np.ndarray]:
    """The reference's ``match(query, data)``: count and sorted table."""
    g = Graph.load_g(data_path)
    q = Query.load_g(query_path)
    with Engine(**engine_kw) as e:
        r = e.match(g, q)
        return r.count(0), r.table(0)


__all__ = ["Graph", "Query", "Engine", "Result", "Plan", "QueryOutcome", "ForkMatchError", "match", "Server",
           "device_count", "FM_MODE_COUNT", "FM_MODE_TABLE"]
