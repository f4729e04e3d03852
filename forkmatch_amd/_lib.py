"""ctypes bindings of libforkmatch.so (declarations in include/forkmatch.h).

The library must have been built in-tree (``python -m forkmatch_amd.build``);
loading fails loudly otherwise. There is no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libforkmatch.so"

u32, u64, i32, f64 = C.c_uint32, C.c_uint64, C.c_int32, C.c_double
P = C.POINTER
vp = C.c_void_p
cstr = C.c_char_p

MAXN = 32


class GraphInfo(C.Structure):
    _fields_ = [("vertices", u32), ("labels", u32), ("edges", u64), ("symmetric", u32),
                ("header_edges", u32), ("device_bytes", u64)]


class PlanInfo(C.Structure):
    _fields_ = [("vertices", u32), ("nreqs", u32), ("max_reqs_per_depth", u32)]


class EngineConfig(C.Structure):
    _fields_ = [("soln_prealloc", u64), ("output_prealloc", u64), ("bank_size", u32),
                ("output_page_bytes", u32), ("queue_capacity", u32), ("max_queries", u32),
                ("device", i32), ("reserved0", u32), ("prealloc", u64), ("reserved", u32 * 4)]


class EngineInfo(C.Structure):
    _fields_ = [("compute_units", u32), ("waves_per_cu", u32), ("grid_waves", u32), ("banks", u32),
                ("bank_bytes", u64), ("output_pages", u32), ("output_page_bytes", u32),
                ("device_name", C.c_char * 64), ("arena_bytes", u64)]


class MatchOpts(C.Structure):
    _fields_ = [("mode", u32), ("instrument", u32), ("max_rows", u64), ("max_count", u64),
                ("timeout_ms", u32), ("grid_waves", u32), ("task_rows", u32), ("root_chunk", u32),
                ("plan", u32), ("max_partials", u64), ("symmetry", u32), ("waves_per_query", u32),
                ("tile_parents", u32), ("reserved", u32 * 1)]


class QueryResult(C.Structure):
    _fields_ = [("count", u64), ("rows", u64), ("status", u32), ("flags", u32), ("vertices", u32),
                ("reserved", u32), ("alg_bytes", u64), ("row_lookups", u64), ("partials", u64 * MAXN),
                ("candidates", u64), ("probes", u64)]


class LaunchStats(C.Structure):
    _fields_ = [("kernel_ms", f64), ("total_ms", f64), ("tasks_pushed", u64), ("root_claims", u64),
                ("banks_allocated", u64), ("banks_reused", u64), ("alloc_failures", u64),
                ("pages_allocated", u64), ("tasks_split", u64), ("requeues", u64), ("grid_waves", u32),
                ("status", u32), ("busy_frac", f64)]


class ServerConfig(C.Structure):
    _fields_ = [("slots", u32), ("slot_pages", u32), ("bank_budget", u32), ("idle_ms", u32), ("root_chunk", u32),
                ("task_rows", u32), ("table", u32), ("max_launch_ms", u32), ("plan", u32), ("slice_ms", u32),
                ("symmetry", u32), ("reserved", u32 * 5)]


class ServerCompletion(C.Structure):
    _fields_ = [("ticket", u32), ("status", u32), ("flags", u32), ("vertices", u32), ("count", u64), ("rows", u64),
                ("submit_ms", f64), ("done_ms", f64)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", u64), ("completed", u64), ("refused", u64), ("launches", u64), ("in_flight", u32),
                ("live_banks", u32), ("dropped_tasks", u64), ("refused_memory", u64)]


# name -> (restype, argtypes); every function declared in include/forkmatch.h
SIGNATURES = {
    "fm_last_error": (cstr, []),
    "fm_version": (cstr, []),
    "fm_device_count": (C.c_int, []),
    "fm_graph_parse_g": (C.c_int, [cstr, P(vp)]),
    "fm_graph_from_edges": (C.c_int, [u32, u64, P(u32), P(u32), P(u32), P(vp)]),
    "fm_graph_upload": (C.c_int, [vp]),
    "fm_graph_load_g": (C.c_int, [cstr, P(vp)]),
    "fm_graph_info_get": (C.c_int, [vp, P(GraphInfo)]),
    "fm_graph_copy_csr": (C.c_int, [vp, C.c_int, P(u32), P(u32), P(u32)]),
    "fm_graph_copy_device_csr": (C.c_int, [vp, C.c_int, P(u32), P(u32)]),
    "fm_graph_save_bin": (C.c_int, [vp, cstr]),
    "fm_graph_load_bin": (C.c_int, [cstr, P(vp)]),
    "fm_graph_write_g": (C.c_int, [vp, cstr]),
    "fm_graph_destroy": (None, [vp]),
    "fm_gen_erdos_renyi": (C.c_int, [u32, u64, u32, C.c_int, f64, u32, u64, P(vp)]),
    "fm_gen_rmat": (C.c_int, [u32, u64, f64, f64, f64, u32, C.c_int, f64, u32, u64, P(vp)]),
    "fm_gen_query_walk": (C.c_int, [vp, u32, C.c_int, u32, u64, P(vp)]),
    "fm_query_load_g": (C.c_int, [cstr, P(vp)]),
    "fm_query_from_edges": (C.c_int, [u32, u32, P(u32), P(u32), P(u32), P(vp)]),
    "fm_query_write_g": (C.c_int, [vp, cstr]),
    "fm_query_info": (C.c_int, [vp, P(u32), P(u32)]),
    "fm_query_edges": (C.c_int, [vp, P(u32), P(u32), P(u32)]),
    "fm_query_plan": (C.c_int, [vp, P(PlanInfo), P(u32), P(u32), P(u32), P(u32), P(u32)]),
    "fm_query_plan_on": (C.c_int, [vp, vp, u32, P(PlanInfo), P(u32), P(u32), P(u32), P(u32), P(u32)]),
    "fm_query_symmetry": (C.c_int, [vp, u32, P(u32), P(u32), u32, P(u32), P(u32), u32]),
    "fm_query_set_count_only": (C.c_int, [vp, u32]),
    "fm_query_destroy": (None, [vp]),
    "fm_engine_create": (C.c_int, [P(EngineConfig), P(vp)]),
    "fm_engine_destroy": (None, [vp]),
    "fm_engine_info_get": (C.c_int, [vp, P(EngineInfo)]),
    "fm_match": (C.c_int, [vp, vp, vp, P(MatchOpts), P(vp)]),
    "fm_match_batch": (C.c_int, [vp, vp, P(vp), u32, P(MatchOpts), P(vp)]),
    "fm_result_queries": (u32, [vp]),
    "fm_result_query": (C.c_int, [vp, u32, P(QueryResult)]),
    "fm_result_stats": (C.c_int, [vp, P(LaunchStats)]),
    "fm_result_copy_table": (C.c_int, [vp, u32, P(u32), u64, C.c_int, P(u64)]),
    "fm_result_destroy": (None, [vp]),
    "fm_server_create": (C.c_int, [vp, vp, P(ServerConfig), P(vp)]),
    "fm_server_submit": (C.c_int, [vp, vp, u64, P(u32)]),
    "fm_server_poll": (C.c_int, [vp, P(ServerCompletion), u32, P(u32)]),
    "fm_server_copy_table": (C.c_int, [vp, u32, P(u32), u64, C.c_int, P(u64)]),
    "fm_server_release": (C.c_int, [vp, u32]),
    "fm_server_stats_get": (C.c_int, [vp, P(ServerStats)]),
    "fm_server_destroy": (None, [vp]),
    "fm_intersect_pairs": (C.c_int, [vp, vp, vp, vp, vp, vp, u32, vp, vp, vp, P(f64)]),
    "fm_intersect_pairs_noalloc": (C.c_int, [vp, vp, vp, vp, vp, vp, u32, vp, vp, vp, P(f64)]),
    "fm_device_alloc": (C.c_int, [u64, P(vp)]),
    "fm_device_free": (C.c_int, [vp]),
    "fm_copy_to_device": (C.c_int, [vp, vp, u64]),
    "fm_copy_to_host": (C.c_int, [vp, vp, u64]),
}

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is not built; run `python -m forkmatch_amd.build`")
        handle = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib
