"""ctypes binding of the gfx950 RBSC library (include/rbsc.h, include/rbsc_gen.h).

The shared library is built in-tree (``rbsc_amd/librbsc.so``) by
``python -c "import __graft_entry__ as g; g.build()"`` or ``make -C rbsc_amd/csrc``.
Loading fails loudly when it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librbsc.so")
WEBGRAPH_TO_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bin", "webgraph-to")
CLI_PATH = os.path.join(_HERE, "bin", "rbsc")

RBSC_OK = 0
ERRORS = {
    -1: "RBSC_ERR_INVALID",
    -2: "RBSC_ERR_NOMEM",
    -3: "RBSC_ERR_HIP",
    -4: "RBSC_ERR_BUDGET",
    -5: "RBSC_ERR_INPUT",
    -6: "RBSC_ERR_KERNEL",
    -7: "RBSC_ERR_NODEVICE",
}
MODE_STREAM, MODE_QUOTIENT, MODE_QUOTIENT_AGGRESSIVE = 0, 1, 2
FLAG_HASHED64 = 1
FLAG_XCD = 2
FLAG_SPLIT = 4
FLAG_SPLIT16 = 8
FLAG_TIESMALL = 16
FLAG_FULLSORT = 32
FLAG_NOBINS = 64
FLAG_COLOUR_ORDER = 128
FLAG_PACKED = 512
FLAG_UNPACKED = 1024
FLAG_SORTRANK = 2048
FLAG_FULLROUNDS = 4096
FLAG_GATHER_FIRST = 8192
BIN_SLOT_ORDER = 1
BIN_HASHED = 2


class BinOpts(C.Structure):
    _fields_ = [("hot_heads", C.c_uint32), ("rows_per_bin", C.c_uint32),
                ("chunk_slots", C.c_uint64), ("scatter_lines", C.c_uint32),
                ("flags", C.c_uint32), ("dense_heads", C.c_uint32), ("reserved", C.c_uint32)]

u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)


class Graph(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_uint32),
        ("n_edges", C.c_uint64),
        ("row_offsets", u64p),
        ("src", u32p),
        ("dst", u32p),
        ("weight", u64p),
        ("init_color", u32p),
    ]


class Opts(C.Structure):
    _fields_ = [
        ("max_batch_edges", C.c_uint64),
        ("mode", C.c_uint32),
        ("max_rounds", C.c_uint32),
        ("verify", C.c_uint32),
        ("flags", C.c_uint32),
    ]


class Result(C.Structure):
    _fields_ = [
        ("nodes", C.c_uint64),
        ("done", C.c_uint32),
        ("rounds", C.c_uint32),
        ("time_s", C.c_double),
        ("device_ms", C.c_double),
        ("verified", C.c_int32),
        ("passes", C.c_uint32),
        ("edges_per_round", C.c_uint64),
    ]


class HostGraph(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_uint32),
        ("n_edges", C.c_uint64),
        ("row_offsets", u64p),
        ("dst", u32p),
        ("weight", u64p),
        ("init_color", u32p),
    ]


class Batch(C.Structure):
    _fields_ = [
        ("n_graphs", C.c_uint32),
        ("node_offsets", u64p),
        ("line_offsets", u64p),
        ("src", u32p),
        ("dst", u32p),
        ("weight", u64p),
        ("init_color", u32p),
    ]


class BatchOutputs(C.Structure):
    _fields_ = [("blocks", u32p), ("rounds", u32p), ("labels", u32p), ("colors", u32p),
                ("block_sizes", u32p)]


class HostBatch(C.Structure):
    _fields_ = [
        ("n_graphs", C.c_uint32),
        ("node_offsets", u64p),
        ("line_offsets", u64p),
        ("src", u32p),
        ("dst", u32p),
    ]


class ServerOpts(C.Structure):
    _fields_ = [
        ("device", C.c_int),
        ("hbm_budget", C.c_uint64),
        ("workers", C.c_uint32),
        ("batch_max", C.c_uint32),
        ("keep_partition", C.c_uint32),
    ]


class RequestResult(C.Structure):
    _fields_ = [
        ("ticket", C.c_uint64),
        ("status", C.c_int32),
        ("nodes", C.c_uint32),
        ("rounds", C.c_uint32),
        ("batched", C.c_uint32),
        ("n_nodes", C.c_uint32),
        ("n_edges", C.c_uint64),
        ("latency_s", C.c_double),
        ("queue_s", C.c_double),
        ("service_s", C.c_double),
        ("arrival_s", C.c_double),
        ("done_s", C.c_double),
    ]


class ServerStats(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in (
        "submitted", "completed", "failed", "batches", "batched_graphs", "in_flight",
        "admitted_bytes", "peak_admitted_bytes", "budget")]


PENDING = 1

# Every exported symbol of the C headers, with (restype, argtypes).
SIGNATURES = {
    "rbsc_last_error": (C.c_char_p, []),
    "rbsc_version": (C.c_char_p, []),
    "rbsc_host_alloc": (C.c_int, [C.c_uint64, C.POINTER(C.c_void_p)]),
    "rbsc_host_free": (C.c_int, [C.c_void_p]),
    "rbsc_ctx_create": (C.c_int, [C.c_int, C.c_uint64, C.POINTER(C.c_void_p)]),
    "rbsc_ctx_destroy": (C.c_int, [C.c_void_p]),
    "rbsc_refine": (C.c_int, [C.c_void_p, C.POINTER(Graph), C.POINTER(Opts), C.POINTER(Result), u32p]),
    "rbsc_graph_upload": (C.c_int, [C.c_void_p, C.POINTER(Graph), C.POINTER(C.c_void_p)]),
    "rbsc_graph_free": (C.c_int, [C.c_void_p, C.c_void_p]),
    "rbsc_graph_bytes": (C.c_uint64, [C.c_void_p]),
    "rbsc_graph_refine": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Opts), C.POINTER(Result), u32p, u32p]),
    "rbsc_graph_relayout": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32]),
    "rbsc_graph_lines_to_host": (C.c_int, [C.c_void_p, C.c_void_p]),
    "rbsc_ctx_profile": (C.c_int, [C.c_void_p, C.c_int]),
    "rbsc_ctx_profile_read": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), u64p, C.c_int]),
    "rbsc_ctx_stats": (C.c_int, [C.c_void_p, u64p, C.c_int]),
    "rbsc_graph_check_partition": (C.c_int, [C.c_void_p, C.c_void_p, u32p, C.POINTER(C.c_uint64)]),
    "rbsc_server_create": (C.c_int, [C.POINTER(ServerOpts), C.POINTER(C.c_void_p)]),
    "rbsc_server_destroy": (C.c_int, [C.c_void_p]),
    "rbsc_server_submit": (C.c_int, [C.c_void_p, C.POINTER(Graph), C.c_double,
                                     C.POINTER(C.c_uint64)]),
    "rbsc_server_submit_rmat": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint64, C.c_double,
                                          C.c_double, C.c_double, C.c_double, C.c_uint64,
                                          C.c_double, C.POINTER(C.c_uint64)]),
    "rbsc_server_submit_resident": (C.c_int, [C.c_void_p, C.c_void_p, C.c_double,
                                              C.POINTER(C.c_uint64)]),
    "rbsc_server_wait": (C.c_int, [C.c_void_p, C.c_uint64, C.c_double,
                                   C.POINTER(RequestResult), u32p]),
    "rbsc_server_wait_labels": (C.c_int, [C.c_void_p, C.c_uint64, C.c_double,
                                          C.POINTER(RequestResult), C.POINTER(C.c_void_p)]),
    "rbsc_server_labels_free": (C.c_int, [C.c_void_p]),
    "rbsc_server_get_stats": (C.c_int, [C.c_void_p, C.POINTER(ServerStats)]),
    "rbsc_batch_upload": (C.c_int, [C.c_void_p, C.POINTER(Batch), C.POINTER(C.c_void_p)]),
    "rbsc_batch_free": (C.c_int, [C.c_void_p, C.c_void_p]),
    "rbsc_batch_refine": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Opts), u32p, u32p, u32p, u32p,
                                    C.POINTER(Result)]),
    "rbsc_batch_refine_ex": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Opts),
                                       C.POINTER(BatchOutputs), C.POINTER(Result)]),
    "rbsc_batch_refine_device": (C.c_int, [C.c_void_p, C.POINTER(Batch), C.c_uint64, C.c_uint64,
                                           C.POINTER(Opts), C.c_void_p, C.c_void_p, C.c_void_p,
                                           C.c_void_p]),
    "rbsc_batch_autotune": (C.c_int, [C.c_void_p, C.c_void_p, u32p, u32p, C.POINTER(C.c_double)]),
    "rbsc_batch_set_launch": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32]),
    "rbsc_batch_launch_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32)]),
    "rbsc_batch_set_digits": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32)]),
    "rbsc_refine_batch": (C.c_int, [C.c_void_p, C.POINTER(Batch), C.POINTER(Opts), u32p, u32p, u32p,
                                    C.POINTER(Result)]),
    "rbsc_bvgraph_load": (C.c_int, [C.c_char_p, C.POINTER(HostGraph)]),
    "rbsc_bvgraph_check_offsets": (C.c_int, [C.c_char_p, u64p, u64p]),
    "rbsc_host_graph_write_text": (C.c_int, [C.POINTER(HostGraph), C.c_char_p, C.c_int]),
    "rbsc_host_batch_free": (None, [C.POINTER(HostBatch)]),
    "rbsc_gen_batch_mixed": (C.c_int, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64,
                                       C.POINTER(HostBatch)]),
    "rbsc_gen_rmat_device": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint64, C.c_double, C.c_double,
                                       C.c_double, C.c_double, C.c_uint64, C.POINTER(C.c_void_p)]),
    "rbsc_graph_dims": (C.c_int, [C.c_void_p, u32p, u64p]),
    "rbsc_graph_rows": (C.c_uint32, [C.c_void_p]),
    "rbsc_graph_stored_lines": (C.c_uint64, [C.c_void_p]),
    "rbsc_ctx_set_dedup": (C.c_int, [C.c_void_p, C.c_int]),
    "rbsc_graph_split_hot": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32]),
    "rbsc_graph_bin": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "rbsc_graph_unbin": (C.c_int, [C.c_void_p, C.c_void_p]),
    "rbsc_graph_bin_info": (C.c_int, [C.c_void_p, u64p]),
    "rbsc_graph_download": (C.c_int, [C.c_void_p, C.c_void_p, u64p, u32p]),
    "rbsc_host_graph_free": (None, [C.POINTER(HostGraph)]),
    "rbsc_gen_rmat": (C.c_int, [C.c_uint32, C.c_uint64, C.c_double, C.c_double, C.c_double,
                                C.c_double, C.c_uint64, C.POINTER(HostGraph)]),
    "rbsc_gen_ba": (C.c_int, [C.c_uint32, C.c_uint32, C.c_double, C.c_uint32, C.c_uint64,
                              C.POINTER(HostGraph)]),
    "rbsc_gen_er_undirected": (C.c_int, [C.c_uint32, C.c_double, C.c_uint64, C.POINTER(HostGraph)]),
    "rbsc_gen_ba_undirected": (C.c_int, [C.c_uint32, C.c_uint32, C.c_uint64, C.POINTER(HostGraph)]),
    "rbsc_gen_regular_undirected": (C.c_int, [C.c_uint32, C.c_uint32, C.c_uint64, C.POINTER(HostGraph)]),
}

_lib = None


def load() -> C.CDLL:
    """Load librbsc.so (once).  Raises RuntimeError when it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `make -C rbsc_amd/csrc` "
            "(there is no CPU fallback for the RBSC path)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class RbscError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


def check(code: int) -> None:
    if code != RBSC_OK:
        msg = load().rbsc_last_error().decode(errors="replace")
        raise RbscError(code, msg)
