"""rbsc_amd -- MI355X-native RBSC refinement (coarsest backward-equivalence /
1-WL colour refinement with edge multiplicities), a drop-in for the
reference ``rbsc`` tool (src/rbsc/rbsc.cu of f-biondi/rbsc).

Host-side mirror of the reference's interface:

* ``Context.refine(graph, max_batch_edges=..., mode=...)`` is the reference
  binary's run: same input (lines ``src weight dst`` grouped by ``src``), the
  same edge-batch parameter (rbsc.cu:204) and the same three results
  ``time``, ``done``, ``nodes`` (rbsc.cu:492-494), plus the canonical
  first-occurrence partition (rbsc.cu:431-445).
* ``Context.upload`` keeps a graph resident in HBM for repeated refinement.

All arithmetic runs in the HIP library ``librbsc.so``; this module only
marshals buffers.  It raises when the library is missing.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import _lib
from ._lib import (FLAG_COLOUR_ORDER, FLAG_GATHER_FIRST, FLAG_PACKED, FLAG_UNPACKED, FLAG_SORTRANK, FLAG_FULLROUNDS, FLAG_FULLSORT, FLAG_HASHED64, FLAG_NOBINS, FLAG_SPLIT, FLAG_XCD, MODE_QUOTIENT, MODE_QUOTIENT_AGGRESSIVE, MODE_STREAM,
                   RbscError, check, load)

__all__ = ["Context", "DeviceGraph", "HostGraph", "HostBatch", "DeviceBatch", "BatchResult",
           "PinnedArray",
           "Server", "RequestResult",
           "RefineResult", "RbscError", "load",
           "MODE_STREAM", "MODE_QUOTIENT", "MODE_QUOTIENT_AGGRESSIVE", "FLAG_HASHED64", "FLAG_XCD", "FLAG_SPLIT", "FLAG_FULLSORT", "FLAG_NOBINS", "FLAG_COLOUR_ORDER", "FLAG_PACKED", "FLAG_UNPACKED", "FLAG_SORTRANK", "FLAG_FULLROUNDS", "FLAG_GATHER_FIRST", "gen"]

_MODES = {"stream": MODE_STREAM, "quotient": MODE_QUOTIENT,
          "aggressive": MODE_QUOTIENT_AGGRESSIVE}


def _ptr(a: Optional[np.ndarray], ctype):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ctype))


@dataclass
class HostGraph:
    """A graph in the reference's line form.

    ``dst`` holds the head of every line; lines are grouped by source and
    given either as ``row_offsets`` (CSR, n_nodes + 1 entries) or as ``src``
    (COO).  ``weight`` None means all weights are 1; ``init_color`` None
    means the reference's initial colouring (rbsc.cu:51-63).
    """
    n_nodes: int
    dst: np.ndarray
    row_offsets: Optional[np.ndarray] = None
    src: Optional[np.ndarray] = None
    weight: Optional[np.ndarray] = None
    init_color: Optional[np.ndarray] = None

    def __post_init__(self):
        self.dst = np.ascontiguousarray(self.dst, dtype=np.uint32)
        if self.row_offsets is not None:
            self.row_offsets = np.ascontiguousarray(self.row_offsets, dtype=np.uint64)
        if self.src is not None:
            self.src = np.ascontiguousarray(self.src, dtype=np.uint32)
        if self.weight is not None:
            self.weight = np.ascontiguousarray(self.weight, dtype=np.uint64)
        if self.init_color is not None:
            self.init_color = np.ascontiguousarray(self.init_color, dtype=np.uint32)
        if self.row_offsets is None and self.src is None and len(self.dst):
            raise ValueError("HostGraph needs src or row_offsets")

    @property
    def n_edges(self) -> int:
        return int(len(self.dst))

    def sources(self) -> np.ndarray:
        """Source of every line (COO view)."""
        if self.src is not None:
            return self.src
        deg = np.diff(self.row_offsets.astype(np.int64))
        return np.repeat(np.arange(self.n_nodes, dtype=np.uint32), deg)

    def as_struct(self) -> _lib.Graph:
        g = _lib.Graph()
        g.n_nodes = self.n_nodes
        g.n_edges = self.n_edges
        g.row_offsets = _ptr(self.row_offsets, C.c_uint64)
        g.src = _ptr(self.src, C.c_uint32)
        g.dst = _ptr(self.dst, C.c_uint32)
        g.weight = _ptr(self.weight, C.c_uint64)
        g.init_color = _ptr(self.init_color, C.c_uint32)
        return g

    def to_text(self) -> bytes:
        """The reference's stdin format (rbsc.cu:168-180): N, E, lines "u w v"."""
        src = self.sources().astype(np.uint64)
        w = self.weight if self.weight is not None else np.ones(self.n_edges, np.uint64)
        body = "".join(f"{a} {b} {c}\n" for a, b, c in zip(src.tolist(), w.tolist(), self.dst.tolist()))
        return f"{self.n_nodes}\n{self.n_edges}\n{body}".encode()

    def write_text(self, path: str, fmt: str = "lines") -> None:
        """Writes the graph as the reference's stdin text ("lines") or as
        Valmari's MDP text ("valmari"), in the library (fast for large graphs)."""
        ro = self.row_offsets
        if ro is None:
            src = self.sources()
            if len(src) and np.any(np.diff(src.astype(np.int64)) < 0):
                raise ValueError("write_text needs lines in ascending source order")
            ro = np.concatenate([[0], np.cumsum(np.bincount(src, minlength=self.n_nodes))])
        ro = np.ascontiguousarray(ro, np.uint64)
        hg = _lib.HostGraph()
        hg.n_nodes = self.n_nodes
        hg.n_edges = self.n_edges
        hg.row_offsets = _ptr(ro, C.c_uint64)
        hg.dst = _ptr(self.dst, C.c_uint32)
        hg.weight = _ptr(self.weight, C.c_uint64)
        hg.init_color = _ptr(self.init_color, C.c_uint32)
        check(load().rbsc_host_graph_write_text(C.byref(hg), path.encode(),
                                                0 if fmt == "lines" else 1))

    def to_valmari_text(self) -> bytes:
        """Valmari's MDP format, 1-indexed (valmari-adapter:9-16; MDP_read.cc:51-64),
        with the initial colouring as initial blocks (MDP_read.cc:140-150)."""
        src = self.sources().astype(np.uint64) + 1
        w = self.weight if self.weight is not None else np.ones(self.n_edges, np.uint64)
        dst = self.dst.astype(np.uint64) + 1
        blocks = []
        if self.init_color is not None:
            labels = np.unique(self.init_color)
            for lab in labels[1:]:
                members = np.nonzero(self.init_color == lab)[0] + 1
                blocks.append(" ".join(map(str, members.tolist())) + " 0\n")
        nb = 1 + len(blocks)
        body = "".join(f"{a} {b} {c}\n" for a, b, c in zip(src.tolist(), w.tolist(), dst.tolist()))
        return (f"{self.n_nodes} 0 0 {self.n_edges} {nb}\n" + body + "".join(blocks)).encode()


@dataclass
class RefineResult:
    nodes: int
    done: int
    rounds: int
    time_s: float
    device_ms: float
    verified: int
    passes: int
    partition: Optional[np.ndarray] = None
    colors: Optional[np.ndarray] = None


def _result(r: _lib.Result, partition=None, colors=None) -> RefineResult:
    return RefineResult(int(r.nodes), int(r.done), int(r.rounds), float(r.time_s),
                        float(r.device_ms), int(r.verified), int(r.passes), partition, colors)


def _opts(max_batch_edges=0, mode="stream", max_rounds=0, verify=False, flags=0) -> _lib.Opts:
    o = _lib.Opts()
    o.flags = int(flags)
    o.max_batch_edges = int(max_batch_edges)
    o.mode = _MODES[mode] if isinstance(mode, str) else int(mode)
    o.max_rounds = int(max_rounds)
    o.verify = 1 if verify else 0
    return o


class DeviceGraph:
    """A graph resident in HBM (handle to an ``rbsc_dgraph``)."""

    def __init__(self, ctx: "Context", handle: C.c_void_p):
        self.ctx = ctx
        self.handle = handle

    @property
    def dims(self):
        n = C.c_uint32()
        e = C.c_uint64()
        check(load().rbsc_graph_dims(self.handle, C.byref(n), C.byref(e)))
        return int(n.value), int(e.value)

    @property
    def rows(self) -> int:
        """Rows the signature pass visits (identical rows stored once)."""
        return int(load().rbsc_graph_rows(self.handle))

    @property
    def stored_lines(self) -> int:
        """Lines the signature pass visits."""
        return int(load().rbsc_graph_stored_lines(self.handle))

    @property
    def nbytes(self) -> int:
        return int(load().rbsc_graph_bytes(self.handle))

    def refine(self, max_rounds=0, verify=False, partition=True, colors=False, mode="stream",
               max_batch_edges=0, flags=0) -> RefineResult:
        n, _ = self.dims
        part = np.empty(n, np.uint32) if partition else None
        col = np.empty(n, np.uint32) if colors else None
        r = _lib.Result()
        o = _opts(max_batch_edges, mode, max_rounds, verify, flags)
        check(load().rbsc_graph_refine(self.ctx.handle, self.handle, C.byref(o), C.byref(r),
                                       _ptr(part, C.c_uint32), _ptr(col, C.c_uint32)))
        return _result(r, part, col)

    def check_partition(self, colors: np.ndarray) -> int:
        """Exact, hash-free stability check of a colouring (one label per
        node).  Returns the number of violations: 0 iff stable."""
        col = np.ascontiguousarray(colors, np.uint32)
        n, _ = self.dims
        if len(col) != n:
            raise ValueError("one colour per node expected")
        bad = C.c_uint64()
        check(load().rbsc_graph_check_partition(self.ctx.handle, self.handle,
                                                _ptr(col, C.c_uint32), C.byref(bad)))
        return int(bad.value)

    def relayout(self, n_parts: int):
        """Group lines by destination (head mod n_parts) for cache-resident
        gathers; results are unchanged."""
        check(load().rbsc_graph_relayout(self.ctx.handle, self.handle, n_parts))

    def split_hot(self, hot_heads: int):
        """Two-pass layout: lines into the hot_heads most-gathered nodes
        first (their colours stay cache resident); results are unchanged."""
        check(load().rbsc_graph_split_hot(self.ctx.handle, self.handle, int(hot_heads)))

    def bin(self, hot_heads: int = 0, rows_per_bin: int = 0, chunk_slots: int = 0,
            scatter_lines: int = 0, slot_order: bool = False, hashed: bool = False,
            dense_heads: int = 0):
        """Build the binned signature layout (row sums in LDS per bin of
        rows, hot heads gathered in head order, the other colours scattered
        to their lines' slots in chunks of lines); later refinements use it.
        0 picks the defaults (2^22 hot heads, 8192 rows per bin, 2^19 slots
        per chunk, 16384 lines per scatter workgroup staged through LDS, 2^21
        dense heads).  Hot heads from dense_heads on are warm: their colours
        are gathered per XCD share in a pass of their own.
        hashed: keep per-node hashed u64 colours (no hashing per line)
        instead of gathering raw u32 colours.  Results are unchanged."""
        o = _lib.BinOpts(int(hot_heads), int(rows_per_bin), int(chunk_slots), int(scatter_lines),
                         (_lib.BIN_SLOT_ORDER if slot_order else 0) | (_lib.BIN_HASHED if hashed else 0),
                         int(dense_heads), 0)
        check(load().rbsc_graph_bin(self.ctx.handle, self.handle, C.byref(o)))

    def unbin(self):
        check(load().rbsc_graph_unbin(self.ctx.handle, self.handle))

    @property
    def bin_info(self) -> dict:
        """Sizes of the binned layout (all zero when the graph has none)."""
        v = (C.c_uint64 * 10)()
        check(load().rbsc_graph_bin_info(self.handle, v))
        return {"binned": bool(v[0]), "bins": int(v[1]), "hot_lines": int(v[2]),
                "cold_lines": int(v[3]), "chunks": int(v[4]), "split_bins": int(v[5]),
                "bytes": int(v[6]), "hot_heads": int(v[7] & 0xFFFFFFFF),
                "rows_per_bin": 1 << int((v[7] >> 32) & 0xFF) if v[0] else 0,
                "hashed": bool((v[7] >> 40) & 1), "dense_heads": int(v[8]),
                "warm_lines": int(v[9])}

    def lines_to_host(self):
        """Keep the line arrays in pinned host memory; refinements then stream
        them in slices of ``max_batch_edges`` lines."""
        check(load().rbsc_graph_lines_to_host(self.ctx.handle, self.handle))

    def download(self) -> HostGraph:
        n, e = self.dims
        ro = np.empty(n + 1, np.uint64)
        dst = np.empty(max(e, 1), np.uint32)
        check(load().rbsc_graph_download(self.ctx.handle, self.handle, _ptr(ro, C.c_uint64),
                                         _ptr(dst, C.c_uint32)))
        return HostGraph(n, dst[:e], row_offsets=ro)

    def free(self):
        """Give the graph's device memory back (through the context's stream,
        or, once the context is closed, the null stream)."""
        if self.handle:
            load().rbsc_graph_free(self.ctx.handle, self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


@dataclass
class HostBatch:
    """Independent graphs in one set of arrays: graph g owns nodes
    [node_offsets[g], node_offsets[g+1]) and lines [line_offsets[g],
    line_offsets[g+1]); src/dst are local ids; lines grouped by source."""
    node_offsets: np.ndarray
    line_offsets: np.ndarray
    src: np.ndarray
    dst: np.ndarray
    weight: Optional[np.ndarray] = None
    init_color: Optional[np.ndarray] = None

    def __post_init__(self):
        self.node_offsets = np.ascontiguousarray(self.node_offsets, np.uint64)
        self.line_offsets = np.ascontiguousarray(self.line_offsets, np.uint64)
        self.src = np.ascontiguousarray(self.src, np.uint32)
        self.dst = np.ascontiguousarray(self.dst, np.uint32)
        if self.weight is not None:
            self.weight = np.ascontiguousarray(self.weight, np.uint64)
        if self.init_color is not None:
            self.init_color = np.ascontiguousarray(self.init_color, np.uint32)

    @property
    def n_graphs(self) -> int:
        return len(self.node_offsets) - 1

    @property
    def n_nodes(self) -> int:
        return int(self.node_offsets[-1])

    def graph(self, g: int) -> HostGraph:
        a, b = int(self.line_offsets[g]), int(self.line_offsets[g + 1])
        na, nb = int(self.node_offsets[g]), int(self.node_offsets[g + 1])
        return HostGraph(nb - na, self.dst[a:b], src=self.src[a:b],
                         weight=None if self.weight is None else self.weight[a:b],
                         init_color=None if self.init_color is None else self.init_color[na:nb])

    @staticmethod
    def from_graphs(graphs) -> "HostBatch":
        no, lo, src, dst, w, ic = [0], [0], [], [], [], []
        for g in graphs:
            src.append(g.sources())
            dst.append(g.dst)
            w.append(g.weight)
            ic.append(g.init_color)
            no.append(no[-1] + g.n_nodes)
            lo.append(lo[-1] + g.n_edges)
        weight = None if all(x is None for x in w) else np.concatenate(
            [x if x is not None else np.ones(len(d), np.uint64) for x, d in zip(w, dst)])
        if any(x is None for x in ic) and not all(x is None for x in ic):
            raise ValueError("initial colours must be given for every graph of a batch or none")
        init = None if not ic or ic[0] is None else np.concatenate(ic)
        return HostBatch(np.array(no), np.array(lo),
                         np.concatenate(src) if src else np.zeros(0, np.uint32),
                         np.concatenate(dst) if dst else np.zeros(0, np.uint32), weight, init)

    def as_struct(self) -> _lib.Batch:
        b = _lib.Batch()
        b.n_graphs = self.n_graphs
        b.node_offsets = _ptr(self.node_offsets, C.c_uint64)
        b.line_offsets = _ptr(self.line_offsets, C.c_uint64)
        b.src = _ptr(self.src, C.c_uint32)
        b.dst = _ptr(self.dst, C.c_uint32)
        b.weight = _ptr(self.weight, C.c_uint64)
        b.init_color = _ptr(self.init_color, C.c_uint32)
        return b


@dataclass
class BatchResult:
    blocks: np.ndarray
    rounds: np.ndarray
    labels: Optional[np.ndarray]
    colors: Optional[np.ndarray]
    total: RefineResult
    block_sizes: Optional[np.ndarray] = None
    node_offsets: Optional[np.ndarray] = None

    def histogram(self, g: int) -> np.ndarray:
        """Graph g's final colour histogram: its block sizes in descending
        order (one entry per block)."""
        if self.block_sizes is None:
            raise ValueError("refine(histogram=True) computes the histograms")
        a = int(self.node_offsets[g])
        return self.block_sizes[a:a + int(self.blocks[g])]


class _PinnedBlock:
    """Owner of one ``rbsc_host_alloc`` block, exposed to numpy through
    ``__array_interface__``: every array made from it (and every view of
    those) keeps the block as its base, so the block is released by
    ``rbsc_host_free`` only when the last array over it is gone."""

    def __init__(self, n: int, dtype=np.uint32):
        n = max(int(n), 1)
        dt = np.dtype(dtype)
        p = C.c_void_p()
        check(load().rbsc_host_alloc(dt.itemsize * n, C.byref(p)))
        self._p = p
        self.__array_interface__ = {"shape": (n,), "typestr": dt.str, "data": (p.value, False),
                                    "version": 3}

    def __del__(self):
        try:
            if self._p:
                load().rbsc_host_free(self._p)
                self._p = None
        except Exception:
            pass


class PinnedArray:
    """A numpy array (uint32 unless ``dtype`` says otherwise) over page-locked
    host memory (``rbsc_host_alloc``): copies between it and the device are
    not staged through pageable memory.

    ``free()`` drops this object's hold on the memory; arrays handed out
    earlier (``.array`` and its views) keep it alive until they are gone."""

    def __init__(self, n: int, dtype=np.uint32):
        self.array = np.asarray(_PinnedBlock(n, dtype))

    def free(self):
        self.array = None

    def __del__(self):
        self.free()


class DeviceBatch:
    """A batch resident in HBM (``rbsc_dbatch``)."""

    def __init__(self, ctx: "Context", handle, host: HostBatch):
        self.ctx = ctx
        self.handle = handle
        self.n_graphs = host.n_graphs
        self.n_nodes = host.n_nodes
        self.node_offsets = host.node_offsets
        self._pinned = None

    def refine(self, max_rounds=0, labels=True, colors=False, pinned=False,
               flags=0, histogram=False) -> BatchResult:
        """pinned: download the labels into page-locked memory shared by the
        batch and the returned array (faster); that array is reused, and
        overwritten, by the next pinned refine of this batch, and stays
        valid after the batch is freed (the memory is released with the
        last array over it)."""
        blocks = np.empty(max(self.n_graphs, 1), np.uint32)
        rounds = np.empty(max(self.n_graphs, 1), np.uint32)
        if labels and pinned:
            if self._pinned is None:
                self._pinned = PinnedArray(self.n_nodes)
            lab = self._pinned.array
        else:
            lab = np.empty(max(self.n_nodes, 1), np.uint32) if labels else None
        col = np.empty(max(self.n_nodes, 1), np.uint32) if colors else None
        sizes = np.empty(max(self.n_nodes, 1), np.uint32) if histogram else None
        r = _lib.Result()
        o = _opts(max_rounds=max_rounds, flags=flags)
        out = _lib.BatchOutputs(_ptr(blocks, C.c_uint32), _ptr(rounds, C.c_uint32),
                                _ptr(lab, C.c_uint32), _ptr(col, C.c_uint32),
                                _ptr(sizes, C.c_uint32))
        check(load().rbsc_batch_refine_ex(self.ctx.handle, self.handle, C.byref(o), C.byref(out),
                                          C.byref(r)))
        n, g = self.n_nodes, self.n_graphs
        return BatchResult(blocks[:g], rounds[:g], None if lab is None else lab[:n],
                           None if col is None else col[:n], _result(r),
                           None if sizes is None else sizes[:n], self.node_offsets)

    def autotune(self):
        t = C.c_uint32()
        w = C.c_uint32()
        ms = C.c_double()
        check(load().rbsc_batch_autotune(self.ctx.handle, self.handle, C.byref(t), C.byref(w),
                                         C.byref(ms)))
        return int(t.value), int(w.value), float(ms.value)

    def set_launch(self, threads: int, wgs: int):
        check(load().rbsc_batch_set_launch(self.ctx.handle, self.handle, threads, wgs))

    def set_digits(self, digit_bits):
        """LDS radix digit width (4 or 8 bits) of each size tier, largest first."""
        v = (C.c_uint32 * 4)(*[int(x) for x in digit_bits])
        check(load().rbsc_batch_set_digits(self.handle, v))

    @property
    def launch_info(self) -> dict:
        """The persistent launches' parameters (set or autotuned)."""
        v = (C.c_uint32 * 10)()
        check(load().rbsc_batch_launch_info(self.handle, v))
        return {"threads": int(v[0]), "workgroups": int(v[1]),
                "digit_bits": [int(x) for x in v[2:6]], "tier_graphs": [int(x) for x in v[6:10]]}

    def free(self):
        if self._pinned is not None:
            self._pinned.free()
            self._pinned = None
        if self.handle:
            load().rbsc_batch_free(self.ctx.handle, self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """One HIP device context (``rbsc_ctx``)."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        lib = load()
        h = C.c_void_p()
        check(lib.rbsc_ctx_create(device, hbm_budget, C.byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            load().rbsc_ctx_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    PHASES = ("signature", "gather", "sort", "rank", "setup", "finish", "scatter")

    def profile(self, enable: bool = True):
        check(load().rbsc_ctx_profile(self.handle, 1 if enable else 0))

    def profile_read(self, reset: bool = True) -> dict:
        ms = (C.c_double * len(self.PHASES))()
        n = (C.c_uint64 * len(self.PHASES))()
        check(load().rbsc_ctx_profile_read(self.handle, ms, n, 1 if reset else 0))
        return {p: (float(ms[i]), int(n[i])) for i, p in enumerate(self.PHASES)}

    def stats(self, reset: bool = False) -> dict:
        """Node-side counters: rounds redone with a full sort, rounds with tie
        runs too long for the in-kernel sorts, and last rounds confirmed
        without ranking (see FLAG_FULLROUNDS)."""
        v = (C.c_uint64 * 3)()
        check(load().rbsc_ctx_stats(self.handle, v, 1 if reset else 0))
        return {"full_sort_redos": int(v[0]), "long_run_rounds": int(v[1]),
                "confirmed_rounds": int(v[2])}

    def refine(self, g: HostGraph, max_batch_edges=0, mode="stream", max_rounds=0,
               verify=False, partition=True, flags=0, partition_out=None) -> RefineResult:
        """The drop-in's whole run (rbsc_refine).  partition_out: an array of
        at least n_nodes uint32 to receive the partition (a PinnedArray's is
        written by the device directly), else one is allocated."""
        st = g.as_struct()
        o = _opts(max_batch_edges, mode, max_rounds, verify, flags)
        if partition and partition_out is not None:
            if (partition_out.dtype != np.uint32 or not partition_out.flags.c_contiguous
                    or len(partition_out) < g.n_nodes):
                raise ValueError("partition_out: contiguous uint32, n_nodes entries")
            part = partition_out[:g.n_nodes]
        else:
            part = np.empty(g.n_nodes, np.uint32) if partition else None
        r = _lib.Result()
        check(load().rbsc_refine(self.handle, C.byref(st), C.byref(o), C.byref(r),
                                 _ptr(part, C.c_uint32)))
        return _result(r, part)

    def upload(self, g: HostGraph) -> DeviceGraph:
        st = g.as_struct()
        h = C.c_void_p()
        check(load().rbsc_graph_upload(self.handle, C.byref(st), C.byref(h)))
        return DeviceGraph(self, h)

    def set_dedup(self, on: bool) -> None:
        """Store identical rows once in graphs built from now on (default on)."""
        check(load().rbsc_ctx_set_dedup(self.handle, 1 if on else 0))

    def upload_batch(self, b: HostBatch) -> DeviceBatch:
        st = b.as_struct()
        h = C.c_void_p()
        check(load().rbsc_batch_upload(self.handle, C.byref(st), C.byref(h)))
        return DeviceBatch(self, h, b)

    def gen_rmat(self, scale, n_draws, a=0.57, b=0.19, c=0.19, clone_frac=0.0,
                 seed=1) -> DeviceGraph:
        h = C.c_void_p()
        check(load().rbsc_gen_rmat_device(self.handle, scale, n_draws, a, b, c, clone_frac,
                                          seed, C.byref(h)))
        return DeviceGraph(self, h)


from . import gen  # noqa: E402  (uses the classes above)


@dataclass
class RequestResult:
    ticket: int
    status: int
    nodes: int
    rounds: int
    batched: bool
    n_nodes: int
    n_edges: int
    latency_s: float
    queue_s: float
    service_s: float
    arrival_s: float
    done_s: float
    partition: Optional[np.ndarray] = None


class _ServerLabels:
    """Owner of a partition buffer from ``rbsc_server_wait_labels`` (page-locked
    memory the refinement wrote the labels into), exposed to numpy through
    ``__array_interface__``: the buffer goes back to the library's pool with
    ``rbsc_server_labels_free`` once the last array over it is gone."""

    def __init__(self, p: int, n: int):
        self._p = p
        self.__array_interface__ = {"shape": (n,), "typestr": "<u4", "data": (p, False),
                                    "version": 3}

    def __del__(self):
        try:
            if self._p:
                load().rbsc_server_labels_free(C.c_void_p(self._p))
                self._p = None
        except Exception:
            pass


class Server:
    """Continuous-batching request loop on one device (``rbsc_server``):
    FIFO admission under an HBM budget, small graphs batched into persistent
    launches, larger graphs on concurrent worker contexts."""

    def __init__(self, device=0, hbm_budget=0, workers=0, batch_max=0, keep_partition=True):
        o = _lib.ServerOpts()
        o.device = device
        o.hbm_budget = int(hbm_budget)
        o.workers = int(workers)
        o.batch_max = int(batch_max)
        o.keep_partition = 1 if keep_partition else 0
        self.keep_partition = keep_partition
        self.handle = C.c_void_p()
        check(load().rbsc_server_create(C.byref(o), C.byref(self.handle)))
        self._nodes = {}

    def submit(self, g: HostGraph, delay_s=0.0) -> int:
        st = g.as_struct()
        t = C.c_uint64()
        check(load().rbsc_server_submit(self.handle, C.byref(st), float(delay_s), C.byref(t)))
        self._nodes[t.value] = g.n_nodes
        return int(t.value)

    def submit_rmat(self, scale, n_draws, a=0.57, b=0.19, c=0.19, clone_frac=0.0, seed=1,
                    delay_s=0.0) -> int:
        t = C.c_uint64()
        check(load().rbsc_server_submit_rmat(self.handle, scale, n_draws, a, b, c, clone_frac, seed,
                                             float(delay_s), C.byref(t)))
        self._nodes[t.value] = 1 << scale
        return int(t.value)

    def submit_resident(self, dg: DeviceGraph, delay_s=0.0) -> int:
        t = C.c_uint64()
        check(load().rbsc_server_submit_resident(self.handle, dg.handle, float(delay_s),
                                                 C.byref(t)))
        self._nodes[t.value] = dg.dims[0]
        return int(t.value)

    def wait(self, ticket: int, timeout_s=-1.0) -> Optional[RequestResult]:
        """The request's result (raises on its error), or None if it is not
        finished within timeout_s.  The partition is the page-locked buffer
        the refinement wrote, not a copy; it returns to the library's pool
        when the last array over it is gone."""
        r = _lib.RequestResult()
        lab = C.c_void_p()
        rc = load().rbsc_server_wait_labels(self.handle, ticket, float(timeout_s), C.byref(r),
                                            C.byref(lab))
        if rc == _lib.PENDING:
            return None
        self._nodes.pop(ticket, None)
        check(rc)
        n = int(r.n_nodes)
        part = None
        if self.keep_partition:
            part = (np.asarray(_ServerLabels(lab.value, n)) if lab.value
                    else np.empty(0, np.uint32))
        return RequestResult(int(r.ticket), int(r.status), int(r.nodes), int(r.rounds),
                             bool(r.batched), n, int(r.n_edges), float(r.latency_s),
                             float(r.queue_s), float(r.service_s), float(r.arrival_s),
                             float(r.done_s), part)

    def stats(self) -> dict:
        st = _lib.ServerStats()
        check(load().rbsc_server_get_stats(self.handle, C.byref(st)))
        return {k: int(getattr(st, k)) for k, _ in _lib.ServerStats._fields_}

    def close(self):
        if self.handle:
            load().rbsc_server_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
