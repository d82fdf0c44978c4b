"""Allocation-free batched refinement (rbsc_batch_refine_device) against the
batched entry point that allocates (rbsc_batch_refine), through the C ABI.

Inputs and outputs live in device memory that the test allocates with the
HIP runtime (test plumbing); the entry point itself builds its launch layout
in the library's fixed scratch array.  Every graph's block count, round
count, labels and raw colours must equal those of rbsc_batch_refine on the
same batch, on calls in a row with the input shrinking, weighted and with
initial colours; batches beyond the fixed limits and malformed batches are
rejected cleanly.
"""
import ctypes as C

import numpy as np
import pytest

import rbsc_amd
from rbsc_amd import HostBatch, _lib, gen
from rbsc_amd._lib import check, load

pytestmark = pytest.mark.gpu

MAX_GRAPHS, MAX_NODES, MAX_LINES = 1 << 15, 1 << 21, 1 << 22


class _Device:
    """Device buffers through the HIP runtime (hipMalloc / hipMemcpy)."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so")
        self.ptrs = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), C.c_size_t(max(int(nbytes), 8))) == 0
        self.ptrs.append(p)
        return p

    def put(self, a):
        if a is None:
            return None
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        if a.nbytes:
            assert self.hip.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0
        return p

    def get(self, p, n, dtype=np.uint32):
        out = np.empty(n, dtype)
        if n:
            assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), p,
                                      C.c_size_t(out.nbytes), 2) == 0
        return out

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def _sorted_by_source(hb: HostBatch) -> HostBatch:
    """The same batch with each graph's lines sorted by (source, head)."""
    src, dst = hb.src.copy(), hb.dst.copy()
    w = None if hb.weight is None else hb.weight.copy()
    lo = hb.line_offsets.astype(np.int64)
    for g in range(hb.n_graphs):
        a, z = lo[g], lo[g + 1]
        o = np.lexsort((dst[a:z], src[a:z])) + a
        src[a:z], dst[a:z] = src[o], dst[o]
        if w is not None:
            w[a:z] = w[o]
    return HostBatch(hb.node_offsets, hb.line_offsets, src, dst, w, hb.init_color)


def _refine_device(ctx, hb: HostBatch, flags=0, max_rounds=0):
    dev = _Device()
    try:
        b = _lib.Batch()
        b.n_graphs = hb.n_graphs
        ptrs = {k: dev.put(getattr(hb, k)) for k in
                ("node_offsets", "line_offsets", "src", "dst", "weight", "init_color")}
        b.node_offsets = C.cast(ptrs["node_offsets"], _lib.u64p)
        b.line_offsets = C.cast(ptrs["line_offsets"], _lib.u64p)
        b.src = C.cast(ptrs["src"], _lib.u32p)
        b.dst = C.cast(ptrs["dst"], _lib.u32p)
        b.weight = C.cast(ptrs["weight"], _lib.u64p) if ptrs["weight"] is not None else None
        b.init_color = C.cast(ptrs["init_color"], _lib.u32p) if ptrs["init_color"] is not None else None
        G, n, L = hb.n_graphs, hb.n_nodes, int(hb.line_offsets[-1])
        d_blocks, d_rounds = dev.alloc(4 * G), dev.alloc(4 * G)
        d_labels, d_colors = dev.alloc(4 * n), dev.alloc(4 * n)
        o = _lib.Opts()
        o.max_rounds = max_rounds
        o.flags = flags
        rc = load().rbsc_batch_refine_device(ctx.handle, C.byref(b), n, L, C.byref(o), d_blocks,
                                             d_rounds, d_labels, d_colors)
        if rc:
            return rc, None
        return 0, (dev.get(d_blocks, G), dev.get(d_rounds, G), dev.get(d_labels, n),
                   dev.get(d_colors, n))
    finally:
        dev.free()


def _check_equal(ctx, hb: HostBatch, flags=0, max_rounds=0):
    G, n, L = hb.n_graphs, hb.n_nodes, int(hb.line_offsets[-1])
    assert G <= MAX_GRAPHS and n <= MAX_NODES and L <= MAX_LINES
    db = ctx.upload_batch(hb)
    ref = db.refine(labels=True, colors=True, flags=flags, max_rounds=max_rounds)
    db.free()
    rc, got = _refine_device(ctx, hb, flags, max_rounds)
    assert rc == 0, load().rbsc_last_error()
    blocks, rounds, labels, colors = got
    assert np.array_equal(blocks, ref.blocks)
    assert np.array_equal(rounds, ref.rounds)
    assert np.array_equal(labels, ref.labels)
    assert np.array_equal(colors, ref.colors)


def test_no_alloc_batch_equals_allocating_entry_point(ctx):
    big = _sorted_by_source(gen.batch_mixed(900, 10, 5000, seed=51))
    small = _sorted_by_source(gen.batch_mixed(120, 10, 5000, seed=52))
    assert small.n_nodes < big.n_nodes and small.line_offsets[-1] < big.line_offsets[-1]
    _check_equal(ctx, big)
    _check_equal(ctx, small)        # second call, smaller input: nothing left over
    _check_equal(ctx, small, max_rounds=2)


def test_no_alloc_batch_weighted_initial_colours_and_edges(ctx):
    rng = np.random.default_rng(53)
    from rbsc_amd import HostGraph
    gs = []
    for i, n in enumerate((1, 2, 10, 639, 640, 641, 1280, 1281, 2560, 2561, 5120)):
        gs.append(gen.er_undirected(n, 3.0, seed=600 + i) if n > 2 else
                  HostGraph(n, np.zeros(0, np.uint32), src=np.zeros(0, np.uint32)))
    gs.append(HostGraph(0, np.zeros(0, np.uint32), src=np.zeros(0, np.uint32)))
    gs.append(HostGraph(7, np.array([1, 2, 0], np.uint32), src=np.array([0, 0, 5], np.uint32)))
    for g in gs:
        g.weight = rng.integers(1, 5, g.n_edges).astype(np.uint64)
        g.init_color = rng.integers(0, 3, g.n_nodes).astype(np.uint32)
    hb = _sorted_by_source(HostBatch.from_graphs(gs))
    _check_equal(ctx, hb)
    _check_equal(ctx, hb, flags=8)  # 16-bit split: colliding runs common


def test_no_alloc_batch_rejections(ctx):
    # beyond the fixed limits: rejected before any device work
    b = _lib.Batch()
    b.n_graphs = MAX_GRAPHS + 1
    one = np.zeros(2, np.uint64)
    b.node_offsets = one.ctypes.data_as(_lib.u64p)
    b.line_offsets = one.ctypes.data_as(_lib.u64p)
    out = C.c_void_p(1)
    rc = load().rbsc_batch_refine_device(ctx.handle, C.byref(b), 10, 0, None, out, out, None, None)
    assert rc == -1 and b"exceeds the fixed scratch" in load().rbsc_last_error()
    # a graph above 5120 nodes, a head out of range, lines out of source order:
    # the layout build flags them and the refinement is not launched
    from rbsc_amd import HostGraph
    big = HostBatch.from_graphs([gen.er_undirected(5121, 2.0, seed=1)])
    rc, _ = _refine_device(ctx, big)
    assert rc == -1
    bad_id = HostBatch(np.array([0, 3]), np.array([0, 2]), np.array([0, 1], np.uint32),
                       np.array([1, 3], np.uint32))
    rc, _ = _refine_device(ctx, bad_id)
    assert rc == -5
    unsorted = HostBatch(np.array([0, 3]), np.array([0, 3]), np.array([1, 0, 2], np.uint32),
                         np.array([0, 1, 2], np.uint32))
    rc, _ = _refine_device(ctx, unsorted)
    assert rc == -5
    # and a valid call right after still gives the allocating path's results
    _check_equal(ctx, _sorted_by_source(gen.batch_mixed(50, 10, 5000, seed=54)))
