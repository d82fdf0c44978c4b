"""Seeded synthetic workloads (host side), wrapping include/rbsc_gen.h.

Each function returns a :class:`rbsc_amd.HostGraph` in CSR form with lines
sorted by (src, dst).  Seeds make every workload reproducible.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import HostBatch, HostGraph, _lib
from ._lib import check, load


def _take(hg: _lib.HostGraph) -> HostGraph:
    n, e = int(hg.n_nodes), int(hg.n_edges)
    try:
        ro = np.ctypeslib.as_array(hg.row_offsets, shape=(n + 1,)).copy()
        dst = np.ctypeslib.as_array(hg.dst, shape=(max(e, 1),))[:e].copy()
        w = np.ctypeslib.as_array(hg.weight, shape=(e,)).copy() if hg.weight else None
        ic = np.ctypeslib.as_array(hg.init_color, shape=(n,)).copy() if hg.init_color else None
    finally:
        load().rbsc_host_graph_free(C.byref(hg))
    return HostGraph(n, dst, row_offsets=ro, weight=w, init_color=ic)


def rmat(scale: int, n_draws: int, a=0.57, b=0.19, c=0.19, clone_frac=0.0, seed=1) -> HostGraph:
    """Directed R-MAT, duplicates kept, optional cloned nodes (same graph as
    ``Context.gen_rmat`` builds on the device)."""
    hg = _lib.HostGraph()
    check(load().rbsc_gen_rmat(scale, n_draws, a, b, c, clone_frac, seed, C.byref(hg)))
    return _take(hg)


def ba(n: int, m: int, dup_frac=0.0, n_init_colors=0, seed=1) -> HostGraph:
    """Directed Barabasi-Albert in-degree growth with multiplicities."""
    hg = _lib.HostGraph()
    check(load().rbsc_gen_ba(n, m, dup_frac, n_init_colors, seed, C.byref(hg)))
    return _take(hg)


def er_undirected(n: int, avg_deg: float, seed=1) -> HostGraph:
    hg = _lib.HostGraph()
    check(load().rbsc_gen_er_undirected(n, avg_deg, seed, C.byref(hg)))
    return _take(hg)


def ba_undirected(n: int, m: int, seed=1) -> HostGraph:
    hg = _lib.HostGraph()
    check(load().rbsc_gen_ba_undirected(n, m, seed, C.byref(hg)))
    return _take(hg)


def regular_undirected(n: int, k: int, seed=1) -> HostGraph:
    hg = _lib.HostGraph()
    check(load().rbsc_gen_regular_undirected(n, k, seed, C.byref(hg)))
    return _take(hg)


# Workload presets, named by what they are (sizes in the names).
def rmat_2p20_clone30(seed=1001) -> HostGraph:
    """R-MAT (.57,.19,.19,.05), 2^20 nodes, 2^24 draws, 30% cloned nodes."""
    return rmat(20, 1 << 24, 0.57, 0.19, 0.19, 0.30, seed)


def ba_8m_m6_init3(seed=1002) -> HostGraph:
    """BA growth, 8M nodes, m = 6, 10% lines with multiplicity 2-4, 3 initial colours."""
    return ba(8_000_000, 6, 0.10, 3, seed)


def batch_mixed(n_graphs: int, n_min=10, n_max=5000, seed=1004) -> HostBatch:
    """Undirected ER / BA / random-regular graphs (i mod 3), node counts
    log-uniform in [n_min, n_max]."""
    hb = _lib.HostBatch()
    check(load().rbsc_gen_batch_mixed(n_graphs, n_min, n_max, seed, C.byref(hb)))
    try:
        g = int(hb.n_graphs)
        no = np.ctypeslib.as_array(hb.node_offsets, shape=(g + 1,)).copy()
        lo = np.ctypeslib.as_array(hb.line_offsets, shape=(g + 1,)).copy()
        L = int(lo[-1])
        src = np.ctypeslib.as_array(hb.src, shape=(max(L, 1),))[:L].copy()
        dst = np.ctypeslib.as_array(hb.dst, shape=(max(L, 1),))[:L].copy()
    finally:
        load().rbsc_host_batch_free(C.byref(hb))
    return HostBatch(no, lo, src, dst)


def bvgraph(basename: str) -> HostGraph:
    """A WebGraph BVGraph file (BASENAME.graph + .properties), decoded by the
    library, lines in (src, dst) order."""
    hg = _lib.HostGraph()
    check(load().rbsc_bvgraph_load(basename.encode(), C.byref(hg)))
    return _take(hg)


def bvgraph_check_offsets(basename: str):
    """(offsets read from BASENAME.offsets, nodes whose decoded position differs)."""
    n = C.c_uint64()
    m = C.c_uint64()
    check(load().rbsc_bvgraph_check_offsets(basename.encode(), C.byref(n), C.byref(m)))
    return int(n.value), int(m.value)
