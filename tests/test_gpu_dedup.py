"""Identical rows stored once: results are bit-identical to the layout that
stores every row (and to the oracle) on every path that reads the layout:
rounds, streaming, partitioned layouts, download, the exact check."""
import numpy as np
import pytest

import oracle
import rbsc_amd
from rbsc_amd import HostGraph, gen

pytestmark = pytest.mark.gpu


def _coo(n, lines, weight=None):
    order = sorted(range(len(lines)), key=lambda i: lines[i][0])  # group by source, keep row order
    src = np.array([lines[i][0] for i in order], np.uint32)
    dst = np.array([lines[i][1] for i in order], np.uint32)
    w = None if weight is None else np.array([weight[i] for i in order], np.uint64)
    return HostGraph(n, dst, src=src, weight=w)


def test_rows_counted_once(ctx):
    # rows 0, 1, 2 identical; row 3 same heads, other weight; row 4 same
    # multiset, other order; row 5 distinct
    lines = [(0, 6), (0, 7), (1, 6), (1, 7), (2, 6), (2, 7), (3, 6), (3, 7), (4, 7), (4, 6),
             (5, 6)]
    w = [1, 1, 1, 1, 1, 1, 2, 1, 1, 1, 1]
    g = _coo(8, lines, w)
    dg = ctx.upload(g)
    assert dg.rows == 4 and dg.stored_lines == 7 and dg.dims == (8, 11)
    c, K, R, tr = oracle.rbsc_refine(8, g.sources(), g.dst, g.weight, trace_rounds=16)
    for r in range(1, R + 1):
        assert np.array_equal(dg.refine(max_rounds=r, partition=False, colors=True).colors, tr[r - 1])
    back = dg.download()
    assert np.array_equal(back.sources(), g.sources()) and np.array_equal(back.dst, g.dst)
    dg.free()


CASES = {
    "rmat_s14_c30": lambda: gen.rmat(14, 1 << 18, clone_frac=0.3, seed=21),
    "rmat_s12_w": lambda: _w(gen.rmat(12, 1 << 16, clone_frac=0.3, seed=22), 3),
    "ba_init3": lambda: gen.ba(20000, 2, 0.1, 3, seed=23),
}


def _w(g, k):
    g.weight = np.random.default_rng(3).integers(1, k + 1, g.n_edges).astype(np.uint64)
    return g


@pytest.mark.parametrize("name", sorted(CASES))
def test_dedup_equals_full_layout(ctx, name):
    g = CASES[name]()
    c, K, R, tr = oracle.rbsc_refine(g.n_nodes, g.sources(), g.dst, g.weight, g.init_color,
                                     trace_rounds=64)
    outs = []
    for dedup in (True, False):
        ctx.set_dedup(dedup)
        dg = ctx.upload(g)
        if dedup:
            assert dg.rows < int(np.count_nonzero(np.diff(g.row_offsets.astype(np.int64))))
        else:
            assert dg.stored_lines == g.n_edges
        for r in range(1, R + 1):
            res = dg.refine(max_rounds=r, partition=False, colors=True)
            assert np.array_equal(res.colors, tr[r - 1]), f"dedup={dedup} round {r}"
        full = dg.refine(verify=True, colors=True)
        assert (full.nodes, full.rounds, full.verified) == (K, R, 1)
        outs.append(full.colors)
        # partitioned layout over the stored rows
        dg.relayout(4)
        assert np.array_equal(dg.refine(colors=True).colors, full.colors)
        dg.free()
        # streamed lines
        dg = ctx.upload(g)
        dg.lines_to_host()
        st = dg.refine(colors=True, max_batch_edges=g.n_edges // 4, verify=True)
        assert np.array_equal(st.colors, full.colors) and st.verified == 1
        dg.free()
    ctx.set_dedup(True)
    assert np.array_equal(outs[0], outs[1])


def test_repeated_rows_in_the_exact_check(ctx):
    # X = 0: (5, 6); Y = 1: (6, 5) (same multiset, other order: stored apart);
    # Z = 2 repeats X (stored once); W = 3: (5, 5)
    lines = [(0, 5), (0, 6), (1, 6), (1, 5), (2, 5), (2, 6), (3, 5), (3, 5)]
    g = _coo(7, lines)
    dg = ctx.upload(g)
    assert dg.rows == 3
    r = dg.refine(verify=True)
    c, K, _, _ = oracle.rbsc_refine(7, g.sources(), g.dst)
    assert r.nodes == K and r.verified == 1
    assert np.array_equal(r.partition, oracle.canonical(c))
    # {X, Y, Z} {W} {5} {6} {4}: stable, Z compared against Y's stored row
    assert dg.check_partition(np.array([0, 0, 0, 1, 2, 3, 4], np.uint32)) == 0
    # Z with W: Z's row (X's) differs from W's
    assert dg.check_partition(np.array([0, 0, 1, 1, 2, 3, 4], np.uint32)) > 0
    dg.free()


def test_repeated_row_with_other_initial_colour(ctx):
    # rows 0 and 1 identical, initial colours differ: different blocks
    lines = [(0, 3), (0, 4), (1, 3), (1, 4), (2, 3)]
    g = _coo(5, lines)
    g.init_color = np.array([0, 1, 0, 0, 0], np.uint32)
    dg = ctx.upload(g)
    r = dg.refine(verify=True)
    c, K, _, _ = oracle.rbsc_refine(5, g.sources(), g.dst, None, g.init_color)
    assert r.nodes == K and r.verified == 1
    assert np.array_equal(r.partition, oracle.canonical(c))
    bad = r.partition.copy()
    bad[2] = bad[0]  # row 2 (one line) merged with row 0 (two lines): unstable
    assert dg.check_partition(bad) > 0
    merged = r.partition.copy()
    merged[merged == merged[1]] = merged[0]  # merging 0 and 1 mixes initial colours
    assert dg.check_partition(merged) > 0
    dg.free()
