"""Exact (hash-free) stability check on the GPU.

Pinned by what the mathematics fixes: the final partition of the round
function (the oracle's, equal to the exact refiner's) is stable, so the check
reports 0 violations; every strict coarsening of the coarsest stable
partition is unstable, so merging any two of its blocks must be reported; the
discrete partition is always stable; a block that mixes initial colours is
always reported.
"""
import numpy as np
import pytest

import oracle
import rbsc_amd
from rbsc_amd import HostGraph, gen

pytestmark = pytest.mark.gpu


def _weighted(g, wmax, seed=1):
    g.weight = np.random.default_rng(seed).integers(1, wmax + 1, g.n_edges).astype(np.uint64)
    return g


CASES = {
    "rmat_s12_c30": lambda: gen.rmat(12, 1 << 16, clone_frac=0.3, seed=3),
    "rmat_s12_w": lambda: _weighted(gen.rmat(12, 1 << 16, clone_frac=0.3, seed=4), 5),
    "ba_init3": lambda: gen.ba(20000, 4, 0.1, 3, seed=9),
    "reg3": lambda: gen.regular_undirected(3000, 3, seed=4),
    "er_und": lambda: gen.er_undirected(5000, 3.0, seed=2),
}


def _merge_two_blocks(part, rng):
    labels = np.unique(part)
    a, b = rng.choice(labels, 2, replace=False)
    out = part.copy()
    out[out == b] = a
    return out


@pytest.mark.parametrize("name", sorted(CASES))
def test_refine_verify_and_coarsenings(ctx, name):
    g = CASES[name]()
    dg = ctx.upload(g)
    r = dg.refine(verify=True)
    assert r.verified == 1
    c, K, _, _ = oracle.rbsc_refine(g.n_nodes, g.sources(), g.dst, g.weight, g.init_color)
    assert np.array_equal(r.partition, oracle.canonical(c))
    assert dg.check_partition(r.partition) == 0
    rng = np.random.default_rng(7)
    if K > 1:
        for _ in range(3):
            assert dg.check_partition(_merge_two_blocks(r.partition, rng)) > 0
    assert dg.check_partition(np.arange(g.n_nodes, dtype=np.uint32)) == 0  # discrete
    if g.init_color is not None and len(np.unique(g.init_color)) > 1:
        assert dg.check_partition(np.zeros(g.n_nodes, np.uint32)) > 0
    dg.free()


def test_verify_streamed_lines(ctx):
    g = _weighted(gen.rmat(13, 1 << 17, clone_frac=0.2, seed=8), 3)
    dg = ctx.upload(g)
    dg.lines_to_host()
    r = dg.refine(verify=True, max_batch_edges=g.n_edges // 4)
    assert r.verified == 1
    rng = np.random.default_rng(3)
    assert dg.check_partition(_merge_two_blocks(r.partition, rng)) > 0
    dg.free()


def test_verify_one_shot_and_edge_cases(ctx):
    r = ctx.refine(gen.rmat(10, 1 << 14, clone_frac=0.3, seed=5), verify=True)
    assert r.verified == 1
    empty = HostGraph(7, np.zeros(0, np.uint32), src=np.zeros(0, np.uint32))
    r = ctx.refine(empty, verify=True)
    assert r.verified == 1 and r.nodes == 1
    dg = ctx.upload(empty)
    assert dg.check_partition(np.zeros(7, np.uint32)) == 0
    dg.free()
    # a 2-cycle plus a dangling node: {0,1} vs {2} is stable, {0,1,2} is not
    g = HostGraph(3, np.array([1, 0], np.uint32), src=np.array([0, 1], np.uint32))
    dg = ctx.upload(g)
    assert dg.check_partition(np.array([0, 0, 1], np.uint32)) == 0
    assert dg.check_partition(np.array([0, 0, 0], np.uint32)) > 0
    dg.free()


def test_verify_rejects_relaid_layout(ctx):
    g = gen.rmat(11, 1 << 15, seed=6)
    dg = ctx.upload(g)
    dg.relayout(4)
    with pytest.raises(rbsc_amd.RbscError):
        dg.check_partition(np.zeros(g.n_nodes, np.uint32))
    dg.free()
