"""Batched refinement of many small graphs vs the oracle and the single-graph
path, through the C ABI.

Every graph of a batch must get exactly the raw colours, block count, round
count and canonical labels it gets alone (oracle.rbsc_refine restates
rbsc.cu's round function; DeviceGraph.refine is the single-graph HIP path),
whatever the workgroup width and workgroup count of the persistent launch,
and whether it sits in LDS (n <= 5120) or goes through the single-graph path.
"""
import numpy as np
import pytest

import oracle
import rbsc_amd
from rbsc_amd import HostBatch, HostGraph, gen

pytestmark = pytest.mark.gpu


def _coo(n, lines):
    lines = sorted(lines)
    src = np.array([a for a, _ in lines], np.uint32)
    dst = np.array([b for _, b in lines], np.uint32)
    return HostGraph(n, dst, src=src)


def _check_against_oracle(b: HostBatch, res, max_rounds=0, rounds_exact=True):
    for g in range(b.n_graphs):
        hg = b.graph(g)
        a, z = int(b.node_offsets[g]), int(b.node_offsets[g + 1])
        if hg.n_nodes == 0:
            assert res.blocks[g] == 0
            continue
        c, K, R, _ = oracle.rbsc_refine(hg.n_nodes, hg.sources(), hg.dst, hg.weight, hg.init_color,
                                        max_rounds=max_rounds)
        assert res.blocks[g] == K, f"graph {g} (n={hg.n_nodes}): blocks"
        if rounds_exact:
            assert res.rounds[g] == R, f"graph {g} (n={hg.n_nodes}): rounds"
        if res.colors is not None:
            assert np.array_equal(res.colors[a:z], c), f"graph {g} (n={hg.n_nodes}): colours"
        if res.labels is not None:
            assert np.array_equal(res.labels[a:z], oracle.canonical(c)), f"graph {g}: labels"
        if res.block_sizes is not None:
            # final colour histogram: block sizes in descending order, then zeros
            want = np.zeros(z - a, np.uint32)
            want[:K] = np.sort(np.bincount(c))[::-1][:K] if K else []
            assert np.array_equal(res.block_sizes[a:z], want), f"graph {g}: block sizes"
            assert np.array_equal(res.histogram(g), want[:K])


def _mixed_graphs(seed):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(60):
        n = int(np.exp(rng.uniform(np.log(2), np.log(5120))))
        k = i % 4
        if k == 0:
            g = gen.er_undirected(n, float(rng.uniform(1, 6)), seed=seed * 1000 + i)
        elif k == 1:
            g = gen.ba_undirected(max(n, 6), int(rng.integers(1, 4)), seed=seed * 1000 + i)
        elif k == 2:
            nn = max(n, 8) + (max(n, 8) & 1)
            g = gen.regular_undirected(nn, 3, seed=seed * 1000 + i)
        else:
            g = gen.rmat(max(1, int(np.log2(max(n, 2)))), 4 * n, clone_frac=0.3, seed=seed * 1000 + i)
        out.append(g)
    return out


def test_batch_mixed_matches_oracle(ctx):
    b = gen.batch_mixed(400, 10, 5000, seed=21)
    db = ctx.upload_batch(b)
    res = db.refine(labels=True, colors=True, histogram=True)
    _check_against_oracle(b, res)
    assert res.total.nodes == int(res.blocks.sum())
    assert res.total.rounds == int(res.rounds.max())
    db.free()


def test_batch_sort_collisions_fixed(ctx):
    # With the 16-bit split hook (flag 8) neighbours equal in the sorted
    # bits and different below are common; the in-LDS fix-up must still give
    # every graph exactly the oracle's colours.
    b = gen.batch_mixed(200, 10, 5000, seed=23)
    db = ctx.upload_batch(b)
    res = db.refine(labels=True, colors=True, flags=8)
    _check_against_oracle(b, res)
    ref = db.refine(labels=True, colors=True)
    assert np.array_equal(res.colors, ref.colors)
    db.free()


def test_batch_labels_into_pinned_memory(ctx):
    # labels downloaded into page-locked memory owned by the batch equal the
    # pageable download; the pinned array is reused by the next pinned call
    b = gen.batch_mixed(300, 10, 5000, seed=22)
    db = ctx.upload_batch(b)
    ref = db.refine(labels=True)
    pin = db.refine(labels=True, pinned=True)
    assert np.array_equal(pin.labels, ref.labels)
    assert np.array_equal(pin.blocks, ref.blocks)
    first = pin.labels.copy()
    again = db.refine(labels=True, pinned=True)
    assert np.array_equal(again.labels, first)
    db.free()


def test_pinned_labels_valid_after_batch_is_freed(ctx):
    # the batch is a temporary: its pinned labels must outlive it
    b = gen.batch_mixed(120, 10, 5000, seed=24)
    want = ctx.upload_batch(b).refine(labels=True).labels.copy()
    res = ctx.upload_batch(b).refine(labels=True, pinned=True)
    import gc
    gc.collect()
    assert int(res.labels.sum()) == int(want.sum())
    assert np.array_equal(res.labels, want)


def test_batch_directed_sizes_and_edges(ctx):
    gs = _mixed_graphs(3)
    gs += [HostGraph(0, np.zeros(0, np.uint32), src=np.zeros(0, np.uint32)),
           HostGraph(9, np.zeros(0, np.uint32), src=np.zeros(0, np.uint32)),
           _coo(1, [(0, 0)]),
           _coo(2, [(0, 1)]),
           _coo(50, [(i, (i + 1) % 50) for i in range(50)]),
           _coo(50, [(i, i + 1) for i in range(49)]),
           gen.er_undirected(5120, 3.0, seed=77),     # largest LDS graph
           gen.er_undirected(5121, 3.0, seed=78),     # smallest single-path graph
           gen.rmat(14, 1 << 17, clone_frac=0.2, seed=79)]
    b = HostBatch.from_graphs(gs)
    db = ctx.upload_batch(b)
    res = db.refine(labels=True, colors=True, histogram=True)
    _check_against_oracle(b, res)
    db.free()


def test_batch_size_tiers(ctx):
    # graphs at both edges of the four launch tiers (above 2560 nodes, above
    # 1280, above 640, the rest), a few of each, weighted and with initial
    # colours too: every graph's colours, rounds and labels equal the oracle's
    gs = []
    for k, n in enumerate((639, 640, 641, 1279, 1280, 1281, 2559, 2560, 2561, 5119, 5120)):
        for j in range(3):
            gs.append(gen.er_undirected(n, 2.0 + j, seed=500 + 10 * k + j))
    rng = np.random.default_rng(12)
    for g in gs[::4]:
        g.weight = rng.integers(1, 4, g.n_edges).astype(np.uint64)
    b = HostBatch.from_graphs(gs)
    db = ctx.upload_batch(b)
    res = db.refine(labels=True, colors=True)
    _check_against_oracle(b, res)
    db.free()
    for g in gs:
        g.init_color = rng.integers(0, 3, g.n_nodes).astype(np.uint32)
    b = HostBatch.from_graphs(gs)
    db = ctx.upload_batch(b)
    res = db.refine(labels=True, colors=True, histogram=True)
    _check_against_oracle(b, res)
    db.free()


def test_batch_weighted_and_initial_colours(ctx):
    gs = _mixed_graphs(5)
    rng = np.random.default_rng(9)
    for g in gs:
        g.weight = rng.integers(1, 6, g.n_edges).astype(np.uint64)
        g.init_color = rng.integers(0, 3, g.n_nodes).astype(np.uint32) * np.uint32(1000003)
    big = gen.ba(9000, 3, 0.1, 4, seed=3)
    big.weight = rng.integers(1, 4, big.n_edges).astype(np.uint64)
    gs.append(big)
    b = HostBatch.from_graphs(gs)
    db = ctx.upload_batch(b)
    res = db.refine(labels=True, colors=True, histogram=True)
    _check_against_oracle(b, res)
    db.free()


def test_batch_per_round_colours(ctx):
    gs = _mixed_graphs(7)[:30]
    b = HostBatch.from_graphs(gs)
    db = ctx.upload_batch(b)
    for r in (1, 2, 3):
        res = db.refine(max_rounds=r, labels=False, colors=True)
        _check_against_oracle(b, res, max_rounds=r)
    db.free()


@pytest.mark.parametrize("threads", [256, 512, 1024])
def test_batch_launch_shapes_identical(ctx, threads):
    b = gen.batch_mixed(300, 10, 5000, seed=31)
    db = ctx.upload_batch(b)
    ref = db.refine(labels=True, colors=True)
    for wgs in (1, 7, 256, 1024):
        db.set_launch(threads, wgs)
        res = db.refine(labels=True, colors=True)
        assert np.array_equal(res.colors, ref.colors), (threads, wgs)
        assert np.array_equal(res.labels, ref.labels)
        assert np.array_equal(res.blocks, ref.blocks)
        assert np.array_equal(res.rounds, ref.rounds)
    db.free()


def test_batch_autotune_keeps_results(ctx):
    b = gen.batch_mixed(500, 10, 5000, seed=41)
    db = ctx.upload_batch(b)
    ref = db.refine(labels=True, colors=True)
    t, w, ms = db.autotune()
    assert t in (256, 512, 1024) and w >= 1 and ms > 0
    info = db.launch_info
    assert (info["threads"], info["workgroups"]) == (t, w)
    assert all(d in (4, 8) for d in info["digit_bits"])
    assert sum(info["tier_graphs"]) == b.n_graphs
    res = db.refine(labels=True, colors=True)
    assert np.array_equal(res.colors, ref.colors)
    assert np.array_equal(res.labels, ref.labels)
    db.free()


@pytest.mark.parametrize("digits", [(4, 4, 4, 4), (8, 4, 8, 4), (4, 8, 4, 8)])
def test_batch_lds_sort_digit_widths(ctx, digits):
    # the LDS radix sort with 4-bit digits (8 passes) in some tiers and
    # 8-bit digits (4 passes) in others: every graph's colours, rounds and
    # labels equal the oracle's, also under the 16-bit split (flag 8), and
    # with initial colours and weights (their first sort is by colour)
    b = gen.batch_mixed(300, 10, 5000, seed=42)
    db = ctx.upload_batch(b)
    db.set_digits(digits)
    assert db.launch_info["digit_bits"] == list(digits)
    for flags in (0, 8):
        res = db.refine(labels=True, colors=True, flags=flags)
        _check_against_oracle(b, res)
    db.free()
    gs = _mixed_graphs(13)[:40]
    rng = np.random.default_rng(14)
    for g in gs:
        g.weight = rng.integers(1, 4, g.n_edges).astype(np.uint64)
        g.init_color = rng.integers(0, 3, g.n_nodes).astype(np.uint32)
    hb = HostBatch.from_graphs(gs)
    db = ctx.upload_batch(hb)
    db.set_digits(digits)
    _check_against_oracle(hb, db.refine(labels=True, colors=True))
    db.free()


def test_batch_alone_equals_in_mix(ctx):
    gs = _mixed_graphs(11)
    b = HostBatch.from_graphs(gs)
    db = ctx.upload_batch(b)
    mix = db.refine(labels=True, colors=True)
    db.free()
    for g in (0, 5, 17, 42):
        one = HostBatch.from_graphs([gs[g]])
        d1 = ctx.upload_batch(one)
        r1 = d1.refine(labels=True, colors=True)
        d1.free()
        a, z = int(b.node_offsets[g]), int(b.node_offsets[g + 1])
        assert np.array_equal(r1.colors, mix.colors[a:z])
        assert (r1.blocks[0], r1.rounds[0]) == (mix.blocks[g], mix.rounds[g])
        dg = ctx.upload(gs[g])
        single = dg.refine(colors=True)
        dg.free()
        assert np.array_equal(single.colors, mix.colors[a:z])
        assert (single.nodes, single.rounds) == (mix.blocks[g], mix.rounds[g])


def test_batch_input_errors(ctx):
    bad = HostBatch(np.array([0, 3]), np.array([0, 2]), np.array([0, 1], np.uint32),
                    np.array([1, 3], np.uint32))
    with pytest.raises(rbsc_amd.RbscError):
        ctx.upload_batch(bad)
    ungrouped = HostBatch(np.array([0, 3]), np.array([0, 3]), np.array([0, 1, 0], np.uint32),
                          np.array([1, 2, 2], np.uint32))
    with pytest.raises(rbsc_amd.RbscError):
        ctx.upload_batch(ungrouped)
