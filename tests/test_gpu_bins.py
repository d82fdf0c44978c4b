"""Binned signature layout (rbsc_graph_bin): every round's colours equal the
oracle's and the plain kernel's, for every bin size, hot range and chunk
split, with weights, initial colours, repeated rows and rows whose lines all
go to nodes without lines."""
import numpy as np
import pytest

import oracle
import rbsc_amd
from rbsc_amd import HostGraph, gen

pytestmark = pytest.mark.gpu


def _weighted(g, k, seed=3):
    g.weight = np.random.default_rng(seed).integers(1, k + 1, g.n_edges).astype(np.uint64)
    return g


def _wide_weighted(g, seed=4):
    # full 64-bit weights: products and sums wrap mod 2^64
    g.weight = np.random.default_rng(seed).integers(0, 2**63, g.n_edges, dtype=np.uint64) * 2 + 1
    return g


CASES = {
    "rmat_s14_c30": lambda: gen.rmat(14, 1 << 18, clone_frac=0.3, seed=31),
    "rmat_s12_w3": lambda: _weighted(gen.rmat(12, 1 << 16, clone_frac=0.2, seed=32), 3),
    "rmat_s11_w64": lambda: _wide_weighted(gen.rmat(11, 1 << 14, seed=33)),
    "ba_init3": lambda: gen.ba(20000, 3, 0.1, 3, seed=34),
}

# (hot heads, rows per bin, slots per chunk, cold lines per scatter
# workgroup, scatter records in slot order, hashed form, dense heads):
# defaults, everything cold, everything hot, small bins split over many
# chunks, the largest bins, scatter chunks of a few lines (not multiples of
# 8) or above the staged scatter's LDS capacity (taken as that capacity);
# warm heads (dense heads below the hot ones): all hot heads warm, a few
# dense ones with split bins, with slot-order scatter, with cold lines in the
# same bins; the raw form on all of them, the hashed form on seven
SHAPES = [(0, 0, 0, 0, False, False, 0), (1, 256, 0, 0, True, False, 0),
          (1 << 30, 512, 0, 0, False, False, 0), (64, 256, 300, 13, False, False, 0),
          (4096, 16384, 1000, 1 << 20, True, False, 0), (16, 1024, 0, 1, False, False, 0),
          (3, 256, 0, 77, True, False, 0), (2, 256, 0, 1 << 20, False, False, 0),
          (1 << 30, 256, 0, 0, False, False, 1), (4096, 512, 300, 13, False, False, 5),
          (2048, 256, 0, 0, True, False, 64), (300, 1024, 0, 0, False, False, 40),
          (0, 0, 0, 0, False, True, 0), (1, 256, 0, 0, True, True, 0),
          (64, 256, 300, 13, False, True, 0), (3, 256, 0, 77, True, True, 0),
          (2, 256, 0, 1 << 20, False, True, 0), (1 << 30, 256, 0, 0, False, True, 1),
          (300, 1024, 300, 0, False, True, 40)]


def _trace(g, rounds=64):
    return oracle.rbsc_refine(g.n_nodes, g.sources(), g.dst, g.weight, g.init_color,
                              trace_rounds=rounds)


@pytest.mark.parametrize("name", sorted(CASES))
def test_binned_rounds_match_oracle(ctx, name):
    g = CASES[name]()
    c, K, R, tr = _trace(g)
    dg = ctx.upload(g)
    plain = dg.refine(colors=True)
    assert (plain.nodes, plain.rounds) == (K, R)
    for shape in SHAPES:
        hot, rows, chunk, scatter, slot_order, hashed, dense = shape
        dg.bin(hot, rows, chunk, scatter, slot_order, hashed, dense)
        info = dg.bin_info
        assert info["binned"] and info["hashed"] == hashed
        assert info["hot_lines"] + info["cold_lines"] <= dg.stored_lines
        assert info["warm_lines"] <= info["hot_lines"]
        if dense and dense < min(hot, 64):
            assert info["warm_lines"] > 0
        if chunk and chunk < 1000:
            assert info["split_bins"] > 0
        for r in range(1, R + 1):
            res = dg.refine(max_rounds=r, partition=False, colors=True)
            assert np.array_equal(res.colors, tr[r - 1]), f"{shape} round {r}"
        full = dg.refine(verify=True, colors=True)
        assert (full.nodes, full.rounds, full.verified) == (K, R, 1)
        assert np.array_equal(full.colors, plain.colors)
        assert np.array_equal(full.partition, oracle.canonical(c))
    # the plain kernel on the binned graph
    res = dg.refine(colors=True, flags=rbsc_amd.FLAG_NOBINS)
    assert np.array_equal(res.colors, plain.colors)
    dg.free()


def _coo(n, lines, weight=None):
    order = sorted(range(len(lines)), key=lambda i: lines[i][0])
    src = np.array([lines[i][0] for i in order], np.uint32)
    dst = np.array([lines[i][1] for i in order], np.uint32)
    w = None if weight is None else np.array([weight[i] for i in order], np.uint64)
    return HostGraph(n, dst, src=src, weight=w)


def test_rows_into_nodes_without_lines_only(ctx):
    # nodes 0..3 have lines; 4..8 have none.  Row 0 reaches only line-less
    # nodes (its whole signature is the per-row term), row 1 mixes, row 2 is
    # a multi-line row into line-less nodes with weights, row 3 is all lined.
    lines = [(0, 4), (0, 5), (1, 4), (1, 2), (2, 6), (2, 7), (2, 8), (3, 0), (3, 1), (3, 2)]
    w = [1, 2, 1, 1, 3, 1, 5, 1, 1, 2]
    for weight in (None, w):
        g = _coo(9, lines, weight)
        c, K, R, tr = _trace(g)
        dg = ctx.upload(g)
        for hot, hashed in ((1, False), (2, False), (1 << 20, False), (1, True), (1 << 20, True)):
            dg.bin(hot, 256, 0, hashed=hashed)
            for r in range(1, R + 1):
                res = dg.refine(max_rounds=r, partition=False, colors=True)
                assert np.array_equal(res.colors, tr[r - 1])
            assert dg.refine(verify=True).verified == 1
        dg.free()


def test_bins_follow_the_line_layout(ctx):
    g = gen.rmat(12, 1 << 15, clone_frac=0.2, seed=35)
    c, K, R, _ = _trace(g)
    dg = ctx.upload(g)
    dg.bin()
    n0 = dg.nbytes
    info = dg.bin_info
    assert info["bytes"] > 0 and n0 >= info["bytes"]
    dg.unbin()
    assert not dg.bin_info["binned"] and dg.nbytes == n0 - info["bytes"]
    dg.bin()
    dg.lines_to_host()  # streamed rounds read the plain lines: bins dropped
    assert not dg.bin_info["binned"]
    st = dg.refine(max_batch_edges=dg.stored_lines // 3 + 1, colors=True)
    assert (st.nodes, st.rounds) == (K, R)
    dg.free()
    dg = ctx.upload(g)
    dg.bin()
    dg.relayout(4)  # a partitioned layout replaces the bins
    assert not dg.bin_info["binned"]
    assert dg.refine().nodes == K
    dg.free()


def test_hashed64_on_a_binned_graph(ctx):
    # the hashed form serves hashed colours; the raw form leaves them to the
    # plain kernel
    g = gen.rmat(12, 1 << 15, clone_frac=0.2, seed=36)
    c, K, R, tr = _trace(g)
    dg = ctx.upload(g)
    for hashed in (False, True):
        dg.bin(256, 256, 0, hashed=hashed)
        res = dg.refine(colors=True, flags=rbsc_amd.FLAG_HASHED64)
        assert (res.nodes, res.rounds) == (K, R) and np.array_equal(res.colors, tr[R - 1])
    dg.free()


def test_hashed_form_bin_boundaries(ctx):
    # rows into a few heavily gathered heads only, rows into rarely gathered
    # heads only and mixed rows of 1-11 lines: bins whose cold regions start
    # and end inside aligned groups of 8, bins with few or no cold slots,
    # chunks of a few slots, a short last bin; every round against the oracle
    rng = np.random.default_rng(37)
    n = 3000
    lines = []
    for u in range(2000):
        k = int(rng.integers(1, 12))
        for _ in range(k):
            if u % 7 == 0:
                v = int(rng.integers(0, 40))        # hot only
            elif u % 11 == 0:
                v = int(rng.integers(40, 2000))     # cold only
            else:
                v = int(rng.integers(0, n))
            lines.append((u, v))
    g = _coo(n, lines)
    c, K, R, tr = _trace(g)
    dg = ctx.upload(g)
    # (dense heads 3 or 17: warm runs that start and end inside groups of 8)
    for hot, rows, chunk, hashed, dense in ((40, 256, 0, True, 0), (40, 256, 37, True, 0),
                                            (5, 256, 9, True, 0), (1 << 20, 256, 0, True, 0),
                                            (40, 256, 37, False, 0), (5, 256, 9, False, 0),
                                            (40, 256, 0, False, 3), (1 << 20, 256, 37, False, 17),
                                            (40, 256, 0, True, 17)):
        dg.bin(hot, rows, chunk, 3, hashed=hashed, dense_heads=dense)
        for r in range(1, R + 1):
            res = dg.refine(max_rounds=r, partition=False, colors=True)
            assert np.array_equal(res.colors, tr[r - 1]), \
                f"{(hot, rows, chunk, hashed, dense)} round {r}"
    dg.free()


def test_staged_scatter_sparse_tiles(ctx):
    # u % 3 lines per node u into random heads among 300,000: a group's cold
    # lines are so sparse in heads that a tile of 16384 lines spans about
    # 200,000 heads and touches each bin with a run of a few lines; tiles of
    # 64 lines, runs of one line; every round against the oracle, raw and
    # hashed
    rng = np.random.default_rng(38)
    n = 300000
    heads = rng.integers(0, n, n)
    lines = [(u, int(heads[(u + k) % n])) for u in range(n) for k in range(u % 3)]
    g = _coo(n, lines)
    c, K, R, tr = _trace(g)
    dg = ctx.upload(g)
    for scatter, hashed in ((0, False), (64, False), (0, True), (64, True)):
        dg.bin(1, 1024, 0, scatter, hashed=hashed)
        assert dg.bin_info["cold_lines"] > 0
        for r in range(1, R + 1):
            res = dg.refine(max_rounds=r, partition=False, colors=True)
            assert np.array_equal(res.colors, tr[r - 1]), f"{(scatter, hashed)} round {r}"
    dg.free()


@pytest.mark.parametrize("name", ["rmat_s14_c30", "rmat_s12_w3", "rmat_s11_w64"])
def test_first_round_closed_form_equals_gathered(ctx, name):
    # From a uniform start the first round's signatures come from per-row
    # weights (no line read); with RBSC_FLAG_GATHER_FIRST the line kernels
    # gather them.  Both equal the oracle's colours round by round on the
    # binned layout (raw and hashed forms), the plain layout, and lines
    # streamed from host memory in two and in five slices, with repeated
    # rows, weights up to 2^64 and rows whose lines all go to nodes without
    # lines.
    g = CASES[name]()
    c, K, R, tr = _trace(g)
    want = oracle.canonical(c)
    dg = ctx.upload(g)

    def check(label, flags=0, batch=0):
        for f in (flags, flags | rbsc_amd.FLAG_GATHER_FIRST):
            for r in (1, 2):
                res = dg.refine(max_rounds=r, partition=False, colors=True, flags=f,
                                max_batch_edges=batch)
                assert np.array_equal(res.colors, tr[r - 1]), f"{label} flags={f} round {r}"
            full = dg.refine(verify=not batch, flags=f, max_batch_edges=batch)
            assert (full.nodes, full.rounds) == (K, R), f"{label} flags={f}"
            assert batch or full.verified == 1
            assert np.array_equal(full.partition, want), f"{label} flags={f}"

    for hashed in (False, True):
        dg.bin(0, 1024, 0, hashed=hashed)
        check(f"binned hashed={hashed}")
    check("plain", rbsc_amd.FLAG_NOBINS)
    dg.lines_to_host()
    for parts in (2, 5):
        check(f"streamed in {parts}", batch=dg.stored_lines // parts + 1)
    dg.free()
