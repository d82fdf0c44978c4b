"""Pins the CPU oracle before it is trusted as the checker.

* randomize: the published SplitMix64 known answer and an explicit inverse of
  every step (bijectivity, rbsc.cu:65-71 / rsc.cpp:63-69).
* rbsc restatement: final partition equal to the exact hash-free refiner, and
  block counts equal to the reference's own compiled rsc and Valmari tools on
  the committed fixtures (tests/golden/fixtures.json).
* closed forms: cycle, path, star, regular graphs, clones, disjoint unions.
* a planted mistake in the oracle is caught by these checks.
"""
import json
import os

import numpy as np
import pytest

import oracle
from rbsc_amd import HostGraph, gen

M64 = (1 << 64) - 1
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "fixtures.json")


def _inv_xorshift(y, s):
    x = y
    for _ in range(64 // s + 1):
        x = y ^ (x >> s)
    return x & M64


def _inv_randomize(y):
    y = (y * pow(9, -1, 1 << 64)) & M64
    y = ((y >> 7) | (y << 57)) & M64          # rotr 7
    y = (y * pow(5, -1, 1 << 64)) & M64
    y = _inv_xorshift(y, 31)
    y = (y * pow(0x94d049bb133111eb, -1, 1 << 64)) & M64
    y = _inv_xorshift(y, 27)
    y = (y * pow(0xbf58476d1ce4e5b9, -1, 1 << 64)) & M64
    y = _inv_xorshift(y, 30)
    return (y - 0x9e3779b97f4a7c15) & M64


def test_randomize_splitmix_known_answer():
    # The first SplitMix64 output from state 0 is 0xe220a8397b1dcdaf
    # (Vigna's published generator); randomize(0) is that value through the
    # xoshiro** scrambler rotl(x * 5, 7) * 9.
    sm0 = 0xE220A8397B1DCDAF
    t = (sm0 * 5) & M64
    want = ((((t << 7) | (t >> 57)) & M64) * 9) & M64
    assert oracle.randomize(0) == want


def test_randomize_is_a_bijection():
    rng = np.random.default_rng(3)
    for v in [0, 1, 2, M64, 1 << 63] + [int(x) for x in rng.integers(0, 2**63, 200)]:
        assert _inv_randomize(oracle.randomize(v)) == v


LARGE = ("rmat_2p20_clone30", "ba_8m_m6_init3", "cnr-2000")  # checked elsewhere


def _golden(large=False):
    with open(GOLDEN) as f:
        return [r for r in json.load(f)["fixtures"] if (r["name"] in LARGE) == large]


def _build(name):
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "make_golden", os.path.join(os.path.dirname(__file__), "golden", "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    for n, fn in mg.SPECS:
        if n == name:
            return fn(), mg.graph_hash
    raise KeyError(name)


@pytest.mark.parametrize("rec", _golden(), ids=lambda r: r["name"])
def test_oracle_matches_reference_tools(rec):
    g, graph_hash = _build(rec["name"])
    assert graph_hash(g) == rec["sha256"], "generator output changed"
    src = g.sources()
    colors, K, rounds, _ = oracle.rbsc_refine(g.n_nodes, src, g.dst, g.weight, g.init_color)
    assert K == rec["valmari_nodes"]
    if "rsc_nodes" in rec:
        assert K == rec["rsc_nodes"]
        w32 = None if g.weight is None else g.weight.astype(np.uint32)
        rc, rK, _ = oracle.rsc_refine(g.n_nodes, src, g.dst, w32)
        assert rK == rec["rsc_nodes"]
        assert np.array_equal(oracle.canonical(rc), oracle.canonical(colors))
    ec, eK, _ = oracle.exact_refine(g.n_nodes, src, g.dst, g.weight, g.init_color)
    assert eK == K
    assert np.array_equal(oracle.canonical(ec), oracle.canonical(colors))


def test_oracle_on_rmat_2p20_clone30_matches_reference_tools():
    """The 2^20-node R-MAT workload (16.8M lines): the oracle's rbsc and rsc
    restatements give the block count the reference's rsc and Valmari
    printed for the same graph (recorded by tests/golden/make_golden.py)."""
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "make_golden", os.path.join(os.path.dirname(__file__), "golden", "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    rec = {r["name"]: r for r in _golden(large=True)}["rmat_2p20_clone30"]
    g = gen.rmat_2p20_clone30()
    assert mg.graph_hash(g) == rec["sha256"]
    src = g.sources()
    _, K, _, _ = oracle.rbsc_refine(g.n_nodes, src, g.dst)
    _, rK, _ = oracle.rsc_refine(g.n_nodes, src, g.dst)
    assert K == rK == rec["rsc_nodes"] == rec["valmari_nodes"]


@pytest.mark.skipif(not oracle.have_ref(), reason="reference tools not built (no reference tree)")
def test_reference_tools_live_small():
    g = gen.rmat(9, 1 << 12, clone_frac=0.3, seed=99)
    _, K, _, _ = oracle.rbsc_refine(g.n_nodes, g.sources(), g.dst)
    assert oracle.run_ref_rsc(g.to_text())[2] == K
    assert oracle.run_ref_valmari(g.to_valmari_text())[2] == K


def _coo(n, pairs):
    pairs = sorted(pairs)
    src = np.array([p[0] for p in pairs], np.uint32)
    dst = np.array([p[1] for p in pairs], np.uint32)
    return n, src, dst


@pytest.mark.parametrize("n", [1, 2, 7, 64])
def test_closed_form_cycle_and_path(n):
    N, s, d = _coo(n, [(i, (i + 1) % n) for i in range(n)])
    _, K, _, _ = oracle.rbsc_refine(N, s, d)
    assert K == 1
    if n > 1:
        N, s, d = _coo(n, [(i, i + 1) for i in range(n - 1)])
        _, K, _, _ = oracle.rbsc_refine(N, s, d)
        assert K == n


def test_closed_form_star_isolated_clone_union():
    N, s, d = _coo(9, [(0, i) for i in range(1, 9)])
    assert oracle.rbsc_refine(N, s, d)[1] == 2
    assert oracle.rbsc_refine(5, np.zeros(0, np.uint32), np.zeros(0, np.uint32))[1] == 1
    # a clone of a node (same lines) joins its block
    base = [(0, 1), (1, 2), (2, 0), (3, 4), (4, 3), (4, 0)]
    n, s, d = _coo(6, base + [(5, 3), (5, 0)])  # 5 copies the lines of 4
    c, _, _, _ = oracle.rbsc_refine(n, s, d)
    assert c[5] == c[4]
    # disjoint union of isomorphic copies has the block count of one copy
    g = gen.rmat(8, 1 << 11, clone_frac=0.2, seed=4)
    src = g.sources()
    K1 = oracle.rbsc_refine(g.n_nodes, src, g.dst)[1]
    src2 = np.concatenate([src, src + g.n_nodes])
    dst2 = np.concatenate([g.dst, g.dst + g.n_nodes])
    assert oracle.rbsc_refine(2 * g.n_nodes, src2, dst2)[1] == K1


def test_weights_are_linear():
    g = gen.rmat(9, 1 << 12, clone_frac=0.3, seed=8)
    src = g.sources()
    w = np.random.default_rng(1).integers(1, 4, g.n_edges).astype(np.uint64)
    c1, K1, _, _ = oracle.rbsc_refine(g.n_nodes, src, g.dst, w)
    # the same multigraph with every line repeated w times
    rs = np.repeat(src, w.astype(np.int64))
    rd = np.repeat(g.dst, w.astype(np.int64))
    c2, K2, _, _ = oracle.rbsc_refine(g.n_nodes, rs, rd)
    assert K1 == K2 and np.array_equal(c1, c2)


def test_trace_rounds_consistent():
    g = gen.rmat(9, 1 << 12, clone_frac=0.3, seed=10)
    src = g.sources()
    c, K, R, tr = oracle.rbsc_refine(g.n_nodes, src, g.dst, trace_rounds=16)
    assert np.array_equal(tr[R - 1], c)
    for r in range(1, R + 1):
        cr, Kr, Rr, _ = oracle.rbsc_refine(g.n_nodes, src, g.dst, max_rounds=r)
        assert Rr == r and np.array_equal(cr, tr[r - 1])
    # every round refines the previous one (monotone refinement)
    for r in range(1, R):
        prev, cur = tr[r - 1], tr[r]
        m = {}
        for a, b in zip(cur.tolist(), prev.tolist()):
            assert m.setdefault(a, b) == b


def test_hash_partition_never_finer_than_exact():
    for seed in range(5):
        g = gen.er_undirected(300, 3.0, seed=seed)
        src = g.sources()
        K = oracle.rbsc_refine(g.n_nodes, src, g.dst)[1]
        eK = oracle.exact_refine(g.n_nodes, src, g.dst)[1]
        assert K == eK


def test_multibatch_reduces_to_single_batch():
    g = gen.rmat(10, 1 << 13, clone_frac=0.3, seed=12)
    src = g.sources()
    K = oracle.rbsc_refine(g.n_nodes, src, g.dst)[1]
    nodes, done, passes = oracle.rbsc_multibatch(g.n_nodes, src, g.dst, None, g.n_edges)
    assert (nodes, done, passes) == (K, 1, 1)
    nodes, done, passes = oracle.rbsc_multibatch(g.n_nodes, src, g.dst, None, 10 * g.n_edges)
    assert (nodes, done) == (K, 1)


def test_multibatch_hand_worked():
    # Lines 0->1, 0->2, 1->3 (x2), batch = 2 lines.  Pass 1: batch 0 holds
    # node 0's group (node 1 frozen as a singleton, 2 and 3 dangling) and
    # labels {0}, {2,3}; batch 1 holds node 1's group and labels {1}.  The
    # quotient 0->2, 0->1, 2->1 (weight 2) still needs 2 batches; pass 2
    # splits nothing more, so the loop stops at 3 nodes with done = 0
    # (rbsc.cu:487, 493) -- here the coarsest partition also has 3 blocks.
    n, s, d = _coo(4, [(0, 1), (0, 2), (1, 3), (1, 3)])
    nodes, done, passes = oracle.rbsc_multibatch(n, s, d, None, 2)
    assert (nodes, done, passes) == (3, 0, 2)
    assert oracle.rbsc_refine(n, s, d)[1] == 3
    # with a batch of 4 lines everything is one batch
    assert oracle.rbsc_multibatch(n, s, d, None, 4) == (3, 1, 1)
    # a group larger than the batch is never processed: it stays a singleton
    # and the run does not reach one batch (rbsc.cu:315-334, 460)
    n, s, d = _coo(3, [(0, 1), (0, 2), (0, 1), (1, 2)])
    nodes, done, passes = oracle.rbsc_multibatch(n, s, d, None, 2)
    assert done in (0, 1) and 1 <= nodes <= 3


def test_planted_mistake_is_caught():
    """A restatement that ranks by signature only (dropping the colour term,
    as rsc does) but starts from rbsc's dangling split would still agree with
    the exact refiner here; dropping a line of the signature does not."""
    g = gen.rmat(10, 1 << 13, clone_frac=0.3, seed=14)
    src, dst = g.sources(), g.dst
    K = oracle.rbsc_refine(g.n_nodes, src, dst)[1]
    bad_src, bad_dst = src[:-1], dst[:-1]          # the planted mistake
    Kbad = oracle.rbsc_refine(g.n_nodes, bad_src, bad_dst)[1]
    eK = oracle.exact_refine(g.n_nodes, src, dst)[1]
    assert K == eK
    assert Kbad != eK or not np.array_equal(
        oracle.canonical(oracle.rbsc_refine(g.n_nodes, bad_src, bad_dst)[0]),
        oracle.canonical(oracle.exact_refine(g.n_nodes, src, dst)[0]))
