"""Pins the CPU oracle before it is trusted as the checker.

* randomize: the published SplitMix64 known answer and an explicit inverse of
  every step (bijectivity, rbsc.cu:65-71 / rsc.cpp:63-69).
* rbsc restatement: final partition equal to the exact hash-free refiner, and
  block counts equal to the reference's own compiled rsc and Valmari tools on
  the committed fixtures (tests/golden/fixtures.json).
* closed forms: cycle, path, star, regular graphs, clones, disjoint unions.
* per-round colours: equal to a second restatement that follows the
  reference's own steps (rbsc.cu:352-425) with numpy's stable sorts.
* multi-batch quotient passes: hand-worked cases (a group larger than the
  batch, a group straddling a batch boundary, the aggressive stop).
* mistakes planted in compiled copies of the oracle (colour-before-signature
  rank order, no dangling split) are caught by these checks.
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
    # one batch: the classes are the single-batch refinement's
    c = oracle.rbsc_refine(g.n_nodes, src, g.dst)[0]
    part = oracle.rbsc_multibatch(g.n_nodes, src, g.dst, None, g.n_edges, partition=True)[3]
    assert np.array_equal(part, oracle.canonical(c))
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
    # The aggressive stop (rbsc_aggressive.cu:487) also needs the batch count
    # to shrink: after pass 1 the quotient has 3 lines, ceil(3 / 2) = 2 is
    # not below the 2 batches of pass 1, so it stops after one pass.
    assert oracle.rbsc_multibatch(n, s, d, None, 2, aggressive=True) == (3, 0, 1)
    # with a batch of 4 lines everything is one batch
    assert oracle.rbsc_multibatch(n, s, d, None, 4) == (3, 1, 1)
    # The classes: pass 1 labels node 0 -> 0, nodes 2 and 3 -> 1 (dangling,
    # both colour 0 and signature 0 in batch 0), node 1 -> 2; pass 2 keeps
    # every quotient node apart, so the composition is [0, 2, 1, 1],
    # canonical [0, 1, 2, 2].
    part = oracle.rbsc_multibatch(n, s, d, None, 2, partition=True)[3]
    assert part.tolist() == [0, 1, 2, 2]


def test_multibatch_group_larger_than_the_batch():
    # Lines 0->1, 0->1, 0->2, 1->2, batch = 2 lines (rbsc.cu:302-336).
    # Batch 0 = lines [0, 2): node 0's group starts at 0 but does not end
    # inside it (line 2 is node 0's too), so no end is found and the batch is
    # skipped (332).  Batch 1 = lines [2, 4): line 2 continues node 0's
    # group, so the batch starts at line 3 (node 1's group) and ends at 4;
    # node 1 is the only node in the batch, nodes 0 and 2 are off-batch
    # singletons (colours 1 and 3, signature UINT64_MAX), node 1 starts at
    # N+1 = 4 with signature randomize(3 + 0): the pairs (H, 4), (MAX, 1),
    # (MAX, 3) are 3 distinct = the 3 of the start, so one round; node 1 gets
    # label 0 (431-445).  Nodes 0 and 2 were never labelled and get 1 and 2
    # (457-461): 3 nodes from 3, so the pass loop stops (487) with
    # batches = 2: (nodes, done, passes) = (3, 0, 1).
    n, s, d = _coo(3, [(0, 1), (0, 2), (0, 1), (1, 2)])
    assert oracle.rbsc_multibatch(n, s, d, None, 2) == (3, 0, 1)
    # labels [1, 0, 2] in node order, canonical [0, 1, 2]
    assert oracle.rbsc_multibatch(n, s, d, None, 2, partition=True)[3].tolist() == [0, 1, 2]
    assert oracle.rbsc_multibatch(n, s, d, None, 2, aggressive=True) == (3, 0, 1)


def test_multibatch_group_straddling_a_batch_boundary():
    # Lines 0->1, 1->2, 1->3, 2->3, batch = 2 lines.  Batch 0 = lines
    # [0, 2): node 0's group ends at line 1, node 1's continues past the
    # batch, so the batch is [0, 1) and node 1 is marked off-batch
    # (326-328); node 3 stays dangling (batch 0 starts all nodes dangling).
    # Colours: node 0 at N+1 = 5, nodes 1 and 2 off-batch at 2 and 3, node 3
    # at 0.  Round 0: node 0 has signature randomize(2 + 0), giving 4
    # distinct pairs against the 3 of the start; round 1 keeps 4, so the
    # batch labels node 0 -> 0 and node 3 -> 1.  Batch 1 = lines [2, 4):
    # line 2 continues node 1's group, so the batch is [3, 4), node 2 alone,
    # in one round (4 distinct pairs from 4), label 2.  Node 1 was in no
    # batch and gets 3 (457-461): 4 nodes from 4, stop.  (4, 0, 1).
    n, s, d = _coo(4, [(0, 1), (1, 2), (1, 3), (2, 3)])
    assert oracle.rbsc_multibatch(n, s, d, None, 2) == (4, 0, 1)
    # labels [0, 3, 2, 1], canonical [0, 1, 2, 3]
    assert oracle.rbsc_multibatch(n, s, d, None, 2, partition=True)[3].tolist() == [0, 1, 2, 3]
    # the coarsest partition has 4 blocks too (a path of length 3 into a
    # dangling node, and node 1 with lines into both)
    assert oracle.exact_refine(n, s, d)[1] == 4


def _reference_round_steps(n, src, dst, weight=None, max_rounds=64):
    """rbsc.cu:352-425 for one batch holding every line, step by step, with
    numpy's stable sort standing in for CUB's stable radix sorts:
    init_partition (51-63: N+1 for nodes with a line, 0 for the dangling
    ones), set_values (80-87), ReduceByKey (366-377) + reorder (379) +
    mend_partition (382), SortPairs of the indexes by the colour's low 32
    bits (386), gather (388), stable SortPairs by the 64-bit signature
    (390), compute_thruths_lex (118-123: full 64-bit colours compared),
    InclusiveSum (394) and reorder (396); the salt advances once per round
    (404-419 run once); stop when the count repeats (425).  Returns the
    colour vector after every round."""
    src = np.asarray(src, np.int64)
    dst = np.asarray(dst, np.int64)
    wt = np.ones(len(dst), np.uint64) if weight is None else np.asarray(weight, np.uint64)
    has = np.zeros(n, bool)
    has[src] = True
    col = np.where(has, n + 1, 0).astype(np.uint64)
    K, it, trace = 1, 0, []
    while True:
        K_old = K
        vals = np.array([(oracle.randomize(int(col[v]) + it) * int(w)) & M64
                         for v, w in zip(dst.tolist(), wt.tolist())], np.uint64)
        sig = np.zeros(n, np.uint64)
        np.add.at(sig, src, vals)                      # mod 2^64; dangling stay 0
        o1 = np.argsort(col & np.uint64(0xFFFFFFFF), kind="stable")
        z1 = sig[o1]
        o2 = np.argsort(z1, kind="stable")
        idx, z = o1[o2], z1[o2]
        truths = np.zeros(n, np.int64)
        truths[1:] = (z[1:] != z[:-1]) | (col[idx[1:]] != col[idx[:-1]])
        cs = np.cumsum(truths)
        new = np.empty(n, np.uint64)
        new[idx] = cs
        K = int(cs[-1]) + 1 if n else 0
        col = new
        it += 1
        trace.append(col.astype(np.uint32))
        if K == K_old or it >= max_rounds:
            return trace


def _step_graphs():
    path = _coo(6, [(i, i + 1) for i in range(5)])
    w = gen.rmat(7, 1 << 10, clone_frac=0.3, seed=61)
    rng = np.random.default_rng(62)
    wts = rng.integers(1, 4, w.n_edges).astype(np.uint64)
    return [path[:3] + (None,),
            (w.n_nodes, w.sources(), w.dst, wts),
            (lambda g: (g.n_nodes, g.sources(), g.dst, None))(gen.rmat(8, 1 << 11, 0.57, 0.19,
                                                                       0.19, 0.2, seed=63)),
            (lambda g: (g.n_nodes, g.sources(), g.dst, None))(gen.ba(400, 2, 0.1, 0, seed=64))]


def _trace_of(n, s, d, w, lib=None):
    _, _, R, tr = oracle.rbsc_refine(n, s, d, w, trace_rounds=64, _from=lib)
    return [tr[r] for r in range(R)]


def test_per_round_colours_follow_the_reference_steps():
    """The oracle's colour vector after every round equals the reference's
    own sequence of steps restated with numpy's stable sorts (a second,
    independent restatement of the per-round colours, not a retyping of the
    oracle's comparator)."""
    for n, s, d, w in _step_graphs():
        want = _reference_round_steps(n, s, d, w)
        got = _trace_of(n, s, d, w)
        assert len(got) == len(want)
        for r, (a, b) in enumerate(zip(got, want)):
            assert np.array_equal(a, b), f"round {r + 1}"


MUTATIONS = {
    # the rank order with the colour before the signature
    "colour_before_sig": (
        """    if (x->sig != y->sig) return x->sig < y->sig ? -1 : 1;
    u32 cx = (u32)x->col, cy = (u32)y->col;
    if (cx != cy) return cx < cy ? -1 : 1;""",
        """    u32 cx = (u32)x->col, cy = (u32)y->col;
    if (cx != cy) return cx < cy ? -1 : 1;
    if (x->sig != y->sig) return x->sig < y->sig ? -1 : 1;"""),
    # no dangling split: every node starts at N+1 (as rsc starts all at 1)
    "no_dangling_split": (
        "col[i] = init ? (u64)init[i] : (has[i] ? (u64)N + 1 : 0);",
        "col[i] = init ? (u64)init[i] : (u64)N + 1;"),
}


def _build_copy(tmp_path, name, mutation):
    import shutil
    import subprocess
    src = os.path.join(os.path.dirname(oracle.__file__), "rbsc_oracle.c")
    text = open(src).read()
    if mutation is not None:
        old, new = mutation
        assert text.count(old) == 1, f"mutation {name} no longer matches the oracle source"
        text = text.replace(old, new)
    c = tmp_path / f"{name}.c"
    c.write_text(text)
    so = tmp_path / f"lib{name}.so"
    subprocess.run([shutil.which("gcc") or "gcc", "-O2", "-shared", "-fPIC", "-o", str(so),
                    str(c), "-lm"], check=True, capture_output=True)
    return oracle.load_lib(str(so))


def _caught(lib):
    """True iff the per-round check above rejects this compiled copy of the oracle on
    one of the graphs (its final partitions may still all be right)."""
    for n, s, d, w in _step_graphs():
        want = _reference_round_steps(n, s, d, w)
        got = _trace_of(n, s, d, w, lib)
        if len(got) != len(want) or any(not np.array_equal(a, b) for a, b in zip(got, want)):
            return True
    return False


def test_planted_mistakes_in_a_copy_of_the_oracle_are_caught(tmp_path):
    """Copies of the oracle with a mistake planted (the rank order with the
    colour first; no dangling split) still give the right final partitions
    -- dense ranks of equal pairs and a start that the first round refines
    anyway -- so block counts and the exact refiner cannot see them.  The
    per-round check against the reference's steps must reject both, and
    accept the unchanged copy."""
    assert not _caught(_build_copy(tmp_path, "unchanged", None))
    for name, mutation in MUTATIONS.items():
        lib = _build_copy(tmp_path, name, mutation)
        for n, s, d, w in _step_graphs():  # the final partitions still agree
            a = oracle.rbsc_refine(n, s, d, w, _from=lib)
            b = oracle.rbsc_refine(n, s, d, w)
            assert a[1] == b[1] and np.array_equal(oracle.canonical(a[0]), oracle.canonical(b[0]))
        assert _caught(lib), f"planted mistake {name} not caught"


def test_dropped_line_changes_the_partition():
    """Dropping a line of the input (not a mistake in the oracle) moves the
    result away from the exact refiner's partition of the full graph."""
    g = gen.rmat(10, 1 << 13, clone_frac=0.3, seed=14)
    src, dst = g.sources(), g.dst
    K = oracle.rbsc_refine(g.n_nodes, src, dst)[1]
    bad_src, bad_dst = src[:-1], dst[:-1]          # one line fewer
    Kbad = oracle.rbsc_refine(g.n_nodes, bad_src, bad_dst)[1]
    eK = oracle.exact_refine(g.n_nodes, src, dst)[1]
    assert K == eK
    assert Kbad != eK or not np.array_equal(
        oracle.canonical(oracle.rbsc_refine(g.n_nodes, bad_src, bad_dst)[0]),
        oracle.canonical(oracle.exact_refine(g.n_nodes, src, dst)[0]))
