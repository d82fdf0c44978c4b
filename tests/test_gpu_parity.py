"""HIP path vs the oracle, through the C ABI (librbsc.so).

Bit-exact per-round colours (the dense ranks rbsc.cu:384-396 produces each
round), final canonical partitions, block counts and round counts, on seeded
graphs, golden fixtures, closed forms and edge cases; plus the device R-MAT
generator against the host one and the rbsc command end to end.
"""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle
import rbsc_amd
from rbsc_amd import HostGraph, gen
from rbsc_amd._lib import CLI_PATH

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(__file__)


def _oracle_full(g, trace_rounds=0):
    return oracle.rbsc_refine(g.n_nodes, g.sources(), g.dst, g.weight, g.init_color,
                              trace_rounds=trace_rounds)


CASES = {
    "rmat_s10_c30": lambda: gen.rmat(10, 1 << 14, clone_frac=0.3, seed=5),
    "rmat_s14_c20": lambda: gen.rmat(14, 1 << 18, clone_frac=0.2, seed=11),
    "rmat_s12_w": lambda: _weighted(gen.rmat(12, 1 << 16, clone_frac=0.3, seed=3), 7),
    "ba_init3": lambda: gen.ba(20000, 4, 0.1, 3, seed=9),
    "er_und": lambda: gen.er_undirected(5000, 5.0, seed=2),
    "reg3": lambda: gen.regular_undirected(3000, 3, seed=4),
}


def _weighted(g, wmax):
    g.weight = np.random.default_rng(1).integers(1, wmax + 1, g.n_edges).astype(np.uint64)
    return g


@pytest.mark.parametrize("name", sorted(CASES))
def test_per_round_colors_bit_exact(ctx, name):
    g = CASES[name]()
    c, K, R, tr = _oracle_full(g, trace_rounds=64)
    dg = ctx.upload(g)
    for r in range(1, R + 1):
        res = dg.refine(max_rounds=r, partition=False, colors=True)
        assert res.rounds == r
        assert np.array_equal(res.colors, tr[r - 1]), f"round {r} colours differ"
    full = dg.refine()
    assert (full.nodes, full.rounds, full.done) == (K, R, 1)
    assert np.array_equal(full.partition, oracle.canonical(c))
    dg.free()


def _make_golden():
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_golden",
                                                  os.path.join(HERE, "golden", "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    return mg


def _golden_specs(large=False):
    mg = _make_golden()
    with open(os.path.join(HERE, "golden", "fixtures.json")) as f:
        recs = {r["name"]: r for r in json.load(f)["fixtures"]}
    return [(n, fn, recs[n]) for n, fn in (mg.LARGE_SPECS if large else mg.SPECS)]


@pytest.mark.parametrize("spec", _golden_specs(large=True), ids=lambda s: s[0])
def test_named_workloads_match_reference_tools(ctx, spec):
    """The 2^20-node R-MAT (30% clones) and the 8M-node BA multigraph with 3
    initial colours: block counts equal to those the reference's rsc and
    Valmari printed for the same graphs, the exact stability check passes,
    and the partition equals the oracle's."""
    name, fn, rec = spec
    g = fn()
    assert _make_golden().graph_hash(g) == rec["sha256"], "generator output changed"
    dg = ctx.upload(g)
    r = dg.refine(verify=True)
    dg.free()
    assert r.nodes == rec["valmari_nodes"] and r.verified == 1
    if "rsc_nodes" in rec:
        assert r.nodes == rec["rsc_nodes"]
    c, K, R, _ = _oracle_full(g)
    assert (r.nodes, r.rounds) == (K, R)
    assert np.array_equal(r.partition, oracle.canonical(c))


@pytest.mark.parametrize("spec", _golden_specs(), ids=lambda s: s[0])
def test_golden_fixture_block_counts(ctx, spec):
    name, fn, rec = spec
    g = fn()
    r = ctx.refine(g)
    assert r.nodes == rec["valmari_nodes"]
    if "rsc_nodes" in rec:
        assert r.nodes == rec["rsc_nodes"]
    c, K, R, _ = _oracle_full(g)
    assert r.rounds == R
    assert np.array_equal(r.partition, oracle.canonical(c))


def _coo_graph(n, pairs, weight=None):
    pairs = sorted(pairs)
    src = np.array([p[0] for p in pairs], np.uint32)
    dst = np.array([p[1] for p in pairs], np.uint32)
    return HostGraph(n, dst, src=src, weight=weight)


def test_edge_cases(ctx):
    # no lines at all: one block after one round
    r = ctx.refine(HostGraph(7, np.zeros(0, np.uint32), src=np.zeros(0, np.uint32)))
    assert (r.nodes, r.rounds) == (1, 1)
    assert np.array_equal(r.partition, np.zeros(7, np.uint32))
    # one node with a self loop
    r = ctx.refine(_coo_graph(1, [(0, 0)]))
    assert r.nodes == 1
    # a star whose centre row spans many wave tiles (carry fix-up path)
    n = 40000
    g = _coo_graph(n, [(0, i) for i in range(1, n)] + [(i, 0) for i in range(1, n, 2)])
    c, K, R, _ = _oracle_full(g)
    r = ctx.refine(g)
    assert (r.nodes, r.rounds) == (K, R)
    assert np.array_equal(r.partition, oracle.canonical(c))
    # directed cycle and path
    assert ctx.refine(_coo_graph(50, [(i, (i + 1) % 50) for i in range(50)])).nodes == 1
    assert ctx.refine(_coo_graph(50, [(i, i + 1) for i in range(49)])).nodes == 50


def test_long_rows_mixed_with_short(ctx):
    rng = np.random.default_rng(5)
    n = 6000
    deg = rng.integers(0, 4, n)
    deg[rng.integers(0, n, 12)] = rng.integers(600, 5000, 12)
    src = np.repeat(np.arange(n), deg).astype(np.uint32)
    dst = rng.integers(0, n, len(src)).astype(np.uint32)
    g = HostGraph(n, dst, src=src)
    c, K, R, tr = _oracle_full(g, trace_rounds=64)
    dg = ctx.upload(g)
    for r in range(1, R + 1):
        assert np.array_equal(dg.refine(max_rounds=r, partition=False, colors=True).colors, tr[r - 1])
    dg.free()


def test_grouped_but_unsorted_sources(ctx):
    # groups must be contiguous, not ascending (ReduceByKey contiguity)
    g = _coo_graph(5, [(0, 1), (1, 2), (2, 0)])
    src = np.array([3, 3, 1, 0, 4], np.uint32)
    dst = np.array([1, 2, 0, 3, 3], np.uint32)
    g = HostGraph(5, dst, src=src)
    c, K, R, _ = oracle.rbsc_refine(5, src, dst)
    r = ctx.refine(g)
    assert r.nodes == K and np.array_equal(r.partition, oracle.canonical(c))


def test_input_errors(ctx):
    with pytest.raises(rbsc_amd.RbscError) as e:
        ctx.refine(HostGraph(3, np.array([1, 2, 0], np.uint32), src=np.array([0, 1, 0], np.uint32)))
    assert e.value.code == -5  # ungrouped
    with pytest.raises(rbsc_amd.RbscError) as e:
        ctx.refine(HostGraph(3, np.array([7], np.uint32), src=np.array([0], np.uint32)))
    assert e.value.code == -5  # id out of range


@pytest.mark.parametrize("where", ["first", "last"])
def test_large_upload_with_a_bad_head_is_refused(ctx, where):
    # line arrays from 2^20 lines cross the link in chunks whose heads the
    # build checks as they land; a bad head in any chunk is refused before
    # any kernel indexes by head, and the context keeps working
    g = gen.rmat(16, 1 << 21, seed=41)
    assert g.n_edges >= 1 << 20
    dst = g.dst.copy()
    dst[3 if where == "first" else g.n_edges - 2] = g.n_nodes + 5
    bad = HostGraph(g.n_nodes, dst, row_offsets=g.row_offsets) if g.row_offsets is not None \
        else HostGraph(g.n_nodes, dst, src=g.src)
    with pytest.raises(rbsc_amd.RbscError) as e:
        ctx.refine(bad)
    assert e.value.code == -5
    with pytest.raises(rbsc_amd.RbscError) as e:
        ctx.upload(bad)
    assert e.value.code == -5
    c, K, R, _ = _oracle_full(g)
    r = ctx.refine(g)
    assert (r.nodes, r.rounds) == (K, R) and np.array_equal(r.partition, oracle.canonical(c))


def test_large_weighted_upload_in_chunks(ctx):
    g = _weighted(gen.rmat(16, 1 << 21, clone_frac=0.2, seed=42), 4)
    assert g.n_edges >= 1 << 20
    c, K, R, _ = _oracle_full(g)
    r = ctx.refine(g)
    assert (r.nodes, r.rounds) == (K, R) and np.array_equal(r.partition, oracle.canonical(c))
    dg = ctx.upload(g)
    r2 = dg.refine()
    assert np.array_equal(r2.partition, r.partition)
    dg.free()


def test_device_rmat_equals_host_rmat(ctx):
    h = gen.rmat(13, 1 << 17, clone_frac=0.3, seed=21)
    d = ctx.gen_rmat(13, 1 << 17, clone_frac=0.3, seed=21)
    back = d.download()
    assert back.n_nodes == h.n_nodes
    assert np.array_equal(back.row_offsets, h.row_offsets)
    assert np.array_equal(back.dst, h.dst)
    c, K, R, _ = _oracle_full(h)
    r = d.refine()
    assert (r.nodes, r.rounds) == (K, R)
    assert np.array_equal(r.partition, oracle.canonical(c))
    d.free()


def test_device_out_of_memory_is_an_error_and_the_context_recovers(ctx):
    # 2^39 draws need 4 TiB of sort keys: the first device allocation fails
    # (after the memory pool's cached blocks are handed back and the
    # allocation retried) and is reported as RBSC_ERR_NOMEM; nothing is
    # launched before it.
    with pytest.raises(rbsc_amd.RbscError) as e:
        ctx.gen_rmat(12, 1 << 39, seed=5)
    assert e.value.code == -2
    h = gen.rmat(10, 1 << 13, clone_frac=0.2, seed=22)
    c, K, R, _ = _oracle_full(h)
    r = ctx.refine(h)
    assert (r.nodes, r.rounds) == (K, R)
    assert np.array_equal(r.partition, oracle.canonical(c))


def test_graph_and_batch_outlive_their_context():
    # Memory freed after its context is closed goes back through the null
    # stream; later contexts allocate and refine as before.
    h = gen.rmat(12, 1 << 16, clone_frac=0.2, seed=31)
    c, K, R, _ = _oracle_full(h)
    hb = gen.batch_mixed(40, 10, 300, seed=32)
    c1 = rbsc_amd.Context(0)
    g = c1.upload(h)
    db = c1.upload_batch(hb)
    r1 = g.refine()
    b1 = db.refine()
    c1.close()
    g.free()
    db.free()
    with rbsc_amd.Context(0) as c2:
        g2 = c2.upload(h)
        r2 = g2.refine()
        db2 = c2.upload_batch(hb)
        b2 = db2.refine()
        for r in (r1, r2):
            assert (r.nodes, r.rounds) == (K, R)
            assert np.array_equal(r.partition, oracle.canonical(c))
        assert np.array_equal(b1.blocks, b2.blocks) and np.array_equal(b1.labels, b2.labels)
        g2.free()
        db2.free()


def test_rmat_2p20_clone30_bit_exact(ctx):
    """The 2^20-node / 2^24-draw R-MAT with 30% clones, checked bit-exact
    against the oracle (final colours and partition)."""
    g = gen.rmat_2p20_clone30()
    c, K, R, _ = _oracle_full(g)
    dg = ctx.upload(g)
    res = dg.refine(colors=True)
    assert (res.nodes, res.rounds) == (K, R)
    assert np.array_equal(res.colors, c)
    assert np.array_equal(res.partition, oracle.canonical(c))
    dg.free()


def test_cli_end_to_end(tmp_path):
    g = gen.rmat(10, 1 << 13, clone_frac=0.3, seed=6)
    c, K, R, _ = _oracle_full(g)
    part = tmp_path / "part.txt"
    p = subprocess.run([CLI_PATH, str(g.n_edges), "1", "--partition-out", str(part)],
                       input=g.to_text(), stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                       timeout=120)
    assert p.returncode == 0, p.stderr.decode()
    lines = p.stdout.decode().split("\n")
    float(lines[0])
    assert int(lines[1]) == 1 and int(lines[2]) == K
    got = np.loadtxt(part, dtype=np.uint32)
    assert np.array_equal(got, oracle.canonical(c))


@pytest.mark.parametrize("name", ["ba_n5000_m3_dup10_init3", "ba_8m_m6_init3"])
def test_cli_init_colors(tmp_path, name):
    """Initial colours through the drop-in command (``--init-colors``, the
    counterpart of Valmari's initial blocks, MDP_read.cc:140-150): the BA
    multigraphs with three initial labels give the block counts Valmari
    printed for them (tests/golden/fixtures.json) and the oracle's
    partition."""
    g = gen.ba(5000, 3, 0.10, 3, seed=13) if name == "ba_n5000_m3_dup10_init3" \
        else gen.ba_8m_m6_init3()
    with open(os.path.join(HERE, "golden", "fixtures.json")) as f:
        want_nodes = {r["name"]: r for r in json.load(f)["fixtures"]}[name]["valmari_nodes"]
    text, init, part = tmp_path / "g.txt", tmp_path / "init.txt", tmp_path / "part.txt"
    g.write_text(str(text), "lines")
    init.write_text("\n".join(map(str, g.init_color.tolist())) + "\n")
    with open(text, "rb") as f:
        p = subprocess.run([CLI_PATH, str(g.n_edges), "1", "--init-colors", str(init),
                            "--partition-out", str(part), "--verify"], stdin=f,
                           stdout=subprocess.PIPE, stderr=subprocess.PIPE, timeout=240)
    assert p.returncode == 0, p.stderr.decode()
    lines = p.stdout.decode().split("\n")
    float(lines[0])
    assert int(lines[1]) == 1 and int(lines[2]) == want_nodes
    assert "verified=1" in p.stderr.decode()
    c, K, _, _ = _oracle_full(g)
    assert K == want_nodes
    got = np.fromfile(part, dtype=np.uint32, sep="\n")
    assert np.array_equal(got, oracle.canonical(c))
    # without the initial colours the count differs: the file is read
    with open(text, "rb") as f:
        q = subprocess.run([CLI_PATH, str(g.n_edges), "1"], stdin=f, stdout=subprocess.PIPE,
                           stderr=subprocess.PIPE, timeout=240)
    assert q.returncode == 0 and int(q.stdout.decode().split("\n")[2]) != want_nodes


@pytest.mark.parametrize("frac", [0.5, 0.25, 0.1])
def test_streaming_equals_resident(ctx, frac):
    """Edge-batch streaming (lines in pinned host memory, slices of
    max_batch_edges lines) computes the exact full-graph rounds."""
    g = _weighted(gen.rmat(14, 1 << 18, clone_frac=0.2, seed=15), 3) if frac == 0.1 else \
        gen.rmat(14, 1 << 18, clone_frac=0.2, seed=15)
    c, K, R, tr = _oracle_full(g, trace_rounds=64)
    B = int(g.n_edges * frac)
    r = ctx.refine(g, max_batch_edges=B, mode="stream")
    assert (r.nodes, r.rounds, r.done) == (K, R, 1)
    assert np.array_equal(r.partition, oracle.canonical(c))
    # the same through a resident graph whose lines were moved to the host
    dg = ctx.upload(g)
    dg.lines_to_host()
    for rr in range(1, R + 1):
        res = dg.refine(max_rounds=rr, partition=False, colors=True, max_batch_edges=B)
        assert np.array_equal(res.colors, tr[rr - 1])
    dg.free()


@pytest.mark.parametrize("name", sorted(CASES))
def test_last_round_confirmed_equals_ranked(ctx, name):
    """The last round confirmed from the previous round's multi-node blocks
    (no sort, no rank) gives the partition, block count and rounds of the
    ranked last round (RBSC_FLAG_FULLROUNDS), which equal the oracle's; the
    confirmation fires on the node-order cases with a packed previous round,
    and asking for colours gets the ranked round's colours."""
    g = CASES[name]()
    dg = ctx.upload(g)
    a = dg.refine(flags=rbsc_amd.FLAG_FULLROUNDS)
    ctx.stats(reset=True)
    b = dg.refine()
    st = ctx.stats(reset=True)
    assert (a.nodes, a.rounds) == (b.nodes, b.rounds)
    assert np.array_equal(a.partition, b.partition)
    c, K, R, _ = _oracle_full(g)
    assert (b.nodes, b.rounds) == (K, R)
    assert np.array_equal(b.partition, oracle.canonical(c))
    if name.startswith("rmat"):
        assert st["confirmed_rounds"] == 1, st
    col = dg.refine(colors=True)
    assert np.array_equal(col.colors, c)
    dg.bin()
    bb = dg.refine(verify=True)
    assert (bb.nodes, bb.rounds, bb.verified) == (K, R, 1)
    assert np.array_equal(bb.partition, b.partition)
    dg.free()


def test_partition_into_page_locked_memory(ctx):
    """rbsc_refine writing its labels straight into an rbsc_host_alloc buffer
    (no staging copy) gives the same partition as into pageable memory, and
    leaves the buffer's tail untouched."""
    g = gen.rmat(12, 1 << 16, clone_frac=0.3, seed=19)
    want = ctx.refine(g)
    out = rbsc_amd.PinnedArray(g.n_nodes + 7).array
    out[:] = 0xDEADBEEF
    r = ctx.refine(g, partition_out=out)
    assert (r.nodes, r.rounds) == (want.nodes, want.rounds)
    assert np.array_equal(out[:g.n_nodes], want.partition)
    assert np.all(out[g.n_nodes:] == 0xDEADBEEF)
    c, K, _, _ = _oracle_full(g)
    assert np.array_equal(want.partition, oracle.canonical(c))


@pytest.mark.parametrize("frac,aggr", [(0.5, False), (0.25, False), (0.25, True), (0.1, False)])
def test_quotient_mode_matches_oracle(ctx, frac, aggr):
    """Reference multi-batch quotient passes (rbsc.cu:274-487): nodes, done and
    passes equal the oracle's restatement."""
    g = gen.rmat(12, 1 << 16, clone_frac=0.3, seed=17)
    src = g.sources()
    B = int(g.n_edges * frac)
    want = oracle.rbsc_multibatch(g.n_nodes, src, g.dst, None, B, aggressive=aggr, partition=True)
    r = ctx.refine(g, max_batch_edges=B, mode="aggressive" if aggr else "quotient")
    assert (r.nodes, r.done, r.passes) == want[:3]
    assert np.array_equal(r.partition, want[3])


def test_cli_batch_modes(tmp_path):
    g = gen.rmat(10, 1 << 13, clone_frac=0.3, seed=6)
    src = g.sources()
    c, K, R, _ = _oracle_full(g)
    B = g.n_edges // 4
    p = subprocess.run([CLI_PATH, str(B), "1", "--mode", "stream"], input=g.to_text(),
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, timeout=120)
    assert p.returncode == 0, p.stderr.decode()
    lines = p.stdout.decode().split("\n")
    assert int(lines[1]) == 1 and int(lines[2]) == K
    want = oracle.rbsc_multibatch(g.n_nodes, src, g.dst, None, B)
    # the reference's semantics are the default for B < E
    for extra in ([], ["--mode", "quotient"]):
        p = subprocess.run([CLI_PATH, str(B), "1"] + extra, input=g.to_text(),
                           stdout=subprocess.PIPE, stderr=subprocess.PIPE, timeout=120)
        assert p.returncode == 0, p.stderr.decode()
        lines = p.stdout.decode().split("\n")
        assert (int(lines[2]), int(lines[1])) == want[:2]
    want_a = oracle.rbsc_multibatch(g.n_nodes, src, g.dst, None, B, aggressive=True)
    p = subprocess.run([CLI_PATH, str(B), "1", "--mode", "aggressive"], input=g.to_text(),
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, timeout=120)
    assert p.returncode == 0, p.stderr.decode()
    lines = p.stdout.decode().split("\n")
    assert (int(lines[2]), int(lines[1])) == want_a[:2]


@pytest.mark.parametrize("parts", [2, 4, 8, 16])
def test_destination_partitioned_layout(ctx, parts):
    """Lines regrouped by head mod P (padding lines, segment rows, carries)
    give the same per-round colours."""
    for g in (gen.rmat(14, 1 << 18, clone_frac=0.2, seed=15),
              _weighted(gen.rmat(12, 1 << 16, clone_frac=0.3, seed=3), 5),
              gen.ba(20000, 4, 0.1, 3, seed=9)):
        c, K, R, tr = _oracle_full(g, trace_rounds=64)
        dg = ctx.upload(g)
        dg.relayout(parts)
        for r in range(1, R + 1):
            res = dg.refine(max_rounds=r, partition=False, colors=True)
            assert np.array_equal(res.colors, tr[r - 1]), f"P={parts} round {r}"
        full = dg.refine()
        assert (full.nodes, full.rounds) == (K, R)
        assert np.array_equal(full.partition, oracle.canonical(c))
        dg.free()


@pytest.mark.parametrize("hot", [1, 64, 1000, 1 << 30])
def test_hot_cold_split_layout(ctx, hot):
    """Lines into the `hot` most-gathered heads in a first pass and the rest
    in a second give the same per-round colours (also hashed64)."""
    for g in (gen.rmat(14, 1 << 18, clone_frac=0.2, seed=15),
              _weighted(gen.rmat(12, 1 << 16, clone_frac=0.3, seed=3), 5),
              gen.ba(20000, 4, 0.1, 3, seed=9)):
        c, K, R, tr = _oracle_full(g, trace_rounds=64)
        dg = ctx.upload(g)
        dg.split_hot(hot)
        for r in range(1, R + 1):
            res = dg.refine(max_rounds=r, partition=False, colors=True)
            assert np.array_equal(res.colors, tr[r - 1]), f"hot={hot} round {r}"
        full = dg.refine(flags=rbsc_amd.FLAG_HASHED64)
        assert (full.nodes, full.rounds) == (K, R)
        assert np.array_equal(full.partition, oracle.canonical(c))
        dg.free()


@pytest.mark.parametrize("parts", [1, 4])
def test_hashed64_per_node_hashing(ctx, parts):
    """Per-node hashed colours (instead of per-line hashing of raw colours)
    give the same colours."""
    for g in (gen.rmat(14, 1 << 18, clone_frac=0.2, seed=15),
              _weighted(gen.rmat(12, 1 << 16, clone_frac=0.3, seed=3), 5),
              gen.ba(20000, 4, 0.1, 3, seed=9)):
        c, K, R, tr = _oracle_full(g, trace_rounds=64)
        dg = ctx.upload(g)
        if parts > 1:
            dg.relayout(parts)
        for r in range(1, R + 1):
            res = dg.refine(max_rounds=r, partition=False, colors=True, flags=rbsc_amd.FLAG_HASHED64)
            assert np.array_equal(res.colors, tr[r - 1]), f"round {r}"
        dg.free()
    r = ctx.refine(gen.rmat(13, 1 << 17, clone_frac=0.2, seed=8), max_batch_edges=30000,
                   flags=rbsc_amd.FLAG_HASHED64)
    g = gen.rmat(13, 1 << 17, clone_frac=0.2, seed=8)
    c, K, R, _ = _oracle_full(g)
    assert (r.nodes, r.rounds) == (K, R) and np.array_equal(r.partition, oracle.canonical(c))


@pytest.mark.parametrize("flags", [rbsc_amd.FLAG_XCD, rbsc_amd.FLAG_XCD | rbsc_amd.FLAG_HASHED64])
def test_xcd_partitioned_single_launch(ctx, flags):
    """Eight destination partitions served concurrently (atomic row sums)."""
    for g in (gen.rmat(14, 1 << 18, clone_frac=0.2, seed=15),
              _weighted(gen.rmat(12, 1 << 16, clone_frac=0.3, seed=3), 5),
              gen.ba(20000, 4, 0.1, 3, seed=9)):
        c, K, R, tr = _oracle_full(g, trace_rounds=64)
        dg = ctx.upload(g)
        dg.relayout(8)
        for r in range(1, R + 1):
            res = dg.refine(max_rounds=r, partition=False, colors=True, flags=flags)
            assert np.array_equal(res.colors, tr[r - 1]), f"round {r}"
        dg.free()


@pytest.mark.parametrize("flags", [0, 32, 8, 8 | 16, 16, 128, 128 | 32, 128 | 8, 128 | 8 | 16,
                                   512, 512 | 8, 512 | 8 | 16, 1024, 1024 | 8, 2048, 2048 | 8])
def test_sort_split_and_tie_fixup(ctx, flags):
    """Top-half sort + colliding-run fix-up (default, 0), full 64-bit sort
    (32), frequent collisions via a 16-bit split (8), and the full-sort
    fallback forced by a 2-entry run list (16); on the R-MAT (no initial
    colours) the default sorts nodes from id order, where the forced
    fallback takes the exact redo from colour order, and 128 keeps
    positions in colour order instead; 512 sorts packed 8-byte keys in
    every node-order round (low signature bits and colours read by node
    where top bits tie), 1024 in none; 2048 sorts in the first round too
    (instead of the hash-table rank); the BA graph has initial colours."""
    for g in (gen.rmat(15, 1 << 19, clone_frac=0.2, seed=15),
              gen.ba(30000, 3, 0.1, 3, seed=9)):
        c, K, R, tr = _oracle_full(g, trace_rounds=64)
        dg = ctx.upload(g)
        for r in range(1, R + 1):
            res = dg.refine(max_rounds=r, partition=False, colors=True, flags=flags)
            assert np.array_equal(res.colors, tr[r - 1]), f"flags={flags} round {r}"
        dg.free()


@pytest.mark.parametrize("flags", [8, 8 | 128, 8 | 512])
def test_tie_runs_of_every_length(ctx, flags):
    """A class of 6000 rows with one signature (each row one line into its own
    line-less node) beside an R-MAT with many distinct signatures: under the
    16-bit split (8) the class's run collides with other signatures, so the
    long-run path (sorted in a compact array, rank redone) runs next to the
    in-kernel small and mid sorts; every round equals the oracle's, in node
    order and in colour order (128), with packed keys (512)."""
    base = gen.rmat(16, 1 << 20, clone_frac=0.1, seed=41)
    n0 = base.n_nodes
    src = [base.sources()]
    dst = [base.dst]
    k = 6000
    src.append(np.arange(n0, n0 + k, dtype=np.uint32))
    dst.append(np.arange(n0 + k, n0 + 2 * k, dtype=np.uint32))
    g = HostGraph(n0 + 2 * k, np.concatenate(dst), src=np.concatenate(src))
    c, K, R, tr = _oracle_full(g, trace_rounds=64)
    dg = ctx.upload(g)
    ctx.stats(reset=True)
    for r in range(1, R + 1):
        res = dg.refine(max_rounds=r, partition=False, colors=True, flags=flags)
        assert np.array_equal(res.colors, tr[r - 1]), f"flags={flags} round {r}"
    st = ctx.stats(reset=True)
    assert st["long_run_rounds"] > 0, st
    dg.free()


@pytest.mark.parametrize("flags", [0, rbsc_amd.FLAG_SORTRANK, 8])
def test_first_round_hash_rank(ctx, flags):
    """The first round from a uniform start (one old colour) ranks the
    distinct signatures through a hash table (default, also under the
    16-bit split 8) or sorts all nodes (FLAG_SORTRANK): every round equals
    the oracle's on a weighted R-MAT (first-round signatures nearly all
    distinct), a sparse R-MAT (many line-less nodes, colour 0 kept apart)
    and a 3-regular graph (one signature: the fixpoint after one round)."""
    for g in (_weighted(gen.rmat(14, 1 << 18, clone_frac=0.2, seed=21), 1 << 20),
              gen.rmat(12, 1 << 13, clone_frac=0.3, seed=22),
              gen.regular_undirected(3000, 3, seed=4)):
        c, K, R, tr = _oracle_full(g, trace_rounds=64)
        dg = ctx.upload(g)
        for r in range(1, R + 1):
            res = dg.refine(max_rounds=r, partition=False, colors=True, flags=flags)
            assert np.array_equal(res.colors, tr[r - 1]), f"flags={flags} round {r}"
        res = dg.refine(partition=True, flags=flags)
        assert (res.nodes, res.rounds) == (K, R)
        dg.free()
