"""The reference's multi-batch quotient mode (rbsc.cu:274-487, stop rule of
rbsc_aggressive.cu:487) on the device, through the C ABI (rbsc_refine with
RBSC_MODE_QUOTIENT / _AGGRESSIVE): node count, done, passes and the
canonical partition of the original nodes equal the oracle's restatement,
on small graphs with many batches (skipped batches, groups straddling a
batch boundary, groups longer than a batch, weights, CSR input), on the
2^20-node R-MAT at 50% and 25% batches against committed golden values,
and malformed input is refused as in the other modes."""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle
import rbsc_amd
from rbsc_amd import HostGraph, gen

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(__file__)


@pytest.fixture(scope="module")
def ctx():
    c = rbsc_amd.Context(0)
    yield c
    c.close()


def _grouped(rng, n, m):
    """m lines with sources grouped in a random group order."""
    s = rng.permutation(n)[np.sort(rng.integers(0, n, m))]
    return s.astype(np.uint32), rng.integers(0, n, m).astype(np.uint32)


def _check(ctx, g, B, aggr):
    src = g.sources()
    want = oracle.rbsc_multibatch(g.n_nodes, src, g.dst, g.weight, B, aggressive=aggr,
                                  partition=True)
    r = ctx.refine(g, max_batch_edges=B, mode="aggressive" if aggr else "quotient")
    assert (r.nodes, r.done, r.passes) == want[:3], (B, aggr)
    assert np.array_equal(r.partition, want[3]), (B, aggr)
    return r


def test_many_small_batches(ctx):
    rng = np.random.default_rng(21)
    for trial in range(60):
        n = int(rng.integers(1, 40))
        m = int(rng.integers(2, 120))
        s, d = _grouped(rng, n, m)
        w = rng.integers(1, 5, m).astype(np.uint64) if trial % 3 == 0 else None
        g = HostGraph(n, d, src=s, weight=w)
        for B in sorted({1, 2, 3, 5, max(1, m // 3), m - 1}):
            if 1 <= B < m:
                for aggr in (False, True):
                    _check(ctx, g, B, aggr)


def test_csr_input_equals_coo(ctx):
    g = gen.rmat(9, 1 << 12, clone_frac=0.3, seed=23)
    csr = HostGraph(g.n_nodes, g.dst, row_offsets=np.concatenate(
        [[0], np.cumsum(np.bincount(g.sources(), minlength=g.n_nodes))]).astype(np.uint64))
    assert np.array_equal(csr.sources(), g.sources())
    for frac in (3, 7):
        a = _check(ctx, g, g.n_edges // frac, False)
        b = ctx.refine(csr, max_batch_edges=g.n_edges // frac, mode="quotient")
        assert (a.nodes, a.done, a.passes) == (b.nodes, b.done, b.passes)
        assert np.array_equal(a.partition, b.partition)


def test_weighted_rmat(ctx):
    g = gen.rmat(11, 1 << 15, clone_frac=0.3, seed=29)
    g.weight = np.random.default_rng(3).integers(1, 4, g.n_edges).astype(np.uint64)
    for frac in (2, 5):
        for aggr in (False, True):
            _check(ctx, g, g.n_edges // frac, aggr)


def test_rmat_2p20_half_and_quarter_batches(ctx):
    with open(os.path.join(HERE, "golden", "quotient_c1.json")) as f:
        gold = json.load(f)
    g = gen.rmat_2p20_clone30()
    assert (g.n_nodes, g.n_edges) == (gold["n_nodes"], gold["n_lines"])
    for c in gold["cases"]:
        r = ctx.refine(g, max_batch_edges=c["max_batch_edges"],
                       mode="aggressive" if c["aggressive"] else "quotient")
        print(f"B={c['max_batch_edges']} aggressive={c['aggressive']}: {r.time_s:.3f} s, "
              f"{r.nodes} nodes, {r.passes} passes", flush=True)
        assert (r.nodes, r.done, r.passes) == (c["nodes"], c["done"], c["passes"])
        h = hashlib.sha256(np.ascontiguousarray(r.partition, "<u4").tobytes()).hexdigest()
        assert h == c["partition_sha256"]


def test_malformed_lines_are_refused(ctx):
    ungrouped = HostGraph(3, np.array([1, 2, 0], np.uint32), src=np.array([0, 1, 0], np.uint32))
    with pytest.raises(rbsc_amd.RbscError) as e:
        ctx.refine(ungrouped, max_batch_edges=2, mode="quotient")
    assert "not grouped" in str(e.value)
    bad_id = HostGraph(3, np.array([1, 5, 0], np.uint32), src=np.array([0, 1, 2], np.uint32))
    with pytest.raises(rbsc_amd.RbscError) as e:
        ctx.refine(bad_id, max_batch_edges=2, mode="quotient")
    assert "node id" in str(e.value)
    bad_csr = HostGraph(3, np.array([1, 2, 0], np.uint32),
                        row_offsets=np.array([0, 2, 1, 3], np.uint64))
    with pytest.raises(rbsc_amd.RbscError):
        ctx.refine(bad_csr, max_batch_edges=2, mode="quotient")
    # the context still works afterwards
    _check(ctx, gen.rmat(8, 1 << 11, clone_frac=0.3, seed=31), 700, False)
