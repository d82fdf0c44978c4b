"""The benchmark workloads at their stated sizes, through the C ABI.

* R-MAT scale 27 (134,217,728 nodes, 2^31 draws, 20 % clones, seed 1003):
  the exact stability check passes on the final partition, its block count
  equals the one the reference's own rsc (src/rsc/rsc.cpp, compiled from
  source) printed for the same graph, and the binned layout, the plain line
  layout, the sort-ranked first round and the 25 % edge batches streamed
  from pinned host memory give the same partition.
* 50,000 mixed graphs (n log-uniform in [10, 5000], seed 1004): every
  graph's raw colours, block count, round count, labels and block-size
  histogram equal the oracle's restatement of rbsc.cu.
* Serving with requests of more than 10^8 lines: each served partition
  equals the one the graph gets alone, which passes the exact check.
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle
import rbsc_amd
from rbsc_amd import gen

pytestmark = pytest.mark.gpu

# Block count of the scale-27 R-MAT printed by the reference rsc (1 core,
# its own timer), recorded in profiles/rsc_reference_rmat_s27_d31_c20.md.
RMAT_S27_BLOCKS = 31_824_702


def test_rmat_s27_full_size(ctx):
    dg = ctx.gen_rmat(27, 1 << 31, 0.57, 0.19, 0.19, 0.20, 1003)
    N, E = dg.dims
    assert N == 1 << 27
    dg.bin()
    assert dg.bin_info["binned"]
    ref = dg.refine(verify=True)
    assert (ref.nodes, ref.done, ref.verified) == (RMAT_S27_BLOCKS, 1, 1)
    want = ref.partition
    assert int(want.max()) + 1 == RMAT_S27_BLOCKS
    plain = dg.refine(flags=rbsc_amd.FLAG_NOBINS)
    assert (plain.nodes, plain.rounds) == (ref.nodes, ref.rounds)
    assert np.array_equal(plain.partition, want), "plain layout differs from binned"
    del plain
    srt = dg.refine(flags=rbsc_amd.FLAG_SORTRANK)
    assert (srt.nodes, srt.rounds) == (ref.nodes, ref.rounds)
    assert np.array_equal(srt.partition, want), "sort-ranked first round differs"
    del srt
    dg.lines_to_host()
    st = dg.refine(max_batch_edges=E // 4)
    assert (st.nodes, st.rounds, st.done) == (ref.nodes, ref.rounds, 1)
    assert np.array_equal(st.partition, want), "25% streamed batches differ"
    dg.free()


def _histogram(labels):
    return np.sort(np.bincount(labels))[::-1]


def test_batch_50000_mixed_full_size(ctx):
    hb = gen.batch_mixed(50000, 10, 5000, seed=1004)
    db = ctx.upload_batch(hb)
    res = db.refine(labels=True, colors=True, histogram=True)
    db.free()
    no = hb.node_offsets.astype(np.int64)
    lo = hb.line_offsets.astype(np.int64)

    def check(g):
        a, z = no[g], no[g + 1]
        la, lz = lo[g], lo[g + 1]
        c, K, R, _ = oracle.rbsc_refine(int(z - a), hb.src[la:lz], hb.dst[la:lz])
        lab = oracle.canonical(c)
        ok = (res.blocks[g] == K and res.rounds[g] == R and np.array_equal(res.colors[a:z], c)
              and np.array_equal(res.labels[a:z], lab)
              and np.array_equal(res.histogram(g), _histogram(lab)))
        return None if ok else g

    with ThreadPoolExecutor(min(16, os.cpu_count() or 4)) as ex:
        bad = [g for g in ex.map(check, range(hb.n_graphs), chunksize=256) if g is not None]
    assert not bad, f"{len(bad)} graphs differ from the oracle, first {bad[:5]}"
    assert res.total.nodes == int(res.blocks.sum())


def test_serving_requests_above_1e8_lines(ctx):
    rng = np.random.default_rng(1005)
    small = [gen.er_undirected(int(n), 3.0, seed=i)
             for i, n in enumerate(rng.integers(20, 5000, 24))]
    # a resident binned graph of 2.7e8 lines and an R-MAT made at admission
    # of 1.3e8 lines, beside medium and small requests
    big = ctx.gen_rmat(24, 1 << 28, 0.57, 0.19, 0.19, 0.0, 2001)
    big.bin()
    assert big.dims[1] > 10**8
    spec = (23, 1 << 27, 0.57, 0.19, 0.19, 0.0, 2002)
    alone_big = big.refine(verify=True)
    dspec = ctx.gen_rmat(*spec)
    assert dspec.dims[1] > 10**8
    alone_spec = dspec.refine(verify=True)
    dspec.free()
    assert alone_big.verified == 1 and alone_spec.verified == 1
    medium = ctx.gen_rmat(18, 1 << 22, 0.57, 0.19, 0.19, 0.2, 2003)
    alone_med = medium.refine(verify=True)
    assert alone_med.verified == 1
    alone_small = [ctx.refine(g) for g in small]
    with rbsc_amd.Server(workers=4, keep_partition=True) as srv:
        t_big = srv.submit_resident(big, delay_s=0.0)
        t_spec = srv.submit_rmat(*spec[:6], seed=spec[6], delay_s=0.01)
        t_med = [srv.submit_resident(medium, delay_s=0.005 * k) for k in range(3)]
        t_small = [srv.submit(g, delay_s=float(rng.uniform(0, 0.1))) for g in small]
        r_big, r_spec = srv.wait(t_big), srv.wait(t_spec)
        r_med = [srv.wait(t) for t in t_med]
        r_small = [srv.wait(t) for t in t_small]
        st = srv.stats()
    big.free()
    medium.free()
    for r, a in [(r_big, alone_big), (r_spec, alone_spec)] + [(r, alone_med) for r in r_med] + \
            list(zip(r_small, alone_small)):
        assert r.status == 0
        assert (r.nodes, r.rounds) == (a.nodes, a.rounds)
        assert np.array_equal(r.partition, a.partition)
    assert not r_big.batched and not r_spec.batched
    assert st["failed"] == 0 and st["completed"] == 2 + 3 + len(small)
