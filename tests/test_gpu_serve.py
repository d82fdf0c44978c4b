"""The serving loop: every request's result equals the graph refined alone
(single-graph HIP path, itself pinned to the oracle), whatever the mix,
arrival order and batching; admission never lets in-flight HBM estimates
exceed the budget; requests that can never fit fail cleanly."""
import numpy as np
import pytest

import oracle
import rbsc_amd
from rbsc_amd import HostGraph, gen

pytestmark = pytest.mark.gpu


def _mix(seed):
    rng = np.random.default_rng(seed)
    gs = []
    for i in range(40):  # small: batch launches
        n = int(np.exp(rng.uniform(np.log(10), np.log(5000))))
        k = i % 3
        if k == 0:
            gs.append(gen.er_undirected(n, float(rng.uniform(2, 6)), seed=seed * 100 + i))
        elif k == 1:
            gs.append(gen.ba_undirected(max(n, 8), int(rng.integers(1, 4)), seed=seed * 100 + i))
        else:
            gs.append(gen.rmat(max(3, int(np.log2(n))), 8 * n, clone_frac=0.3, seed=seed * 100 + i))
    # large: worker contexts
    gs.append(gen.rmat(14, 1 << 18, clone_frac=0.2, seed=seed + 1))
    gs.append(gen.rmat(15, 1 << 19, clone_frac=0.3, seed=seed + 2))
    gs.append(gen.ba(30000, 4, 0.1, 3, seed=seed + 3))
    w = gen.rmat(13, 1 << 17, clone_frac=0.3, seed=seed + 4)
    w.weight = rng.integers(1, 5, w.n_edges).astype(np.uint64)
    gs.append(w)
    gs.append(gen.er_undirected(6000, 3.0, seed=seed + 5))
    return gs


def _alone(ctx, g):
    dg = ctx.upload(g)
    r = dg.refine()
    dg.free()
    return r


def test_serve_results_equal_alone(ctx):
    gs = _mix(3)
    alone = [_alone(ctx, g) for g in gs]
    rng = np.random.default_rng(5)
    with rbsc_amd.Server(workers=3, keep_partition=True) as srv:
        order = rng.permutation(len(gs))
        tickets = {}
        for i in order:
            tickets[int(i)] = srv.submit(gs[i], delay_s=float(rng.uniform(0, 0.05)))
        res = {i: srv.wait(t) for i, t in tickets.items()}
        st = srv.stats()
    for i, g in enumerate(gs):
        r, a = res[i], alone[i]
        assert r.status == 0
        assert (r.nodes, r.rounds) == (a.nodes, a.rounds), f"graph {i} n={g.n_nodes}"
        assert np.array_equal(r.partition, a.partition), f"graph {i}"
        assert r.batched == (g.n_nodes <= 5120)
        assert r.latency_s >= r.service_s >= 0
    # the oracle on a few
    for i in (0, 7, 40, 43):
        g = gs[i]
        c, K, R, _ = oracle.rbsc_refine(g.n_nodes, g.sources(), g.dst, g.weight, g.init_color)
        assert res[i].nodes == K and np.array_equal(res[i].partition, oracle.canonical(c))
    assert st["completed"] == len(gs) and st["failed"] == 0 and st["in_flight"] == 0
    assert st["batched_graphs"] == 40


def test_serve_batches_arrivals_together(ctx):
    gs = [gen.er_undirected(200 + 10 * i, 3.0, seed=i) for i in range(300)]
    with rbsc_amd.Server(workers=1) as srv:
        ts = [srv.submit(g) for g in gs]
        rs = [srv.wait(t) for t in ts]
        st = srv.stats()
    assert all(r.batched for r in rs)
    assert st["batched_graphs"] == 300 and st["batches"] < 300  # many graphs per launch
    for g, r in zip(gs[::50], rs[::50]):
        assert r.nodes == _alone(ctx, g).nodes


def test_serve_budget_admission(ctx):
    gs = [gen.rmat(14, 1 << 16, seed=20 + i) for i in range(8)]
    need = 9 * gs[0].n_edges + 160 * gs[0].n_nodes + 4096
    budget = 2 * need + need // 2  # two in flight at most
    alone = [_alone(ctx, g) for g in gs]
    with rbsc_amd.Server(workers=4, hbm_budget=budget) as srv:
        ts = [srv.submit(g) for g in gs]
        big = gen.rmat(17, 1 << 19, seed=3)
        tb = srv.submit(big)
        rs = [srv.wait(t) for t in ts]
        with pytest.raises(rbsc_amd.RbscError):
            srv.wait(tb)
        st = srv.stats()
    assert st["peak_admitted_bytes"] <= budget
    assert st["failed"] == 1
    for r, a in zip(rs, alone):
        assert (r.nodes, r.rounds) == (a.nodes, a.rounds)
        assert np.array_equal(r.partition, a.partition)
    # later arrivals waited for memory
    assert max(r.queue_s for r in rs) > 0


def test_serve_rmat_and_resident(ctx):
    with rbsc_amd.Server(workers=2) as srv:
        t_small = srv.submit_rmat(10, 1 << 14, clone_frac=0.3, seed=11)
        t_large = srv.submit_rmat(15, 1 << 19, clone_frac=0.2, seed=12)
        dg = ctx.gen_rmat(14, 1 << 18, 0.57, 0.19, 0.19, 0.2, 13)
        ref = dg.refine()
        tr = [srv.submit_resident(dg, delay_s=0.01 * k) for k in range(4)]
        rs = srv.wait(t_small), srv.wait(t_large)
        rr = [srv.wait(t) for t in tr]
    for r in rr:
        assert (r.nodes, r.rounds) == (ref.nodes, ref.rounds)
        assert np.array_equal(r.partition, ref.partition)
    dg.free()
    hs = _alone(ctx, gen.rmat(10, 1 << 14, clone_frac=0.3, seed=11))
    assert rs[0].batched and (rs[0].nodes, rs[0].rounds) == (hs.nodes, hs.rounds)
    assert np.array_equal(rs[0].partition, hs.partition)
    dl = ctx.gen_rmat(15, 1 << 19, 0.57, 0.19, 0.19, 0.2, 12)
    hl = dl.refine()
    dl.free()
    assert not rs[1].batched and (rs[1].nodes, rs[1].rounds) == (hl.nodes, hl.rounds)
    assert np.array_equal(rs[1].partition, hl.partition)


def test_served_partitions_outlive_the_server(ctx):
    """Partitions are the page-locked buffers the refinements wrote (no copy);
    they stay valid after the server is destroyed and go back to the
    library's pool when the last array over them is gone, for both the
    batched and the single-graph path; a later server reuses the pool."""
    graphs = [gen.rmat(9, 1 << 13, clone_frac=0.3, seed=41),
              gen.rmat(14, 1 << 18, clone_frac=0.2, seed=42)]
    want = [_alone(ctx, g) for g in graphs]
    for rep in range(2):
        with rbsc_amd.Server(workers=2) as srv:
            ts = [srv.submit(g) for g in graphs]
            rs = [srv.wait(t) for t in ts]
        import gc
        gc.collect()
        assert rs[0].batched and not rs[1].batched
        for r, a in zip(rs, want):
            assert (r.nodes, r.rounds) == (a.nodes, a.rounds)
            assert np.array_equal(r.partition, a.partition)
        views = [r.partition[1:5] for r in rs]
        del rs
        gc.collect()
        for v, a in zip(views, want):
            assert np.array_equal(v, a.partition[1:5])
        del views
        gc.collect()


@pytest.mark.parametrize("binned", [False, True])
def test_serve_same_resident_graph_concurrently(ctx, binned):
    # rows longer than a 512-line tile carry partial sums between kernels,
    # and a binned graph's cold slots take new colours every round; both live
    # in each worker's context, so concurrent refinements of one resident
    # graph cannot disturb each other
    dg = ctx.gen_rmat(15, 1 << 21, 0.57, 0.19, 0.19, 0.1, 17)
    if binned:
        dg.bin(256, 1024, 0)  # few hot heads: most lines are cold
        assert dg.bin_info["cold_lines"] > dg.bin_info["hot_lines"]
    ref = dg.refine(colors=True)
    with rbsc_amd.Server(workers=4) as srv:
        ts = [srv.submit_resident(dg) for _ in range(16)]
        rs = [srv.wait(t) for t in ts]
    dg.free()
    for r in rs:
        assert (r.nodes, r.rounds) == (ref.nodes, ref.rounds)
        assert np.array_equal(r.partition, ref.partition)
