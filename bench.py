"""Benchmark of the RBSC refinement round on one MI355X.

Workload (default): directed R-MAT (a, b, c, d) = (.57, .19, .19, .05), scale
27 (134,217,728 nodes), 2^31 draws, uint32 ids, 20% of the nodes cloned,
initial colour 0 everywhere, built on the device with a fixed seed and
resident in HBM.  A step is one refinement of that graph to its fixpoint
(every round of rbsc.cu:354-425); the metric is edge visits per second
(lines processed per round x rounds / time).

Prints one JSON line (rank 0) with the measured value, a live roofline for
the round (HIP events on the library's stream) and a single-core CPU
baseline on a bounded sample of the same R-MAT family: the reference's rsc
compiled from its sources (oracle/_ref) when that is built, else the
oracle's rsc restatement.  Further legs, each in its own JSON block: the
one-shot drop-in call, the rbsc command on the workload's text, the
multi-batch quotient mode, the 50,000-graph batch, the serving loop and the
streamed (pinned host memory) refinement.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s

CONFIGS = {
    # name: (scale, log2 draws, clone fraction, seed)
    "rmat_s27_d31_c20": (27, 31, 0.20, 1003),
    "rmat_s20_d24_c30": (20, 24, 0.30, 1001),
    "rmat_s24_d28_c20": (24, 28, 0.20, 1003),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(scale: int, log_draws: int, clone: float, seed: int) -> dict:
    """Single-core CPU refinement of a bounded R-MAT sample of the same
    family: the reference's own rsc and Valmari tools compiled from source
    (oracle/_ref) when present, timed by their own timers (the figure the
    reference's harness records), beside the oracle's rsc restatement."""
    import subprocess
    import tempfile
    import oracle
    from rbsc_amd import gen
    g = gen.rmat(scale, 1 << log_draws, clone_frac=clone, seed=seed)
    E = g.n_edges
    src = g.sources()
    t = time.perf_counter()
    _, K, R = oracle.rsc_refine(g.n_nodes, src, g.dst)
    dt = time.perf_counter() - t
    sample = (f"R-MAT scale {scale}, 2^{log_draws} draws, {int(clone * 100)}% clones: {E} lines, "
              f"{R} rsc rounds, {K} blocks")
    host = {"host_threads": os.cpu_count(), "cpu_model": _cpu_model()}
    port = {"value": R * E / dt, "seconds": dt, "blocks": K}
    # the oracle's exact lumping refiner (the Valmari tool's partition, by
    # splitters), one core, on the same sample
    t = time.perf_counter()
    _, KL = oracle.lump_refine(g.n_nodes, src, g.dst)
    lump = {"seconds": time.perf_counter() - t, "blocks": KL,
            "method": "splitter refinement, oracle/lump.c"}
    out = {"value": port["value"], "unit": "edge-visits/s", "cores": 1, "kind": "port",
           "sample": f"rsc restatement (oracle/rbsc_oracle.c) on {sample}", **host,
           "lumping_port": lump}
    if not oracle.have_ref():
        return out
    with tempfile.TemporaryDirectory() as d:
        lines, mdp = os.path.join(d, "g.txt"), os.path.join(d, "g.mdp")
        g.write_text(lines, "lines")
        g.write_text(mdp, "valmari")
        runs = {}
        for name, exe, path in (("rsc", oracle.REF_RSC, lines), ("valmari", oracle.REF_VALMARI, mdp)):
            with open(path, "rb") as f:
                p = subprocess.run([exe], stdin=f, stdout=subprocess.PIPE, timeout=600, check=True)
            v = p.stdout.decode().split("\n")
            runs[name] = {"seconds": float(v[0]), "blocks": int(v[2])}
    out.update({"value": R * E / runs["rsc"]["seconds"], "kind": "reference",
                "sample": f"reference rsc (src/rsc/rsc.cpp, compiled from source, its own timer) on "
                          f"{sample}",
                "port": port, "valmari": runs["valmari"],
                "blocks_agree": runs["rsc"]["blocks"] == runs["valmari"]["blocks"] == K == KL})
    return out


# The reference's rsc (src/rsc/rsc.cpp, compiled from source, one core) on
# the full workload, once: its own timer, its block count and the host, as
# recorded in profiles/rsc_reference_rmat_s27_d31_c20.md.  Too long (about 15
# minutes of refinement plus 40 GB of text) for every bench run, so the live
# leg above times a bounded sample of the same family.
RSC_FULL_SIZE = {
    "rmat_s27_d31_c20": {"seconds": 862.642872, "blocks": 31824702, "done": 1, "rounds": 7,
                         "edge_visits_per_s": 7 * 2139161279 / 862.642872, "cores": 1,
                         "cpu_model": "Intel(R) Xeon(R) Processor @ 2.70GHz",
                         "record": "profiles/rsc_reference_rmat_s27_d31_c20.md"},
}


def full_size_reference(config: str, blocks: int, refine_s: float, one_shot_s,
                        command_s=None) -> dict | None:
    """The recorded full-size rsc run beside this run's time to the fixpoint
    (one refinement from resident lines, the drop-in's one-shot span from
    host arrays, and the rbsc command's own reported time), each the span
    the tool's own timer covers."""
    rec = RSC_FULL_SIZE.get(config)
    if rec is None:
        return None
    out = dict(rec)
    out["blocks_agree"] = rec["blocks"] == blocks
    out["speedup_refine"] = rec["seconds"] / refine_s
    if one_shot_s:
        out["speedup_one_shot"] = rec["seconds"] / one_shot_s
    if command_s:
        out["speedup_rbsc_command"] = rec["seconds"] / command_s
    return out


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def batch_leg(ctx, n_graphs: int, seed: int, loop_sample: int) -> dict:
    """Many small graphs: one persistent launch over a resident batch vs the
    single-graph path called once per graph (on a sample of the batch)."""
    import numpy as np
    import rbsc_amd
    from rbsc_amd import gen
    t0 = time.perf_counter()
    hb = gen.batch_mixed(n_graphs, 10, 5000, seed=seed)
    db = ctx.upload_batch(hb)
    log(f"[bench] batch: {n_graphs} graphs, {hb.n_nodes} nodes, {len(hb.dst)} lines, "
        f"built and uploaded in {time.perf_counter() - t0:.1f}s")
    threads, wgs, tune_ms = db.autotune()
    launch = db.launch_info
    log(f"[bench] batch autotune: {launch} ({tune_ms:.1f} ms)")
    db.refine(labels=True, pinned=True)
    reps = 5
    t = time.perf_counter()
    dev = 0.0
    for _ in range(reps):
        r = db.refine(labels=True, pinned=True)
        dev += r.total.device_ms
    wall = (time.perf_counter() - t) / reps
    blocks = int(r.blocks.sum())
    db.free()
    # per-graph loop on the first loop_sample graphs (resident single-graph path)
    idx = np.linspace(0, n_graphs - 1, loop_sample).astype(int)
    dgs = [ctx.upload(hb.graph(int(i))) for i in idx]
    for g in dgs[:4]:
        g.refine(partition=True)
    t = time.perf_counter()
    for g in dgs:
        g.refine(partition=True)
    loop_s = time.perf_counter() - t
    for g in dgs:
        g.free()
    loop_rate = loop_sample / loop_s
    return {"workload": f"batch_mixed_{n_graphs}_n10_5000",
            "graphs": n_graphs, "nodes": hb.n_nodes, "lines": int(len(hb.dst)),
            "graphs_per_s": n_graphs / wall, "ms_per_batch": wall * 1e3,
            "device_ms_per_batch": dev / reps, "blocks": blocks,
            "launch": {"threads": threads, "workgroups": wgs,
                       "lds_sort_digit_bits_per_tier": launch["digit_bits"],
                       "graphs_per_tier": launch["tier_graphs"],
                       "autotuned": True},
            "per_graph_loop_graphs_per_s": loop_rate,
            "per_graph_loop_sample": loop_sample,
            "speedup_vs_per_graph_loop": (n_graphs / wall) / loop_rate}


def serving_leg(ctx, rates, n_req: int, pool_size: int, seed: int, workers: int = 16) -> dict:
    """Poisson arrivals of refinement requests over a pool of R-MAT graphs
    (E log-uniform in [1e3, 5e8], N = E/16, colour 0): small graphs as host
    graphs (batched), larger ones resident in HBM (binned from 2^22 lines).  Reports p50/p99 latency
    (arrival -> result) and graphs/s per rate, against a plain loop that
    refines the same requests one after another, and checks that every
    served partition equals the loop's."""
    import numpy as np
    import rbsc_amd
    from rbsc_amd import gen
    rng = np.random.default_rng(seed)
    E = np.exp(rng.uniform(np.log(1e3), np.log(5e8), pool_size)).astype(np.int64)
    pool, pool_bytes = [], 0
    t0 = time.perf_counter()
    for k, e in enumerate(E):
        scale = int(max(3, round(np.log2(e / 16))))
        if (1 << scale) <= 5120:
            pool.append(("host", gen.rmat(scale, int(e), seed=seed + k)))
        else:
            dg = ctx.gen_rmat(scale, int(e), 0.57, 0.19, 0.19, 0.0, seed + k)
            if e >= (1 << 22):  # served many times: binned once, like the main workload
                dg.bin()
            pool_bytes += dg.nbytes
            pool.append(("resident", dg))
    log(f"[bench] serving pool: {pool_size} graphs, {int(E.sum())} lines, "
        f"{pool_bytes / 2**30:.1f} GiB resident, built in {time.perf_counter() - t0:.1f}s")

    def submit(srv, item, delay):
        kind, g = item
        return srv.submit(g, delay_s=delay) if kind == "host" else srv.submit_resident(g, delay_s=delay)

    # plain loop: the same kind of requests, one at a time, no overlap
    picks0 = rng.integers(0, pool_size, n_req)
    loop_part = {}
    t = time.perf_counter()
    for i in picks0:
        kind, g = pool[i]
        r = ctx.refine(g) if kind == "host" else g.refine()
        loop_part[int(i)] = r.partition
    loop_s = time.perf_counter() - t
    out = {"workload": f"rmat_pool{pool_size}_E1e3_5e8_poisson", "requests_per_rate": n_req,
           "pool_lines": int(E.sum()), "plain_loop_graphs_per_s": n_req / loop_s, "rates": []}
    budget = (250 << 30) - pool_bytes
    # one long-lived server takes every rate (its workers and contexts stay up
    # between rates, as in a serving process); batch counts are per rate, the
    # admitted-bytes peak is the server's so far
    srv = rbsc_amd.Server(workers=workers, hbm_budget=budget, keep_partition=True)
    try:
        out["rates"] = _serve_rates(srv, pool, rates, n_req, pool_size, rng, loop_part, submit)
    finally:
        srv.close()
    for kind, g in pool:
        if kind == "resident":
            g.free()
    return out


def _serve_rates(srv, pool, rates, n_req, pool_size, rng, loop_part, submit) -> list:
    """One record per arrival rate, all served by `srv` (see serving_leg)."""
    import numpy as np
    recs = []
    st = srv.stats()
    for lam in rates:
        gaps = rng.exponential(1.0 / lam, n_req)
        arrivals = np.cumsum(gaps) - gaps[0]
        picks = rng.integers(0, pool_size, n_req)
        st0 = st
        t = time.perf_counter()
        ts = [submit(srv, pool[i], float(a)) for i, a in zip(picks, arrivals)]
        rs = [srv.wait(tk) for tk in ts]
        wall = time.perf_counter() - t
        st = srv.stats()
        lat = np.array([r.latency_s for r in rs])
        first = min(r.arrival_s for r in rs)
        last = max(r.done_s for r in rs)
        mism = 0
        for i, r in zip(picks, rs):
            i = int(i)
            if i in loop_part and not np.array_equal(r.partition, loop_part[i]):
                mism += 1
        rec = {"rate_per_s": lam, "p50_latency_ms": float(np.percentile(lat, 50) * 1e3),
               "p99_latency_ms": float(np.percentile(lat, 99) * 1e3),
               "graphs_per_s": n_req / (last - first), "wall_s": wall,
               "batches": st["batches"] - st0["batches"],
               "batched_graphs": st["batched_graphs"] - st0["batched_graphs"],
               "peak_admitted_gib": st["peak_admitted_bytes"] / 2**30,
               "partition_mismatches_vs_loop": mism}
        log(f"[bench] serving rate {lam}/s: p50 {rec['p50_latency_ms']:.1f} ms, "
            f"p99 {rec['p99_latency_ms']:.1f} ms, {rec['graphs_per_s']:.0f} graphs/s, "
            f"mismatches {mism}")
        recs.append(rec)
    return recs


def measured_traffic(workload: str, kernel: str):
    """HBM bytes per launch of `kernel` (or per round of the signature pass)
    on `workload` from the committed counter summary (FETCH_SIZE and
    WRITE_SIZE runs, calibrated against a copy of known size), or None if
    there is none."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_hbm_traffic.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(line.split(":", 1)[1].strip())
    except OSError:
        pass
    return None


def pinned_copy_peak(nbytes: int = 1 << 30, reps: int = 5) -> float:
    """Host-to-device rate (bytes/s) of hipMemcpyAsync from page-locked
    memory on this box: the attainable peak of the streamed path."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    h, d, st = C.c_void_p(), C.c_void_p(), C.c_void_p()
    e0, e1 = C.c_void_p(), C.c_void_p()
    assert hip.hipHostMalloc(C.byref(h), C.c_size_t(nbytes), 0) == 0
    assert hip.hipMalloc(C.byref(d), C.c_size_t(nbytes)) == 0
    assert hip.hipStreamCreate(C.byref(st)) == 0
    assert hip.hipEventCreate(C.byref(e0)) == 0 and hip.hipEventCreate(C.byref(e1)) == 0
    C.memset(h, 1, nbytes)
    assert hip.hipMemcpyAsync(d, h, C.c_size_t(nbytes), 1, st) == 0  # warm
    assert hip.hipEventRecord(e0, st) == 0
    for _ in range(reps):
        assert hip.hipMemcpyAsync(d, h, C.c_size_t(nbytes), 1, st) == 0
    assert hip.hipEventRecord(e1, st) == 0
    assert hip.hipEventSynchronize(e1) == 0
    ms = C.c_float()
    assert hip.hipEventElapsedTime(C.byref(ms), e0, e1) == 0
    hip.hipEventDestroy(e0)
    hip.hipEventDestroy(e1)
    hip.hipStreamDestroy(st)
    hip.hipFree(d)
    hip.hipHostFree(h)
    return reps * nbytes / (ms.value / 1e3)


def streaming_leg(dg, E: int, frac: float, steps: int) -> dict:
    """The same graph with its lines in pinned host memory, refined in edge
    batches of frac * E lines streamed over PCIe each round (double-buffered
    copies on a side stream, overlapped with the signature kernel).  The
    rate over the link is priced against the pinned host-to-device copy
    rate measured on the same box."""
    peak = pinned_copy_peak()
    t0 = time.perf_counter()
    dg.lines_to_host()
    log(f"[bench] lines moved to pinned host memory in {time.perf_counter() - t0:.1f}s")
    B = int(E * frac)
    dg.refine(partition=False, max_batch_edges=B)
    rounds, dev_ms = 0, 0.0
    t = time.perf_counter()
    for _ in range(steps):
        r = dg.refine(partition=False, max_batch_edges=B)
        rounds += r.rounds
        dev_ms += r.device_ms
    wall = time.perf_counter() - t
    # every round but the first (whose signatures come in closed form from
    # per-row weights) streams the stored line array (unweighted heads)
    pcie = 4 * dg.stored_lines
    line_rounds = rounds - steps
    rate = pcie * line_rounds / (dev_ms / 1e3)
    return {"workload": "same graph, lines in pinned host memory", "batch_fraction": frac,
            "max_batch_edges": B, "value": rounds * E / wall, "unit": "edge-visits/s",
            "ms_per_round": dev_ms / rounds, "pcie_bytes_per_round": pcie,
            "streamed_rounds_per_step": line_rounds // steps,
            "pcie_GBps": rate / 1e9,
            "roofline": {"bound": "pcie", "achieved": rate / 1e9, "peak": peak / 1e9,
                         "unit": "GB/s", "frac": rate / peak,
                         "peak_source": "hipMemcpyAsync of 1 GiB from page-locked memory, "
                                        "measured in this run"},
            "blocks": r.nodes}


def one_shot_leg(ctx, dg) -> dict:
    """The drop-in's whole run on the same graph from host arrays
    (rbsc_refine: upload, identical-row merge and relabelling, layout,
    rounds, canonical partition), the span the reference's timer covers
    after parsing (rbsc.cu:272-489).  The graph is first copied back to the
    host in CSR form."""
    t = time.perf_counter()
    hg = dg.download()
    dl = time.perf_counter() - t
    ctx.refine(hg)  # warm: device allocations of this size exist afterwards
    t = time.perf_counter()
    r = ctx.refine(hg)
    wall = time.perf_counter() - t
    # the same from page-locked arrays, as the rbsc command parses into, the
    # partition too (the device writes the labels into it directly)
    import numpy as np
    from rbsc_amd import HostGraph, PinnedArray
    pd = PinnedArray(hg.n_edges).array
    pro = PinnedArray(hg.n_nodes + 1, np.uint64).array
    ppart = PinnedArray(hg.n_nodes).array
    pd[:] = hg.dst
    pro[:] = hg.row_offsets
    hp = HostGraph(hg.n_nodes, pd[:hg.n_edges], row_offsets=pro[:hg.n_nodes + 1])
    t = time.perf_counter()
    rp = ctx.refine(hp, partition_out=ppart)
    wall_p = time.perf_counter() - t
    assert (rp.nodes, rp.rounds) == (r.nodes, r.rounds)
    assert np.array_equal(rp.partition, r.partition)
    del hp, pd, pro, ppart
    return {"workload": "same graph from host CSR arrays (rbsc_refine)", "time_s": rp.time_s,
            "wall_s": wall_p, "host_memory": "page-locked", "rounds": rp.rounds,
            "blocks": rp.nodes, "device_ms": rp.device_ms,
            "pageable": {"time_s": r.time_s, "wall_s": wall}, "download_s": dl}


def cli_leg(dg) -> dict:
    """The rbsc command (rbsc_amd/bin/rbsc, the reference binary's drop-in)
    on the workload's text in the reference's stdin format, piped from the
    text writer: its three output lines (rbsc.cu:492-494; `time` spans
    upload, rounds and canonical partition, as the reference's timer) and
    its parse time, once plain and once with --verify."""
    import subprocess
    from rbsc_amd._lib import CLI_PATH
    hg = dg.download()
    out = {"workload": f"the same graph as text ({hg.n_edges} lines), piped into rbsc_amd/bin/rbsc"}
    for name, extra in (("plain", []), ("verify", ["--verify"])):
        p = subprocess.Popen([CLI_PATH, str(hg.n_edges), "1"] + extra, stdin=subprocess.PIPE,
                             stdout=subprocess.PIPE, stderr=subprocess.PIPE)
        t = time.perf_counter()
        hg.write_text(f"/proc/self/fd/{p.stdin.fileno()}")
        p.stdin.close()
        p.stdin = None
        so, se = p.communicate(timeout=900)
        wall = time.perf_counter() - t
        if p.returncode:
            raise SystemExit(f"bench: rbsc failed: {se.decode()}")
        lines = so.decode().split("\n")
        err = se.decode()
        parse = [x for x in err.split("\n") if "read and parsed" in x]
        ver = [x for x in err.split("\n") if "verified=" in x]
        out[name] = {"time_s": float(lines[0]), "done": int(lines[1]), "nodes": int(lines[2]),
                     "wall_s": wall,
                     "parse_s": float(parse[0].split(" in ")[1].split(" s")[0]) if parse else None,
                     "verified": int(ver[0].split("verified=")[1]) if ver else None}
    return out


def quotient_leg(ctx) -> dict:
    """The reference's multi-batch quotient mode (rbsc.cu:274-487 and the
    stop rule of rbsc_aggressive.cu:487, its cuda_50 / cuda_25 columns) on
    the 2^20-node R-MAT with 30% clones, from host arrays (rbsc_refine: the
    one upload, every pass's batches and contraction on the device, the
    canonical partition back)."""
    from rbsc_amd import gen
    g = gen.rmat_2p20_clone30()
    ctx.refine(g, max_batch_edges=g.n_edges // 2, mode="quotient")  # warm
    out = {"workload": f"rmat_2p20_clone30: {g.n_nodes} nodes, {g.n_edges} lines", "runs": []}
    for frac in (2, 4):
        for mode in ("quotient", "aggressive"):
            B = g.n_edges // frac
            t = time.perf_counter()
            r = ctx.refine(g, max_batch_edges=B, mode=mode)
            wall = time.perf_counter() - t
            out["runs"].append({"batch_frac": 1 / frac, "mode": mode, "max_batch_edges": B,
                                "time_s": r.time_s, "wall_s": wall, "device_ms": r.device_ms,
                                "nodes": r.nodes, "done": r.done, "passes": r.passes,
                                "rounds": r.rounds})
    return out


def copy_probe(nbytes: int) -> None:
    """One plain device-to-device copy of nbytes through the HIP runtime, so a
    counter run can read FETCH_SIZE / WRITE_SIZE of a kernel of known size."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # device to device
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"[bench] copy probe: {nbytes} bytes device to device")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="rmat_s27_d31_c20", choices=sorted(CONFIGS))
    ap.add_argument("--cpu-scale", type=int, default=21)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--parts", type=int, default=1,
                    help="destination partitions of the line layout (power of two)")
    ap.add_argument("--hot", type=int, default=0,
                    help="two-pass layout: lines into the HOT most-gathered heads first")
    ap.add_argument("--fullsort", action="store_true",
                    help="sort all 64 signature bits (default: top 32 bits + colliding runs)")
    ap.add_argument("--hashed64", action="store_true",
                    help="gather per-node hashed colours instead of raw u32 colours")
    ap.add_argument("--xcd", action="store_true",
                    help="with --parts 8: one launch, partition b mod 8 per workgroup b")
    ap.add_argument("--bins", type=int, default=1,
                    help="1: binned signature layout (rbsc_graph_bin) built before timing; "
                         "0: plain line layout")
    ap.add_argument("--bin-hot", type=int, default=0, help="hot heads of the binned layout (0 = default)")
    ap.add_argument("--bin-dense", type=int, default=0,
                    help="hot heads gathered inside the bin pass; the others are warm (0 = default)")
    ap.add_argument("--bin-rows", type=int, default=0, help="rows per bin (0 = default)")
    ap.add_argument("--bin-chunk", type=int, default=0, help="slots per chunk (0 = default)")
    ap.add_argument("--bin-scatter", type=int, default=0,
                    help="cold lines per scatter workgroup (0 = default)")
    ap.add_argument("--bin-hashed", action="store_true",
                    help="hashed binned layout: per-node u64 hashed colours, no hashing per line")
    ap.add_argument("--gather-first", action="store_true",
                    help="gather the first round's colours instead of its closed form (comparison)")
    ap.add_argument("--bin-slot-order", action="store_true",
                    help="scatter records in slot order inside each workgroup's share")
    ap.add_argument("--batch-graphs", type=int, default=50000,
                    help="graphs in the many-small-graphs leg (0 skips it)")
    ap.add_argument("--serve-requests", type=int, default=200,
                    help="requests per arrival rate in the serving leg (0 skips it)")
    ap.add_argument("--serve-rates", default="50,200,1000,2000")
    ap.add_argument("--serve-workers", type=int, default=16,
                    help="worker threads (contexts) of the serving loop")
    ap.add_argument("--stream-frac", type=float, default=0.25,
                    help="also time the pinned-host streaming path at this edge-batch fraction (0 skips)")
    ap.add_argument("--cli", type=int, default=1,
                    help="run the rbsc command on the workload's text (0 = skip)")
    ap.add_argument("--quotient", type=int, default=1,
                    help="time the multi-batch quotient mode at 50%% / 25%% batches (0 = skip)")
    ap.add_argument("--one-shot", type=int, default=1,
                    help="also time the drop-in's whole run (rbsc_refine from host arrays)")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="also run one device-to-device copy of this size (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one MI355X: the RBSC round runs on a single GPU in this "
              "build (no multi-GPU path); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import rbsc_amd
    if args.calib_copy_mib:
        copy_probe(args.calib_copy_mib << 20)
    scale, ldraws, clone, seed = CONFIGS[args.config]
    ctx = rbsc_amd.Context(0)
    t0 = time.perf_counter()
    dg = ctx.gen_rmat(scale, 1 << ldraws, 0.57, 0.19, 0.19, clone, seed)
    N, E = dg.dims
    log(f"[bench] generated {args.config}: N={N} E={E} in {time.perf_counter() - t0:.1f}s")
    if args.parts > 1:
        t0 = time.perf_counter()
        dg.relayout(args.parts)
        log(f"[bench] relayout into {args.parts} destination partitions in {time.perf_counter() - t0:.1f}s")

    if args.hot:
        t0 = time.perf_counter()
        dg.split_hot(args.hot)
        log(f"[bench] hot/cold split at {args.hot} heads in {time.perf_counter() - t0:.1f}s")
    bin_info = None
    if args.bins and args.parts == 1 and not args.hot:
        t0 = time.perf_counter()
        dg.bin(args.bin_hot, args.bin_rows, args.bin_chunk, args.bin_scatter, args.bin_slot_order,
               args.bin_hashed, args.bin_dense)
        bin_info = dict(dg.bin_info, build_s=time.perf_counter() - t0)
        log(f"[bench] binned layout in {bin_info['build_s']:.1f}s: {bin_info}")
    flags = (rbsc_amd.FLAG_FULLSORT if args.fullsort else 0) | (rbsc_amd.FLAG_HASHED64 if args.hashed64 else 0) | (rbsc_amd.FLAG_XCD if args.xcd else 0) | (rbsc_amd.FLAG_GATHER_FIRST if args.gather_first else 0)
    for i in range(args.warmup):
        r = dg.refine(partition=False, flags=flags)
        log(f"[bench] warmup {i}: rounds={r.rounds} blocks={r.nodes} device_ms={r.device_ms:.2f}")

    # Exactness of the measured path, outside the timed region: the exact
    # stability check on the final partition, and the same refinement with
    # the first round ranked by a full sort instead of the hash table.
    chk = dg.refine(verify=True, flags=flags)
    srt = dg.refine(partition=False, flags=flags | rbsc_amd.FLAG_SORTRANK)
    checks = {"verified": chk.verified, "blocks": chk.nodes, "rounds": chk.rounds,
              "sortrank_blocks": srt.nodes, "sortrank_rounds": srt.rounds,
              "sortrank_same": (srt.nodes, srt.rounds) == (chk.nodes, chk.rounds)}
    log(f"[bench] checks: {checks}")
    if chk.verified != 1 or not checks["sortrank_same"]:
        raise SystemExit(f"bench: the measured refinement failed its checks: {checks}")
    del chk

    ctx.profile(True)
    ctx.profile_read(reset=True)
    rounds = 0
    dev_ms = 0.0
    t = time.perf_counter()
    for i in range(args.steps):
        r = dg.refine(partition=False, flags=flags)
        rounds += r.rounds
        dev_ms += r.device_ms
    wall = time.perf_counter() - t
    phases = ctx.profile_read(reset=True)
    ctx.profile(False)
    slow = ctx.stats(reset=True)
    log(f"[bench] timed {args.steps} steps: rounds={rounds} wall={wall:.3f}s device={dev_ms:.1f}ms")

    visits = rounds * E
    value = visits / wall
    ms_round = dev_ms / rounds
    U = dg.rows            # rows the signature pass visits (identical rows once)
    E_st = dg.stored_lines  # their lines
    # Whole round: 8 B per line the pass reads (head id + neighbour colour;
    # identical rows are stored once, so E_st lines) and 32 B per node (row
    # offset, signature write and read, old and new colour).  Beside it the
    # reference-equivalent figure over every input line (8 E + 32 N), the
    # bytes a pass over all lines would move.
    b_round = 8 * E_st + 32 * N
    b_round_ref = 8 * E + 32 * N
    # The signature pass (the reference's set_values + ReduceByKey): 8 B per
    # stored line (head id + neighbour colour) plus 16 B per stored row (row
    # offset or per-row term, signature write), once per round.  On the
    # binned layout the pass is three launches, warm_gather_kernel and
    # cold_stage_kernel (phase "scatter": they fill the buffer of warm and
    # cold colours) and bin_sig_kernel (phase "signature"); on the plain
    # layout it is sig_kernel alone.
    b_sig = 8 * E_st + 16 * U
    # The first round from a uniform start takes its signatures in closed
    # form from per-row weights (no line is read; timed as set-up), so the
    # line pass runs in rounds - steps of the rounds: its kernels are timed
    # per launch, the round as the average over all rounds.
    per_round = {k: phases[k][0] / rounds for k in phases}
    per_launch = {k: phases[k][0] / max(phases[k][1], 1) for k in phases}
    line_pass_rounds = phases["signature"][1]
    ROUND = ("scatter", "signature", "finish", "gather", "sort", "rank", "setup")
    round_ms = sum(per_round[k] for k in ROUND)
    dom = max(ROUND, key=lambda k: per_round[k])
    binned = bin_info is not None and (bin_info["hashed"] or not (flags & rbsc_amd.FLAG_HASHED64))
    sig_ms = per_launch["signature"] + (per_launch["scatter"] if binned else 0.0)
    pass_name = "signature_pass" if binned else "sig_kernel"
    kernels = ({"warm_gather_kernel + cold_stage_kernel": per_launch["scatter"],
                "bin_sig_kernel": per_launch["signature"]}
               if binned else {"sig_kernel": per_launch["signature"]})
    # bytes of the average round: the line pass's in the rounds that run it,
    # the node side's in every round
    b_round_avg = (line_pass_rounds * 8 * E_st + rounds * 32 * N) / rounds
    traffic = measured_traffic(args.config, pass_name) if args.parts == 1 and not args.hot else None
    out = {
        "metric": "edge-visits/s per WL round (% HBM roofline); batched graphs/s, p99 latency",
        "value": value,
        "unit": "edge-visits/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": args.config, "n_nodes": N, "n_lines": E, "rows": U,
                   "stored_lines": E_st,
                   "parts": args.parts, "hot": args.hot, "bins": bin_info, "fullsort": bool(args.fullsort), "hashed64": bool(args.hashed64), "xcd": bool(args.xcd),
                   "rounds_per_step": rounds // args.steps,
                   "line_pass_rounds_per_step": line_pass_rounds // args.steps,
                   "node_side_slow_paths": slow,
                   "blocks": r.nodes, "ms_per_round": ms_round},
        "roofline": {"bound": "hbm", "kernel": " + ".join(kernels),
                     "achieved": b_sig / (sig_ms / 1e3) / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": b_sig / (sig_ms / 1e3) / HBM_PEAK,
                     "traffic": traffic,
                     "alg_bytes_per_launch": b_sig, "launch_ms": sig_ms,
                     "kernel_ms": kernels,
                     "dominant_phase": dom,
                     "round": {"alg_bytes": b_round_avg, "ms": round_ms,
                               "achieved": b_round_avg / (round_ms / 1e3) / 1e9,
                               "frac": b_round_avg / (round_ms / 1e3) / HBM_PEAK,
                               "phase_ms": per_round,
                               "line_pass_bytes": b_round,
                               "reference_equivalent": {
                                   "alg_bytes": b_round_ref,
                                   "achieved": b_round_ref / (round_ms / 1e3) / 1e9,
                                   "frac": b_round_ref / (round_ms / 1e3) / HBM_PEAK}}},
        "checks": checks,
    }
    if bin_info is not None:
        out["config"]["layout_build_s"] = bin_info["build_s"]
    if args.one_shot:
        out["one_shot"] = one_shot_leg(ctx, dg)
        log(f"[bench] one-shot: {out['one_shot']}")
    if args.cli and args.parts == 1 and not args.hot:
        out["cli"] = cli_leg(dg)
        log(f"[bench] rbsc command: {out['cli']}")
    if args.quotient:
        out["quotient"] = quotient_leg(ctx)
        log(f"[bench] quotient: {out['quotient']}")
    if args.batch_graphs:
        out["batched"] = batch_leg(ctx, args.batch_graphs, 1004, 500)
        log(f"[bench] batched: {out['batched']['graphs_per_s']:.0f} graphs/s "
            f"(per-graph loop {out['batched']['per_graph_loop_graphs_per_s']:.0f} graphs/s)")
    if args.serve_requests:
        rates = [float(x) for x in args.serve_rates.split(",") if x]
        out["serving"] = serving_leg(ctx, rates, args.serve_requests, 64, 1005, args.serve_workers)
    if args.stream_frac and args.parts == 1 and not args.hot:
        out["streamed"] = streaming_leg(dg, E, args.stream_frac, max(1, min(args.steps, 2)))
        log(f"[bench] streamed at {args.stream_frac}: {out['streamed']['value']:.3g} edge-visits/s, "
            f"{out['streamed']['pcie_GBps']:.1f} GB/s over PCIe")
    if not args.no_cpu:
        log("[bench] cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(args.cpu_scale, args.cpu_scale + 4, 0.20, seed)
        full = full_size_reference(args.config, r.nodes, wall / args.steps,
                                   out.get("one_shot", {}).get("time_s"),
                                   out.get("cli", {}).get("plain", {}).get("time_s"))
        if full is not None:
            out["cpu_baseline"]["full_size"] = full
    dg.free()
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
