"""Write tests/golden/fixtures.json: block counts of the reference's own CPU
tools on seeded graphs.

The tools are the reference's rsc (src/rsc/rsc.cpp) and Valmari's
Markov-lumping minimizer (src/part_ref/MDPmin.cc), compiled from the
reference sources by oracle/Makefile into oracle/_ref/.  Graphs come from the
seeded generators of rbsc_amd.gen (host side, no GPU), identified by their
arguments and a SHA-256 of their arrays so a generator change is detected.

Run from the repository root:  python tests/golden/make_golden.py [NAMES...]
(with names, only those are recomputed and the other records are kept).
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from rbsc_amd import HostGraph, gen  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fixtures.json")


def weighted_rmat(scale, draws, clone, seed, wmax):
    g = gen.rmat(scale, draws, clone_frac=clone, seed=seed)
    rng = np.random.default_rng(seed)
    g.weight = rng.integers(1, wmax + 1, size=g.n_edges).astype(np.uint64)
    return g


def closed_form(name):
    if name == "cycle100":
        n = 100
        src = np.arange(n); dst = (src + 1) % n
    elif name == "path50":
        n = 50
        src = np.arange(n - 1); dst = src + 1
    elif name == "star_out20":
        n = 21
        src = np.zeros(20, np.int64); dst = np.arange(1, 21)
    elif name == "isolated10":
        n = 10
        src = np.zeros(0, np.int64); dst = np.zeros(0, np.int64)
    else:
        raise KeyError(name)
    order = np.lexsort((dst, src))
    return HostGraph(n, dst[order], src=src[order])


SPECS = [
    ("rmat_s10_d14_c30", lambda: gen.rmat(10, 1 << 14, clone_frac=0.3, seed=5)),
    ("rmat_s12_d16_c30", lambda: gen.rmat(12, 1 << 16, clone_frac=0.3, seed=7)),
    ("rmat_s14_d18_c20", lambda: gen.rmat(14, 1 << 18, clone_frac=0.2, seed=11)),
    ("rmat_s13_d17_c00", lambda: gen.rmat(13, 1 << 17, clone_frac=0.0, seed=29)),
    ("rmat_s11_d15_w5", lambda: weighted_rmat(11, 1 << 15, 0.3, 31, 5)),
    ("ba_n5000_m3_dup10_init3", lambda: gen.ba(5000, 3, 0.10, 3, seed=13)),
    ("ba_n4000_m6_dup10", lambda: gen.ba(4000, 6, 0.10, 0, seed=37)),
    ("er_n2000_d4", lambda: gen.er_undirected(2000, 4.0, seed=17)),
    ("reg3_n1000", lambda: gen.regular_undirected(1000, 3, seed=19)),
    ("reg4_n998", lambda: gen.regular_undirected(998, 4, seed=41)),
    ("ba_und_n3000_m2", lambda: gen.ba_undirected(3000, 2, seed=23)),
    ("cycle100", lambda: closed_form("cycle100")),
    ("path50", lambda: closed_form("path50")),
    ("star_out20", lambda: closed_form("star_out20")),
]


# The two named benchmark workloads beside the R-MAT scale-27 one: checked
# on the GPU against these counts (Valmari alone takes about a minute on C2).
LARGE_SPECS = [
    ("rmat_2p20_clone30", gen.rmat_2p20_clone30),
    ("ba_8m_m6_init3", gen.ba_8m_m6_init3),
]
# The LAW graph the reference's tests hold (BVGraph, decoded by the library).
BVGRAPH_SPECS = [
    ("cnr-2000", lambda: gen.bvgraph(os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                  "bvgraph", "cnr-2000"))),
]


def graph_hash(g: HostGraph) -> str:
    h = hashlib.sha256()
    h.update(np.uint64(g.n_nodes).tobytes())
    h.update(g.sources().astype(np.uint32).tobytes())
    h.update(g.dst.astype(np.uint32).tobytes())
    if g.weight is not None:
        h.update(g.weight.astype(np.uint64).tobytes())
    if g.init_color is not None:
        h.update(g.init_color.astype(np.uint32).tobytes())
    return h.hexdigest()


def run_tool(exe, g: HostGraph, fmt: str) -> int:
    """Block count printed by a reference tool on g (written by the library's
    text writer, byte-identical to HostGraph.to_text / to_valmari_text)."""
    import subprocess
    import tempfile
    with tempfile.NamedTemporaryFile(suffix="." + fmt) as f:
        g.write_text(f.name, fmt)
        with open(f.name, "rb") as inp:
            p = subprocess.run([exe], stdin=inp, stdout=subprocess.PIPE, timeout=3600, check=True)
    return int(p.stdout.decode().split("\n")[2])


def main():
    if not oracle.have_ref():
        raise SystemExit("build oracle/_ref first: make -C oracle ref")
    out = []
    old = {}
    if os.path.exists(OUT):
        with open(OUT) as f:
            old = {r["name"]: r for r in json.load(f)["fixtures"]}
    only = set(sys.argv[1:])  # names to (re)compute; others are kept as recorded
    for name, fn in SPECS + LARGE_SPECS + BVGRAPH_SPECS:
        if only and name not in only and name in old:
            out.append(old[name])
            continue
        g = fn()
        rec = {"name": name, "n_nodes": g.n_nodes, "n_edges": g.n_edges, "sha256": graph_hash(g)}
        if g.init_color is None:
            rec["rsc_nodes"] = run_tool(oracle.REF_RSC, g, "lines")
        rec["valmari_nodes"] = run_tool(oracle.REF_VALMARI, g, "valmari")
        out.append(rec)
        print(rec, flush=True)
    with open(OUT, "w") as f:
        json.dump({"tools": {"rsc": "src/rsc/rsc.cpp (g++ -std=c++17 -O2)",
                             "valmari": "src/part_ref/MDPmin.cc (g++ -std=c++11 -O2)"},
                   "fixtures": out}, f, indent=1)


if __name__ == "__main__":
    main()
