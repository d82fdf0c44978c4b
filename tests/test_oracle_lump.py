"""The oracle's exact lumping refiner by splitters (oracle/lump.c), the CPU
counterpart of the reference's Valmari tool (src/part_ref/MDPmin.cc, whose
lumping definition is MDP_minimize.cc:55-65).

It is pinned by the block counts the compiled Valmari tool printed for every
committed fixture and for the large workloads (tests/golden/fixtures.json),
by the hash-free round-based refiner of rbsc_oracle.c (same partition on
random graphs with weights and initial colours), by the stability and
coarseness of its result checked from the definition, and by a planted
mistake in a compiled copy (the worklist rule applied to waiting blocks)
that these checks reject."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle
from rbsc_amd import gen

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "fixtures.json")


def _fixtures():
    with open(GOLDEN) as f:
        return {r["name"]: r for r in json.load(f)["fixtures"]}


def _random_graphs(count, seed):
    rng = np.random.default_rng(seed)
    for t in range(count):
        n = int(rng.integers(1, 40))
        E = int(rng.integers(0, 120))
        src = np.sort(rng.integers(0, n, E)).astype(np.uint32)
        dst = rng.integers(0, n, E).astype(np.uint32)
        w = None if t % 3 else rng.integers(1, 4, E).astype(np.uint64)
        init = None if t % 2 else rng.integers(0, 3, n).astype(np.uint32)
        yield n, src, dst, w, init


def _stable(n, src, dst, w, init, labels):
    """From the definition: every block's members send equal total weight
    into every block, and no block mixes initial colours."""
    K = int(labels.max()) + 1 if n else 0
    wt = np.ones(len(dst), np.uint64) if w is None else w
    M = np.zeros((n, K), dtype=object)
    for s, d, x in zip(src.tolist(), dst.tolist(), wt.tolist()):
        M[s, labels[d]] += x
    for b in range(K):
        members = np.flatnonzero(labels == b)
        for u in members[1:]:
            if not all(M[u, c] == M[members[0], c] for c in range(K)):
                return False
        if init is not None and len(set(init[members].tolist())) != 1:
            return False
    return True


def test_lump_equals_the_round_based_exact_refiner():
    for n, src, dst, w, init in _random_graphs(300, 1):
        c, K, _ = oracle.exact_refine(n, src, dst, w, init)
        lab, K2 = oracle.lump_refine(n, src, dst, w, init)
        assert K2 == K
        assert np.array_equal(lab, oracle.canonical(c))


def test_lump_result_is_stable_by_definition():
    for n, src, dst, w, init in _random_graphs(60, 2):
        lab, _ = oracle.lump_refine(n, src, dst, w, init)
        assert _stable(n, src, dst, w, init, lab)


def test_lump_closed_forms():
    # a directed cycle: one block; a path: every node its own block (its
    # distance to the end); a star into a hub: leaves together
    n = 12
    cyc = np.arange(n, dtype=np.uint32)
    assert oracle.lump_refine(n, cyc, (cyc + 1) % n)[1] == 1
    path_s = np.arange(n - 1, dtype=np.uint32)
    assert oracle.lump_refine(n, path_s, path_s + 1)[1] == n
    leaves = np.arange(1, n, dtype=np.uint32)
    lab, K = oracle.lump_refine(n, leaves, np.zeros(n - 1, np.uint32))
    assert K == 2 and len(set(lab[1:].tolist())) == 1


@pytest.mark.parametrize("name", sorted(r for r in _fixtures()
                                        if r not in ("ba_8m_m6_init3",)))
def test_lump_block_counts_equal_the_valmari_tool(name):
    rec = _fixtures()[name]
    if name == "rmat_2p20_clone30":
        g = gen.rmat_2p20_clone30()
    elif name == "cnr-2000":
        pytest.skip("covered by the BVGraph tests (decoder needed)")
    else:
        import importlib.util
        spec = importlib.util.spec_from_file_location(
            "make_golden", os.path.join(os.path.dirname(__file__), "golden", "make_golden.py"))
        mg = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mg)
        g = dict(mg.SPECS)[name]()
        assert mg.graph_hash(g) == rec["sha256"]
    lab, K = oracle.lump_refine(g.n_nodes, g.sources(), g.dst, g.weight, g.init_color)
    assert K == rec["valmari_nodes"]
    c, Kr, _, _ = oracle.rbsc_refine(g.n_nodes, g.sources(), g.dst, g.weight, g.init_color)
    assert np.array_equal(lab, oracle.canonical(c))


# Hopcroft's rule applied wrongly: the largest part of a block left out of
# the worklist even when the block was still waiting there (its weights are
# then never made uniform)
MISTAKE = ("if ((was_waiting || b != largest_b) && !waiting[b]) {",
           "if (b != largest_b && !waiting[b]) {")


def test_planted_worklist_mistake_is_caught(tmp_path):
    src = os.path.join(os.path.dirname(oracle.__file__), "lump.c")
    text = open(src).read()
    assert text.count(MISTAKE[0]) == 1, "the planted mistake no longer matches lump.c"
    libs = {}
    for name, t in (("unchanged", text), ("mistake", text.replace(*MISTAKE))):
        c = tmp_path / f"{name}.c"
        c.write_text(t)
        so = tmp_path / f"lib{name}.so"
        base = os.path.join(os.path.dirname(oracle.__file__), "rbsc_oracle.c")
        subprocess.run([shutil.which("gcc") or "gcc", "-O2", "-shared", "-fPIC", "-o", str(so),
                        base, str(c), "-lm"], check=True, capture_output=True)
        libs[name] = oracle.load_lib(str(so))

    def wrong(lib):
        for n, s, d, w, init in _random_graphs(300, 3):
            c, K, _ = oracle.exact_refine(n, s, d, w, init)
            lab, _ = oracle.lump_refine(n, s, d, w, init, _from=lib)
            if not np.array_equal(lab, oracle.canonical(c)):
                return True
        return False

    assert not wrong(libs["unchanged"])
    assert wrong(libs["mistake"]), "planted worklist mistake not caught"
