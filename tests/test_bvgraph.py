"""The BVGraph decoder against the data the reference's own tests hold
(tests/golden/bvgraph, from webgraph-rs/data): every node's first bit
equals its .offsets entry, the decoded transpose file is the transpose of
the decoded graph, the first 55 successor lists match cnr-2000_edges.txt,
the strongly connected components number 100977 (webgraph-rs
algo/tests/test_sccs.rs:177), and the converter prints the reference
producer's formats."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from rbsc_amd import gen
from rbsc_amd._lib import CLI_PATH, WEBGRAPH_TO_PATH

D = os.path.join(os.path.dirname(__file__), "golden", "bvgraph")
CNR = os.path.join(D, "cnr-2000")


@pytest.fixture(scope="module")
def cnr():
    return gen.bvgraph(CNR)


def test_counts_and_first_rows(cnr):
    assert (cnr.n_nodes, cnr.n_edges) == (325557, 3216152)
    with open(os.path.join(D, "cnr-2000_edges.txt")) as f:
        rows = [list(map(int, l.split())) for l in f.read().strip("\n").split("\n")]
    assert len(rows) == 55
    for i, want in enumerate(rows):
        a, b = int(cnr.row_offsets[i]), int(cnr.row_offsets[i + 1])
        assert cnr.dst[a:b].tolist() == want
    # successor lists are sorted and in range
    src = cnr.sources().astype(np.int64)
    key = (src << 32) | cnr.dst.astype(np.int64)
    assert np.all(np.diff(key) > 0) and cnr.dst.max() < cnr.n_nodes


@pytest.mark.parametrize("name", ["cnr-2000", "cnr-2000-t"])
def test_bit_offsets_match_offsets_file(name):
    n, bad = gen.bvgraph_check_offsets(os.path.join(D, name))
    assert n == 325557 + 1 and bad == 0


def test_transpose_file_is_the_transpose(cnr):
    t = gen.bvgraph(CNR + "-t")
    src = cnr.sources().astype(np.int64)
    dst = cnr.dst.astype(np.int64)
    o = np.lexsort((src, dst))
    assert np.array_equal(t.sources().astype(np.int64), dst[o])
    assert np.array_equal(t.dst.astype(np.int64), src[o])


def test_strongly_connected_components(cnr):
    import scipy.sparse as sp
    from scipy.sparse.csgraph import connected_components
    n = cnr.n_nodes
    a = sp.csr_matrix((np.ones(cnr.n_edges, np.int8), cnr.dst, cnr.row_offsets.astype(np.int64)),
                      shape=(n, n))
    k, _ = connected_components(a, directed=True, connection="strong")
    assert k == 100977


def test_small_graph_and_errors(tmp_path):
    g = gen.bvgraph(os.path.join(D, "test"))
    assert (g.n_nodes, g.n_edges) == (24, 28)
    with pytest.raises(Exception):
        gen.bvgraph(str(tmp_path / "missing"))
    # a truncated bit stream is reported, not decoded past its end
    for ext in ("properties",):
        (tmp_path / f"cut.{ext}").write_bytes(open(CNR + f".{ext}", "rb").read())
    (tmp_path / "cut.graph").write_bytes(open(CNR + ".graph", "rb").read()[:100000])
    with pytest.raises(Exception):
        gen.bvgraph(str(tmp_path / "cut"))


def test_converter_formats(cnr, tmp_path):
    g = gen.bvgraph(os.path.join(D, "test"))
    out = subprocess.run([WEBGRAPH_TO_PATH, "arcs-cuda", os.path.join(D, "test.graph")],
                         stdout=subprocess.PIPE, check=True).stdout
    assert out == g.to_text()
    out = subprocess.run([WEBGRAPH_TO_PATH, "arcs-valmari", os.path.join(D, "test")],
                         stdout=subprocess.PIPE, check=True).stdout
    assert out == g.to_valmari_text()
    out = subprocess.run([WEBGRAPH_TO_PATH, "gdata", CNR], stdout=subprocess.PIPE,
                         check=True).stdout
    assert out == b"325557\n3216152\n"


def test_cnr_refinement_matches_reference_tools(cnr):
    """The oracle on cnr-2000 gives the block count the reference's rsc and
    Valmari printed (golden fixture, tests/golden/make_golden.py)."""
    import json
    with open(os.path.join(os.path.dirname(__file__), "golden", "fixtures.json")) as f:
        rec = {r["name"]: r for r in json.load(f)["fixtures"]}["cnr-2000"]
    _, K, _, _ = oracle.rbsc_refine(cnr.n_nodes, cnr.sources(), cnr.dst)
    assert K == rec["rsc_nodes"] == rec["valmari_nodes"]


@pytest.mark.gpu
def test_cnr_through_the_rbsc_command():
    import json
    with open(os.path.join(os.path.dirname(__file__), "golden", "fixtures.json")) as f:
        rec = {r["name"]: r for r in json.load(f)["fixtures"]}["cnr-2000"]
    cmd = f"{WEBGRAPH_TO_PATH} arcs-cuda {CNR} | {CLI_PATH} 3216152 1 --verify"
    p = subprocess.run(cmd, shell=True, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                       timeout=300)
    assert p.returncode == 0, p.stderr.decode()
    lines = p.stdout.decode().split("\n")
    assert int(lines[1]) == 1 and int(lines[2]) == rec["valmari_nodes"]
    assert b"verified=1" in p.stderr
