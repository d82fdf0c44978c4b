"""The harness counterpart of the reference's bench.py / benchlarge.py:
table schemas, row life cycle (status 0 -> 1 or -1, resume), JSON cells,
time-out records and the shell pipelines, with stand-in tools on the CPU
and the real rbsc command on the GPU."""
import json
import os
import sqlite3
import stat

import numpy as np
import pytest

import oracle
from rbsc_amd import harness

# The reference's tables (bench.db and benchlarge.db as shipped).
REF_BENCH_SQL = ("CREATE TABLE results(name TEXT, tool TEXT, nodes INT, edges INT, valmari TEXT, "
                 "cpu TEXT, cuda TEXT, cuda_50 TEXT, cuda_25 TEXT, mc_32 TEXT, mc_64 TEXT, status INT)")
REF_LARGE_SQL = ("CREATE TABLE results(name TEXT, tool TEXT, nodes INT, edges INT, valmari TEXT, "
                 "cpu TEXT, cuda TEXT, status INT)")


def _tool(tmp_path, name, body):
    p = tmp_path / name
    p.write_text("#!/bin/sh\n" + body + "\n")
    p.chmod(p.stat().st_mode | stat.S_IEXEC)
    return str(p)


def _schema(db):
    con = sqlite3.connect(db)
    sql = con.execute("SELECT sql FROM sqlite_master WHERE name='results'").fetchone()[0]
    con.close()
    return sql


@pytest.mark.parametrize("large", [False, True])
def test_schema_matches_reference(tmp_path, large):
    db = str(tmp_path / "b.db")
    harness.init_db(db, large, names=("rmat_s12_d16_c30_seed1011",))
    assert _schema(db) == (REF_LARGE_SQL if large else REF_BENCH_SQL)
    con = sqlite3.connect(db)
    row = con.execute("SELECT name, tool, status FROM results").fetchall()
    assert row == [("rmat_s12_d16_c30_seed1011", "gen", 0)]


def test_run_life_cycle_with_stand_in_tools(tmp_path):
    db = str(tmp_path / "b.db")
    harness.init_db(db, False, names=("rmat_s12_d16_c30_seed1011", "reg3_n20000_seed1013"))
    con = sqlite3.connect(db)
    con.execute("INSERT INTO results VALUES ('missing', 'file', 1, 1, '', '', '', '', '', '', '', 0)")
    con.execute("INSERT INTO results VALUES ('web', 'law', 1, 2, '', '', '', '', '', '', '', 0)")
    con.execute("INSERT INTO results VALUES ('old', 'gen', 1, 3, 'x', '', '', '', '', '', '', 1)")
    con.commit()
    log = tmp_path / "calls.log"
    # stand-ins: consume stdin, print time / done / nodes; rbsc logs its argv
    part_ref = _tool(tmp_path, "part_ref", "cat > /dev/null; printf '1.5\\n1\\n7\\n'")
    rsc = _tool(tmp_path, "rsc", "cat > /dev/null; echo garbage")
    rbsc = _tool(tmp_path, "rbsc", f"cat > /dev/null; echo \"$@\" >> {log}; printf '0.25\\n1\\n7\\n'")
    n = harness.run_db(db, 1, datasets=str(tmp_path / "datasets"), rbsc=rbsc, rsc=rsc,
                       part_ref=part_ref, timeout=60)
    assert n == 2
    rows = {r[0]: r for r in con.execute("SELECT * FROM results")}
    r = rows["reg3_n20000_seed1013"]
    assert r[11] == 1 and r[2] == 20000 and r[3] == 60000
    assert json.loads(r[4]) == {"done": 1, "nodes": 7, "time": 1.5}
    assert json.loads(r[5]) == {"done": 0, "nodes": 0, "time": 60.0}  # unparsable -> time-out
    for col in (6, 7, 8):
        assert json.loads(r[col]) == {"done": 1, "nodes": 7, "time": 0.25}
    assert json.loads(r[9]) == json.loads(r[10]) == {"done": 0, "nodes": 0, "time": 0.0}
    assert rows["missing"][11] == -1 and rows["web"][11] == -1  # no such file / graph here
    assert rows["old"][4] == "x" and rows["old"][11] == 1  # finished rows are not rerun
    calls = log.read_text().split("\n")
    assert "60000 1" in calls and "30000 1" in calls and "15000 1" in calls  # 100/50/25 %
    # rows run in order of edges: the small graph first
    assert harness.run_db(db, 1, rbsc=rbsc, rsc=rsc, part_ref=part_ref, timeout=60) == 0


def test_file_rows_and_valmari_conversion(tmp_path):
    from rbsc_amd import gen
    ds = tmp_path / "datasets"
    ds.mkdir()
    g = gen.rmat(9, 3000, clone_frac=0.3, seed=4)
    (ds / "small.txt").write_bytes(g.to_text())
    mdp = tmp_path / "small.mdp"
    harness._lines_to_mdp(str(ds / "small.txt"), str(mdp))
    assert mdp.read_bytes() == g.to_valmari_text()


@pytest.mark.skipif(not oracle.have_ref(), reason="reference tools not built")
def test_real_reference_tools_in_the_harness(tmp_path):
    db = str(tmp_path / "b.db")
    harness.init_db(db, True, names=("rmat_s12_d16_c30_seed1011",))
    rbsc = _tool(tmp_path, "rbsc", "cat > /dev/null; printf '0\\n1\\n0\\n'")
    harness.run_db(db, 1, large=True, rbsc=rbsc, rsc=oracle.REF_RSC, part_ref=oracle.REF_VALMARI,
                   timeout=600)
    con = sqlite3.connect(db)
    r = con.execute("SELECT valmari, cpu, status FROM results").fetchone()
    from rbsc_amd import gen
    g = gen.rmat(12, 1 << 16, clone_frac=0.3, seed=1011)
    _, K, _, _ = oracle.rbsc_refine(g.n_nodes, g.sources(), g.dst)
    assert r[2] == 1
    assert json.loads(r[0])["nodes"] == K  # Valmari
    assert json.loads(r[1])["nodes"] == K  # rsc


@pytest.mark.gpu
def test_harness_with_the_rbsc_command(tmp_path):
    from rbsc_amd import gen
    db = str(tmp_path / "b.db")
    harness.init_db(db, False, names=("rmat_s12_d16_c30_seed1011",))
    stand_in = _tool(tmp_path, "cpu_tool", "cat > /dev/null; printf '1\\n1\\n1\\n'")
    harness.run_db(db, 1, rsc=stand_in, part_ref=stand_in, timeout=600)
    con = sqlite3.connect(db)
    r = con.execute("SELECT cuda, cuda_50, cuda_25, status FROM results").fetchone()
    assert r[3] == 1
    g = gen.rmat(12, 1 << 16, clone_frac=0.3, seed=1011)
    c, K, R, _ = oracle.rbsc_refine(g.n_nodes, g.sources(), g.dst)
    assert json.loads(r[0])["nodes"] == K and json.loads(r[0])["done"] == 1
    for col, p in ((1, 0.50), (2, 0.25)):
        nodes, done, _ = oracle.rbsc_multibatch(g.n_nodes, g.sources(), g.dst, None,
                                                int(g.n_edges * p))
        cell = json.loads(r[col])
        assert (cell["nodes"], cell["done"]) == (nodes, done)


@pytest.mark.skipif(not oracle.have_ref(), reason="reference tools not built")
def test_law_row_through_the_bvgraph_decoder(tmp_path):
    """A 'law' row as the reference lays it out (NAME/NAME.graph), read by
    webgraph-to: Valmari and rsc give the recorded cnr-2000 block count."""
    src = os.path.join(os.path.dirname(__file__), "golden", "bvgraph")
    (tmp_path / "cnr-2000").mkdir()
    for ext in ("graph", "properties", "offsets"):
        os.symlink(os.path.join(src, f"cnr-2000.{ext}"), tmp_path / "cnr-2000" / f"cnr-2000.{ext}")
    db = str(tmp_path / "b.db")
    con = sqlite3.connect(db)
    harness.ensure_table(con, True)
    con.execute("INSERT INTO results VALUES ('cnr-2000', 'law', 1, 1, '', '', '', 0)")
    con.commit()
    rbsc = _tool(tmp_path, "rbsc", "cat > /dev/null; printf '0\\n1\\n0\\n'")
    assert harness.run_db(db, 1, large=True, rbsc=rbsc, rsc=oracle.REF_RSC,
                          part_ref=oracle.REF_VALMARI, timeout=600, law_dir=str(tmp_path)) == 1
    r = con.execute("SELECT nodes, edges, valmari, cpu, status FROM results").fetchone()
    assert r[:2] == (325557, 3216152) and r[4] == 1
    assert json.loads(r[2])["nodes"] == 85418 and json.loads(r[3])["nodes"] == 85418
