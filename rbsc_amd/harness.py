"""Counterpart of the reference's benchmark harness (bench.py, benchlarge.py).

Same sqlite tables, same row life cycle and the same cells, so a database
filled here reads like one filled by the reference:

* ``bench.db``: ``results(name, tool, nodes, edges, valmari, cpu, cuda,
  cuda_50, cuda_25, mc_32, mc_64, status)`` (reference bench.py:55-61, 105-117);
* ``benchlarge.db``: ``results(name, tool, nodes, edges, valmari, cpu, cuda,
  status)`` (reference benchlarge.py:55-60, 82-92).

Rows with ``status = 0`` are run in order of ``edges``; a finished row gets
``status = 1``, a row whose input fails gets ``-1`` (bench.py:118-122), so a
rerun resumes where the last one stopped.  Each tool cell is the JSON of
``{"done", "nodes", "time"}`` (bench.py:15-19); a tool whose output does not
parse (time-out, crash) records ``time = TIMEOUT, done = 0, nodes = 0``
(bench.py:41-54).  Tools run as the same shell pipelines (bench.py:7-12):
``{input} | timeout T ./rbsc {edges} {gpu}``, ``./rsc``, ``./part_ref``;
``rbsc`` is the library's command (rbsc_amd/bin/rbsc), ``rsc`` and
``part_ref`` the reference's CPU tools, which the caller provides
(``--rsc``, ``--part-ref``; by default ``./rsc`` and ``./part_ref`` in the
working directory, as the reference's harness runs them).  The ``cuda``
columns run at edge batches of 100 %, 50 % and 25 % (bench.py:83-89).

Inputs: ``tool = 'file'`` reads ``datasets/{name}.txt`` as the reference
does (bench.py:9, 36-37); ``tool = 'law'`` reads the BVGraph
``{name}/{name}.graph`` (where the reference's download.sh puts it) through
the library's decoder, ``webgraph-to arcs-cuda|arcs-valmari`` in place of
``./webgraph to ...`` (bench.py:8); nothing is downloaded, so a missing
graph fails its row; ``tool = 'gen'`` names a seeded workload of
:data:`WORKLOADS`, written as text by the library.  The sigref columns stay empty results, as in the
reference, which removed that experiment (bench.py:92-103).

Usage::

    python -m rbsc_amd.harness init bench.db [--large]   # table + generated rows
    python -m rbsc_amd.harness run bench.db GPUS          # like: python bench.py GPUS
    python -m rbsc_amd.harness run benchlarge.db GPUS --large --batch EDGES
"""
from __future__ import annotations

import argparse
import json
import os
import shlex
import sqlite3
import subprocess
import sys
import tempfile
from dataclasses import dataclass
from typing import Callable, Dict, Optional

TIMEOUT = 3600          # bench.py:7
TIMEOUT_LARGE = 10800   # benchlarge.py:7

BENCH_COLUMNS = ("name TEXT, tool TEXT, nodes INT, edges INT, valmari TEXT, cpu TEXT, cuda TEXT, "
                 "cuda_50 TEXT, cuda_25 TEXT, mc_32 TEXT, mc_64 TEXT, status INT")
LARGE_COLUMNS = "name TEXT, tool TEXT, nodes INT, edges INT, valmari TEXT, cpu TEXT, cuda TEXT, status INT"

HERE = os.path.dirname(os.path.abspath(__file__))
RBSC = os.path.join(HERE, "bin", "rbsc")
# The CPU tools are the caller's: like the reference's harness, which runs
# ./rsc and ./part_ref from its working directory (bench.py:11-12), unless
# the caller names their paths.
RSC = "./rsc"
PART_REF = "./part_ref"


def _workloads() -> Dict[str, Callable]:
    from . import gen
    return {
        "rmat_s20_d24_c30": gen.rmat_2p20_clone30,
        "ba_8m_m6_init3": gen.ba_8m_m6_init3,
        "rmat_s16_d20_c20_seed1010": lambda: gen.rmat(16, 1 << 20, clone_frac=0.2, seed=1010),
        "rmat_s12_d16_c30_seed1011": lambda: gen.rmat(12, 1 << 16, clone_frac=0.3, seed=1011),
        "er_n100000_d6_seed1012": lambda: gen.er_undirected(100000, 6.0, seed=1012),
        "reg3_n20000_seed1013": lambda: gen.regular_undirected(20000, 3, seed=1013),
    }


WEBGRAPH_TO = os.path.join(HERE, "bin", "webgraph-to")

WORKLOADS = ("rmat_s20_d24_c30", "ba_8m_m6_init3", "rmat_s16_d20_c20_seed1010",
             "rmat_s12_d16_c30_seed1011", "er_n100000_d6_seed1012", "reg3_n20000_seed1013")


@dataclass
class Result:
    """One tool run, stored as the reference stores it (bench.py:15-19)."""
    done: int = 0
    nodes: int = 0
    time: float = 0.0

    def cell(self) -> str:
        return json.dumps({"done": self.done, "nodes": self.nodes, "time": self.time})


def experiment(command: str, timeout: float, runs: int = 1) -> Result:
    """Runs a shell pipeline and parses its first three stdout lines
    (time, done, nodes); anything else records a time-out (bench.py:41-54)."""
    res = Result()
    print(command, flush=True)
    for _ in range(runs):
        out = subprocess.run([command], stdout=subprocess.PIPE, shell=True).stdout.decode()
        try:
            lines = out.split("\n")[:-1]
            res.time += float(lines[0])
            res.done = int(lines[1])
            res.nodes = int(lines[2])
        except (ValueError, IndexError):
            return Result(0, 0, float(timeout))
    res.time /= runs
    return res


class Inputs:
    """Text inputs of one row: the reference's line format for rbsc and rsc,
    the MDP format for part_ref (what valmari-adapter writes)."""

    def __init__(self, name: str, tool: str, workdir: str, datasets: str, law_dir: str = "."):
        self.name, self.tool = name, tool
        self.dir = workdir
        self.init = None
        self.cmd_lines = self.cmd_mdp = None
        if tool == "law":
            base = os.path.join(law_dir, name, name)
            if not os.path.exists(base + ".graph"):
                raise FileNotFoundError(f"{base}.graph (LAW graphs are not downloaded here)")
            q = shlex.quote
            out = subprocess.run([WEBGRAPH_TO, "gdata", base], stdout=subprocess.PIPE, check=True)
            self.nodes, self.edges = (int(v) for v in out.stdout.split())
            self.cmd_lines = f"{q(WEBGRAPH_TO)} arcs-cuda {q(base)}"
            self.cmd_mdp = f"{q(WEBGRAPH_TO)} arcs-valmari {q(base)}"
        elif tool == "file":
            self.lines = os.path.join(datasets, f"{name}.txt")
            if not os.path.exists(self.lines):
                raise FileNotFoundError(self.lines)
            with open(self.lines) as f:
                self.nodes = int(f.readline())
                self.edges = int(f.readline())
            self.mdp = os.path.join(workdir, f"{name}.mdp")
            _lines_to_mdp(self.lines, self.mdp)
        elif tool == "gen":
            wl = _workloads()
            if name not in wl:
                raise KeyError(f"unknown generated workload {name!r}")
            g = wl[name]()
            self.nodes, self.edges = g.n_nodes, g.n_edges
            self.lines = os.path.join(workdir, f"{name}.txt")
            self.mdp = os.path.join(workdir, f"{name}.mdp")
            g.write_text(self.lines, "lines")
            g.write_text(self.mdp, "valmari")
            if g.init_color is not None:
                self.init = os.path.join(workdir, f"{name}.init")
                with open(self.init, "w") as f:
                    f.write("\n".join(map(str, g.init_color.tolist())) + "\n")
        else:
            raise RuntimeError(f"row {name!r}: unknown input tool {tool!r}")
        if self.cmd_lines is None:
            self.cmd_lines = f"cat {shlex.quote(self.lines)}"
            self.cmd_mdp = f"cat {shlex.quote(self.mdp)}"


def _lines_to_mdp(src: str, dst: str) -> None:
    """valmari-adapter:9-16: 1-based ids, header "N 0 0 E 1"."""
    with open(src) as f, open(dst, "w") as out:
        n = int(f.readline())
        e = int(f.readline())
        out.write(f"{n} 0 0 {e} 1\n")
        for line in f:
            a, w, b = line.split()
            out.write(f"{int(a) + 1} {w} {int(b) + 1}\n")


def ensure_table(con: sqlite3.Connection, large: bool) -> None:
    con.execute(f"CREATE TABLE IF NOT EXISTS results({LARGE_COLUMNS if large else BENCH_COLUMNS})")
    con.commit()


def init_db(path: str, large: bool, names=WORKLOADS) -> None:
    """Creates the reference's table and queues the generated workloads."""
    con = sqlite3.connect(path)
    ensure_table(con, large)
    have = {r[0] for r in con.execute("SELECT name FROM results")}
    wl = _workloads()
    for name in names:
        if name in have:
            continue
        g = wl[name]()
        empty = ("",) * (3 if large else 7)
        con.execute(f"INSERT INTO results VALUES (?, ?, ?, ?, {', '.join('?' * len(empty))}, 0)",
                    (name, "gen", g.n_nodes, g.n_edges) + empty)
    con.commit()
    con.close()


def run_db(path: str, gpus: int, large: bool = False, batch: Optional[int] = None,
           datasets: str = "datasets", rbsc: str = RBSC, rsc: str = RSC, part_ref: str = PART_REF,
           timeout: Optional[int] = None, law_dir: str = ".") -> int:
    """Runs every row with status 0 (in order of edges); returns rows done."""
    T = timeout if timeout is not None else (TIMEOUT_LARGE if large else TIMEOUT)
    con = sqlite3.connect(path)
    con.row_factory = sqlite3.Row
    cur = con.cursor()
    cur.execute("SELECT * FROM results WHERE status = 0 ORDER BY edges")
    done = 0
    for rc in cur.fetchall():
        name = rc["name"]
        try:
            with tempfile.TemporaryDirectory() as wd:
                inp = Inputs(name, rc["tool"], wd, datasets, law_dir)
                extra = f" --init-colors {shlex.quote(inp.init)}" if inp.init else ""
                q = shlex.quote
                valmari = experiment(f"{inp.cmd_mdp} | timeout {T} {q(part_ref)}", T)
                # rsc has no initial-partition input: it runs on the uncoloured
                # graph, as the reference would
                cpu = experiment(f"{inp.cmd_lines} | timeout {T} {q(rsc)}", T)
                if large:
                    edges = batch if batch else inp.edges
                    cuda = {1: experiment(f"{inp.cmd_lines} | timeout {T} {q(rbsc)} {edges} "
                                          f"{gpus}{extra}", T)}
                    con.execute("UPDATE results SET nodes=?,edges=?,valmari=?,cpu=?,cuda=?,status=? "
                                "WHERE name=?", (inp.nodes, inp.edges, valmari.cell(), cpu.cell(),
                                                 cuda[1].cell(), 1, name))
                else:
                    cuda = {p: experiment(f"{inp.cmd_lines} | timeout {T} {q(rbsc)} "
                                          f"{int(inp.edges * p)} {gpus}{extra}", T)
                            for p in (1, 0.50, 0.25)}
                    mc = Result()
                    con.execute("UPDATE results SET nodes=?,edges=?,valmari=?,cpu=?,cuda=?,cuda_50=?,"
                                "cuda_25=?,mc_32=?,mc_64=?,status=? WHERE name=?",
                                (inp.nodes, inp.edges, valmari.cell(), cpu.cell(), cuda[1].cell(),
                                 cuda[0.50].cell(), cuda[0.25].cell(), mc.cell(), mc.cell(), 1,
                                 name))
                con.commit()
                done += 1
        except Exception as e:  # the reference marks the row failed and goes on
            print(f"FAIL: {e!r}", flush=True)
            con.execute("UPDATE results SET status=? WHERE name=?", (-1, name))
            con.commit()
    con.close()
    return done


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m rbsc_amd.harness")
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("init", help="create the results table and queue generated workloads")
    a.add_argument("db")
    a.add_argument("--large", action="store_true")
    a.add_argument("--only", nargs="*", default=None, help="workload names (default: all)")
    b = sub.add_parser("run", help="run every row with status 0")
    b.add_argument("db")
    b.add_argument("gpus", type=int)
    b.add_argument("--large", action="store_true")
    b.add_argument("--batch", type=int, default=None, help="edge batch of the large table")
    b.add_argument("--datasets", default="datasets")
    b.add_argument("--timeout", type=int, default=None)
    b.add_argument("--law-dir", default=".", help="directory holding NAME/NAME.graph")
    b.add_argument("--rsc", default=RSC, help="the RSC CPU tool (default ./rsc)")
    b.add_argument("--part-ref", default=PART_REF, help="Valmari's tool (default ./part_ref)")
    b.add_argument("--rbsc", default=RBSC, help="the rbsc command (default: the library's)")
    args = ap.parse_args(argv)
    if args.cmd == "init":
        init_db(args.db, args.large, tuple(args.only) if args.only else WORKLOADS)
        return 0
    run_db(args.db, args.gpus, args.large, args.batch, args.datasets, rbsc=args.rbsc,
           rsc=args.rsc, part_ref=args.part_ref, timeout=args.timeout, law_dir=args.law_dir)
    return 0


if __name__ == "__main__":
    sys.exit(main())
