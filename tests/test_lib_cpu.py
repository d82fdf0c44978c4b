"""CPU-side checks of the library boundary: librbsc.so loads and exports every
function declared in include/*.h, host generators are deterministic and
well-formed, device R-MAT spec matches the host one, and the rbsc command
fails cleanly (nothing on stdout) when there is no device or bad input."""
import os
import re
import subprocess

import numpy as np
import pytest

import rbsc_amd
from rbsc_amd import HostGraph, gen
from rbsc_amd._lib import CLI_PATH, SIGNATURES

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(rbsc_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = rbsc_amd.load()
    declared = _declared_functions()
    assert declared, "no declarations found"
    for name in declared:
        assert hasattr(lib, name), f"{name} declared but not exported"
        assert name in SIGNATURES, f"{name} has no ctypes signature"


def test_version():
    assert rbsc_amd.load().rbsc_version().startswith(b"rbsc-gfx950")


def _check_csr(g: HostGraph):
    ro = g.row_offsets
    assert ro[0] == 0 and ro[-1] == g.n_edges
    assert np.all(np.diff(ro.astype(np.int64)) >= 0)
    assert np.all(g.dst < g.n_nodes)
    src = g.sources()
    key = src.astype(np.uint64) << np.uint64(32) | g.dst.astype(np.uint64)
    assert np.all(np.diff(key.astype(np.int64)) >= 0), "lines not sorted by (src, dst)"


def test_generators_deterministic_and_sorted():
    a = gen.rmat(12, 1 << 15, clone_frac=0.3, seed=3)
    b = gen.rmat(12, 1 << 15, clone_frac=0.3, seed=3)
    c = gen.rmat(12, 1 << 15, clone_frac=0.3, seed=4)
    assert np.array_equal(a.dst, b.dst) and np.array_equal(a.row_offsets, b.row_offsets)
    assert not (len(a.dst) == len(c.dst) and np.array_equal(a.dst, c.dst))
    for g in [a, gen.ba(3000, 3, 0.1, 3, seed=1), gen.er_undirected(500, 4, seed=1),
              gen.ba_undirected(500, 2, seed=1), gen.regular_undirected(400, 3, seed=1)]:
        _check_csr(g)


def test_rmat_clones_copy_their_source():
    g = gen.rmat(11, 1 << 14, clone_frac=0.3, seed=8)
    ro = g.row_offsets.astype(np.int64)
    rows = {}
    for u in range(g.n_nodes):
        rows.setdefault(tuple(g.dst[ro[u]:ro[u + 1]].tolist()), []).append(u)
    # about 30% of nodes share their line list with another node
    shared = sum(len(v) for k, v in rows.items() if len(v) > 1 and len(k) > 0)
    assert shared > 0.2 * g.n_nodes


def test_ba_structure():
    g = gen.ba(2000, 4, 0.0, 3, seed=2)
    deg = np.diff(g.row_offsets.astype(np.int64))
    assert np.all(deg[5:] == 4) and np.all(deg[:5] == 4)
    assert g.init_color is not None and set(np.unique(g.init_color)) <= {0, 1, 2}
    g2 = gen.ba(2000, 4, 0.5, 0, seed=2)
    assert g2.n_edges > g.n_edges


def test_regular_is_regular():
    g = gen.regular_undirected(1000, 4, seed=3)
    assert np.all(np.diff(g.row_offsets.astype(np.int64)) == 4)


def test_text_formats():
    g = HostGraph(3, np.array([1, 2, 0], np.uint32), src=np.array([0, 0, 2], np.uint32),
                  weight=np.array([1, 5, 2], np.uint64), init_color=np.array([0, 1, 0], np.uint32))
    assert g.to_text() == b"3\n3\n0 1 1\n0 5 2\n2 2 0\n"
    assert g.to_valmari_text() == b"3 0 0 3 2\n1 1 2\n1 5 3\n3 2 1\n2 0\n"


def _has_gpu():
    try:
        rbsc_amd.Context(0).close()
        return True
    except rbsc_amd.RbscError:
        return False


@pytest.mark.skipif(_has_gpu(), reason="checks the no-device path")
def test_no_device_is_a_clean_error():
    with pytest.raises(rbsc_amd.RbscError) as e:
        rbsc_amd.Context(0)
    assert e.value.code == -7
    p = subprocess.run([CLI_PATH], input=b"2\n1\n0 1 1\n", stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, timeout=60)
    assert p.returncode != 0 and p.stdout == b""


@pytest.mark.parametrize("text", [b"", b"3\n", b"3\n2\n0 1 1\n", b"2\n1\n0 1 5\n", b"2\n1\nx y z\n"])
def test_cli_rejects_bad_input(text):
    p = subprocess.run([CLI_PATH], input=text, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                       timeout=60)
    assert p.returncode != 0
    assert p.stdout == b""


def test_batch_generator_shapes():
    from rbsc_amd import gen
    b = gen.batch_mixed(200, 10, 5000, seed=5)
    b2 = gen.batch_mixed(200, 10, 5000, seed=5)
    assert np.array_equal(b.dst, b2.dst) and np.array_equal(b.node_offsets, b2.node_offsets)
    n = np.diff(b.node_offsets.astype(np.int64))
    assert b.n_graphs == 200 and n.min() >= 10 and n.max() <= 5000
    for g in range(0, 200, 13):
        hg = b.graph(g)
        s = hg.sources()
        assert np.all(np.diff(s.astype(np.int64)) >= 0)
        assert hg.dst.max() < hg.n_nodes
        fwd = set(zip(s.tolist(), hg.dst.tolist()))
        assert fwd == set(zip(hg.dst.tolist(), s.tolist()))  # undirected: both directions


def test_host_batch_from_graphs_roundtrip():
    from rbsc_amd import HostBatch, gen
    gs = [gen.er_undirected(30 + i, 3.0, seed=i) for i in range(5)]
    b = HostBatch.from_graphs(gs)
    for i, g in enumerate(gs):
        h = b.graph(i)
        assert h.n_nodes == g.n_nodes
        assert np.array_equal(h.dst, g.dst) and np.array_equal(h.sources(), g.sources())
    gs[1].init_color = np.zeros(gs[1].n_nodes, np.uint32)
    with pytest.raises(ValueError):
        HostBatch.from_graphs(gs)


def test_write_text_matches_python_formats(tmp_path):
    from rbsc_amd import gen
    g = gen.ba(300, 3, 0.2, 3, seed=4)
    g.weight = np.random.default_rng(2).integers(1, 9, g.n_edges).astype(np.uint64)
    for fmt, want in (("lines", g.to_text()), ("valmari", g.to_valmari_text())):
        p = tmp_path / f"g.{fmt}"
        g.write_text(str(p), fmt)
        assert p.read_bytes() == want
    h = gen.rmat(8, 2000, clone_frac=0.3, seed=3)  # unweighted, no colours
    p = tmp_path / "h.lines"
    h.write_text(str(p))
    assert p.read_bytes() == h.to_text()


def test_pinned_labels_outlive_their_batch(monkeypatch):
    """Labels handed out over page-locked memory stay readable after the
    batch (the memory's other holder) is freed: the memory is released with
    the last array over it, not before.  The allocator is replaced by libc
    malloc/free so this runs without a device."""
    import ctypes as C
    libc = C.CDLL(None)
    libc.malloc.restype = C.c_void_p
    libc.malloc.argtypes = [C.c_size_t]
    libc.free.argtypes = [C.c_void_p]
    freed = []

    class FakeLib:
        def rbsc_host_alloc(self, nbytes, ref):
            ref._obj.value = libc.malloc(int(nbytes))
            return 0

        def rbsc_host_free(self, p):
            freed.append(p.value)
            libc.free(p)
            return 0

    fake = FakeLib()
    monkeypatch.setattr(rbsc_amd, "load", lambda: fake)
    pa = rbsc_amd.PinnedArray(1000)
    pa.array[:] = np.arange(1000, dtype=np.uint32)
    view = pa.array[10:20]
    pa.free()                   # what DeviceBatch.free() does
    del pa
    assert freed == []          # the view still holds the memory
    assert np.array_equal(view, np.arange(10, 20, dtype=np.uint32))
    del view
    import gc
    gc.collect()
    assert len(freed) == 1      # released with the last array


def test_server_labels_free_checks_its_argument():
    """rbsc_server_labels_free takes back only buffers rbsc_server_wait_labels
    handed out (a null pointer is a no-op); no device is touched."""
    import ctypes as C
    L = rbsc_amd.load()
    assert L.rbsc_server_labels_free(None) == 0
    assert L.rbsc_server_labels_free(C.c_void_p(0x1000)) == -1  # RBSC_ERR_INVALID
