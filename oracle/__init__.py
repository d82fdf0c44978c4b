"""CPU oracle for the RBSC refinement round -- TEST INFRASTRUCTURE ONLY.

Only tests, the smoke check and the CPU-baseline leg of the bench may import
this package; the product path (rbsc_amd, librbsc.so) never does.

* liboracle.so (rbsc_oracle.c, lump.c): plain-C restatements of rbsc.cu /
  rsc.cpp, an exact hash-free refiner, and an exact lumping refiner by
  splitters (the Valmari tool's partition), each citing the reference
  file:line or the published method.
* _ref/rsc and _ref/part_ref: the reference's own CPU tools compiled from
  src/rsc/rsc.cpp and src/part_ref/MDPmin.cc by oracle/Makefile (only where
  the reference tree is present; never shipped as source).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_RSC = os.path.join(HERE, "_ref", "rsc")
REF_VALMARI = os.path.join(HERE, "_ref", "part_ref")

u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)
_lib = None


def build():
    subprocess.run(["make", "-C", HERE, "all"], check=True, capture_output=True)


def load_lib(path: str) -> C.CDLL:
    """The oracle's functions from the shared library at `path` (the built
    oracle, or a copy a test compiled)."""
    L = C.CDLL(path)
    _declare(L)
    return L


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = load_lib(LIB)
    return _lib


def _declare(L):
    L.oracle_randomize.restype = C.c_uint64
    L.oracle_randomize.argtypes = [C.c_uint64]
    L.oracle_rbsc_refine.restype = C.c_int
    L.oracle_rbsc_refine.argtypes = [C.c_uint32, C.c_uint64, u32p, u64p, u32p, u32p, C.c_uint32,
                                     u32p, u32p, u32p, u32p, C.c_uint32]
    L.oracle_canonicalize.restype = C.c_uint32
    L.oracle_canonicalize.argtypes = [C.c_uint32, u32p, u32p]
    L.oracle_rsc_refine.restype = C.c_int
    L.oracle_rsc_refine.argtypes = [C.c_uint32, C.c_uint64, u32p, u32p, u32p, u32p, u32p, u32p]
    L.oracle_exact_refine.restype = C.c_int
    L.oracle_exact_refine.argtypes = [C.c_uint32, C.c_uint64, u32p, u64p, u32p, u32p, u32p,
                                      u32p, u32p]
    if hasattr(L, "oracle_lump_refine"):  # copies built from rbsc_oracle.c alone lack it
        L.oracle_lump_refine.restype = C.c_int
        L.oracle_lump_refine.argtypes = [C.c_uint32, C.c_uint64, u32p, u64p, u32p, u32p, u32p,
                                         u32p]
    L.oracle_rbsc_multibatch.restype = C.c_int
    L.oracle_rbsc_multibatch.argtypes = [C.c_uint32, C.c_uint64, u32p, u64p, u32p, C.c_uint64,
                                         C.c_int, u32p, u32p, u32p, u32p]


def _p(a, t):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


def _arrs(src, dst, weight, wdtype=np.uint64):
    src = np.ascontiguousarray(src, np.uint32)
    dst = np.ascontiguousarray(dst, np.uint32)
    w = None if weight is None else np.ascontiguousarray(weight, wdtype)
    return src, dst, w


def randomize(v: int) -> int:
    return int(lib().oracle_randomize(C.c_uint64(v & (2**64 - 1))))


def canonical(colors: np.ndarray) -> np.ndarray:
    """First-occurrence relabel (rbsc.cu:431-445)."""
    colors = np.ascontiguousarray(colors, np.uint32)
    out = np.empty_like(colors)
    lib().oracle_canonicalize(len(colors), _p(colors, C.c_uint32), _p(out, C.c_uint32))
    return out


def rbsc_refine(n, src, dst, weight=None, init=None, max_rounds=0, trace_rounds=0, _from=None):
    """Restatement of rbsc.cu single-batch refinement.  Returns
    (colors, K, rounds, trace) where trace[r] = colours after round r+1.
    _from: another build of the oracle (load_lib), for tests of the checks."""
    src, dst, w = _arrs(src, dst, weight)
    init = None if init is None else np.ascontiguousarray(init, np.uint32)
    colors = np.empty(max(n, 1), np.uint32)
    K = C.c_uint32()
    R = C.c_uint32()
    trace = np.zeros((trace_rounds, max(n, 1)), np.uint32) if trace_rounds else None
    L = _from or lib()
    rc = L.oracle_rbsc_refine(n, len(dst), _p(src, C.c_uint32), _p(w, C.c_uint64),
                              _p(dst, C.c_uint32), _p(init, C.c_uint32), max_rounds,
                              _p(colors, C.c_uint32), C.byref(K), C.byref(R),
                              _p(trace, C.c_uint32), trace_rounds)
    if rc:
        raise MemoryError("oracle_rbsc_refine failed")
    return colors[:n], int(K.value), int(R.value), trace


def rsc_refine(n, src, dst, weight=None):
    """Restatement of rsc.cpp.  Returns (colors, K, rounds)."""
    src, dst, w = _arrs(src, dst, weight, np.uint32)
    colors = np.empty(max(n, 1), np.uint32)
    K = C.c_uint32()
    R = C.c_uint32()
    rc = lib().oracle_rsc_refine(n, len(dst), _p(src, C.c_uint32), _p(w, C.c_uint32),
                                 _p(dst, C.c_uint32), _p(colors, C.c_uint32), C.byref(K), C.byref(R))
    if rc:
        raise MemoryError("oracle_rsc_refine failed")
    return colors[:n], int(K.value), int(R.value)


def exact_refine(n, src, dst, weight=None, init=None):
    """Hash-free coarsest stable partition.  Returns (colors, K, rounds)."""
    src, dst, w = _arrs(src, dst, weight)
    init = None if init is None else np.ascontiguousarray(init, np.uint32)
    colors = np.empty(max(n, 1), np.uint32)
    K = C.c_uint32()
    R = C.c_uint32()
    rc = lib().oracle_exact_refine(n, len(dst), _p(src, C.c_uint32), _p(w, C.c_uint64),
                                   _p(dst, C.c_uint32), _p(init, C.c_uint32),
                                   _p(colors, C.c_uint32), C.byref(K), C.byref(R))
    if rc:
        raise MemoryError("oracle_exact_refine failed")
    return colors[:n], int(K.value), int(R.value)


def lump_refine(n, src, dst, weight=None, init=None, _from=None):
    """Exact coarsest lumping by splitter refinement (lump.c): the partition
    of the Valmari tool's definition, computed in O(m log n) line scans.
    Returns (labels in first-occurrence order, K)."""
    src, dst, w = _arrs(src, dst, weight)
    init = None if init is None else np.ascontiguousarray(init, np.uint32)
    labels = np.empty(max(n, 1), np.uint32)
    K = C.c_uint32()
    rc = (_from or lib()).oracle_lump_refine(n, len(dst), _p(src, C.c_uint32), _p(w, C.c_uint64),
                                  _p(dst, C.c_uint32), _p(init, C.c_uint32),
                                  _p(labels, C.c_uint32), C.byref(K))
    if rc:
        raise MemoryError("oracle_lump_refine failed")
    return labels[:n], int(K.value)


def rbsc_multibatch(n, src, dst, weight, max_batch, aggressive=False, partition=False):
    """Restatement of rbsc.cu's multi-batch quotient passes.
    Returns (nodes, done, passes), and with partition=True also the
    canonical partition of the original nodes into their final classes."""
    src, dst, w = _arrs(src, dst, weight)
    nodes = C.c_uint32()
    done = C.c_uint32()
    passes = C.c_uint32()
    part = np.empty(max(n, 1), np.uint32) if partition else None
    rc = lib().oracle_rbsc_multibatch(n, len(dst), _p(src, C.c_uint32), _p(w, C.c_uint64),
                                      _p(dst, C.c_uint32), max_batch, int(aggressive),
                                      C.byref(nodes), C.byref(done), C.byref(passes),
                                      _p(part, C.c_uint32))
    if rc:
        raise MemoryError("oracle_rbsc_multibatch failed")
    out = (int(nodes.value), int(done.value), int(passes.value))
    return out + (part[:n],) if partition else out


def have_ref() -> bool:
    return os.path.exists(REF_RSC) and os.path.exists(REF_VALMARI)


def _parse3(out: bytes):
    lines = out.decode().split("\n")
    return float(lines[0]), int(lines[1]), int(lines[2])


def run_ref_rsc(text: bytes, timeout=600):
    """Run the compiled reference rsc on rbsc-format text; (time, done, nodes)."""
    p = subprocess.run([REF_RSC], input=text, stdout=subprocess.PIPE, timeout=timeout, check=True)
    return _parse3(p.stdout)


def run_ref_valmari(text: bytes, timeout=600):
    """Run the compiled reference Valmari tool on MDP-format text."""
    p = subprocess.run([REF_VALMARI], input=text, stdout=subprocess.PIPE, timeout=timeout, check=True)
    return _parse3(p.stdout)
