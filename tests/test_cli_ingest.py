"""The rbsc command's stdin reader (the reference's "N E" + E lines
"src weight dst", rbsc.cu:156-180) without a device: `rbsc --check-input`
parses with its threaded two-pass reader and prints digests of what it
read, which must equal numpy's digests of the same lines -- through a pipe
and from a file (mapped), at sizes that cut the text into many ranges,
with any white space between tokens and bytes after the last line ignored;
sorted sources go over as row offsets.  Malformed text is refused."""
import os
import subprocess

import numpy as np
import pytest

from rbsc_amd._lib import CLI_PATH

pytestmark = pytest.mark.skipif(not os.path.exists(CLI_PATH), reason="rbsc command not built")

A, B, C, D = (0x9e3779b97f4a7c15, 0xbf58476d1ce4e5b9, 0x94d049bb133111eb, 0x2545f4914f6cdd1d)


def _digests(n, src, w, dst):
    e = np.arange(len(dst), dtype=np.uint64)
    with np.errstate(over="ignore"):
        line = (src.astype(np.uint64) * np.uint64(A) + w.astype(np.uint64) * np.uint64(B) +
                dst.astype(np.uint64) * np.uint64(C) + e * np.uint64(D)).sum(dtype=np.uint64)
        ro_d = 0
        sorted_ = len(src) == 0 or bool(np.all(np.diff(src.astype(np.int64)) >= 0))
        if sorted_ and len(src):
            ro = np.concatenate([[0], np.cumsum(np.bincount(src, minlength=n))]).astype(np.uint64)
            ro_d = int((ro * np.arange(1, n + 2, dtype=np.uint64)).sum(dtype=np.uint64))
    return int(line), ("csr" if sorted_ and len(src) else "coo"), ro_d


def _text(n, src, w, dst, sep=" ", eol="\n", tail=""):
    body = "".join(f"{a}{sep}{b}{sep}{c}{eol}" for a, b, c in zip(src.tolist(), w.tolist(),
                                                                   dst.tolist()))
    return f"{n}{eol}{len(dst)}{eol}{body}{tail}".encode()


def _run(data, tmp_path=None, chunk=None):
    if tmp_path is None:
        env = dict(os.environ)
        if chunk:
            env["RBSC_READ_CHUNK"] = str(chunk)
        return subprocess.run([CLI_PATH, "--check-input"], input=data, capture_output=True,
                              timeout=300, env=env)
    f = tmp_path / "in.txt"
    f.write_bytes(data)
    with open(f, "rb") as fh:
        return subprocess.run([CLI_PATH, "--check-input"], stdin=fh, capture_output=True,
                              timeout=300)


def _check(out, n, src, w, dst):
    assert out.returncode == 0, out.stderr.decode()
    f = out.stdout.decode().split()
    line, kind, ro_d = _digests(n, src, w, dst)
    unit = "unit" if np.all(w == 1) else "weighted"
    assert f == [str(n), str(len(dst)), f"{line:016x}", kind, unit, f"{ro_d:016x}"]


@pytest.mark.parametrize("sort", [True, False])
@pytest.mark.parametrize("weights", [True, False])
def test_large_text_pipe_and_file(tmp_path, sort, weights):
    rng = np.random.default_rng(3 + sort + 2 * weights)
    n, m = 50000, 400000  # ~5 MB of text: many ranges
    src = rng.integers(0, n, m).astype(np.uint32)
    if sort:
        src = np.sort(src)
    dst = rng.integers(0, n, m).astype(np.uint32)
    w = rng.integers(1, 1 << 40, m).astype(np.uint64) if weights else np.ones(m, np.uint64)
    data = _text(n, src, w, dst)
    _check(_run(data), n, src, w, dst)
    _check(_run(data, tmp_path), n, src, w, dst)
    # a pipe read in chunks much smaller than the text: tokens cut at chunk ends
    for chunk in (4096, 65537, 1 << 20):
        _check(_run(data, chunk=chunk), n, src, w, dst)


def test_any_white_space_and_trailing_bytes(tmp_path):
    rng = np.random.default_rng(9)
    n, m = 3000, 300000
    src = np.sort(rng.integers(0, n, m)).astype(np.uint32)
    dst = rng.integers(0, n, m).astype(np.uint32)
    w = rng.integers(1, 5, m).astype(np.uint64)
    for sep, eol, tail in ((" \t ", "\r\n", ""), ("\t", "\n\n", "junk after the last line\n"),
                           ("  ", " ", "7 7 7 7")):
        data = _text(n, src, w, dst, sep, eol, tail)
        _check(_run(data), n, src, w, dst)
        _check(_run(data, tmp_path), n, src, w, dst)
        _check(_run(data, chunk=5000), n, src, w, dst)


def test_small_and_empty():
    _check(_run(b"3\n4\n0 1 1\n0 2 2\n1 1 0\n2 1 2\n"), 3, np.array([0, 0, 1, 2], np.uint32),
           np.array([1, 2, 1, 1], np.uint64), np.array([1, 2, 0, 2], np.uint32))
    out = _run(b"5\n0\n")
    assert out.returncode == 0 and out.stdout.decode().split()[:2] == ["5", "0"]


def test_malformed_text_is_refused():
    rng = np.random.default_rng(11)
    n, m = 1000, 300000
    src = np.sort(rng.integers(0, n, m)).astype(np.uint32)
    dst = rng.integers(0, n, m).astype(np.uint32)
    w = np.ones(m, np.uint64)
    good = _text(n, src, w, dst)
    cases = []
    for at in (len(good) // 7, len(good) // 2, len(good) - 10):  # a stray byte inside the lines
        i = good.index(b" ", at)
        cases.append(good[:i] + b"x" + good[i + 1:])
    cases.append(good[:len(good) // 3])                                  # truncated
    cases.append(good.replace(b"\n1 1 ", b"\n-1 1 ", 1))                 # a sign
    cases.append(f"{n}\n{m}\n".encode() + good.split(b"\n", 2)[2].replace(
        f"\n{src[5]} 1 ".encode(), f"\n{n} 1 ".encode(), 1))            # id out of range
    cases.append(b"3\n1\n0 99999999999999999999999 1\n")                 # overflow
    cases.append(b"4294967295\n0\n")                                     # N too large
    cases.append(b"x\n")
    for data in cases:
        for chunk in (None, 4096):
            out = _run(data, chunk=chunk)
            assert out.returncode != 0 and out.stdout == b"", (data[:40], chunk)
