"""The host C++ of libkcore (both edge-file readers with their chunked
parallel parse, the symmetrize/dedup CSR builder, the generators, binary CSR
files, the edge-list writer) built with g++ -fsanitize=address,undefined
into tests/sanitize/host_selftest and run on the reader quirk fixtures and
on generated files, cut into many parse chunks on several threads.  Any
sanitizer report fails the test; the results must equal libkcore's own.
No GPU, no HIP: these sources are plain C++."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import pico_amd as pa

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = [os.path.join(HERE, "sanitize", "host_selftest.cpp"),
       os.path.join(ROOT, "pico_amd", "host", "kc_host.cpp"),
       os.path.join(ROOT, "pico_amd", "host", "kc_cache.cpp")]
RDIR = os.path.join(HERE, "golden", "readers")
ENV = {"ASAN_OPTIONS": "detect_leaks=1:halt_on_error=1:abort_on_error=0",
       "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1"}


@pytest.fixture(scope="module")
def selftest(tmp_path_factory):
    if not shutil.which("g++"):
        pytest.skip("no g++")
    exe = str(tmp_path_factory.mktemp("san") / "host_selftest")
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
                        "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
                        "-pthread"] + SRC + ["-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return exe


def _run(exe, *args, chunk=None, threads="4"):
    env = dict(os.environ, **ENV, KC_HOST_THREADS=threads)
    if chunk:
        env["KC_PARSE_CHUNK"] = str(chunk)
    r = subprocess.run([exe] + list(args), capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0 and "ERROR: AddressSanitizer" not in r.stderr and \
        "runtime error" not in r.stderr, r.stderr[-4000:]
    return r.stdout.strip()


def _fnv(rowptr, col):
    h = 1469598103934665603
    for x in list(rowptr) + list(col):
        h = ((h ^ int(x)) * 1099511628211) & ((1 << 64) - 1)
    return h


def _lib_line(reader, path):
    try:
        g = pa.HostGraph.read_coo(path) if reader == "coo" else pa.HostGraph.read_dedup(path)
    except pa.KCoreError:
        return None
    return "V=%u nnz=%d hash=%016x" % (g.V, g.nnz, _fnv(g.rowptr, g.col))


FILES = sorted(os.path.join(RDIR, f) for f in os.listdir(RDIR) if f.endswith(".txt")) + \
    [os.path.join(HERE, "golden", "example.txt")]


@pytest.mark.parametrize("reader", ["coo", "dedup"])
def test_readers_under_sanitizers(selftest, reader, tmp_path):
    rng = np.random.default_rng(3)
    gen = str(tmp_path / "gen.txt")
    lines = []
    for _ in range(4000):
        a, b = rng.integers(30, 500, 2)
        k = rng.integers(8)
        lines.append("" if k == 0 else "%d" % a if k == 1 else "%d %d 1" % (a, b))
    with open(gen, "w") as f:
        f.write("%%x\n500 500 4000\n" + "\n".join(lines) + "\n")
    for path in FILES + [gen]:
        want = _lib_line(reader, path)
        for chunk, threads in ((None, "1"), (7, "4"), (64, "3")):
            out = _run(selftest, reader, path, chunk=chunk, threads=threads)
            if want is None:
                assert " error " in out, (path, out)
            else:
                assert out.startswith(reader + " ") and want.split()[0] in out and \
                    want.split()[-1] in out, (path, out, want)


def test_generators_cache_writer_under_sanitizers(selftest, tmp_path):
    out = _run(selftest, "gen", str(tmp_path)).splitlines()
    tags = [ln.split()[0] for ln in out]
    assert tags == ["rmat", "chunglu", "er", "load", "reread", "pairs"], out
    h = {ln.split()[0]: ln.split()[-1] for ln in out}
    # save/load, and the symmetrize/dedup builder over the graph's own
    # entries, give the generator's graph back
    assert h["load"] == h["rmat"] and h["pairs"] == h["rmat"]
    g = pa.HostGraph.rmat(12, 16, seed=7)
    assert h["rmat"] == "hash=%016x" % _fnv(g.rowptr, g.col)
