"""The drop-in executables on the GPU: same stdout lines as the reference
programs (gppeel.cu:231-241, util.cpp:6-21, main.cu:84-96), core numbers and
iteration counts from the oracle, and the "i core" result file
(util.cpp:70-89)."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle
import pico_amd as pa
from tests import graphs

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "pico_amd", "bin")
EXAMPLE = os.path.join(ROOT, "tests", "golden", "example.txt")


def _run(name, path, env=None):
    e = dict(os.environ)
    e.update(env or {})
    r = subprocess.run([os.path.join(BIN, name), path], capture_output=True, text=True,
                       timeout=300, env=e)
    assert r.returncode == 0, r.stderr
    return r.stdout


def _field(out, pat):
    m = re.search(pat, out)
    assert m, (pat, out)
    return int(m.group(1))


@pytest.mark.parametrize("name", ["gpp", "peelone", "HistoCore", "CntCore", "NbrCore"])
def test_vgraph_programs_on_example(name, tmp_path):
    save = tmp_path / "cores.txt"
    out = _run(name, EXAMPLE, {"KC_SAVE_RESULTS": str(save)})
    assert "Reading the input graph from the following file:" in out
    assert _field(out, r"Number of nodes = (\d+)") == 5
    assert _field(out, r"Number of edges = (\d+)") == 7
    assert re.search(r"Processing finished in [0-9.e+-]+ \(ms\)\.", out)
    assert _field(out, r" Max K = (\d+)") == 2
    assert _field(out, r" Min K = (\d+)") == 2
    rp, col = graphs.example()
    it = _field(out, r"Number of iterations = (\d+)")
    if name == "gpp":
        assert it == oracle.gpp(rp, col, compat=True)[1]
    elif name == "HistoCore":
        assert it == oracle.histocore(rp, col)[1]
    elif name == "CntCore":
        assert it == oracle.cntcore(rp, col)[1]
    elif name == "NbrCore":
        assert it == oracle.nbrcore(rp, col)[1]
    lines = save.read_text().split()
    assert [int(x) for x in lines[1::2]] == [2] * 5
    assert [int(x) for x in lines[0::2]] == list(range(5))


@pytest.mark.parametrize("name", ["gpp", "peelone", "HistoCore", "CntCore", "NbrCore"])
def test_vgraph_programs_on_rmat_file(name, tmp_path):
    hg = pa.HostGraph.rmat(11, 8, seed=6)
    f = tmp_path / "g.txt"
    hg.write_edgelist(str(f))
    save = tmp_path / "cores.txt"
    out = _run(name, str(f), {"KC_SAVE_RESULTS": str(save)})
    back = pa.HostGraph.read_coo(str(f))
    want = oracle.bz(back.rowptr, back.col)
    if name == "gpp":
        want = oracle.gpp(back.rowptr, back.col, compat=True)[0]
    got = np.array([int(x) for x in save.read_text().split()[1::2]], np.uint32)
    np.testing.assert_array_equal(got, want)
    assert _field(out, r" Max K = (\d+)") == int(want.max())


def test_peelone_dyn_program(tmp_path):
    save = tmp_path / "cores.txt"
    out = _run("peelone_dyn", EXAMPLE, {"KC_SAVE_RESULTS": str(save)})
    assert "Graph loading Started..." in out
    assert _field(out, r"V: (\d+)") == 5
    assert _field(out, r"E: (\d+)") == 7
    assert _field(out, r"Maxdegree = (\d+)") == 3  # example.txt: degrees 3,3,3,2,3
    assert re.search(r"STD: \n0\.4\n", out)
    assert re.search(r"ICDE23_peel: (\d+ms ){10} ave: [0-9.e+-]+", out), out
    assert re.search(r"Our_Paper_peel: (\d+ms ){10} ave: [0-9.e+-]+", out), out
    assert [int(x) for x in re.findall(r"Kmax: (\d+)", out)] == [2, 2]
    assert [int(x) for x in save.read_text().split()[1::2]] == [2] * 5


@pytest.mark.parametrize("name", ["peelone_dyn", "HistoCore"])
def test_programs_through_csr_cache(name, tmp_path):
    # KC_CSR_CACHE: the first run parses and writes the entry, the second
    # loads it; stdout lines and the saved cores are the same as without it
    hg = pa.HostGraph.rmat(11, 8, seed=7)
    f = tmp_path / "g.txt"
    hg.write_edgelist(str(f))
    cache = tmp_path / "cache"
    cache.mkdir()
    outs, cores = [], []
    for env in ({}, {"KC_CSR_CACHE": str(cache)}, {"KC_CSR_CACHE": str(cache)}):
        save = tmp_path / "cores.txt"
        env = dict(env, KC_SAVE_RESULTS=str(save))
        out = _run(name, str(f), env)
        outs.append(re.sub(r"[0-9.e+-]+ ?\(?ms\)?|\d+ms |ave: [0-9.e+-]+", "", out))
        cores.append(save.read_text())
        if env.get("KC_CSR_CACHE"):
            assert len(os.listdir(cache)) == 1
    assert outs[0] == outs[1] == outs[2]
    assert cores[0] == cores[1] == cores[2]
