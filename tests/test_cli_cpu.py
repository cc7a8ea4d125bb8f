"""The drop-in executables' argument contract, checked without a GPU: the
paths that exit before touching the device (argument.cpp:40-43 prints the
help text and exits 0; main.cu:68-74 prints the usage line and exits -1)."""
import os
import subprocess

import pytest

BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "pico_amd", "bin")


def _run(name, *args):
    exe = os.path.join(BIN, name)
    if not os.path.exists(exe):
        pytest.skip("CLI not built (run __graft_entry__.build)")
    return subprocess.run([exe, *args], capture_output=True, text=True, timeout=60)


@pytest.mark.parametrize("name", ["gpp", "peelone", "HistoCore", "CntCore", "NbrCore"])
def test_wrong_argc_prints_help(name):
    for args in ([], ["a", "b"]):
        r = _run(name, *args)
        assert r.returncode == 0
        assert "Required arguments:" in r.stdout and "[--input]" in r.stdout


def test_peelone_dyn_usage():
    r = _run("peelone_dyn")
    assert r.returncode == 255
    assert r.stdout.strip() == "./kcore file deviceID Repeattime"


def test_unknown_program_name():
    exe = os.path.join(BIN, "kcore_cli")
    if not os.path.exists(exe):
        pytest.skip("CLI not built")
    r = subprocess.run([exe, "x"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2
