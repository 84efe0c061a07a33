"""The command line (storm-gpu-nucleotide / storm-gpu-color, the reference's flags, main.c:188-433):
argument errors are reported before any device call."""
import os
import subprocess

import pytest

from conftest import ROOT

BIN = os.path.join(ROOT, "storm_amd", "_build")


@pytest.mark.parametrize("exe", ["storm-gpu-nucleotide", "storm-gpu-color"])
def test_cli_argument_errors(exe, tmp_path):
    path = os.path.join(BIN, exe)
    if not os.path.exists(path):
        pytest.skip("command line not built")
    # no genome / reads: usage
    r = subprocess.run([path], capture_output=True, text=True)
    assert r.returncode == 1 and "usage" in r.stderr
    # -A needs -M first (main.c:377-382)
    r = subprocess.run([path, "-g", "g.fa", "-r", "r.fq", "-A"], capture_output=True, text=True)
    assert r.returncode == 1 and "-M" in r.stderr
    # -M outside 1..4
    r = subprocess.run([path, "-g", "g.fa", "-r", "r.fq", "-M", "5"], capture_output=True, text=True)
    assert r.returncode == 1 and "usage" in r.stderr
