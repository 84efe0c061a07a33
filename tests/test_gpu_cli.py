"""The command line on the golden cases, with the reference's own arguments: the SAM written by
storm-gpu-nucleotide / storm-gpu-color (batched, and plain with -P) equals the reference's, for the
-M cases and for the ordered cases (no -M; -G)."""
import os
import subprocess

import pytest

from conftest import ROOT, sam_body, sam_cases, unpack_case

pytestmark = pytest.mark.gpu

BIN = os.path.join(ROOT, "storm_amd", "_build")


@pytest.mark.parametrize("case", sam_cases())
def test_cli_sam_matches_reference(case, tmp_path):
    meta, paths = unpack_case(case, str(tmp_path))
    exe = os.path.join(BIN, "storm-gpu-nucleotide" if meta["mode"] == "n" else "storm-gpu-color")
    if meta["mode"] == "n":
        rargs = ["-r", paths["reads.fq"]]
    else:
        rargs = ["-r", paths["reads.csfasta"], "-q", paths["reads.qual"]]
    want = sam_body(paths["ref.sam"])
    for extra in ([], ["-P"]):
        out = str(tmp_path / "cli.sam")
        r = subprocess.run([exe, "-g", paths["genome.fa"]] + rargs + meta["args"] + extra + ["-o", out],
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        assert sam_body(out) == want, (case, extra)
