"""The command line on the golden cases, with the reference's own arguments: the SAM written by
storm-gpu-nucleotide / storm-gpu-color (batched, and plain with -P) equals the reference's, for the
-M cases and for the ordered cases (no -M; -G); and with the seeds read from a file (-S, one seed
per line, '%' comment lines, seed.c:140-179) instead of -s."""
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


@pytest.mark.parametrize("case", [c for c in sam_cases() if "ordered" in c or c.endswith("M1")])
def test_cli_seed_file(case, tmp_path):
    meta, paths = unpack_case(case, str(tmp_path))
    exe = os.path.join(BIN, "storm-gpu-nucleotide" if meta["mode"] == "n" else "storm-gpu-color")
    if meta["mode"] == "n":
        rargs = ["-r", paths["reads.fq"]]
    else:
        rargs = ["-r", paths["reads.csfasta"], "-q", paths["reads.qual"]]
    args = list(meta["args"])
    k = args.index("-s")
    seeds = args[k + 1].split(";")
    sf = str(tmp_path / "seeds.txt")
    with open(sf, "w") as f:
        f.write("% seeds of the run\n")
        for sd in seeds:
            f.write(sd + "  \n")
    args[k:k + 2] = ["-S", sf]
    out = str(tmp_path / "cli.sam")
    r = subprocess.run([exe, "-g", paths["genome.fa"]] + rargs + args + ["-o", out], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    assert sam_body(out) == sam_body(paths["ref.sam"])
