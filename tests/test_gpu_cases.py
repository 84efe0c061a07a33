"""Both GPU paths against the oracle on the seeded cases that pin the oracle to the reference
binaries (test_oracle_vs_reference_binary.py): -i 0 .. 20 (every band class of the filter and the
DP), -M 1-4, -u, -A, -F / -R, custom scores,
explicit seeds and position lists, N runs, lowercase, multi-FASTA, repeat filter firing, in
nucleotide and color space; and the ordered mode (no -M: contextual genome map, greedy -G, -u).
Integer and text output: the SAM must be identical."""
import os
import subprocess

import pytest

from conftest import ROOT, sam_body
from test_oracle_vs_reference_binary import CASES, ORDERED_CASES, SEEDS_C, SEEDS_N, make_inputs

pytestmark = pytest.mark.gpu

EXTRA = [
    ("n", "plain", 600, 40, 0, ["-M", "2"], 61),
    ("n", "plain", 600, 40, 1, ["-M", "3", "-t", "90"], 62),
    ("c", "plain", 600, 45, 0, ["-M", "2", "-z", "70"], 63),
    ("c", "plain", 600, 60, 5, ["-M", "3", "-z", "80"], 64),
]

FLAGS = {"-M": ("max_hits", int), "-z": ("z", int), "-t": ("t", int), "-l": ("stdev_distance", int),
         "-u": ("uniqueness", int), "-m": ("match", int), "-x": ("mismatch", int), "-d": ("gap_open", int),
         "-e": ("gap_extend", int), "-s": ("seeds", str)}


def params_of(mode, i, extra):
    kw = dict(color=1 if mode == "c" else 0, allowed_indels=i, seeds=SEEDS_C if mode == "c" else SEEDS_N, max_hits=0)
    k = 0
    while k < len(extra):
        f = extra[k]
        if f == "-A":
            kw["print_all"] = 1
            k += 1
        elif f == "-F":
            kw["sense_mask"] = 1
            k += 1
        elif f == "-G":
            kw["greedy"] = 1
            k += 1
        elif f == "-R":
            kw["sense_mask"] = 2
            k += 1
        else:
            name, conv = FLAGS[f]
            kw[name] = conv(extra[k + 1])
            k += 2
    return kw


@pytest.mark.parametrize("mode,kind,nreads,L,i,extra,seed", CASES + EXTRA + ORDERED_CASES,
                         ids=[f"{c[0]}-{c[1]}-i{c[4]}-{c[6]}" for c in CASES + EXTRA + ORDERED_CASES])
def test_gpu_paths_equal_oracle(mode, kind, nreads, L, i, extra, seed, tmp_path, oracle_built):
    from storm_amd import lib
    tmp = str(tmp_path)
    gpath, rargs = make_inputs(tmp, mode, kind, nreads, L, seed)
    args = ["-i", str(i)] + extra
    if "-s" not in args:
        args = args + ["-s", SEEDS_N if mode == "n" else SEEDS_C]
    orc = os.path.join(ROOT, "oracle", "_build", "oracle-nucleotide" if mode == "n" else "oracle-color")
    subprocess.run([orc, "-g", gpath] + rargs + args + ["-o", os.path.join(tmp, "orc.sam")], check=True,
                   stderr=subprocess.DEVNULL)
    want = sam_body(os.path.join(tmp, "orc.sam"))
    assert len(want) > 5
    reads = rargs[rargs.index("-r") + 1]
    qual = rargs[rargs.index("-q") + 1] if "-q" in rargs else None
    for m in (lib.STORM_MODE_PLAIN, lib.STORM_MODE_BATCHED):
        out = os.path.join(tmp, f"gpu{m}.sam")
        lib.map_files(gpath, reads, out, qual=qual, mode=m, **params_of(mode, i, extra))
        assert sam_body(out) == want, ("plain", "batched")[m]
