"""Staged passes whose raw index hits exceed the per-pass limit are split (the device scans count
items in 32 bits; a 100 Mbp genome with repeats gives billions of hits per million reads).  With
the limit lowered through STORM_PASS_HIT_LIMIT the plain path splits every pass many times and
the staged fallback of the batched path is split too; both must still equal the CPU restatement's
SAM.  Runs in a child process: the limit is read once per process."""
import os
import subprocess
import sys

import pytest

from conftest import ROOT, sam_body

pytestmark = pytest.mark.gpu
SEEDS_N = "####--#-#--##--####;#####--###-#-###;###-##-#######"

CHILD = r'''
import sys
sys.path.insert(0, sys.argv[1])
from storm_amd import lib
for mode in (0, 1):
    st = lib.map_files(sys.argv[2], sys.argv[3], sys.argv[4] + f".{mode}", mode=mode, seeds=sys.argv[5],
                       allowed_indels=7, max_hits=4)
    print(mode, st["n_reads"], st["n_mapped"], st["n_hits"], flush=True)
'''


def test_split_passes_match_oracle(tmp_path, oracle_built):
    from storm_amd import synth
    import numpy as np
    g = synth.repeat_genome(3_000_000, 41)
    gpath = str(tmp_path / "g.fa")
    synth.write_fasta_fast(gpath, "chr1", g)
    pr = synth.fast_nucleotide_reads(g, 600, 100, 42, sub=0.015, indel=0.002, qlo=15, qhi=40)
    rpath = str(tmp_path / "r.fq")
    synth.write_fastq(rpath, synth.packed_to_simreads(pr, np.arange(600)))
    out = str(tmp_path / "gpu.sam")
    env = dict(os.environ, STORM_PASS_HIT_LIMIT="20000")
    print("child: map_files plain and batched with STORM_PASS_HIT_LIMIT=20000", flush=True)
    r = subprocess.run([sys.executable, "-c", CHILD, ROOT, gpath, rpath, out, SEEDS_N], env=env, capture_output=True,
                       text=True, timeout=300)
    print(r.stdout, r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    want = str(tmp_path / "oracle.sam")
    subprocess.run([os.path.join(oracle_built, "oracle-nucleotide"), "-g", gpath, "-r", rpath, "-s", SEEDS_N, "-i", "7",
                    "-M", "4", "-o", want], check=True, stderr=subprocess.DEVNULL)
    assert sam_body(out + ".0") == sam_body(want)
    assert sam_body(out + ".1") == sam_body(want)
