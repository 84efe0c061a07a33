"""-O: the FASTQ of the reads without any hit (hit_map__generate_unmapped_FASTQ_output,
hit_map.c:514-578), written by storm_map_files on both GPU paths, against the oracle's, which is
pinned against the reference binary (test_oracle_vs_reference_binary.py); reads with and without
qualities, -M and ordered mode, both alphabets."""
import os
import subprocess

import pytest

from conftest import ROOT
from test_gpu_cases import params_of
from test_oracle_vs_reference_binary import SEEDS_C, SEEDS_N, UNMAPPED_CASES, make_inputs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode,kind,nreads,L,i,extra,seed,fasta", UNMAPPED_CASES,
                         ids=[f"{c[0]}-{c[6]}{'-fasta' if c[7] else ''}" for c in UNMAPPED_CASES])
def test_gpu_unmapped_fastq_equals_oracle(mode, kind, nreads, L, i, extra, seed, fasta, tmp_path, oracle_built):
    from storm_amd import lib
    tmp = str(tmp_path)
    gpath, rargs = make_inputs(tmp, mode, kind, nreads, L, seed, fasta_reads=fasta)
    orc = os.path.join(ROOT, "oracle", "_build", "oracle-nucleotide" if mode == "n" else "oracle-color")
    ou = os.path.join(tmp, "orc_un.fq")
    subprocess.run([orc, "-g", gpath] + rargs + ["-i", str(i), "-s", SEEDS_N if mode == "n" else SEEDS_C] + extra +
                   ["-o", os.path.join(tmp, "orc.sam"), "-O", ou], check=True, stderr=subprocess.DEVNULL)
    want = open(ou, "rb").read()
    assert want.count(b"\n") >= 40
    reads = rargs[rargs.index("-r") + 1]
    qual = rargs[rargs.index("-q") + 1] if "-q" in rargs else None
    for m in (lib.STORM_MODE_PLAIN, lib.STORM_MODE_BATCHED):
        gu = os.path.join(tmp, f"gpu_un{m}.fq")
        lib.map_files(gpath, reads, os.path.join(tmp, f"gpu{m}.sam"), qual=qual, mode=m, unmapped_path=gu,
                      **params_of(mode, i, extra))
        assert open(gu, "rb").read() == want, ("plain", "batched")[m]
