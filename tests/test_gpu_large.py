"""Genome-scale regime on the GPU: a 20 Mbp reference with segmental duplications and tandem
repeats, 100 nt reads, -i 7 -M 4.  Reads here have hundreds of candidates each (the fused kernel
flushes its candidate buffer several times per task) and some read strands in repeats exceed the
fused kernel's raw-hit buffer (they take the staged kernels).  Batched and plain SAM must both
equal the CPU restatement's SAM."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, sam_body

pytestmark = pytest.mark.gpu
SEEDS_N = "####--#-#--##--####;#####--###-#-###;###-##-#######"
SEEDS_C = "####----##--##-####;####-###--#----####;####-####-####"


@pytest.fixture(scope="module")
def repeat_case(tmp_path_factory):
    from storm_amd import synth
    d = tmp_path_factory.mktemp("large")
    g = synth.repeat_genome(20_000_000, 21)
    gpath = str(d / "g.fa")
    synth.write_fasta_fast(gpath, "chr1", g)
    pr = synth.fast_nucleotide_reads(g, 1500, 100, 22, sub=0.015, indel=0.002, qlo=15, qhi=40)
    rpath = str(d / "r.fq")
    synth.write_fastq(rpath, synth.packed_to_simreads(pr, np.arange(1500)))
    return d, gpath, rpath


@pytest.fixture(scope="module")
def oracle_sam(repeat_case, oracle_built):
    d, gpath, rpath = repeat_case
    out = str(d / "oracle.sam")
    subprocess.run([os.path.join(ROOT, "oracle", "_build", "oracle-nucleotide"), "-g", gpath, "-r", rpath,
                    "-s", SEEDS_N, "-i", "7", "-M", "4", "-o", out], check=True, stderr=subprocess.DEVNULL)
    return sam_body(out)


@pytest.mark.parametrize("mode", [1, 0], ids=["batched", "plain"])
def test_repeat_genome_sam_matches_oracle(repeat_case, oracle_sam, mode):
    from storm_amd import lib
    d, gpath, rpath = repeat_case
    out = str(d / f"gpu_{mode}.sam")
    print(f"map_files mode={mode} 20 Mbp repeat genome, 1500 x 100 nt, -i 7 -M 4", flush=True)
    st = lib.map_files(gpath, rpath, out, mode=mode, seeds=SEEDS_N, allowed_indels=7, max_hits=4)
    print({k: st[k] for k in ("n_reads", "n_mapped", "n_candidates", "n_filter_pass", "n_hits", "n_fallback")},
          flush=True)
    assert st["n_candidates"] > 300 * st["n_reads"]   # the regime this test is for
    assert sam_body(out) == oracle_sam


def test_repeat_genome_color_batched_equals_plain(repeat_case):
    """Color reads on the same 20 Mbp reference: the fused path equals the staged path read for
    read (hits of every rank), including reads handed from one to the other."""
    from storm_amd import lib, synth
    d, gpath, _ = repeat_case
    g = synth.repeat_genome(20_000_000, 21)
    pr = synth.fast_color_reads(g, 4000, (34, 74), 23, snp=0.01, color_err=0.03, qlo=10, qhi=35)
    m = lib.Mapper(gpath, color=1, seeds=SEEDS_C, allowed_indels=3, z=85, t=100, max_hits=4)
    hb, sb = m.map_hits(pr, mode=1)
    hp, sp = m.map_hits(pr, mode=0)
    m.close()
    print({k: sb[k] for k in ("n_reads", "n_mapped", "n_candidates", "n_fallback")}, flush=True)
    assert sb["n_mapped"] > 0.5 * sb["n_reads"]
    for k in hb:
        assert np.array_equal(hb[k], hp[k]), k
