"""Parity at the benchmarked sizes.

* The 1 Mbp / 100k-read nucleotide workload in full (1 Mbp i.i.d. genome of seed 1, 100k reads of
  50 nt, 2 % substitutions, 0.1 % indels, Q20-40, default seeds, -i 3 -z 100 -t 100 -M 1): the
  SAM of both GPU paths equals the CPU restatement's, line for line.
* The 50 Mbp color workload (GC 40-60 % per 10 kb window, reads of 34-74 colors, -i 3 -z 85
  -t 100 -M 1) at 2M reads and the 100 Mbp repeat-genome workload (segmental duplications and
  tandem repeats, 100 nt reads, -i 7 -M 4) at 1M reads: the two GPU paths return the same whole
  hit map and the same traceback bytes for every read (on the repeat genome also with the batched
  call split into three launches, each with reads handed to the staged kernels).  The staged path maps one read length at
  a time, as the reference does (reads_data.c:156-163), and is itself pinned against the CPU
  restatement per length (test_gpu_mixed.py, test_gpu_large.py); at these sizes the comparison
  covers reads with thousands of candidates and, on the repeat genome, reads handed from the fused
  kernel to the staged kernels."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, sam_body

pytestmark = pytest.mark.gpu
SEEDS_N = "####--#-#--##--####;#####--###-#-###;###-##-#######"
SEEDS_C = "####----##--##-####;####-###--#----####;####-####-####"


@pytest.fixture(scope="module")
def config1_case(tmp_path_factory, oracle_built):
    from storm_amd import synth
    d = tmp_path_factory.mktemp("scale1")
    g = synth.random_genome(1_000_000, 1)
    gpath = str(d / "g.fa")
    synth.write_fasta_fast(gpath, "chr1", g)
    pr = synth.fast_nucleotide_reads(g, 100_000, 50, 101, sub=0.02, indel=0.001, qlo=20, qhi=40)
    rpath = str(d / "r.fq")
    synth.write_fastq(rpath, synth.packed_to_simreads(pr, np.arange(100_000)))
    out = str(d / "oracle.sam")
    print("oracle: 100k x 50 nt on 1 Mbp", flush=True)
    subprocess.run([os.path.join(ROOT, "oracle", "_build", "oracle-nucleotide"), "-g", gpath, "-r", rpath, "-i", "3",
                    "-z", "100", "-t", "100", "-M", "1", "-o", out], check=True, stderr=subprocess.DEVNULL)
    return d, gpath, rpath, sam_body(out)


@pytest.mark.parametrize("mode", [1, 0], ids=["batched", "plain"])
def test_config1_full_size_sam_matches_oracle(config1_case, mode):
    from storm_amd import lib
    d, gpath, rpath, want = config1_case
    out = str(d / f"gpu_{mode}.sam")
    print(f"map_files mode={mode}: 100k x 50 nt on 1 Mbp, -i 3 -M 1", flush=True)
    st = lib.map_files(gpath, rpath, out, mode=mode, allowed_indels=3, z=100, t=100, max_hits=1)
    assert st["n_reads"] == 100_000
    got = sam_body(out)
    assert len(got) == len(want)
    assert got == want


def _records_equal(m, pr, lo, hi):
    from storm_amd import lib
    hb, tbb, sb = m.map_records(pr, lib.STORM_MODE_BATCHED, lo, hi)
    hp, tbp, sp = m.map_records(pr, lib.STORM_MODE_PLAIN, lo, hi)
    bad = np.nonzero(~(hb == hp).all(axis=1) | ~(tbb == tbp).reshape(len(tbb), -1).all(axis=1))[0]
    return bad, sb, sp


def test_config3_scale_batched_equals_plain():
    from storm_amd import lib, synth
    import tempfile
    d = tempfile.mkdtemp(prefix="storm_scale3_")
    g = synth.gc_window_genome(50_000_000, 3)
    gpath = os.path.join(d, "g.fa")
    synth.write_fasta_fast(gpath, "chr1", g)
    pr = synth.fast_color_reads(g, 2_000_000, (34, 74), 303, snp=0.01, color_err=0.03, qlo=10, qhi=35)
    m = lib.Mapper(gpath, color=1, seeds=SEEDS_C, allowed_indels=3, z=85, t=100, max_hits=1)
    print("records: 2M color reads of 34-74 colors on 50 Mbp, both paths", flush=True)
    bad, sb, sp = _records_equal(m, pr, 0, 2_000_000)
    m.close()
    print({k: sb[k] for k in ("n_mapped", "n_candidates", "n_filter_pass", "n_fallback")}, flush=True)
    assert sb["n_candidates"] > 500 * 2_000_000 // 1000   # hundreds of candidates per read
    assert len(bad) == 0, f"{len(bad)} reads differ, first {bad[:10]}"
    assert sb["n_mapped"] == sp["n_mapped"]


def test_config4_scale_batched_equals_plain(monkeypatch):
    from storm_amd import lib, synth
    import tempfile
    d = tempfile.mkdtemp(prefix="storm_scale4_")
    g = synth.repeat_genome(100_000_000, 4)
    gpath = os.path.join(d, "g.fa")
    synth.write_fasta_fast(gpath, "chr1", g)
    pr = synth.fast_nucleotide_reads(g, 1_000_000, 100, 404, sub=0.015, indel=0.002, qlo=15, qhi=40)
    m = lib.Mapper(gpath, color=0, seeds=SEEDS_N, allowed_indels=7, max_hits=4)
    print("records: 1M x 100 nt on the 100 Mbp repeat genome, -i 7 -M 4, both paths", flush=True)
    bad, sb, sp = _records_equal(m, pr, 0, 1_000_000)
    # the same batch in three launches over consecutive reads, each with its own staged fallback
    monkeypatch.setenv("STORM_RECORD_CHUNKS", "3")
    hc, tbc, sc = m.map_records(pr, lib.STORM_MODE_BATCHED, 0, 1_000_000)
    hp, tbp, _ = m.map_records(pr, lib.STORM_MODE_PLAIN, 0, 1_000_000)
    m.close()
    print({k: sb[k] for k in ("n_mapped", "n_candidates", "n_filter_pass", "n_fallback")}, flush=True)
    assert sb["n_fallback"] > 0               # reads handed to the staged kernels are covered
    assert len(bad) == 0, f"{len(bad)} reads differ, first {bad[:10]}"
    assert sb["n_mapped"] == sp["n_mapped"]
    assert sc["n_fallback"] == sb["n_fallback"]
    assert np.array_equal(hc, hp) and np.array_equal(tbc, tbp)
