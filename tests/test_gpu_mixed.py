"""Mixed read lengths in one run (the batched path's uneven-length batching) against the CPU
restatement run once per length (the reference maps one read length per run,
reads_data.c:156-163); per-read results must be identical under -M."""
import collections
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu
SEEDS_N = "####--#-#--##--####;#####--###-#-###;###-##-#######"
SEEDS_C = "####----##--##-####;####-###--#----####;####-####-####"


def records_by_read(path):
    d = collections.defaultdict(list)
    for l in open(path):
        if l.startswith("@"):
            continue
        d[l.split("\t", 1)[0]].append(l)
    return d


def _write_subset(mode, reads, idx, path, qpath):
    from storm_amd import synth
    sub = synth.SimReads([reads.seqs[i] for i in idx], [reads.quals[i] for i in idx], reads.origin[idx],
                         reads.strand[idx], None if reads.first_base is None else reads.first_base[idx])
    names = [f"r{i}" for i in idx]
    if mode == "n":
        with open(path, "wb") as f:
            for nm, s, q in zip(names, sub.seqs, sub.quals):
                f.write(b"@" + nm.encode() + b"\n" + synth.BASES[s].tobytes() + b"\n+\n" +
                        (q.astype(np.int64) + 33).astype(np.uint8).tobytes() + b"\n")
    else:
        with open(path, "wb") as f, open(qpath, "wb") as fq:
            for nm, s, q, fb in zip(names, sub.seqs, sub.quals, sub.first_base):
                f.write(b">" + nm.encode() + b"\nT" + bytes([48 + (3 ^ int(fb))]) + synth.COLORS[s].tobytes() + b"\n")
                fq.write(b">" + nm.encode() + b"\n" + " ".join(str(int(x)) for x in q).encode() + b"\n")


@pytest.mark.parametrize("mode,i", [("n", 3), ("c", 3), ("n", 7), ("c", 7)])
def test_mixed_lengths_match_per_length_oracle(mode, i, tmp_path, oracle_built):
    from storm_amd import lib, synth
    g = synth.gc_window_genome(300_000, 5)
    gpath = str(tmp_path / "g.fa")
    synth.write_fasta(gpath, "chr1", g)
    if mode == "n":
        reads = synth.simulate_nucleotide_reads(g, 3000, (35, 75), 6, sub=0.02, indel=0.002)
    else:
        reads = synth.simulate_color_reads(g, 3000, (35, 75), 6, snp=0.01, color_err=0.03)
    lens = np.array([len(s) for s in reads.seqs])
    allp, allq = str(tmp_path / ("all.fq" if mode == "n" else "all.csfasta")), str(tmp_path / "all.qual")
    _write_subset(mode, reads, np.arange(len(lens)), allp, allq)
    seeds = SEEDS_N if mode == "n" else SEEDS_C
    kw = dict(color=1 if mode == "c" else 0, seeds=seeds, allowed_indels=i, max_hits=2, z=85 if mode == "c" else 100,
              t=100)
    out_b = str(tmp_path / "batched.sam")
    out_p = str(tmp_path / "plain.sam")
    q = allq if mode == "c" else None
    lib.map_files(gpath, allp, out_b, qual=q, mode=lib.STORM_MODE_BATCHED, **kw)
    lib.map_files(gpath, allp, out_p, qual=q, mode=lib.STORM_MODE_PLAIN, **kw)
    gb, gp = records_by_read(out_b), records_by_read(out_p)
    assert gb == gp
    orc = os.path.join(oracle_built, "oracle-nucleotide" if mode == "n" else "oracle-color")
    want = {}
    for L in np.unique(lens):
        idx = np.nonzero(lens == L)[0]
        p, qp = str(tmp_path / f"L{L}.{'fq' if mode == 'n' else 'csfasta'}"), str(tmp_path / f"L{L}.qual")
        _write_subset(mode, reads, idx, p, qp)
        args = [orc, "-g", gpath, "-r", p] + (["-q", qp] if mode == "c" else []) + \
               ["-s", seeds, "-i", str(i), "-M", "2", "-z", str(kw["z"]), "-t", "100", "-o", str(tmp_path / "o.sam")]
        subprocess.run(args, check=True, stderr=subprocess.DEVNULL)
        want.update(records_by_read(str(tmp_path / "o.sam")))
    assert len(want) > 0.8 * len(lens)
    assert gb == want


def _concat(parts):
    from storm_amd import synth
    fb = None if parts[0].first_base is None else np.concatenate([p.first_base for p in parts])
    return synth.SimReads(sum([list(p.seqs) for p in parts], []), sum([list(p.quals) for p in parts], []),
                          np.concatenate([p.origin for p in parts]), np.concatenate([p.strand for p in parts]), fb)


@pytest.mark.parametrize("mode,i", [("n", 3), ("n", 7), ("c", 3)])
def test_extreme_lengths_match_per_length_oracle(mode, i, tmp_path, oracle_built):
    """Reads from the filter band's width (the shortest the kernels take) up to 256 codes (the
    longest) in one batched run and one plain run, against the CPU restatement per length: the
    shortest have no seed window and stay unmapped, the longest fill the traceback records."""
    from storm_amd import lib, synth
    g = synth.gc_window_genome(300_000, 9 + i)
    gpath = str(tmp_path / "g.fa")
    synth.write_fasta(gpath, "chr1", g)
    w = 4 if i == 3 else 8
    lengths = [w, w + 1, 13, 20, 33, 64, 100, 150, 200, 240, 254, 255, 256]
    parts = []
    for k, L in enumerate(lengths):
        if mode == "n":
            parts.append(synth.simulate_nucleotide_reads(g, 60, (L, L), 100 + k, sub=0.02, indel=0.002))
        else:
            parts.append(synth.simulate_color_reads(g, 60, (L, L), 100 + k, snp=0.01, color_err=0.03))
    reads = _concat(parts)
    lens = np.array([len(s) for s in reads.seqs])
    assert lens.min() == w and lens.max() == 256
    perm = np.random.default_rng(i).permutation(len(lens))   # lengths interleaved in the file
    allp, allq = str(tmp_path / ("all.fq" if mode == "n" else "all.csfasta")), str(tmp_path / "all.qual")
    _write_subset(mode, reads, perm, allp, allq)
    seeds = SEEDS_N if mode == "n" else SEEDS_C
    kw = dict(color=1 if mode == "c" else 0, seeds=seeds, allowed_indels=i, max_hits=2, z=85 if mode == "c" else 100,
              t=100)
    q = allq if mode == "c" else None
    out_b, out_p = str(tmp_path / "batched.sam"), str(tmp_path / "plain.sam")
    lib.map_files(gpath, allp, out_b, qual=q, mode=lib.STORM_MODE_BATCHED, **kw)
    lib.map_files(gpath, allp, out_p, qual=q, mode=lib.STORM_MODE_PLAIN, **kw)
    gb, gp = records_by_read(out_b), records_by_read(out_p)
    assert gb == gp
    orc = os.path.join(oracle_built, "oracle-nucleotide" if mode == "n" else "oracle-color")
    want = {}
    for L in np.unique(lens):
        idx = np.nonzero(lens == L)[0]
        p, qp = str(tmp_path / f"L{L}.{'fq' if mode == 'n' else 'csfasta'}"), str(tmp_path / f"L{L}.qual")
        _write_subset(mode, reads, idx, p, qp)
        args = [orc, "-g", gpath, "-r", p] + (["-q", qp] if mode == "c" else []) + \
               ["-s", seeds, "-i", str(i), "-M", "2", "-z", str(kw["z"]), "-t", "100", "-o", str(tmp_path / "o.sam")]
        subprocess.run(args, check=True, stderr=subprocess.DEVNULL)
        want.update(records_by_read(str(tmp_path / "o.sam")))
    long_mapped = sum(1 for r in want if lens[int(r[1:])] >= 200)
    assert long_mapped > 0.8 * 60 * 5
    assert gb == want
