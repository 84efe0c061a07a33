"""FASTQ reads parsed, packed and ordered by tag on the device (storm_reads_load, k_ingest.hip)
against the host loader, which is pinned against the oracle's loader (test_reads_load.py) and,
for quality lines the reference reads as comments, against the reference binary
(test_oracle_vs_reference_binary.py).  Byte arrays must be identical: names, packed codes, packed
quality levels, lengths, tags and their order, color primer fields, reads without quality, the
read-group name; files the reference rejects as corrupt fail the same way on both paths.
"""
import os
import subprocess
import time

import numpy as np
import pytest

from storm_amd import lib, synth
from conftest import ROOT, sam_body
from reads_files import cases, corrupt_cases, varlen_case, write_fastq_hash_quals

pytestmark = pytest.mark.gpu


def same(a, b, where):
    assert a["n"] == b["n"], where
    assert a["names"] == b["names"], where
    assert a["db_name"] == b["db_name"], where
    assert np.array_equal(a["len"], b["len"]), where
    assert np.array_equal(a["tags"], b["tags"]), where
    assert np.array_equal(a["codes"], b["codes"]), where
    for k in ("qual", "first_base", "first_qual", "no_qual"):
        if a[k] is None or b[k] is None:
            assert a[k] is None and b[k] is None, (where, k)
        else:
            assert np.array_equal(a[k], b[k]), (where, k)


def test_device_loader_equals_host_loader(tmp_path):
    files = cases(str(tmp_path)) + [(varlen_case(str(tmp_path), 0), 0), (varlen_case(str(tmp_path), 1), 1)]
    for path, color in files:
        d = lib.load_reads(path, use_device=True, color=color)
        h = lib.load_reads(path, use_device=False, color=color)
        print(f"{os.path.basename(path)}: {d['n']} reads, device {d['device_ms']:.2f} ms", flush=True)
        assert d["on_device"] and not h["on_device"], path
        same(d, h, path)


def test_corrupt_files_fail_on_both_paths(tmp_path):
    for path in corrupt_cases(str(tmp_path)):
        msgs = []
        for use_device in (True, False):
            with pytest.raises(lib.StormError) as e:
                lib.load_reads(path, use_device=use_device, color=0)
            msgs.append(str(e.value))
        assert msgs[0] == msgs[1] and "corrupt" in msgs[0], (path, msgs)


def test_other_formats_load_on_the_host(tmp_path):
    g = synth.random_genome(20000, 3)
    r = synth.simulate_color_reads(g, 200, 40, 5)
    path, qpath = os.path.join(str(tmp_path), "r.csfasta"), os.path.join(str(tmp_path), "r.qual")
    synth.write_csfasta(path, qpath, r)
    d = lib.load_reads(path, qual=qpath, use_device=True, color=1)
    assert not d["on_device"] and d["n"] == 200


def test_map_files_with_hash_quality_lines_equals_oracle(tmp_path, oracle_built):
    """End to end through storm_map_files (FASTQ loaded on the device): reads whose quality line
    starts with '#' map without quality and print QUAL '*', as the reference and the oracle do."""
    tmp = str(tmp_path)
    g = synth.random_genome(60000, 41)
    gpath = os.path.join(tmp, "g.fa")
    synth.write_fasta(gpath, "chr1", g)
    r = synth.simulate_nucleotide_reads(g, 1500, 50, 48, sub=0.03, indel=0.01)
    rpath = os.path.join(tmp, "r.fq")
    write_fastq_hash_quals(rpath, r)
    orc = os.path.join(ROOT, "oracle", "_build", "oracle-nucleotide")
    seeds = "####--#-#--##--####;#####--###-#-###;###-##-#######"
    subprocess.run([orc, "-g", gpath, "-r", rpath, "-i", "3", "-M", "2", "-s", seeds, "-o",
                    os.path.join(tmp, "orc.sam")], check=True, stderr=subprocess.DEVNULL)
    want = sam_body(os.path.join(tmp, "orc.sam"))
    for mode in (lib.STORM_MODE_PLAIN, lib.STORM_MODE_BATCHED):
        out = os.path.join(tmp, f"gpu{mode}.sam")
        lib.map_files(gpath, rpath, out, mode=mode, color=0, allowed_indels=3, max_hits=2, seeds=seeds)
        got = sam_body(out)
        assert got == want, mode
    assert any(l.split("\t")[10] == "*" for l in want if not l.startswith("@"))


def test_device_loader_two_million_reads(tmp_path):
    """2M reads of 100 nt (a 500 MB FASTQ): identical arrays, and the device path's time."""
    g = synth.random_genome(1_000_000, 9)
    pr = synth.fast_nucleotide_reads(g, 2_000_000, 100, 19, sub=0.015, indel=0.002, qlo=2, qhi=40)
    path = os.path.join(str(tmp_path), "big.fq")
    L = 100
    codes = synth.unpack_rows(pr.codes, L)
    letters = synth.BASES[codes]
    rng = np.random.default_rng(4)
    qchars = (33 + rng.integers(2, 41, size=(len(pr.len), L))).astype(np.uint8)
    with open(path, "wb") as f:
        for s in range(0, len(pr.len), 200_000):
            e = min(len(pr.len), s + 200_000)
            block = bytearray()
            for k in range(s, e):
                block += b"@r%d\n" % k + letters[k].tobytes() + b"\n+\n" + qchars[k].tobytes() + b"\n"
            f.write(block)
    t0 = time.perf_counter()
    d = lib.load_reads(path, use_device=True, color=0)
    td = time.perf_counter() - t0
    t0 = time.perf_counter()
    h = lib.load_reads(path, use_device=False, color=0)
    th = time.perf_counter() - t0
    print(f"2M x 100 nt FASTQ: device path {td:.2f} s ({d['device_ms']:.1f} ms of kernels and copies on the "
          f"device), host path {th:.2f} s", flush=True)
    assert d["on_device"]
    same(d, h, path)
