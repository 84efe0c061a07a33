"""Writes the golden fixtures of the mapping path from the compiled reference.

Run in the container that holds the reference sources (it builds oracle/_ref from them first):

    python tests/golden/make_golden.py

Outputs, all small and committed:
  sam_<case>/   genome.fa.gz, reads.*.gz, ref.sam.gz (the reference's SAM: -M, or the ordered
                genome-map output when the case has no -M), args.json
  filter_W<w>_L<l>.npz   packed genome/reads, window starts and the u8 score of every candidate
                         from the reference SSE2 routine (threshold sweep over t = 0..254)

The reference runs with MALLOC_PERTURB_=255 so that malloc returns zeroed memory: the color
SAM writer ORs base symbols into a malloc'd buffer (hit_map.c:240-246), and zeroed memory is
what makes that output deterministic.
"""
from __future__ import annotations

import ctypes as C
import gzip
import hashlib
import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from storm_amd import synth  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref")
SEEDS_N = "####--#-#--##--####;#####--###-#-###;###-##-#######"
SEEDS_C = "####----##--##-####;####-###--#----####;####-####-####"

# (name, mode, genome kind, genome size, reads, read length, -i, extra args, seed)
SAM_CASES = [
    ("n_i3_M1", "n", "plain", 60000, 1500, 50, 3, ["-M", "1", "-z", "100", "-t", "100"], 101),
    ("n_i7_M4_rep", "n", "repeat", 60000, 1500, 75, 7, ["-M", "4", "-l", "3"], 102),
    ("n_i1_M2_N", "n", "masked", 40000, 1000, 36, 1, ["-M", "2"], 103),
    ("c_i3_M1", "c", "plain", 60000, 1500, 50, 3, ["-M", "1", "-z", "85", "-t", "100"], 104),
    ("c_i7_M4_rep", "c", "repeat", 60000, 1500, 60, 7, ["-M", "4", "-l", "3"], 105),
    ("c_i2_M3_u", "c", "masked", 40000, 1000, 35, 2, ["-M", "3", "-u", "4", "-A"], 106),
    # ordered mode (no -M): the genome map output, contextual by default, greedy with -G
    ("n_i7_ordered_rep", "n", "repeat", 60000, 1500, 75, 7, ["-l", "3"], 107),
    ("n_i3_ordered_G_u", "n", "masked", 40000, 1000, 50, 3, ["-G", "-u", "3"], 108),
    ("c_i3_ordered_rep", "c", "repeat", 60000, 1500, 50, 3, ["-z", "85"], 109),
]

FILTER_CASES = [(d, L) for d in (1, 3, 7, 15) for L in (35, 50, 75, 100, 150)]


def sh(cmd, **kw):
    subprocess.run(cmd, check=True, **kw)


def gz_write(path, data: bytes):
    with gzip.GzipFile(path, "wb", mtime=0) as f:
        f.write(data)


def write_genome(path, kind, size, seed):
    g = synth.rng(seed)
    if kind == "repeat":
        codes = synth.repeat_genome(size, seed, 0.2, 0.15)
    else:
        codes = synth.random_genome(size, seed)
    letters = bytearray(synth.BASES[codes].tobytes())
    if kind == "masked":
        for _ in range(6):
            a = int(g.integers(200, size - 400))
            b = a + int(g.integers(1, 150))
            letters[a:b] = b"N" * (b - a)
        a = int(g.integers(0, size - 600))
        letters[a:a + 500] = letters[a:a + 500].lower()
    with open(path, "wb") as f:
        f.write(b">chr%d golden\n" % seed)
        for s in range(0, len(letters), 60):
            f.write(bytes(letters[s:s + 60]) + b"\n")
    return codes


def make_sam_case(name, mode, kind, gsize, nreads, L, i, extra, seed, tmp):
    out = os.path.join(HERE, "sam_" + name)
    os.makedirs(out, exist_ok=True)
    gpath = os.path.join(tmp, "genome.fa")
    codes = write_genome(gpath, kind, gsize, seed)
    if mode == "n":
        r = synth.simulate_nucleotide_reads(codes, nreads, L, seed + 1, sub=0.03, indel=0.01)
        rpath = os.path.join(tmp, "reads.fq")
        synth.write_fastq(rpath, r)
        rargs = ["-r", rpath]
        files = {"reads.fq": rpath}
        binary = os.path.join(REF, "storm-nucleotide")
        seeds = SEEDS_N
    else:
        r = synth.simulate_color_reads(codes, nreads, L, seed + 1, snp=0.02, color_err=0.04)
        rpath, qpath = os.path.join(tmp, "reads.csfasta"), os.path.join(tmp, "reads.qual")
        synth.write_csfasta(rpath, qpath, r)
        rargs = ["-r", rpath, "-q", qpath]
        files = {"reads.csfasta": rpath, "reads.qual": qpath}
        binary = os.path.join(REF, "storm-color")
        seeds = SEEDS_C
    args = ["-s", seeds, "-i", str(i)] + extra
    sam = os.path.join(tmp, "ref.sam")
    sh([binary, "-g", gpath] + rargs + args + ["-o", sam, "-v", "0"],
       env=dict(os.environ, MALLOC_PERTURB_="255"))
    files["genome.fa"] = gpath
    files["ref.sam"] = sam
    digests = {}
    for k, p in files.items():
        data = open(p, "rb").read()
        gz_write(os.path.join(out, k + ".gz"), data)
        digests[k] = hashlib.sha256(data).hexdigest()
    meta = {"mode": mode, "args": args, "reads": list(files.keys()), "sha256": digests,
            "generator": {"genome": kind, "size": gsize, "reads": nreads, "L": L, "seed": seed}}
    with open(os.path.join(out, "args.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)
    print("wrote", out)


def make_filter_case(d, L, lib):
    W = 2 if d <= 1 else 4 if d <= 3 else 8 if d <= 7 else 16
    seed = 1000 + 10 * d + L
    g = synth.random_genome(6000, seed)
    n = 160
    reads, pos = synth.filter_candidates(g, n, L, W, d, seed=seed + 1, sub=0.04)
    pos[:4, 1] = -np.arange(1, 5) * 3 - 1           # windows starting in the leading padding
    pos[4:8, 1] = len(g) - L - np.arange(0, 4)       # windows running into the trailing padding
    pad = 128
    gp = np.zeros(pad + (len(g) + 3) // 4 + pad, np.uint8)
    gp[pad:pad + (len(g) + 3) // 4] = synth.pack_codes(g)
    stride = (L + 3) // 4 + 16
    rp = np.zeros(n * stride + 64, np.uint8)
    for k in range(n):
        rp[k * stride:k * stride + (L + 3) // 4] = synth.pack_codes(reads[k])
    seg = np.arange(0, 2 * n + 1, 2, dtype=np.int64)
    posf = pos.reshape(-1).astype(np.int32)
    scores = np.zeros(2 * n, np.uint8)
    P = C.c_void_p
    rc = lib.ref_filter_scores(0, d, L, 5, -4, -9, -4, P(gp.ctypes.data + pad), n, seg.ctypes.data_as(P),
                               posf.ctypes.data_as(P), rp.ctypes.data_as(P), stride, scores.ctypes.data_as(P), 8)
    assert rc == 0
    np.savez_compressed(os.path.join(HERE, f"filter_W{W}_L{L}.npz"), genome_packed=gp, pad=pad, genome_len=len(g),
                        reads_packed=rp, stride=stride, L=L, W=W, d=d, pos=posf, read_of=np.repeat(np.arange(n), 2),
                        scores=scores)


def main():
    """With arguments, only the SAM cases of those names are written."""
    if not os.path.isdir("/root/reference"):
        raise SystemExit("the reference sources are needed to write golden fixtures")
    sh(["make", "-C", os.path.join(ROOT, "oracle"), "-j8", "ref"])
    tmp = tempfile.mkdtemp(prefix="storm_golden_")
    only = set(sys.argv[1:])
    for c in SAM_CASES:
        if not only or c[0] in only:
            make_sam_case(*c, tmp)
    if only:
        shutil.rmtree(tmp)
        return
    lib = C.CDLL(os.path.join(REF, "libref_filter.so"))
    lib.ref_filter_scores.argtypes = [C.c_int] * 7 + [C.c_void_p, C.c_long, C.c_void_p, C.c_void_p, C.c_void_p,
                                                      C.c_long, C.c_void_p, C.c_int]
    for d, L in FILTER_CASES:
        make_filter_case(d, L, lib)
    print("filter fixtures written")
    shutil.rmtree(tmp)


if __name__ == "__main__":
    main()
