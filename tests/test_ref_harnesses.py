"""The CPU baselines' harnesses around the reference's own code (oracle/_ref) give the reference's
results, so the bench times the reference and not something else:
  * ref-map-harness (the reference program with its sequences and indexes kept across runs) writes,
    run by run and with reads of several lengths, the same SAM as one reference binary run per
    read file;
  * libref_align (alignment__align + alignment__traceback behind ref_align_harness.c) gives, record
    by record, the oracle's DP (orc_align), which the reference's SAM outputs pin.
Skipped where oracle/_ref has not been built."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, ORACLE_BUILD, sam_body

REF = os.path.join(ROOT, "oracle", "_ref")
SEEDS_N = "####--#-#--##--####;#####--###-#-###;###-##-#######"
SEEDS_C = "####----##--##-####;####-###--#----####;####-####-####"

pytestmark = pytest.mark.skipif(not os.path.exists(os.path.join(REF, "ref-map-harness-nucleotide")),
                                reason="reference harnesses not built (oracle/_ref)")


@pytest.mark.parametrize("color", [0, 1], ids=["nucleotide", "color"])
def test_map_harness_equals_reference_binary(color, tmp_path):
    from storm_amd import synth
    d = str(tmp_path)
    g = synth.gc_window_genome(400_000, 7 + color)
    synth.write_fasta_fast(os.path.join(d, "g.fa"), "chr1", g)
    runs = []
    for k, L in enumerate((70, 48, 36)):   # longest first, as the harness needs
        if color:
            pr = synth.fast_color_reads(g, 600, (L - 1, L - 1), 30 + k, snp=0.01, color_err=0.03, qlo=10, qhi=35)
        else:
            pr = synth.fast_nucleotide_reads(g, 600, L, 30 + k, sub=0.02, indel=0.002, qlo=15, qhi=40)
        sr = synth.packed_to_simreads(pr, np.arange(600))
        if color:
            synth.write_csfasta(os.path.join(d, f"r{k}.csfasta"), os.path.join(d, f"r{k}.qual"), sr)
            runs.append(["-r", os.path.join(d, f"r{k}.csfasta"), "-q", os.path.join(d, f"r{k}.qual")])
        else:
            synth.write_fastq(os.path.join(d, f"r{k}.fq"), sr)
            runs.append(["-r", os.path.join(d, f"r{k}.fq")])
    common = ["-g", os.path.join(d, "g.fa"), "-s", SEEDS_C if color else SEEDS_N, "-i", "3", "-z", "90", "-t", "100",
              "-M", "2", "-N", "4", "-v", "0"]
    suffix = "color" if color else "nucleotide"
    args = [os.path.join(REF, f"ref-map-harness-{suffix}"), os.path.join(d, "times.txt")] + common
    for k, r in enumerate(runs):
        args += ["--"] + r + ["-o", os.path.join(d, f"h{k}.sam")]
    env = dict(os.environ, MALLOC_PERTURB_="255")
    subprocess.run(args, check=True, env=env, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    times = [l.split() for l in open(os.path.join(d, "times.txt"))]
    assert [int(t[1]) for t in times] == [0, 1, 2] and all(float(t[3]) > 0 for t in times)
    for k, r in enumerate(runs):
        out = os.path.join(d, f"b{k}.sam")
        subprocess.run([os.path.join(REF, f"storm-{suffix}")] + common + r + ["-o", out], check=True, env=env,
                       stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
        want = sam_body(out)
        assert len(want) > 100
        assert sam_body(os.path.join(d, f"h{k}.sam")) == want, f"run {k}"


class OrcParams(C.Structure):
    _fields_ = [(f, C.c_int) for f in (
        "color", "match", "mismatch", "gap_open", "gap_extend", "allowed_indels", "simd_diags", "t", "z",
        "qual_base", "filter_repeats", "stdev_distance", "n_bytes", "sense_mask", "ignore_lowercase",
        "unordered_M", "uniqueness", "uniqueness_threshold", "print_all", "greedy")]


class OrcAlignOut(C.Structure):
    _fields_ = [("score", C.c_int), ("best_i", C.c_int), ("best_j", C.c_int), ("best_i0", C.c_int),
                ("best_j0", C.c_int), ("tb", C.c_uint8 * 1024), ("tb_len", C.c_int)]


@pytest.mark.parametrize("I,t", [(7, 115), (3, 60)])
def test_align_harness_equals_oracle(I, t, oracle_built):
    from storm_amd import synth
    rng = np.random.default_rng(3 + I)
    L, n, G = 75, 1500, 200_000
    g = rng.integers(0, 4, G).astype(np.uint8)
    diag = rng.integers(40, G - L - 60, n).astype(np.int32)
    codes = np.zeros((n, L), np.uint8)
    for k in range(n):   # reads from their diagonal with substitutions, a third of them random
        seq = g[diag[k]:diag[k] + L].copy()
        sub = rng.random(L) < 0.05
        seq[sub] = (seq[sub] + rng.integers(1, 4, int(sub.sum()))) % 4
        codes[k] = seq if k % 3 else rng.integers(0, 4, L)
    stride = (L + 3) // 4 + 8
    reads = np.zeros((n, stride), np.uint8)
    reads[:, :(L + 3) // 4] = synth.pack_rows(codes, (L + 3) // 4)
    P = C.c_void_p
    lib = C.CDLL(os.path.join(REF, "libref_align.so"))
    lib.ref_align_time.restype = C.c_double
    lib.ref_align_time.argtypes = [P, P, P, C.c_long, C.c_int, P, C.c_long] + [C.c_int] * 8 + [P] * 6 + [C.c_int]
    pad = 1024
    gb = (G + 3) // 4
    gp = np.zeros(pad + gb + pad, np.uint8)
    gp[pad:pad + gb] = synth.pack_codes(g)
    mp = np.full(pad + gb + pad, 0x55, np.uint8)
    mp[pad:pad + gb] = 0
    sc, bj, bi0, bj0, tl = (np.zeros(n, np.int32) for _ in range(5))
    tbr = np.zeros((n, 144), np.uint8)
    lib.ref_align_time(P(gp.ctypes.data + pad), P(mp.ctypes.data + pad), P(reads.ctypes.data), stride, L,
                       P(diag.ctypes.data), n, I, t, 5, -4, -9, -4, 4, 1, P(sc.ctypes.data), P(bj.ctypes.data),
                       P(bi0.ctypes.data), P(bj0.ctypes.data), P(tl.ctypes.data), P(tbr.ctypes.data), 144)
    o = C.CDLL(os.path.join(ORACLE_BUILD, "liboracle.so"))
    o.orc_params_default.argtypes = [C.POINTER(OrcParams), C.c_int]
    o.orc_align.argtypes = [P, P, C.c_int, P, P, C.c_int, C.POINTER(OrcParams), C.c_int, C.POINTER(OrcAlignOut)]
    op = OrcParams()
    o.orc_params_default(C.byref(op), 0)
    op.allowed_indels, op.simd_diags = I, min(I, 15)
    gpad = np.zeros(64 + G + 64, np.uint8)
    gpad[64:64 + G] = g
    mpad = np.ones_like(gpad)
    mpad[64:64 + G] = 0
    q = np.full(L, 3, np.uint8)
    n_acc = 0
    for k in range(n):
        h = int(diag[k])
        w = h - I if h - I > -min(I, 15) else -min(I, 15)
        wr = np.ascontiguousarray(gpad[64 + w:64 + w + L + 2 * I + 1])
        wm = np.ascontiguousarray(mpad[64 + w:64 + w + L + 2 * I + 1])
        c = np.ascontiguousarray(codes[k])
        a = OrcAlignOut()
        s = o.orc_align(P(c.ctypes.data), P(q.ctypes.data), L, P(wr.ctypes.data), P(wm.ctypes.data), I, C.byref(op), t,
                        C.byref(a))
        assert s == sc[k], k
        if s > t:
            n_acc += 1
            assert (a.best_j, a.best_i0, a.best_j0, a.tb_len) == (bj[k], bi0[k], bj0[k], tl[k]), k
            got = bytes(((tbr[k][i >> 1] >> (0 if i & 1 else 4)) & 0xF) for i in range(tl[k]))
            assert got == bytes(a.tb[:a.tb_len]), k
    assert 0.3 * n < n_acc < n
