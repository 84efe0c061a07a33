"""The oracle's CSR index (index.c:19-195) against a brute-force restatement.

Brute force: every window start q in [0, size - span], the key of the seed at q (codes MSB
first, seed.c:203-214), positions grouped by key in ascending order.  Masked regions use the
closed form of the reference's skip loop (index.c:66-72): q is skipped iff
start - span <= q <= end - 2 for some region (the loop jumps to end - 1 and indexes it).  The
oracle runs the loop itself, so agreement over many region layouts (adjacent regions, regions at
the first and last code, gaps shorter than the span) pins the closed form the device kernel uses.
The repeat filter's erased set is checked on inputs where it fires (tandem repeats, -l 3).
"""
import ctypes as C
import os

import numpy as np
import pytest

from storm_amd import synth

LOWER = np.frombuffer(b"acgt", np.uint8)
UPPER = np.frombuffer(b"ACGT", np.uint8)


def oracle(oracle_built):
    lib = C.CDLL(os.path.join(oracle_built, "liboracle.so"))
    P = C.c_void_p
    lib.orc_fasta_sequence.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.POINTER(P), C.POINTER(C.c_int),
                                       C.POINTER(P), C.POINTER(P), C.POINTER(C.c_int)]
    lib.orc_fasta_index.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_char_p, C.c_int, C.POINTER(P),
                                    C.POINTER(C.c_int), C.POINTER(P), C.POINTER(C.c_int)]
    lib.orc_free.argtypes = [P]
    return lib


def _take(lib, p, n, dtype):
    a = np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_int32 if dtype == np.int32 else C.c_uint8)), shape=(max(n, 1),))
    out = a[:n].astype(dtype).copy()
    lib.orc_free(p)
    return out


def oracle_sequence(lib, path, color, lower, ref_no):
    pk, st, en = C.c_void_p(), C.c_void_p(), C.c_void_p()
    size, nr = C.c_int(), C.c_int()
    assert lib.orc_fasta_sequence(path.encode(), color, lower, ref_no, C.byref(pk), C.byref(size), C.byref(st),
                                  C.byref(en), C.byref(nr)) == 0
    packed = _take(lib, pk, (size.value + 3) // 4, np.uint8)
    return packed, size.value, _take(lib, st, nr.value, np.int32), _take(lib, en, nr.value, np.int32)


def oracle_index(lib, path, color, lower, ref_no, seed, l):
    f, m = C.c_void_p(), C.c_void_p()
    nf, nm = C.c_int(), C.c_int()
    assert lib.orc_fasta_index(path.encode(), color, lower, ref_no, seed.encode(), l, C.byref(f), C.byref(nf),
                               C.byref(m), C.byref(nm)) == 0
    return _take(lib, f, nf.value, np.int32), _take(lib, m, nm.value, np.int32)


def unpack(packed, size):
    i = np.arange(size)
    return (packed[i >> 2] >> ((i & 3) << 1)) & 3


def brute_index(codes, regions, seed, l):
    offs = [k for k, c in enumerate(seed) if c == "#"]
    w, span = len(offs), len(seed)
    ks = 1 << (2 * w)
    size = len(codes)
    n = size - span + 1
    first = np.zeros(ks + 1, np.int64)
    if n <= 0 or span >= size:
        return first.astype(np.int32), np.zeros(0, np.int32)
    q = np.arange(n)
    keep = np.ones(n, bool)
    for s, e in regions:
        keep[max(0, s - span):max(0, min(n, e - 1))] = False
    key = np.zeros(n, np.int64)
    for o in offs:
        key = (key << 2) | codes[q + o]
    q, key = q[keep], key[keep]
    counts = np.bincount(key, minlength=ks)
    erased = np.zeros(ks, bool)
    if l > 0:
        count = int(counts.sum())
        c = counts[counts > 0].astype(np.float64)
        # index.c:119-131: sums in key order (np.cumsum accumulates left to right)
        avg = np.cumsum(1.0 / count * c * c)[-1] / count
        sd = np.sqrt(np.cumsum((c - avg) * (c - avg) / count)[-1])
        erased = (counts > 0) & ((counts - avg) > l * sd)
    kept = np.where(erased, 0, counts)
    first[1:] = np.cumsum(kept)
    order = np.argsort(key, kind="stable")
    key_s, q_s = key[order], q[order]
    sel = ~erased[key_s]
    return first.astype(np.int32), q_s[sel].astype(np.int32)


def write_masked_fasta(path, seqs):
    """seqs: list of (name, codes, lowercase mask, N mask)."""
    with open(path, "wb") as f:
        for name, codes, low, nn in seqs:
            letters = np.where(low, LOWER[codes], UPPER[codes])
            letters = np.where(nn, np.uint8(ord("N")), letters).astype(np.uint8)
            f.write(b">" + name.encode() + b"\n")
            for s in range(0, len(letters), 60):
                f.write(letters[s:s + 60].tobytes() + b"\n")


def masked_layout(n, g, n_runs, max_len):
    m = np.zeros(n, bool)
    for _ in range(n_runs):
        s = int(g.integers(0, n))
        m[s:s + int(g.integers(1, max_len))] = True
    return m


def make_cases(tmp_path):
    g = synth.rng(41)
    seqs = []
    # 1: N runs and lowercase runs of every length around the seed spans, adjacent runs
    c = synth.random_genome(60_000, 5)
    low = masked_layout(len(c), g, 120, 40)
    nn = masked_layout(len(c), g, 60, 30)
    nn[1000:1010] = True
    low[1010:1020] = True    # lowercase run right after an N run
    nn[:7] = True            # run at the first code
    low[-5:] = True          # run at the last code
    seqs.append(("chrA", c, low, nn))
    # 2: tandem repeats (repeat filter fires at -l 3) with a few N runs
    c2 = synth.repeat_genome(80_000, 6, seg_dup_frac=0.1, tandem_frac=0.3)
    seqs.append(("chrB", c2, np.zeros(len(c2), bool), masked_layout(len(c2), g, 10, 20)))
    # 3: shorter than every seed span
    c3 = synth.random_genome(12, 7)
    seqs.append(("tiny", c3, np.zeros(12, bool), np.zeros(12, bool)))
    # 4: masked regions separated by gaps shorter than a span
    c4 = synth.random_genome(5_000, 8)
    nn4 = np.zeros(len(c4), bool)
    for s in range(100, 4000, 23):
        nn4[s:s + 3] = True
    seqs.append(("chrC", c4, np.zeros(len(c4), bool), nn4))
    path = os.path.join(str(tmp_path), "masked.fa")
    write_masked_fasta(path, seqs)
    return path, len(seqs)


SEEDS = ["##-#", "#-##--#", "####--#-#--##--####", "####-####-####"]


@pytest.mark.parametrize("lower", [0, 1])
def test_oracle_index_matches_brute_force(tmp_path, oracle_built, lower):
    lib = oracle(oracle_built)
    path, n_refs = make_cases(tmp_path)
    fired = 0
    for ref_no in range(n_refs):
        packed, size, rs, re_ = oracle_sequence(lib, path, 0, lower, ref_no)
        codes = unpack(packed, size).astype(np.int64)
        regions = list(zip(rs.tolist(), re_.tolist()))
        for seed in SEEDS:
            big = seed.count("#") > 8
            for l in ((100,) if big and lower else (100, 3, 0)):
                f_o, m_o = oracle_index(lib, path, 0, lower, ref_no, seed, l)
                f_b, m_b = brute_index(codes, regions, seed, l)
                assert np.array_equal(f_o, f_b), (ref_no, seed, l)
                assert np.array_equal(m_o, m_b), (ref_no, seed, l)
                if l == 3:
                    f0, _ = brute_index(codes, regions, seed, 0)
                    fired += int(f0[-1] != f_b[-1])
    assert fired > 0   # the repeat filter erased keys in some case


def test_oracle_regions_cover_masked_codes(tmp_path, oracle_built):
    """Regions are the maximal runs of non-ACGT codes (reference_data.c:114-130): N runs always,
    lowercase runs only with -j (the loader upper-cases otherwise)."""
    lib = oracle(oracle_built)
    path, _ = make_cases(tmp_path)
    packed, size, rs, re_ = oracle_sequence(lib, path, 0, 0, 3)
    covered = np.zeros(size, bool)
    for s, e in zip(rs, re_):
        assert 0 <= s < e <= size
        covered[s:e] = True
    c4 = synth.random_genome(5_000, 8)
    nn4 = np.zeros(len(c4), bool)
    for s in range(100, 4000, 23):
        nn4[s:s + 3] = True
    assert np.array_equal(covered, nn4)
    assert np.all(rs[1:] >= re_[:-1])
