"""Batch DP + traceback through the C ABI (storm_align_plan_*, the staged k_align kernel) against
the oracle's alignment__align / alignment__traceback restatement (orc_align, orc_align_bound;
alignment.c:326-734), record by record: score (MIN_SCORE on the early exit of alignment.c:613-616
with min_accepted = t), the survival bound, the last-row column, and for every record scoring
above t the traceback's end cell and codes.  The oracle's DP is pinned through the reference's
whole -M SAM outputs (test_oracle_golden.py) and the compiled reference
(test_oracle_vs_reference_binary.py).

Records cover both strands (the oracle reverses a read with orc_read_reverse, the reference's
codes.c:429-439 / read_quality.c:47-64), reads with and without quality levels, mixed lengths,
diagonals whose window is clamped at the sequence start (run.c:667) or runs into the padding
past its end, masked genome stretches, -i from 0 to 31 and both alphabets.
"""
import ctypes as C
import os

import numpy as np
import pytest

from storm_amd import lib
from conftest import ORACLE_BUILD

pytestmark = pytest.mark.gpu

GPAD = 1024
MIN_SCORE = -16383


class OrcParams(C.Structure):
    _fields_ = [(n, C.c_int) for n in (
        "color", "match", "mismatch", "gap_open", "gap_extend", "allowed_indels", "simd_diags", "t", "z",
        "qual_base", "filter_repeats", "stdev_distance", "n_bytes", "sense_mask", "ignore_lowercase",
        "unordered_M", "uniqueness", "uniqueness_threshold", "print_all", "greedy")]


class OrcAlignOut(C.Structure):
    _fields_ = [("score", C.c_int), ("best_i", C.c_int), ("best_j", C.c_int), ("best_i0", C.c_int),
                ("best_j0", C.c_int), ("tb", C.c_uint8 * 1024), ("tb_len", C.c_int)]


class OrcRead(C.Structure):
    _fields_ = [("info", C.c_char_p), ("first_base", C.c_uint8), ("first_qual", C.c_uint8),
                ("qual", C.c_void_p), ("seq", C.c_void_p), ("tag", C.c_ulong)]


def _oracle():
    o = C.CDLL(os.path.join(ORACLE_BUILD, "liboracle.so"))
    P = C.c_void_p
    o.orc_params_default.argtypes = [C.POINTER(OrcParams), C.c_int]
    o.orc_align.argtypes = [P, P, C.c_int, P, P, C.c_int, C.POINTER(OrcParams), C.c_int, C.POINTER(OrcAlignOut)]
    o.orc_align.restype = C.c_int
    o.orc_align_bound.argtypes = [P, P, C.c_int, P, P, C.c_int, C.POINTER(OrcParams), C.POINTER(C.c_int),
                                  C.POINTER(C.c_int)]
    o.orc_read_reverse.argtypes = [C.POINTER(OrcRead), C.POINTER(OrcRead), C.c_int, C.c_int]
    return o


def _pack(codes, nbytes):
    c = np.zeros(nbytes * 4, np.uint8)
    c[:len(codes)] = codes
    c = c.reshape(-1, 4)
    return (c[:, 0] | (c[:, 1] << 2) | (c[:, 2] << 4) | (c[:, 3] << 6)).astype(np.uint8)


def _unpack(packed, n):
    b = np.asarray(packed, np.uint8)
    return np.stack([(b >> s) & 3 for s in (0, 2, 4, 6)], axis=1).reshape(-1)[:n]


def _strand(o, codes, qual, first_qual, L, color):
    """The reverse strand of a read, through the oracle's read reversal."""
    nb = L // 4 + 64
    sp = _pack(codes, nb)
    qp = _pack(qual, nb) if qual is not None else None
    ds, dq = np.zeros(nb, np.uint8), np.zeros(nb, np.uint8)
    src = OrcRead(None, 0, int(first_qual), qp.ctypes.data if qp is not None else None, sp.ctypes.data, 0)
    dst = OrcRead(None, 0, 0, dq.ctypes.data if qp is not None else None, ds.ctypes.data, 0)
    o.orc_read_reverse(C.byref(src), C.byref(dst), L, color)
    return _unpack(ds, L), (_unpack(dq, L) if qp is not None else None)


def _case(seed, color, I, t, with_qual, n_reads=160, G=12000):
    rng = np.random.default_rng(seed)
    genome = rng.integers(0, 4, G).astype(np.uint8)
    mask = np.zeros(G, np.uint8)
    for s in rng.integers(0, G - 300, 4):
        mask[s:s + rng.integers(20, 200)] = 1
    lens = rng.integers(24, 130, n_reads).astype(np.int32)
    lens[:4] = (256, 200, 1 if not color else 2, 3)
    reads, quals, fq, diags, rs = [], [], [], [], []
    for r in range(n_reads):
        L = int(lens[r])
        pos = int(rng.integers(0, G - L - 40))
        seq = genome[pos:pos + L + 40].copy()
        # substitutions, then a few indels, then trimmed to L
        sub = rng.random(len(seq)) < 0.04
        seq[sub] = (seq[sub] + rng.integers(1, 4, int(sub.sum()))) % 4
        for _ in range(int(rng.integers(0, 3)) if L > 4 else 0):
            k = int(rng.integers(1, L))
            if rng.random() < 0.5:
                seq = np.delete(seq, k)
            else:
                seq = np.insert(seq, k, rng.integers(0, 4))
        seq = seq[:L].astype(np.uint8)
        if r % 7 == 3:
            seq = rng.integers(0, 4, L).astype(np.uint8)
        reads.append(seq)
        q = np.where(rng.random(L) < 0.6, 3, rng.integers(0, 4, L)).astype(np.uint8)
        quals.append(q)
        fq.append(int(rng.integers(0, 4)))
        for strand in (0, 1):
            for dh in (0, int(rng.integers(-I - 2, I + 3))):
                rs.append(2 * r + strand)
                diags.append(pos + dh)
    # windows clamped at the start (h - I <= -min(I, 15)) and running into the padding past the end
    for r in range(8):
        L = int(lens[r])
        rs += [2 * r, 2 * r + 1, 2 * r]
        diags += [int(rng.integers(-40, 3)), -I - 5, G - L + int(rng.integers(-5, 30))]
    return genome, mask, reads, (quals if with_qual else None), lens, np.array(fq, np.uint8), \
        np.array(rs, np.int32), np.array(diags, np.int32)


def _oracle_records(o, genome, mask, reads, quals, lens, fq, rs, diags, color, I, t, scores=None):
    op = OrcParams()
    o.orc_params_default(C.byref(op), color)
    if scores:
        op.match, op.mismatch, op.gap_open, op.gap_extend = scores
    op.allowed_indels = I
    op.simd_diags = min(I, 15)
    op.t = t
    d = min(I, 15)
    gp = np.zeros(GPAD + len(genome) + GPAD, np.uint8)   # padding: code 0, masked
    mp = np.ones_like(gp)
    gp[GPAD:GPAD + len(genome)] = genome
    mp[GPAD:GPAD + len(genome)] = mask
    strands = {}
    out = []
    for k in range(len(rs)):
        r, s = int(rs[k]) >> 1, int(rs[k]) & 1
        L = int(lens[r])
        if (r, s) not in strands:
            q = quals[r] if quals is not None else None
            if s == 0:
                strands[(r, s)] = (reads[r], q)
            else:
                strands[(r, s)] = _strand(o, reads[r], q, fq[r], L, color)
        codes, q = strands[(r, s)]
        qlev = np.ascontiguousarray(q if q is not None else np.full(L, 3, np.uint8), np.uint8)
        codes = np.ascontiguousarray(codes, np.uint8)
        h = int(diags[k])
        win = h - I if h - I > -d else -d
        wref = np.ascontiguousarray(gp[GPAD + win:GPAD + win + L + 2 * I + 1])
        wmask = np.ascontiguousarray(mp[GPAD + win:GPAD + win + L + 2 * I + 1])
        al = OrcAlignOut()
        sc = o.orc_align(codes.ctypes.data, qlev.ctypes.data, L, wref.ctypes.data, wmask.ctypes.data, I,
                         C.byref(op), t, C.byref(al))
        full, S = C.c_int(), C.c_int()
        o.orc_align_bound(codes.ctypes.data, qlev.ctypes.data, L, wref.ctypes.data, wmask.ctypes.data, I,
                          C.byref(op), C.byref(full), C.byref(S))
        out.append((sc, S.value, full.value, win, al.best_j, al.best_i0, al.best_j0,
                    bytes(al.tb[:min(al.tb_len, 1024)]), al.tb_len))
    return out


def _tb_codes(row, n):
    """Codes of a TRACEBACK_SEQ_SET packed traceback (alignment.h:81-83): even index, high nibble."""
    return bytes(((row[i >> 1] >> (0 if i & 1 else 4)) & 0xF) for i in range(n))


CASES = [(0, 0, 115, True), (0, 1, 40, True), (0, 3, 115, True), (0, 3, 60, False), (0, 7, 90, True),
         (0, 12, 80, True), (0, 15, 100, False), (0, 20, 70, True), (0, 31, 60, True),
         (1, 0, 115, True), (1, 3, 100, True), (1, 3, 50, False), (1, 7, 80, True), (1, 15, 70, True),
         (1, 31, 60, True)]


# (color, I, t, scores): score sets far from the reference's defaults (larger steps per cell, a
# color refund of twice a large mismatch, gap open equal to the match)
SCORE_CASES = [(0, 3, 200, (20, -12, -25, -6)), (1, 7, 150, (10, -12, -14, -5)), (0, 1, 60, (19, -3, -5, -2)),
               (1, 3, 90, (18, -9, -18, -18))]


@pytest.mark.parametrize("color,I,t,scores", SCORE_CASES, ids=lambda v: str(v))
def test_align_batch_scores_equal_oracle(color, I, t, scores, oracle_built):
    test_align_batch_equals_oracle(color, I, t, True, oracle_built, scores)


@pytest.mark.parametrize("color,I,t,with_qual", CASES, ids=lambda v: str(v))
def test_align_batch_equals_oracle(color, I, t, with_qual, oracle_built, scores=None):
    o = _oracle()
    seed = 1000 * color + 10 * I + (t % 7) + (0 if with_qual else 5) + (sum(scores) if scores else 0)
    genome, mask, reads, quals, lens, fq, rs, diags = _case(seed, color, I, t, with_qual)
    p = lib.params(color, allowed_indels=I, t=t)
    if scores:
        p.match, p.mismatch, p.gap_open, p.gap_extend = scores
    print(f"align batch color={color} -i {I} -t {t} qual={with_qual} scores={scores}: {len(rs)} records", flush=True)
    plan = lib.AlignPlan(genome, mask, reads, quals, lens, rs, diags, p, first_qual=fq if color else None)
    plan.run(1)
    res, tb = plan.fetch()
    plan.run(2)
    res2, tb2 = plan.fetch()
    plan.close()
    assert np.array_equal(res, res2)
    want = _oracle_records(o, genome, mask, reads, quals, lens, fq, rs, diags, color, I, t, scores)
    n_tb = n_exit = 0
    for k, (sc, S, full, win, bj, bi0, bj0, otb, otl) in enumerate(want):
        g = res[k]
        where = (k, int(rs[k]), int(diags[k]))
        assert g["win"] == win, where
        assert g["score"] == sc, (where, int(g["score"]), sc)
        assert g["best_j"] == bj, (where, int(g["best_j"]), bj)
        if S < t:
            # stopped at the first row whose bound fell below t; the full bound is lower still
            n_exit += 1
            assert sc == MIN_SCORE and S <= g["bound"] < t, (where, int(g["bound"]), S)
            assert g["tb_len"] == 0, where
            continue
        # no early stop: the full bound, and the unconditional score (MIN_SCORE when every cell of
        # the last row is masked)
        assert g["bound"] == S and full == sc, (where, int(g["bound"]), S)
        if sc > t:
            n_tb += 1
            assert (g["best_i0"], g["best_j0"], g["tb_len"]) == (bi0, bj0, otl), where
            assert _tb_codes(tb[k], otl) == otb, where
        else:
            assert (g["best_i0"], g["best_j0"], g["tb_len"]) == (-1, -1, 0), where
    print(f"  {n_tb} tracebacks, {n_exit} early exits", flush=True)
    assert n_tb > 0 and n_exit > 0


def test_align_batch_without_records():
    g = np.zeros(1000, np.uint8)
    plan = lib.AlignPlan(g, None, [np.zeros(40, np.uint8)], None, np.array([40], np.int32),
                         np.zeros(0, np.int32), np.zeros(0, np.int32), lib.params(0))
    plan.run(1)
    res, tb = plan.fetch()
    assert len(res) == 0 and tb.shape == (0, lib.TB_BYTES)
    plan.close()
