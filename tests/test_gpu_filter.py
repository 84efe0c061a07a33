"""gfx950 filter kernel vs the reference's per-candidate u8 scores (golden, from the SSE2 routine)
and vs the CPU restatement at the shape of the filter workload (L=75, -i 7: 16 diagonals)."""
import ctypes as C
import glob
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def _unpack(gp, pad, n):
    idx = np.arange(n) + pad * 4
    return ((gp[idx >> 2] >> ((idx & 3) * 2)) & 3).astype(np.uint8)


@pytest.mark.parametrize("fx", sorted(glob.glob(os.path.join(GOLDEN, "filter_*.npz"))), ids=os.path.basename)
def test_filter_matches_reference_scores(fx):
    from storm_amd import lib
    z = np.load(fx)
    gp, pad, glen = z["genome_packed"], int(z["pad"]), int(z["genome_len"])
    L, W, stride = int(z["L"]), int(z["W"]), int(z["stride"])
    genome = _unpack(gp, pad, glen)
    rp = z["reads_packed"]
    n_reads = int(z["read_of"].max()) + 1
    reads = []
    for r in range(n_reads):
        b = rp[r * stride: r * stride + (L + 3) // 4]
        idx = np.arange(L)
        reads.append(((b[idx >> 2] >> ((idx & 3) * 2)) & 3).astype(np.uint8))
    plan = lib.FilterPlan(genome, reads, np.full(n_reads, L), z["read_of"], z["pos"], W)
    plan.run(1)
    s, p = plan.fetch()
    plan.close()
    assert np.array_equal(s, z["scores"])
    assert np.array_equal(p, (z["scores"] > 100).astype(np.uint8))


def test_filter_vs_oracle_config_shape(oracle_built):
    from storm_amd import lib, synth
    ol = C.CDLL(os.path.join(oracle_built, "liboracle.so"))
    ol.orc_filter_score.restype = C.c_int
    ol.orc_filter_score.argtypes = [C.c_void_p, C.c_long, C.c_void_p, C.c_int, C.c_int] + [C.c_int] * 4
    g = synth.random_genome(200_000, 7)
    n_reads, L, W, d = 6000, 75, 8, 7
    reads, pos = synth.filter_candidates(g, n_reads, L, W, d, seed=8)
    plan = lib.FilterPlan(g, list(reads), np.full(n_reads, L), np.repeat(np.arange(n_reads), 2), pos.reshape(-1), W)
    plan.run(2)
    s, _ = plan.fetch()
    plan.close()
    pad = 64
    gp = np.zeros(pad + (len(g) + 3) // 4 + pad, np.uint8)
    gp[pad:pad + (len(g) + 3) // 4] = synth.pack_codes(g)
    stride = (L + 3) // 4 + 8
    rp = np.zeros(n_reads * stride + 64, np.uint8)
    for k in range(n_reads):
        rp[k * stride:k * stride + (L + 3) // 4] = synth.pack_codes(reads[k])
    sel = np.arange(0, 2 * n_reads, 7)   # every 7th candidate: both true-origin and decoy windows
    want = np.array([min(255, ol.orc_filter_score(C.c_void_p(gp.ctypes.data + pad), int(pos.reshape(-1)[c]),
                                                  C.c_void_p(rp.ctypes.data + int(c // 2) * stride), L, W, 5, -4, -9, -4))
                     for c in sel])
    assert np.array_equal(s[sel].astype(np.int64), want)


def test_filter_mixed_lengths_in_one_launch(oracle_built):
    """Candidate pairs whose reads have different lengths share a lane; each keeps its own rows."""
    from storm_amd import lib, synth
    ol = C.CDLL(os.path.join(oracle_built, "liboracle.so"))
    ol.orc_filter_score.restype = C.c_int
    ol.orc_filter_score.argtypes = [C.c_void_p, C.c_long, C.c_void_p, C.c_int, C.c_int] + [C.c_int] * 4
    g = synth.random_genome(50_000, 11)
    rng = synth.rng(12)
    n_reads, W, d = 400, 4, 3
    lens = rng.integers(35, 151, size=n_reads)
    reads, cr, cp = [], [], []
    for k in range(n_reads):
        o = int(rng.integers(10, len(g) - 200))
        reads.append(g[o:o + lens[k]].copy())
        cr += [k, k]
        cp += [o - d, int(rng.integers(0, len(g) - 200))]
    order = rng.permutation(len(cr))           # mixes lengths within lanes
    cr, cp = np.array(cr)[order], np.array(cp)[order]
    plan = lib.FilterPlan(g, reads, lens, cr, cp, W)
    plan.run(1)
    s, _ = plan.fetch()
    plan.close()
    pad = 64
    gp = np.zeros(pad + (len(g) + 3) // 4 + pad, np.uint8)
    gp[pad:pad + (len(g) + 3) // 4] = synth.pack_codes(g)
    want = []
    for c in range(len(cr)):
        rpk = np.zeros((lens[cr[c]] + 3) // 4 + 16, np.uint8)
        rpk[:(lens[cr[c]] + 3) // 4] = synth.pack_codes(reads[cr[c]])
        want.append(min(255, ol.orc_filter_score(C.c_void_p(gp.ctypes.data + pad), int(cp[c]), C.c_void_p(rpk.ctypes.data),
                                                 int(lens[cr[c]]), W, 5, -4, -9, -4)))
    assert np.array_equal(s.astype(np.int64), np.array(want))


@pytest.mark.parametrize("scores", [(5, -4, -9, -4), (4, -3, -7, -2), (3, -2, -5, -2), (2, -5, -3, -4)],
                         ids=["default", "open-ext!=m", "open-ext=m", "open<ext"])
def test_filter_score_sets_vs_oracle(oracle_built, scores):
    """Both forms of the filter cell against the CPU restatement: the short form taken when
    open - extend == match and open >= extend (the default scores, and 3/2/5/2), and the general
    form otherwise (4/3/7/2, and 2/5/3/4 with open < extend), every band class."""
    from storm_amd import lib, synth
    m, x, o, e = scores
    ol = C.CDLL(os.path.join(oracle_built, "liboracle.so"))
    ol.orc_filter_score.restype = C.c_int
    ol.orc_filter_score.argtypes = [C.c_void_p, C.c_long, C.c_void_p, C.c_int, C.c_int] + [C.c_int] * 4
    g = synth.random_genome(100_000, 21)
    pad = 64
    gp = np.zeros(pad + (len(g) + 3) // 4 + pad, np.uint8)
    gp[pad:pad + (len(g) + 3) // 4] = synth.pack_codes(g)
    for W in (2, 4, 8, 16):
        n_reads, L, d = 600, 60 + 3 * W, min(2 * W - 1, 15)
        reads, pos = synth.filter_candidates(g, n_reads, L, W, d, seed=22 + W)
        plan = lib.FilterPlan(g, list(reads), np.full(n_reads, L), np.repeat(np.arange(n_reads), 2), pos.reshape(-1), W,
                              match=m, mismatch=x, gap_open=o, gap_extend=e)
        plan.run(1)
        s, _ = plan.fetch()
        plan.close()
        stride = (L + 3) // 4 + 8
        rp = np.zeros(n_reads * stride + 64, np.uint8)
        for k in range(n_reads):
            rp[k * stride:k * stride + (L + 3) // 4] = synth.pack_codes(reads[k])
        sel = np.arange(0, 2 * n_reads, 3)
        want = np.array([min(255, ol.orc_filter_score(C.c_void_p(gp.ctypes.data + pad), int(pos.reshape(-1)[c]),
                                                      C.c_void_p(rp.ctypes.data + int(c // 2) * stride), L, W, m, x, o, e))
                         for c in sel])
        assert np.array_equal(s[sel].astype(np.int64), want), f"W={W}"
