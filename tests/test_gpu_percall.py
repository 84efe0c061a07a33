"""The drop-in rows of the reference's SIMD filter table (alignment_gfx950__init_<w> /
__align_<w>, same signatures as alignment_simd.h:32-300, one GPU launch per call of simd_mul
windows of one read, run.c:109,660) against the reference SSE2 routine's per-candidate u8 scores
(golden fixtures): bit k of a call's result is score(window k) > min(threshold, 254)."""
import ctypes as C
import glob
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

ROW = {2: ("octa", 8), 4: ("quad", 4), 8: ("pair", 2), 16: ("mono", 1)}


@pytest.mark.parametrize("fx", sorted(glob.glob(os.path.join(GOLDEN, "filter_*_L75.npz")) +
                                      glob.glob(os.path.join(GOLDEN, "filter_*_L35.npz"))), ids=os.path.basename)
def test_percall_rows_match_reference_scores(fx):
    from storm_amd import lib
    so = lib.load()
    z = np.load(fx)
    gp, pad = np.ascontiguousarray(z["genome_packed"]), int(z["pad"])
    L, W, stride = int(z["L"]), int(z["W"]), int(z["stride"])
    rp = np.ascontiguousarray(z["reads_packed"])
    pos, read_of, scores = z["pos"], z["read_of"], z["scores"].astype(np.int64)
    name, mul = ROW[W]
    P = C.c_void_p
    lib.check(so.alignment_gfx950__set_genome(P(gp.ctypes.data), gp.nbytes, pad))
    init = getattr(so, f"alignment_gfx950__init_{name}")
    align = getattr(so, f"alignment_gfx950__align_{name}")
    n_reads = int(read_of.max()) + 1
    checked = 0
    for thr in (60, 100, 300):
        init(5, 4, 9, 4, thr, L)
        for r in range(0, n_reads, 3):
            cands = np.nonzero(read_of == r)[0]
            sel = np.resize(cands, mul)            # simd_mul windows of this read, repeated to fill
            ps = np.ascontiguousarray(pos[sel], dtype=np.int32)
            bits = align(P(gp.ctypes.data + pad), P(ps.ctypes.data), P(rp.ctypes.data + r * stride))
            want = sum(1 << k for k in range(mul) if scores[sel[k]] > min(thr, 254))
            assert bits == want, (fx, thr, r, bits, want)
            checked += mul
    so.alignment_gfx950__clean()
    assert checked > 100


def test_percall_rows_from_many_threads():
    """The rows are reentrant like the reference's (run.c:900 calls them inside an OpenMP loop):
    eight host threads call align_pair at once, every result against the reference's scores."""
    import threading
    from storm_amd import lib
    so = lib.load()
    z = np.load(os.path.join(GOLDEN, "filter_W8_L75.npz"))
    gp, pad = np.ascontiguousarray(z["genome_packed"]), int(z["pad"])
    L, stride = int(z["L"]), int(z["stride"])
    rp = np.ascontiguousarray(z["reads_packed"])
    pos, read_of, scores = z["pos"], z["read_of"], z["scores"].astype(np.int64)
    P = C.c_void_p
    lib.check(so.alignment_gfx950__set_genome(P(gp.ctypes.data), gp.nbytes, pad))
    so.alignment_gfx950__init_pair(5, 4, 9, 4, 100, L)
    n_reads = int(read_of.max()) + 1
    errors, counts = [], [0] * 8

    def worker(t):
        for rep in range(4):
            for r in range(t, n_reads, 8):
                sel = np.resize(np.nonzero(read_of == r)[0], 2)
                ps = np.ascontiguousarray(pos[sel], dtype=np.int32)
                bits = so.alignment_gfx950__align_pair(P(gp.ctypes.data + pad), P(ps.ctypes.data),
                                                       P(rp.ctypes.data + r * stride))
                want = sum(1 << k for k in range(2) if scores[sel[k]] > 100)
                if bits != want or so.alignment_gfx950__status() != 0:
                    errors.append((t, r, bits, want))
                counts[t] += 1

    th = [threading.Thread(target=worker, args=(t,)) for t in range(8)]
    for x in th:
        x.start()
    for x in th:
        x.join()
    so.alignment_gfx950__clean()
    assert not errors, errors[:5]
    assert sum(counts) >= 4 * n_reads


def test_percall_row_reports_errors():
    """A call before any genome is registered returns 0 and sets the thread's status."""
    from storm_amd import lib
    so = lib.load()
    so.alignment_gfx950__clean()
    so.alignment_gfx950__init_pair(5, 4, 9, 4, 100, 75)
    ps = np.zeros(2, np.int32)
    rd = np.zeros(32, np.uint8)
    assert so.alignment_gfx950__align_pair(C.c_void_p(1024), C.c_void_p(ps.ctypes.data), C.c_void_p(rd.ctypes.data)) == 0
    assert so.alignment_gfx950__status() == -1   # STORM_ERR_INVALID
