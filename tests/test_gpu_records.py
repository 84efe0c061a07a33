"""The whole hit map in host memory (storm_mapper_map_records): the plain (staged) and batched
(fused) paths return identical records and tracebacks, the records agree with
storm_mapper_map_hits, and every traceback is a complete glocal path: its read-consuming codes
(match, mismatch, color fix, insertion; alignment.h:81-83) cover the read from read_start to its
last code, and its genome-consuming codes the span ref_start..ref_end (hit_map.h:17-40)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEEDS_N = "####--#-#--##--####;#####--###-#-###;###-##-#######"
SEEDS_C = "####----##--##-####;####-###--#----####;####-####-####"


def _codes(tb_row, n):
    return [(tb_row[i >> 1] >> (0 if i & 1 else 4)) & 0xF for i in range(n)]


@pytest.mark.parametrize("color,I,M", [(0, 7, 4), (0, 3, 1), (1, 3, 2), (1, 7, 4)], ids=lambda v: str(v))
def test_records_plain_equal_batched(color, I, M, tmp_path):
    from storm_amd import lib, synth
    g = synth.gc_window_genome(1_000_000, 41 + I)
    gpath = str(tmp_path / "g.fa")
    synth.write_fasta_fast(gpath, "chr1", g)
    if color:
        pr = synth.fast_color_reads(g, 3000, (35, 149), 7 + I, snp=0.01, color_err=0.03, qlo=10, qhi=35)
    else:
        pr = synth.fast_nucleotide_reads(g, 3000, (36, 150), 7 + I, sub=0.015, indel=0.002, qlo=15, qhi=40)
    m = lib.Mapper(gpath, color=color, allowed_indels=I, max_hits=M, z=85 if color else 100, t=100,
                   seeds=SEEDS_C if color else SEEDS_N)
    print(f"records color={color} -i {I} -M {M}", flush=True)
    hp, tp, _ = m.map_records(pr, mode=lib.STORM_MODE_PLAIN)
    hb, tb, _ = m.map_records(pr, mode=lib.STORM_MODE_BATCHED)
    hh, _ = m.map_hits(pr, mode=lib.STORM_MODE_BATCHED)
    m.close()
    assert np.array_equal(hp, hb)
    assert np.array_equal(tp, tb)   # traceback bytes past tb_len are zero on both paths
    for f in ("score", "ref_id", "ref_start", "sense"):
        assert np.array_equal(hb[f], hh[f]), f
    n_tb = 0
    for r in range(len(pr.len)):
        for k in range(M):
            h = hb[r, k]
            if h["ref_id"] < 0:
                continue
            n = int(h["tb_len"])
            nb = (n + 1) // 2
            assert np.array_equal(tp[r, k, :nb], tb[r, k, :nb]), (r, k)
            c = _codes(tb[r, k], n)
            read_used = sum(1 for x in c if x == 0x3 or (x & 0xC) in (0x4, 0x8, 0xC))
            ref_used = sum(1 for x in c if x == 0x1 or x == 0x3 or (x & 0xC) in (0x4, 0x8))
            assert read_used == int(h["read_end"]) - int(h["read_start"]) + 1, (r, k)
            assert ref_used == int(h["ref_end"]) - int(h["ref_start"]) + 1, (r, k)
            n_tb += 1
    assert n_tb > 0.5 * len(pr.len)


@pytest.mark.parametrize("color,chunks", [(0, "3"), (1, "5"), (1, "")], ids=["nucleotide-3", "color-5", "color-model"])
def test_chunked_records_equal_plain(color, chunks, tmp_path, monkeypatch):
    """A batched records call can run in chunks of consecutive reads (results of chunk k - 1 copied
    out while chunk k runs, each chunk ordered by length on its own; STORM_RECORD_CHUNKS fixes the
    count, unset the library's cost model picks it): its records equal the plain path's with and
    without tracebacks, and a smaller call into the same arrays (not a multiple of the chunk count)
    rewrites its rows and leaves the rows past it untouched."""
    from storm_amd import lib, synth
    if chunks:
        monkeypatch.setenv("STORM_RECORD_CHUNKS", chunks)
    else:
        monkeypatch.delenv("STORM_RECORD_CHUNKS", raising=False)
    g = synth.gc_window_genome(2_000_000, 77 + color)
    gpath = str(tmp_path / "g.fa")
    synth.write_fasta_fast(gpath, "chr1", g)
    n = 150_000
    if color:
        pr = synth.fast_color_reads(g, n, (34, 74), 91, snp=0.01, color_err=0.03, qlo=10, qhi=35)
    else:
        pr = synth.fast_nucleotide_reads(g, n, (36, 120), 91, sub=0.015, indel=0.002, qlo=15, qhi=40)
    m = lib.Mapper(gpath, color=color, allowed_indels=3, max_hits=1, z=85 if color else 100, t=100,
                   seeds=SEEDS_C if color else SEEDS_N)
    print(f"chunked records color={color} chunks={chunks or 'model'}: {n} reads", flush=True)
    hp, tp, sp = m.map_records(pr, mode=lib.STORM_MODE_PLAIN)
    hb, tb, sb = m.map_records(pr, mode=lib.STORM_MODE_BATCHED)
    hn, tn, _ = m.map_records(pr, mode=lib.STORM_MODE_BATCHED, with_tb=False)
    assert tn is None
    assert np.array_equal(hp, hb)
    assert np.array_equal(tp, tb)
    assert np.array_equal(hp, hn)
    assert sb["n_mapped"] == sp["n_mapped"] > 0.9 * n
    # a smaller call into the same arrays: rows [0, k) rewritten, the rest as the first call left them
    k = 99_991
    hb2, tb2 = hb.copy(), tb.copy()
    hb2[:k] = 0
    tb2[:k] = 0
    hb3, tb3, _ = m.map_records(pr, mode=lib.STORM_MODE_BATCHED, lo=0, hi=k, out=(hb2, tb2))
    m.close()
    assert np.array_equal(hb2[:k], hp[:k]) and np.array_equal(tb2[:k], tp[:k])
    assert np.array_equal(hb2[k:], hp[k:]) and np.array_equal(tb2[k:], tp[k:])


def test_length_order_of_large_views(tmp_path, monkeypatch):
    """Views of 300k unsorted mixed-length reads are ordered on the host by several threads (a
    counting sort in parts); the records of both paths, one launch and three chunks, equal those of
    the same reads handed over already sorted longest first (no reordering at all), put back in
    the view's order."""
    from storm_amd import lib, synth
    g = synth.gc_window_genome(2_000_000, 5)
    gpath = str(tmp_path / "g.fa")
    synth.write_fasta_fast(gpath, "chr1", g)
    n = 300_000
    pr = synth.fast_nucleotide_reads(g, n, (36, 100), 17, sub=0.015, indel=0.002, qlo=15, qhi=40)
    perm = np.argsort(-pr.len, kind="stable")
    srt = synth.PackedReads(pr.codes[perm], pr.qual[perm], pr.len[perm])
    m = lib.Mapper(gpath, color=0, allowed_indels=3, max_hits=2, seeds=SEEDS_N)
    print(f"length order: {n} reads", flush=True)
    hs, ts, _ = m.map_records(srt, mode=lib.STORM_MODE_PLAIN)
    want_h = np.empty_like(hs)
    want_t = np.empty_like(ts)
    want_h[perm] = hs
    want_t[perm] = ts
    for mode, chunks in ((lib.STORM_MODE_PLAIN, "1"), (lib.STORM_MODE_BATCHED, "1"), (lib.STORM_MODE_BATCHED, "3")):
        monkeypatch.setenv("STORM_RECORD_CHUNKS", chunks)
        h, t, st = m.map_records(pr, mode=mode)
        assert np.array_equal(h, want_h), (mode, chunks)
        assert np.array_equal(t, want_t), (mode, chunks)
        assert st["n_mapped"] == int((want_h[:, 0]["ref_id"] != -1).sum())
    m.close()


def test_compact_traceback_copy_equals_full_records(tmp_path, monkeypatch):
    """Traceback records leave the device in the call's compact width (the longest read's codes
    plus the indel cap) and are widened to STORM_TB_BYTES on the host: the caller's arrays are the
    same as with whole records copied (STORM_TB_FULL_WIDTH), on both paths, in one launch and in
    chunks, and every byte past a traceback's codes is zero."""
    from storm_amd import lib, synth
    g = synth.gc_window_genome(1_000_000, 23)
    gpath = str(tmp_path / "g.fa")
    synth.write_fasta_fast(gpath, "chr1", g)
    pr = synth.fast_nucleotide_reads(g, 40_000, (40, 90), 29, sub=0.015, indel=0.004, qlo=15, qhi=40)
    m = lib.Mapper(gpath, color=0, allowed_indels=7, max_hits=3, seeds=SEEDS_N)
    got = {}
    for full in (False, True):
        if full:
            monkeypatch.setenv("STORM_TB_FULL_WIDTH", "1")
        else:
            monkeypatch.delenv("STORM_TB_FULL_WIDTH", raising=False)
        for mode, chunks in ((lib.STORM_MODE_PLAIN, "1"), (lib.STORM_MODE_BATCHED, "1"), (lib.STORM_MODE_BATCHED, "4")):
            monkeypatch.setenv("STORM_RECORD_CHUNKS", chunks)
            # arrays filled with 0xff first, so every byte the call leaves is seen
            h = np.full((len(pr.len), 4), -1, lib.HIT_RECORD)
            t = np.full((len(pr.len), 3, lib.TB_BYTES), 0xff, np.uint8)
            m.map_records(pr, mode=mode, out=(h, t))
            got[(full, mode, chunks)] = (h, t)
    m.close()
    ref_h, ref_t = got[(True, lib.STORM_MODE_PLAIN, "1")]
    for k, (h, t) in got.items():
        assert np.array_equal(h, ref_h), k
        assert np.array_equal(t, ref_t), k
    n_tb = 0
    for r in range(len(pr.len)):
        for k in range(3):
            if ref_h[r, k]["ref_id"] < 0:
                assert not ref_t[r, k].any()
                continue
            nb = (int(ref_h[r, k]["tb_len"]) + 1) // 2
            assert not ref_t[r, k, nb:].any(), (r, k)
            n_tb += 1
    assert n_tb > 0.9 * len(pr.len)
