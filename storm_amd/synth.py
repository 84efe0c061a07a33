"""Deterministic synthetic inputs for the mapping hot path (genomes, reads, filter candidates).

Everything is driven by numpy's PCG64 generator with an explicit seed, so the same call gives the
same bytes on every machine.  Codes: A=0 C=1 G=2 T=3 (reference codes.h:16-19); a color is the
XOR of two adjacent bases (codes.h:85-88).
"""
from __future__ import annotations

import os
from dataclasses import dataclass

import numpy as np

BASES = np.frombuffer(b"ACGT", dtype=np.uint8)
COLORS = np.frombuffer(b"0123", dtype=np.uint8)


def rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def random_genome(n: int, seed: int) -> np.ndarray:
    """i.i.d. uniform ACGT codes."""
    return rng(seed).integers(0, 4, size=n, dtype=np.uint8)


def gc_window_genome(n: int, seed: int, window: int = 10_000, gc_lo: float = 0.40, gc_hi: float = 0.60) -> np.ndarray:
    """GC fraction drawn uniformly in [gc_lo, gc_hi] per window of `window` bases."""
    g = rng(seed)
    nw = (n + window - 1) // window
    gc = g.uniform(gc_lo, gc_hi, size=nw)
    p_gc = np.repeat(gc, window)[:n]
    u = g.random(n)
    is_gc = u < p_gc
    low = g.integers(0, 2, size=n, dtype=np.uint8)          # choose within class
    # A/T = 0/3, C/G = 1/2
    codes = np.where(is_gc, 1 + low, np.where(low == 0, 0, 3)).astype(np.uint8)
    return codes


def repeat_genome(n: int, seed: int, seg_dup_frac: float = 0.10, tandem_frac: float = 0.05) -> np.ndarray:
    """i.i.d. background with segmental duplications (copies of 1-20 kb, 1-5 % divergence)
    and tandem repeats (unit 2-50 bp) covering the requested fractions."""
    g = rng(seed)
    codes = g.integers(0, 4, size=n, dtype=np.uint8)
    target = int(seg_dup_frac * n)
    done = 0
    while done < target:
        ln = int(g.integers(1000, 20001))
        if ln >= n // 2:
            break
        src = int(g.integers(0, n - ln))
        dst = int(g.integers(0, n - ln))
        seg = codes[src:src + ln].copy()
        div = g.uniform(0.01, 0.05)
        mut = g.random(ln) < div
        seg[mut] = (seg[mut] + g.integers(1, 4, size=int(mut.sum()), dtype=np.uint8)) & 3
        codes[dst:dst + ln] = seg
        done += ln
    target = int(tandem_frac * n)
    done = 0
    while done < target:
        unit = int(g.integers(2, 51))
        copies = int(g.integers(5, 60))
        ln = unit * copies
        dst = int(g.integers(0, n - ln))
        u = g.integers(0, 4, size=unit, dtype=np.uint8)
        codes[dst:dst + ln] = np.tile(u, copies)
        done += ln
    return codes


def write_fasta(path: str, name: str, codes: np.ndarray, width: int = 70) -> None:
    letters = BASES[codes]
    with open(path, "wb") as f:
        f.write(b">" + name.encode() + b"\n")
        for s in range(0, len(letters), width):
            f.write(letters[s:s + width].tobytes() + b"\n")


@dataclass
class SimReads:
    seqs: list          # list of np.uint8 code arrays (bases, or colors for color reads)
    quals: list         # list of np.int16 phred arrays
    origin: np.ndarray  # true origin (0-based genome position of the first base)
    strand: np.ndarray  # 0 forward, 1 reverse
    first_base: np.ndarray | None = None   # color reads: actual first base


def _mutate(seq: np.ndarray, g: np.random.Generator, sub: float, indel: float, L: int, src: np.ndarray, pos: int):
    """Apply substitutions and 1-3 bp indels to a genome slice, return exactly L codes."""
    out = []
    i = pos
    n_src = len(src)
    while len(out) < L and i < n_src:
        u = g.random()
        if u < indel / 2:          # insertion in the read
            k = int(g.integers(1, 4))
            out.extend(int(x) for x in g.integers(0, 4, size=k))
            continue
        if u < indel:              # deletion from the read
            i += int(g.integers(1, 4))
            continue
        c = int(src[i])
        if g.random() < sub:
            c = (c + int(g.integers(1, 4))) & 3
        out.append(c)
        i += 1
    while len(out) < L:
        out.append(int(g.integers(0, 4)))
    return np.array(out[:L], dtype=np.uint8)


def simulate_nucleotide_reads(genome: np.ndarray, n: int, L, seed: int, sub: float = 0.02, indel: float = 0.001,
                              qlo: int = 20, qhi: int = 40) -> SimReads:
    """Reads of length L (int, or (lo, hi) for uniform lengths) from uniform positions and strands."""
    g = rng(seed)
    G = len(genome)
    lens = np.full(n, L, dtype=np.int64) if np.isscalar(L) else g.integers(L[0], L[1] + 1, size=n)
    span = int(lens.max()) + 16
    origin = g.integers(0, G - span, size=n)
    strand = g.integers(0, 2, size=n)
    seqs, quals = [], []
    for k in range(n):
        Lk = int(lens[k])
        o = int(origin[k])
        sl = genome[o:o + Lk + 16]
        if indel > 0 and g.random() < Lk * indel * 1.5:
            r = _mutate(sl, g, sub, indel * 4, Lk, sl, 0)
        else:
            r = sl[:Lk].copy()
            m = g.random(Lk) < sub
            if m.any():
                r[m] = (r[m] + g.integers(1, 4, size=int(m.sum()), dtype=np.uint8)) & 3
        if strand[k]:
            r = (3 - r[::-1]).astype(np.uint8)
        seqs.append(r)
        quals.append(g.integers(qlo, qhi + 1, size=Lk).astype(np.int16))
    return SimReads(seqs, quals, origin, strand)


def write_fastq(path: str, reads: SimReads, qual_base: int = 33, prefix: str = "r") -> None:
    with open(path, "wb") as f:
        for k, (s, q) in enumerate(zip(reads.seqs, reads.quals)):
            f.write(b"@%s%d\n" % (prefix.encode(), k))
            f.write(BASES[s].tobytes() + b"\n+\n")
            f.write((q.astype(np.int64) + qual_base).astype(np.uint8).tobytes() + b"\n")


def simulate_color_reads(genome: np.ndarray, n: int, L, seed: int, snp: float = 0.01, color_err: float = 0.03,
                         qlo: int = 10, qhi: int = 35) -> SimReads:
    """SOLiD-like reads: L colors (int or (lo, hi)) covering L+1 bases; SNPs in base space,
    then color errors.  `seqs` holds the L colors after the first one, `first_base` the real
    first base (what a csfasta line 'T' + c0 + colors decodes to)."""
    g = rng(seed)
    G = len(genome)
    lens = np.full(n, L, dtype=np.int64) if np.isscalar(L) else g.integers(L[0], L[1] + 1, size=n)
    span = int(lens.max()) + 4
    origin = g.integers(0, G - span, size=n)
    strand = g.integers(0, 2, size=n)
    seqs, quals, firsts = [], [], []
    for k in range(n):
        Lk = int(lens[k])
        o = int(origin[k])
        b = genome[o:o + Lk + 1].copy()
        m = g.random(Lk + 1) < snp
        if m.any():
            b[m] = (b[m] + g.integers(1, 4, size=int(m.sum()), dtype=np.uint8)) & 3
        if strand[k]:
            b = (3 - b[::-1]).astype(np.uint8)
        col = (b[1:] ^ b[:-1]).astype(np.uint8)
        e = g.random(Lk) < color_err
        if e.any():
            col[e] = (col[e] ^ g.integers(1, 4, size=int(e.sum()), dtype=np.uint8)) & 3
        seqs.append(col)
        firsts.append(int(b[0]))
        quals.append(g.integers(qlo, qhi + 1, size=Lk + 1).astype(np.int16))   # primer color + L colors
    return SimReads(seqs, quals, origin, strand, np.array(firsts, dtype=np.uint8))


def write_csfasta(path: str, qual_path: str | None, reads: SimReads, prefix: str = "r") -> None:
    """'T' + first color (T xor first base) + L colors; the .qual holds L+1 integers."""
    T = 3
    with open(path, "wb") as f:
        for k, s in enumerate(reads.seqs):
            c0 = T ^ int(reads.first_base[k])
            f.write(b">%s%d\n" % (prefix.encode(), k))
            f.write(b"T" + bytes([48 + c0]) + COLORS[s].tobytes() + b"\n")
    if qual_path:
        with open(qual_path, "wb") as f:
            for k, q in enumerate(reads.quals):
                f.write(b">%s%d\n" % (prefix.encode(), k))
                f.write(" ".join(str(int(x)) for x in q).encode() + b"\n")


def pack_codes(codes: np.ndarray, nbytes: int | None = None) -> np.ndarray:
    """2-bit packing, code n at bits 2*(n&3) of byte n>>2 (codes.h:125-130)."""
    n = len(codes)
    nb = (n + 3) // 4 if nbytes is None else nbytes
    out = np.zeros(nb, dtype=np.uint8)
    c = np.zeros(((n + 3) // 4) * 4, dtype=np.uint8)
    c[:n] = codes
    c = c.reshape(-1, 4)
    packed = c[:, 0] | (c[:, 1] << 2) | (c[:, 2] << 4) | (c[:, 3] << 6)
    out[:len(packed)] = packed[:nb]
    return out


def filter_candidates(genome: np.ndarray, n_reads: int, L: int, W: int, d: int, seed: int, sub: float = 0.02):
    """Filter workload: each read gets one true-origin window (origin - d) and one uniform random
    decoy window.  Returns (reads codes [n, L], window starts [n, 2])."""
    g = rng(seed)
    G = len(genome)
    origin = g.integers(d + 1, G - (L + 2 * W + 2), size=n_reads)
    idx = origin[:, None] + np.arange(L)[None, :]
    reads = genome[idx]
    m = g.random(reads.shape) < sub
    reads = np.where(m, (reads + g.integers(1, 4, size=reads.shape, dtype=np.uint8)) & 3, reads).astype(np.uint8)
    decoy = g.integers(0, G - (L + 2 * W + 2), size=n_reads)
    pos = np.stack([origin - d, decoy], axis=1).astype(np.int32)
    return reads, pos


def ensure_dir(p: str) -> str:
    os.makedirs(p, exist_ok=True)
    return p


def quality_levels(q: np.ndarray) -> np.ndarray:
    """read_quality.c:10-16: q <= 2 -> 0, <= 7 -> 1, <= 10 -> 2, else 3."""
    return np.where(q <= 2, 0, np.where(q <= 7, 1, np.where(q <= 10, 2, 3))).astype(np.uint8)


def pack_rows(codes: np.ndarray, nbytes: int) -> np.ndarray:
    """Row-wise 2-bit packing of an (n, Lmax) code matrix into (n, nbytes)."""
    n, L = codes.shape
    Lp = ((L + 3) // 4) * 4
    c = np.zeros((n, Lp), np.uint8)
    c[:, :L] = codes
    c = c.reshape(n, Lp // 4, 4)
    packed = c[:, :, 0] | (c[:, :, 1] << 2) | (c[:, :, 2] << 4) | (c[:, :, 3] << 6)
    out = np.zeros((n, nbytes), np.uint8)
    out[:, :min(nbytes, packed.shape[1])] = packed[:, :nbytes]
    return out


@dataclass
class PackedReads:
    """Reads packed for storm_reads_view: codes/levels 2-bit, one row per read."""
    codes: np.ndarray       # (n, stride) uint8
    qual: np.ndarray        # (n, stride) uint8 levels
    len: np.ndarray         # (n,) int32
    first_base: np.ndarray | None = None
    first_qual: np.ndarray | None = None
    origin: np.ndarray | None = None
    strand: np.ndarray | None = None


def _chunk_rngs(seed: int, n_chunks: int):
    return [np.random.Generator(np.random.PCG64(ss)) for ss in np.random.SeedSequence(seed).spawn(n_chunks)]


def _reverse_rows(x: np.ndarray, lens: np.ndarray, extra: int, complement: bool) -> np.ndarray:
    """Row-wise reversal of the first lens[i] + extra codes (complemented when asked)."""
    out = np.zeros_like(x)
    for Lv in np.unique(lens):
        sel = lens == Lv
        seg = x[sel, :Lv + extra][:, ::-1]
        out[sel, :Lv + extra] = (3 - seg) if complement else seg
    return out


def _run_chunks(fn, n: int, chunk: int, seed: int, workers: int):
    from concurrent.futures import ThreadPoolExecutor
    starts = list(range(0, n, chunk))
    gens = _chunk_rngs(seed, len(starts))
    with ThreadPoolExecutor(max_workers=max(1, workers)) as ex:
        list(ex.map(lambda k: fn(starts[k], min(n, starts[k] + chunk), gens[k]), range(len(starts))))


def fast_nucleotide_reads(genome: np.ndarray, n: int, L, seed: int, sub: float, indel: float, qlo: int, qhi: int,
                          chunk: int = 250_000, workers: int = 8) -> PackedReads:
    """Vectorised Illumina-like reads: uniform position and strand, substitutions, and for a
    fraction n*L*indel of the reads one 1-3 bp insertion or deletion at a uniform offset.
    Chunks are generated by a thread pool, each chunk from its own spawned generator."""
    g = rng(seed)
    G = len(genome)
    lens = np.full(n, L, np.int32) if np.isscalar(L) else g.integers(L[0], L[1] + 1, size=n).astype(np.int32)
    Lmax = int(lens.max())
    stride = ((Lmax + 3) // 4 + 7) & ~7
    codes = np.zeros((n, stride), np.uint8)
    qual = np.zeros((n, stride), np.uint8)
    origin = g.integers(0, G - Lmax - 8, size=n)
    strand = g.integers(0, 2, size=n).astype(np.uint8)
    col = np.arange(Lmax)[None, :]

    def one(s, e, g):
        m = e - s
        r = genome[origin[s:e, None] + np.arange(Lmax + 4)[None, :]]
        has = np.nonzero(g.random(m) < (lens[s:e] * indel))[0]
        if len(has):
            Lk = lens[s:e][has].astype(np.int64)
            off = g.integers(1, Lk - 1)[:, None]
            ln = g.integers(1, 4, size=len(has))[:, None]
            ins = (g.random(len(has)) < 0.5)[:, None]
            src = np.where(col < off, col, np.where(ins, col - ln, col + ln))
            rows = np.take_along_axis(r[has], np.clip(src, 0, Lmax + 3), axis=1)
            fill = ins & (col >= off) & (col < off + ln)
            rows = np.where(fill, g.integers(0, 4, size=rows.shape, dtype=np.uint8), rows)
            r[has, :Lmax] = rows
        r = r[:, :Lmax]
        mut = g.random(r.shape, dtype=np.float32) < sub
        r = np.where(mut, (r + g.integers(1, 4, size=r.shape, dtype=np.uint8)) & 3, r).astype(np.uint8)
        rev = strand[s:e] == 1
        if rev.any():
            r[rev] = _reverse_rows(r[rev], lens[s:e][rev], 0, True)
        r = np.where(col < lens[s:e, None], r, 0).astype(np.uint8)
        codes[s:e] = pack_rows(r, stride)
        if qlo > 10:   # every quality maps to level 3 (read_quality.c:10-16)
            lv = np.full((m, Lmax), 3, np.uint8)
        else:
            lv = quality_levels(g.integers(qlo, qhi + 1, size=(m, Lmax)))
        qual[s:e] = pack_rows(np.where(col < lens[s:e, None], lv, 0).astype(np.uint8), stride)

    _run_chunks(one, n, chunk, seed, workers)
    return PackedReads(codes, qual, lens, None, None, origin, strand)


def fast_color_reads(genome: np.ndarray, n: int, L, seed: int, snp: float, color_err: float, qlo: int, qhi: int,
                     chunk: int = 250_000, workers: int = 8) -> PackedReads:
    """Vectorised SOLiD-like reads: L colors covering L+1 bases; SNPs in base space, then color
    errors; qualities uniform in [qlo, qhi] for the primer-side color and the L colors."""
    g = rng(seed)
    G = len(genome)
    lens = np.full(n, L, np.int32) if np.isscalar(L) else g.integers(L[0], L[1] + 1, size=n).astype(np.int32)
    Lmax = int(lens.max())
    stride = ((Lmax + 3) // 4 + 7) & ~7
    codes = np.zeros((n, stride), np.uint8)
    qual = np.zeros((n, stride), np.uint8)
    first_base = np.zeros(n, np.uint8)
    first_qual = np.zeros(n, np.uint8)
    origin = g.integers(0, G - Lmax - 4, size=n)
    strand = g.integers(0, 2, size=n).astype(np.uint8)
    col = np.arange(Lmax)[None, :]

    def one(s, e, g):
        m = e - s
        b = genome[origin[s:e, None] + np.arange(Lmax + 1)[None, :]]
        mut = g.random(b.shape, dtype=np.float32) < snp
        b = np.where(mut, (b + g.integers(1, 4, size=b.shape, dtype=np.uint8)) & 3, b).astype(np.uint8)
        rev = strand[s:e] == 1
        if rev.any():
            b[rev] = _reverse_rows(b[rev], lens[s:e][rev], 1, True)
        c = (b[:, 1:] ^ b[:, :-1]).astype(np.uint8)
        err = g.random(c.shape, dtype=np.float32) < color_err
        c = np.where(err, c ^ g.integers(1, 4, size=c.shape, dtype=np.uint8), c).astype(np.uint8)
        c = np.where(col < lens[s:e, None], c, 0).astype(np.uint8)
        codes[s:e] = pack_rows(c, stride)
        first_base[s:e] = b[:, 0]
        lv = quality_levels(g.integers(qlo, qhi + 1, size=(m, Lmax + 1)))
        first_qual[s:e] = lv[:, 0]
        qual[s:e] = pack_rows(np.where(col < lens[s:e, None], lv[:, 1:], 0).astype(np.uint8), stride)

    _run_chunks(one, n, chunk, seed, workers)
    return PackedReads(codes, qual, lens, first_base, first_qual, origin, strand)


def write_fasta_fast(path: str, name: str, codes: np.ndarray, width: int = 60) -> None:
    letters = BASES[codes]
    n = len(letters)
    full = n // width
    with open(path, "wb") as f:
        f.write(b">" + name.encode() + b"\n")
        if full:
            block = np.concatenate([letters[:full * width].reshape(full, width),
                                    np.full((full, 1), ord("\n"), np.uint8)], axis=1)
            f.write(block.tobytes())
        if n % width:
            f.write(letters[full * width:].tobytes() + b"\n")


def unpack_rows(packed: np.ndarray, L: int) -> np.ndarray:
    """Inverse of pack_rows: (n, nbytes) -> (n, L) codes."""
    p = packed[:, :(L + 3) // 4]
    out = np.stack([p & 3, (p >> 2) & 3, (p >> 4) & 3, (p >> 6) & 3], axis=2).reshape(len(p), -1)
    return out[:, :L].astype(np.uint8)


LEVEL_QUALITY = np.array([2, 7, 10, 40], np.int16)   # one quality value per level (read_quality.c:10-16)


def packed_to_simreads(pr: PackedReads, idx) -> SimReads:
    """The reads `idx` of a PackedReads as SimReads (for the FASTQ / csfasta writers); each
    quality is written as a value of its level, which is all the mapping path reads of it."""
    idx = np.asarray(idx)
    Lmax = int(pr.len[idx].max())
    codes = unpack_rows(pr.codes[idx], Lmax)
    lv = unpack_rows(pr.qual[idx], Lmax)
    seqs, quals = [], []
    for k, i in enumerate(idx):
        L = int(pr.len[i])
        seqs.append(codes[k, :L].copy())
        q = LEVEL_QUALITY[lv[k, :L]]
        if pr.first_qual is not None:
            q = np.concatenate([[LEVEL_QUALITY[pr.first_qual[i]]], q]).astype(np.int16)
        quals.append(q)
    fb = None if pr.first_base is None else pr.first_base[idx].copy()
    return SimReads(seqs, quals, pr.origin[idx], pr.strand[idx], fb)
