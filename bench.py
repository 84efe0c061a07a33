"""Benchmark of the SToRM mapping hot path on one MI355X.

Default workload (one "step" = one launch of the banded filter over the whole batch):
  10M candidate pairs = 5M reads of 75 nt x 2 windows (true origin - 7, uniform random decoy) on a
  1 Mbp i.i.d. genome, -i 7 (W = 8 lanes, 16 diagonals per the reference's geometry), saturating
  u8 scores, z = 100.  Inputs are resident in HBM before the timed region.
Reported: filter GCUPS (cells = candidates x L x 2W), the filter through the per-call drop-in
entry point (one launch per read, as run.c calls the SIMD table) on a sample, whole -M mapping
runs in reads/s (plain per-batch and batched), and the reference's OpenMP SIMD filter timed on
this host's cores as the CPU baseline.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-map] [--no-cpu]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "mapped reads/s + filter GCUPS, batched vs plain GPU; p50/p99 request latency"
OPS_PER_CELL = 13            # compute chain of the reference's SIMD cell (alignment_simd.c:3810-3880)
VALU_PEAK_INT32 = 39.3       # T VALU lane-instructions/s: 256 CU x 64 lanes x 2.4 GHz
VALU_PEAK_PK16 = 78.6        # T 16-bit ops/s with packed VALU instructions (2 per lane-instruction)
# VALU lane-instructions k_filter<8> issues per cell at L=75 (SQ_INSTS_VALU x 64 / cells, rocprofv3 --pmc,
# profiles/k_filter_W8_L75_10Mcand.md); the kernel's own work measure for the VALU-issue roofline
VALU_INSTR_PER_CELL_W8_L75 = 5.438
# VALU lane-instructions k_align<7,false> issues per DP cell (L x 15 band columns) on the DP stage
# workload (score pass of every record + traceback pass of the accepted ones), measured the same way
# (profiles/k_align_I7_L75_2M_records.md)
VALU_INSTR_PER_DP_CELL_I7_L75 = 76.06
# HBM bytes per candidate of the same dispatch: (FETCH_SIZE x 2.00 + WRITE_SIZE x 1.00) x 1024 /
# candidates, the factors calibrated on a device-to-device copy of known size in the same run
# (profiles/k_filter_W8_L75_10Mcand.md)
HBM_BYTES_PER_CAND_W8_L75 = 45.6
# algorithmic bytes per candidate: packed read (19 B at L = 75) + genome window (L + 2W - 1 = 90
# codes, 23 B) + the u8 score and the pass byte
ALG_BYTES_PER_CAND_W8_L75 = 44.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_filter_workload(n_reads: int, L: int, W: int, d: int, seed: int = 2, genome_len: int = 1_000_000):
    """Packed reads (n_reads x stride bytes) and 2 window starts per read, built in chunks."""
    from storm_amd import synth
    g = synth.random_genome(genome_len, 1)
    rng = synth.rng(seed)
    stride = (L + 3) // 4 + 8
    reads = np.zeros((n_reads, stride), np.uint8)
    pos = np.zeros((n_reads, 2), np.int32)
    ch = 250_000
    Lp = ((L + 3) // 4) * 4
    for s in range(0, n_reads, ch):
        n = min(ch, n_reads - s)
        origin = rng.integers(d + 1, genome_len - (L + 2 * W + 2), size=n)
        idx = origin[:, None] + np.arange(L)[None, :]
        r = g[idx]
        m = rng.random(r.shape) < 0.02
        r = np.where(m, (r + rng.integers(1, 4, size=r.shape, dtype=np.uint8)) & 3, r).astype(np.uint8)
        rr = np.zeros((n, Lp), np.uint8)
        rr[:, :L] = r
        rr = rr.reshape(n, Lp // 4, 4)
        reads[s:s + n, :Lp // 4] = rr[:, :, 0] | (rr[:, :, 1] << 2) | (rr[:, :, 2] << 4) | (rr[:, :, 3] << 6)
        pos[s:s + n, 0] = origin - d
        pos[s:s + n, 1] = rng.integers(0, genome_len - (L + 2 * W + 2), size=n)
    return g, reads, pos, stride


def gpu_filter(lib, g, reads, pos, stride, L, W, steps, warmup):
    from storm_amd import synth
    import torch
    P = C.c_void_p
    gp = synth.pack_codes(g)
    n_reads = reads.shape[0]
    rl = np.full(n_reads, L, np.int32)
    cr = np.repeat(np.arange(n_reads, dtype=np.int32), 2)
    cp = np.ascontiguousarray(pos.reshape(-1))
    h = P()
    lib.check(lib.load().storm_filter_plan_create(0, P(gp.ctypes.data), len(g), P(reads.ctypes.data), stride,
                                                  P(rl.ctypes.data), n_reads, P(cr.ctypes.data), P(cp.ctypes.data),
                                                  len(cr), W, 5, -4, -9, -4, 100, C.byref(h)))
    so = lib.load()
    ms = C.c_double()
    lib.check(so.storm_filter_run(h, warmup, C.byref(ms)))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    lib.check(so.storm_filter_run(h, steps, C.byref(ms)))
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / steps
    scores = np.zeros(len(cr), np.uint8)
    passf = np.zeros(len(cr), np.uint8)
    lib.check(so.storm_filter_fetch(h, P(scores.ctypes.data), P(passf.ctypes.data)))
    so.storm_filter_plan_destroy(h)
    return ms.value, wall, scores, passf, cr, cp


def percall_filter(lib, g, reads, pos, stride, L, W, n_calls):
    """The drop-in row: alignment_gfx950__init_pair + __align_pair, one call per read (run.c:660)."""
    from storm_amd import synth
    P = C.c_void_p
    so = lib.load()
    pad = 1024
    gp = np.zeros(pad + (len(g) + 3) // 4 + pad, np.uint8)
    gp[pad:pad + (len(g) + 3) // 4] = synth.pack_codes(g)
    lib.check(so.alignment_gfx950__set_genome(P(gp.ctypes.data), gp.nbytes, pad))
    so.alignment_gfx950__init_pair(5, 4, 9, 4, 100, L)
    bits = np.zeros(n_calls, np.uint32)
    t0 = time.perf_counter()
    for k in range(n_calls):
        win = np.ascontiguousarray(pos[k])
        bits[k] = so.alignment_gfx950__align_pair(P(gp.ctypes.data + pad), P(win.ctypes.data), P(reads[k].ctypes.data))
    dt = time.perf_counter() - t0
    so.alignment_gfx950__clean()
    return dt, bits


def cpu_baseline(g, reads, pos, stride, L, d, n_sample, target_s=10.0):
    """The reference's SSE2 filter (its default Makefile build) with OpenMP on this host."""
    from storm_amd import synth
    path = os.path.join(ROOT, "oracle", "_ref", "libref_filter.so")
    if not os.path.exists(path):
        return None
    lib = C.CDLL(path)
    P = C.c_void_p
    lib.ref_filter_time.restype = C.c_double
    lib.ref_filter_time.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, P, C.c_long, P, P, P, C.c_long, P, C.c_int,
                                    C.c_int]
    cores = min(16, os.cpu_count() or 1)
    pad = 1024
    gp = np.zeros(pad + (len(g) + 3) // 4 + pad, np.uint8)
    gp[pad:pad + (len(g) + 3) // 4] = synth.pack_codes(g)
    n = min(n_sample, reads.shape[0])
    rd = np.ascontiguousarray(reads[:n])
    seg = np.arange(0, 2 * n + 1, 2, dtype=np.int64)
    ps = np.ascontiguousarray(pos[:n].reshape(-1))
    bits = np.zeros(2 * n, np.uint8)
    out = {}
    for isa, name in ((0, "sse2"), (1, "avx2")):
        t = lib.ref_filter_time(isa, d, L, 100, P(gp.ctypes.data + pad), n, P(seg.ctypes.data), P(ps.ctypes.data),
                                P(rd.ctypes.data), stride, P(bits.ctypes.data), cores, 1)
        if t <= 0:
            continue
        reps = max(1, int(target_s / max(t, 1e-3)))
        if name == "avx2":
            reps = max(1, reps // 2)
        t = lib.ref_filter_time(isa, d, L, 100, P(gp.ctypes.data + pad), n, P(seg.ctypes.data), P(ps.ctypes.data),
                                P(rd.ctypes.data), stride, P(bits.ctypes.data), cores, reps)
        out[name] = (2 * n * reps, t, cores)
    return out


def gpu_align_stage(lib, g, reads, pos, stride, L, I, d, n_rec, reps):
    """Stage (d) alone through storm_align_plan_*: the DP + traceback of the filter's true-origin
    passes (read strand 0 at diagonal origin), qualities Q20-40 (all level 3, so no quality
    array), default scores and -t 115; HIP-event time per run over `reps` runs."""
    from storm_amd import synth
    P = C.c_void_p
    so = lib.load()
    gp = synth.pack_codes(g)
    n = min(n_rec, reads.shape[0])
    rl = np.full(n, L, np.int32)
    rs = (2 * np.arange(n, dtype=np.int64)).astype(np.int32)
    diag = np.ascontiguousarray(pos[:n, 0] + d, dtype=np.int32)
    p = lib.params(0, allowed_indels=I)
    h = P()
    lib.check(so.storm_align_plan_create(0, P(gp.ctypes.data), None, len(g), P(reads.ctypes.data), stride, None, 0,
                                         None, P(rl.ctypes.data), n, P(rs.ctypes.data), P(diag.ctypes.data), n,
                                         C.byref(p), C.byref(h)))
    ms = C.c_double()
    lib.check(so.storm_align_run(h, 1, C.byref(ms)))
    lib.check(so.storm_align_run(h, reps, C.byref(ms)))
    res = np.zeros(n, lib.ALIGN_RESULT)
    tb = np.zeros((n, lib.TB_BYTES), np.uint8)
    lib.check(so.storm_align_fetch(h, P(res.ctypes.data), P(tb.ctypes.data)))
    so.storm_align_plan_destroy(h)
    return ms.value, res, tb, rs, diag


def cpu_align_baseline(g, reads, L, I, rs, diag, res, tb, n_sample=20000):
    """The oracle's restatement of alignment__align + alignment__traceback (orc_align, one
    thread, kind "port") on the first n_sample records of the DP workload, timed on this host;
    the GPU records of the sample are compared with it."""
    path = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
    if not os.path.exists(path):
        return None

    class OrcParams(C.Structure):
        _fields_ = [(f, C.c_int) for f in (
            "color", "match", "mismatch", "gap_open", "gap_extend", "allowed_indels", "simd_diags", "t", "z",
            "qual_base", "filter_repeats", "stdev_distance", "n_bytes", "sense_mask", "ignore_lowercase",
            "unordered_M", "uniqueness", "uniqueness_threshold", "print_all", "greedy")]

    class OrcAlignOut(C.Structure):
        _fields_ = [("score", C.c_int), ("best_i", C.c_int), ("best_j", C.c_int), ("best_i0", C.c_int),
                    ("best_j0", C.c_int), ("tb", C.c_uint8 * 1024), ("tb_len", C.c_int)]

    o = C.CDLL(path)
    P = C.c_void_p
    o.orc_params_default.argtypes = [C.POINTER(OrcParams), C.c_int]
    o.orc_align.argtypes = [P, P, C.c_int, P, P, C.c_int, C.POINTER(OrcParams), C.c_int, C.POINTER(OrcAlignOut)]
    o.orc_align.restype = C.c_int
    op = OrcParams()
    o.orc_params_default(C.byref(op), 0)
    op.allowed_indels, op.simd_diags = I, min(I, 15)
    n = min(n_sample, len(rs))
    pad = 64
    gpad = np.zeros(pad + len(g) + pad, np.uint8)
    gpad[pad:pad + len(g)] = g
    mpad = np.ones_like(gpad)
    mpad[pad:pad + len(g)] = 0
    qlev = np.full(L, 3, np.uint8)
    codes = np.stack([(reads[:n, :(L + 3) // 4] >> s) & 3 for s in (0, 2, 4, 6)], axis=2).reshape(n, -1)[:, :L]
    codes = np.ascontiguousarray(codes, dtype=np.uint8)
    d = min(I, 15)
    wins = [int(h - I) if h - I > -d else -d for h in diag[:n]]
    wref = [np.ascontiguousarray(gpad[pad + w:pad + w + L + 2 * I + 1]) for w in wins]
    wmask = [np.ascontiguousarray(mpad[pad + w:pad + w + L + 2 * I + 1]) for w in wins]
    outs = [OrcAlignOut() for _ in range(n)]
    t0 = time.perf_counter()
    for k in range(n):
        o.orc_align(P(codes[k].ctypes.data), P(qlev.ctypes.data), L, P(wref[k].ctypes.data), P(wmask[k].ctypes.data),
                    I, C.byref(op), op.t, C.byref(outs[k]))
    dt = time.perf_counter() - t0
    same = True
    for k in range(n):
        a = outs[k]
        if (int(res["score"][k]), int(res["best_j"][k]), int(res["tb_len"][k])) != (a.score, a.best_j, a.tb_len if a.score > op.t else 0):
            same = False
            break
        if a.score > op.t:
            got = bytes(((tb[k][i >> 1] >> (0 if i & 1 else 4)) & 0xF) for i in range(a.tb_len))
            if got != bytes(a.tb[:a.tb_len]):
                same = False
                break
    return {"value": round(n * L * (2 * I + 1) / dt / 1e9, 4), "unit": "GCUPS", "cores": 1, "kind": "port",
            "sample": f"orc_align (the oracle's restatement of alignment__align + alignment__traceback) on the first "
                      f"{n} records, one thread, {dt:.1f}s", "records_per_s": round(n / dt, 1),
            "gpu_records_identical": same}


SEEDS_N = "####--#-#--##--####;#####--###-#-###;###-##-#######"   # seed.h:29 (nucleotide defaults)
SEEDS_C = "####----##--##-####;####-###--#----####;####-####-####"   # seed.h:31 (color defaults)


def host_cores() -> int:
    try:
        return max(1, min(16, len(os.sched_getaffinity(0))))
    except AttributeError:
        return max(1, min(16, os.cpu_count() or 1))


def sam_body(path):
    with open(path) as f:
        return [l for l in f if not l.startswith("@PG")]


def ref_binary(color: int):
    p = os.path.join(ROOT, "oracle", "_ref", "storm-color" if color else "storm-nucleotide")
    return p if os.path.exists(p) else None


def cli_args(kw) -> list:
    out = ["-i", str(kw["allowed_indels"]), "-z", str(kw["z"]), "-t", str(kw["t"])]
    if kw["max_hits"]:
        out += ["-M", str(kw["max_hits"])]   # no -M: the ordered (genome map) output
    return out


def write_read_files(d, pr, idx, color, tag):
    from storm_amd import synth
    sr = synth.packed_to_simreads(pr, idx)
    if color:
        path, qpath = os.path.join(d, f"{tag}.csfasta"), os.path.join(d, f"{tag}.qual")
        synth.write_csfasta(path, qpath, sr)
        return ["-r", path, "-q", qpath], (path, qpath)
    path = os.path.join(d, f"{tag}.fq")
    synth.write_fastq(path, sr)
    return ["-r", path], (path, None)


def run_reference(color, gpath, rargs, kw, sam, cores):
    exe = ref_binary(color)
    cmd = [exe, "-g", gpath] + rargs + cli_args(kw) + ["-s", SEEDS_C if color else SEEDS_N, "-N", str(cores),
                                                     "-o", sam, "-v", "0"]
    t0 = time.perf_counter()
    # the reference ORs color reads into an uninitialised buffer; a fixed perturbation byte makes it deterministic
    subprocess.run(cmd, check=True, env=dict(os.environ, MALLOC_PERTURB_="255"), stderr=subprocess.DEVNULL,
                   stdout=subprocess.DEVNULL)
    return time.perf_counter() - t0


def cpu_mapping_baseline(lib, d, gpath, pr, kw, color, n_sample, ordered=False):
    """The reference binary (oracle/_ref, default Makefile flags, OpenMP) on a sample of the
    workload's reads, timed on this host: one run on 16 reads gives the fixed cost (genome load,
    index build), one on the sample; reads/s = sample / (difference).  The GPU maps the same files
    and its SAM is compared with the reference's."""
    if ref_binary(color) is None:
        return None
    cores = host_cores()
    # the reference maps one read length per run (reads_data.c:156-163): a mixed-length workload
    # is sampled at its median length
    lens = np.asarray(pr.len)
    if lens.min() != lens.max():
        L_med = int(np.median(lens))
        pick = np.nonzero(lens == L_med)[0]
        length_note = f" (all of length {L_med}, the median: the reference takes one read length per run)"
    else:
        pick = np.arange(len(lens))
        length_note = ""
    n_sample = min(n_sample, len(pick))
    r_tiny, _ = write_read_files(d, pr, pick[:16], color, "tiny")
    r_samp, files = write_read_files(d, pr, pick[:n_sample], color, "sample")
    t_fixed = run_reference(color, gpath, r_tiny, kw, os.path.join(d, "ref_tiny.sam"), cores)
    t_all = run_reference(color, gpath, r_samp, kw, os.path.join(d, "ref_sample.sam"), cores)
    gsam = os.path.join(d, "gpu_sample.sam")
    lib.map_files(gpath, files[0], gsam, qual=files[1], mode=1, color=color, **kw)
    same = sam_body(gsam) == sam_body(os.path.join(d, "ref_sample.sam"))
    dt = max(t_all - t_fixed, 1e-6)
    out = {"value": round(n_sample / dt, 1), "unit": "reads/s", "cores": cores, "kind": "reference",
           "sample": f"{n_sample} reads of the workload{length_note}; reference {os.path.basename(ref_binary(color))} -N {cores}: "
                     f"{t_all:.1f}s total minus {t_fixed:.1f}s fixed cost (load + index, measured on 16 reads)",
           "gpu_sam_identical_to_reference": bool(same)}
    if ordered:
        # the same sample without -M: the ordered output (genome map placement, coordinate-sorted SAM)
        ko = dict(kw, max_hits=0)
        t_ref = run_reference(color, gpath, r_samp, ko, os.path.join(d, "ref_ordered.sam"), cores)
        osam = os.path.join(d, "gpu_ordered.sam")
        t0 = time.perf_counter()
        so = lib.map_files(gpath, files[0], osam, qual=files[1], mode=1, color=color, **ko)
        t_gpu = time.perf_counter() - t0
        same_o = sam_body(osam) == sam_body(os.path.join(d, "ref_ordered.sam"))
        out["ordered"] = {"reads": n_sample, "gpu_sam_identical_to_reference": bool(same_o),
                          "gpu_call_s": round(t_gpu, 3), "gpu_map_s": round(so["seconds_map"], 3),
                          "genome_map_and_sam_s": round(so["seconds_sam"], 3),
                          "reference_s": round(t_ref, 2), "reference_minus_fixed_s": round(t_ref - t_fixed, 2)}
        log(f"ordered mode on {n_sample} reads: GPU call {t_gpu:.2f}s (genome map + SAM {so['seconds_sam']:.2f}s), "
            f"reference {t_ref:.2f}s; SAM identical: {same_o}")
    return out


def map_config(lib, name, gpath, pr, kw, color, n_sample_cpu, sam_reads, skip_plain=False, ordered=False,
               ordered_reads=0):
    """Batched (persistent fused kernel in a hipGraph) and plain (staged kernels, one batch per
    read length) on the same resident mapper; hits returned to host memory; whole-call wall time
    (the mean of repeated calls when one call takes under 0.1 s)."""
    t0 = time.perf_counter()
    m = lib.Mapper(gpath, color=color, **kw)
    t_create = time.perf_counter() - t0
    n = len(pr.len)
    w = min(n, 50_000)
    m.map_hits(pr, mode=1, lo=0, hi=w)   # warm-up: code objects, autotune, graph capture
    if not skip_plain:
        m.map_hits(pr, mode=0, lo=0, hi=w)
    res = {"reads": n, "genome_load_index_s": round(t_create, 2)}

    def timed(mode):
        """One whole call; a call shorter than 0.1 s is repeated for about 0.3 s and averaged."""
        t0 = time.perf_counter()
        h, st = m.map_hits(pr, mode=mode)
        t = time.perf_counter() - t0
        if t < 0.1:
            reps = max(1, min(50, int(0.3 / max(t, 1e-4))))
            t0 = time.perf_counter()
            for _ in range(reps):
                m.map_hits(pr, mode=mode)
            t = (time.perf_counter() - t0) / reps
        return h, st, t

    hb, sb, tb = timed(1)
    res["batched"] = {"reads_per_s": round(n / tb, 1), "seconds": round(tb, 3), "mapped": sb["n_mapped"],
                      "gpu_ms_fused": round(sb["gpu_ms_filter"], 2),
                      "staged_fallback_gpu_ms": round(sb["gpu_ms_seed"] + sb["gpu_ms_align"] + sb["gpu_ms_replay"], 2),
                      "fallback_reads": sb["n_fallback"],
                      "candidates": sb["n_candidates"], "filter_passes": sb["n_filter_pass"],
                      "lookups": sb["n_lookups"], "raw_hits": sb["n_hits"]}
    log(f"{name}: batched {n} reads in {tb:.2f}s = {n / tb:.0f} reads/s, mapped {sb['n_mapped']}, "
        f"candidates {sb['n_candidates']}, fused kernel {sb['gpu_ms_filter']:.1f} ms")
    if not skip_plain:
        hp, sp, tp = timed(0)
        same = all(np.array_equal(hb[k], hp[k]) for k in hb)
        lk_bytes = 8 * sp["n_lookups"]
        res["plain"] = {"reads_per_s": round(n / tp, 1), "seconds": round(tp, 3), "mapped": sp["n_mapped"],
                        "gpu_ms": {k: round(sp["gpu_ms_" + k], 2) for k in ("seed", "lookup", "filter", "align",
                                                                            "replay")},
                        "index_lookup_GBps": round(lk_bytes / (sp["gpu_ms_lookup"] * 1e-3) / 1e9, 1)
                        if sp["gpu_ms_lookup"] > 0 else None,
                        # the same against the 8 TB/s HBM peak (8 B of first[] per (seed, strand, position)
                        # lookup; random gathers, each touching its own cache line)
                        "index_lookup_hbm_frac": round(lk_bytes / (sp["gpu_ms_lookup"] * 1e-3) / 8e12, 4)
                        if sp["gpu_ms_lookup"] > 0 else None}
        res["batched_over_plain"] = round(tp / tb, 2)
        res["plain_equals_batched"] = bool(same)
        log(f"{name}: plain {n / tp:.0f} reads/s (batched/plain = {tp / tb:.2f}x), identical hits: {same}")
    if sam_reads:
        k = min(n, sam_reads)
        st = m.map_sam(pr, "/dev/null", mode=1, lo=0, hi=k)
        res["sam_writer_reads_per_s"] = round(k / max(st["seconds_sam"], 1e-9), 1)
    m.close()
    if ordered_reads:
        # the reference's default output on the same reads: no -M, reads placed on the genome map
        # (host, after the device hit map), coordinate-sorted SAM
        k = min(n, ordered_reads)
        mo = lib.Mapper(gpath, color=color, **dict(kw, max_hits=0))
        st = mo.map_sam(pr, "/dev/null", mode=1, lo=0, hi=k)
        mo.close()
        res["ordered"] = {"reads": k, "gpu_map_s": round(st["seconds_map"], 3),
                          "genome_map_and_sam_s": round(st["seconds_sam"], 3),
                          "reads_per_s": round(k / max(st["seconds_map"] + st["seconds_sam"], 1e-9), 1)}
        log(f"{name}: ordered mode on {k} reads: map {st['seconds_map']:.2f}s, genome map + SAM {st['seconds_sam']:.2f}s")
    if n_sample_cpu:
        cb = cpu_mapping_baseline(lib, os.path.dirname(gpath), gpath, pr, kw, color, n_sample_cpu, ordered)
        res["cpu_baseline"] = cb
        if cb:
            log(f"{name}: reference CPU {cb['value']} reads/s on {cb['cores']} cores; GPU SAM identical on the sample: "
                f"{cb['gpu_sam_identical_to_reference']}")
    return res


def config1(lib, cpu=True):
    """1 Mbp i.i.d. genome (seed 1), 100k Illumina reads of 50 nt (2 % substitutions, 0.1 % 1-3 bp
    indels, Q20-40), default seeds, -i 3 -z 100 -t 100 -M 1."""
    from storm_amd import synth
    d = tempfile.mkdtemp(prefix="storm_c1_")
    g = synth.random_genome(1_000_000, 1)
    gpath = os.path.join(d, "g.fa")
    synth.write_fasta_fast(gpath, "chr1", g)
    pr = synth.fast_nucleotide_reads(g, 100_000, 50, 101, sub=0.02, indel=0.001, qlo=20, qhi=40)
    kw = dict(allowed_indels=3, z=100, t=100, max_hits=1)
    return map_config(lib, "config1", gpath, pr, kw, 0, 100_000 if cpu else 0, 100_000, ordered=True)


def config3(lib, cpu=True, n_reads=5_000_000):
    """50 Mbp genome with GC drawn in 40-60 % per 10 kb window; 5M SOLiD color reads of 34-74
    colors after the primer (35-75 bp reads), 3 % color errors, 1 % SNPs, Q10-35; the three
    default weight-12 color seeds, -i 3 -z 85 -t 100 -M 1."""
    from storm_amd import synth
    d = tempfile.mkdtemp(prefix="storm_c3_")
    g = synth.gc_window_genome(50_000_000, 3)
    gpath = os.path.join(d, "g.fa")
    synth.write_fasta_fast(gpath, "chr1", g)
    t0 = time.perf_counter()
    pr = synth.fast_color_reads(g, n_reads, (34, 74), 303, snp=0.01, color_err=0.03, qlo=10, qhi=35)
    log(f"config3: {n_reads} reads generated in {time.perf_counter() - t0:.1f}s")
    kw = dict(allowed_indels=3, z=85, t=100, max_hits=1)
    return map_config(lib, "config3", gpath, pr, kw, 1, 150_000 if cpu else 0, 500_000)


def config4(lib, cpu=True, n_reads=20_000_000, skip_plain=False):
    """100 Mbp genome: i.i.d. background, 10 % segmental duplications (1-20 kb copies, 1-5 %
    divergence), 5 % tandem repeats (units of 2-50 bp); 20M reads of 100 nt, 1.5 % substitutions,
    0.2 % indels, Q15-40; the three default weight-12 seeds, -i 7 -z 100 -t 115 -M 4."""
    from storm_amd import synth
    d = tempfile.mkdtemp(prefix="storm_c4_")
    g = synth.repeat_genome(100_000_000, 4)
    gpath = os.path.join(d, "g.fa")
    synth.write_fasta_fast(gpath, "chr1", g)
    t0 = time.perf_counter()
    pr = synth.fast_nucleotide_reads(g, n_reads, 100, 404, sub=0.015, indel=0.002, qlo=15, qhi=40)
    log(f"config4: {n_reads} reads generated in {time.perf_counter() - t0:.1f}s")
    kw = dict(allowed_indels=7, z=100, t=115, max_hits=4)
    return map_config(lib, "config4", gpath, pr, kw, 0, 10_000 if cpu else 0, 200_000, skip_plain=skip_plain,
                      ordered_reads=min(n_reads, 2_000_000))


def serve_trace(g, n_req: int, seed: int = 5, lo: int = 1_000, hi: int = 1_000_000):
    """Requests of the serving workload: sizes log-uniform in [lo, hi] reads, read lengths uniform
    in 36-150, half nucleotide and half color reads, -i 3 or 7, drawn from `seed`."""
    from storm_amd import synth
    r = synth.rng(seed)
    sizes = np.exp(r.uniform(np.log(lo), np.log(hi), size=n_req)).astype(np.int64)
    colors = r.integers(0, 2, size=n_req)
    bands = np.where(r.integers(0, 2, size=n_req) == 0, 3, 7)
    reqs = []
    for k in range(n_req):
        n, c, i = int(sizes[k]), int(colors[k]), int(bands[k])
        if c:   # L colors after the primer cover L + 1 bases: 35-149 colors for 36-150 bp reads
            pr = synth.fast_color_reads(g, n, (35, 149), 5000 + k, snp=0.01, color_err=0.03, qlo=10, qhi=35)
        else:
            pr = synth.fast_nucleotide_reads(g, n, (36, 150), 5000 + k, sub=0.015, indel=0.002, qlo=15, qhi=40)
        reqs.append((pr, dict(color=c, allowed_indels=i, z=85 if c else 100, t=100, max_hits=1)))
    return reqs


def serve_run(server, reqs, arrivals):
    """Open loop: request k is submitted at arrivals[k] seconds (Poisson process), a waiter thread
    per request collects it.  Returns per-request latencies and the wall time from the first
    submission to the last result."""
    import threading
    lat = [0.0] * len(reqs)
    info = [None] * len(reqs)
    errs = []

    def waiter(k, rid):
        try:
            _, res = server.wait(rid)
            lat[k] = res["seconds_latency"]
            info[k] = res
        except Exception as e:   # noqa: BLE001 - reported after the run
            errs.append(f"request {k}: {e}")

    th = []
    t0 = time.perf_counter()
    for k, (pr, kw) in enumerate(reqs):
        dt = arrivals[k] - (time.perf_counter() - t0)
        if dt > 0:
            time.sleep(dt)
        rid = server.submit(pr, **kw)
        w = threading.Thread(target=waiter, args=(k, rid))
        w.start()
        th.append(w)
    for w in th:
        w.join()
    wall = time.perf_counter() - t0
    if errs:
        raise RuntimeError("; ".join(errs[:3]))
    return np.array(lat), wall, info


def config5(lib, n_req=40, load=0.8, lanes=3, max_batch=1 << 18):
    """Serving: the config-3 genome (50 Mbp, GC 40-60 % per 10 kb window) with nucleotide and color
    indexes resident; Poisson arrivals of requests of 1k-1M reads (log-uniform), lengths 36-150,
    50/50 nucleotide/color, -i 3 or 7 per request, -M 1; admission by free HBM; continuous
    batching vs the plain per-request path on the same arrival trace."""
    from storm_amd import synth
    d = tempfile.mkdtemp(prefix="storm_c5_")
    g = synth.gc_window_genome(50_000_000, 3)
    gpath = os.path.join(d, "g.fa")
    synth.write_fasta_fast(gpath, "chr1", g)
    t0 = time.perf_counter()
    reqs = serve_trace(g, n_req)
    sizes = np.array([len(pr.len) for pr, _ in reqs])
    total = int(sizes.sum())
    log(f"config5: {n_req} requests, {total} reads generated in {time.perf_counter() - t0:.1f}s")
    res = {"requests": n_req, "reads": total, "request_reads_min_max": [int(sizes.min()), int(sizes.max())]}
    t0 = time.perf_counter()
    srv = lib.Server(gpath, lanes=lanes, max_batch_reads=max_batch)
    res["server_create_s"] = round(time.perf_counter() - t0, 2)
    # warm-up: one small request per class (code objects, autotuner, graph capture on every lane)
    warm = [(synth.fast_nucleotide_reads(g, 20_000, (36, 150), 1, 0.015, 0.002, 15, 40), dict(color=0, allowed_indels=i,
             z=100, t=100)) for i in (3, 7)] + \
           [(synth.fast_color_reads(g, 20_000, (35, 149), 2, 0.01, 0.03, 10, 35), dict(color=1, allowed_indels=i, z=85,
             t=100)) for i in (3, 7)]
    for _ in range(lanes):
        ids = [srv.submit(pr, **kw) for pr, kw in warm]
        for rid in ids:
            srv.wait(rid)
    # capacity: every request submitted at once (closed loop)
    lat_c, wall_c, _ = serve_run(srv, reqs, np.zeros(n_req))
    cap = total / wall_c
    log(f"config5: saturated batched capacity {cap:.0f} reads/s ({wall_c:.2f}s)")
    rate = load * cap / sizes.mean()
    arrivals = np.cumsum(synth.rng(55).exponential(1.0 / rate, size=n_req))
    arrivals -= arrivals[0]
    s0 = srv.stats()
    lat_b, wall_b, info_b = serve_run(srv, reqs, arrivals)
    s1 = srv.stats()
    srv.close()
    nb = s1["batches"] - s0["batches"]
    res["arrival_rate_req_per_s"] = round(rate, 3)
    res["offered_reads_per_s"] = round(rate * sizes.mean(), 1)
    res["batched"] = {"reads_per_s": round(total / wall_b, 1), "wall_s": round(wall_b, 3),
                      "p50_latency_s": round(float(np.percentile(lat_b, 50)), 4),
                      "p99_latency_s": round(float(np.percentile(lat_b, 99)), 4),
                      "saturated_reads_per_s": round(cap, 1),
                      "batches": nb, "mean_reads_per_batch": round((s1["batch_reads"] - s0["batch_reads"]) / max(nb, 1), 1),
                      "mean_requests_per_batch": round((s1["batch_requests"] - s0["batch_requests"]) / max(nb, 1), 2),
                      "max_lanes_busy": s1["max_lanes_busy"], "admission_waits": s1["admission_waits"],
                      "arena_bytes": s1["arena_bytes"], "arena_peak_bytes": s1["arena_peak_bytes"],
                      "mapped": int(sum(i["n_mapped"] for i in info_b))}
    log(f"config5: batched {total / wall_b:.0f} reads/s, p50 {np.percentile(lat_b, 50):.3f}s "
        f"p99 {np.percentile(lat_b, 99):.3f}s")
    t0 = time.perf_counter()
    srv = lib.Server(gpath, mode=lib.STORM_MODE_PLAIN)
    log(f"config5: plain server created in {time.perf_counter() - t0:.1f}s")
    for pr, kw in warm:
        srv.wait(srv.submit(pr, **kw))
    lat_p, wall_p, info_p = serve_run(srv, reqs, arrivals)
    srv.close()
    res["plain"] = {"reads_per_s": round(total / wall_p, 1), "wall_s": round(wall_p, 3),
                    "p50_latency_s": round(float(np.percentile(lat_p, 50)), 4),
                    "p99_latency_s": round(float(np.percentile(lat_p, 99)), 4),
                    "mapped": int(sum(i["n_mapped"] for i in info_p))}
    res["batched_over_plain"] = round(wall_p / wall_b, 2)
    log(f"config5: plain {total / wall_p:.0f} reads/s, p50 {np.percentile(lat_p, 50):.3f}s "
        f"p99 {np.percentile(lat_p, 99):.3f}s; batched/plain {wall_p / wall_b:.2f}x")
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--reads", type=int, default=5_000_000, help="reads of the filter workload (2 windows each)")
    ap.add_argument("--no-map", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--percall", type=int, default=20000)
    ap.add_argument("--dp-records", type=int, default=2_000_000, help="records of the DP stage measurement")
    ap.add_argument("--config", default="default", choices=["default", "1", "3", "4", "5"],
                    help="default: filter headline + config-1 mapping; 1/3/4: that mapping workload alone")
    ap.add_argument("--reads-c3", type=int, default=5_000_000)
    ap.add_argument("--reads-c4", type=int, default=20_000_000)
    ap.add_argument("--no-plain", action="store_true", help="mapping workloads: batched path only")
    ap.add_argument("--requests", type=int, default=40, help="config 5: requests in the arrival trace")
    ap.add_argument("--calib-copy", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device to device 3 times (calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {a.gpus}); the path shards reads across GPUs by design "
              f"but this build measures one GPU.", file=sys.stderr)
        sys.exit(2)
    import torch
    from storm_amd import lib
    assert torch.cuda.is_available()
    torch.zeros(1, device="cuda:0")
    lib.load()
    if a.calib_copy:
        src = torch.ones(a.calib_copy << 18, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration: copied {a.calib_copy} MiB device to device 3 times")
        del src, dst
    if a.config == "5":
        res = config5(lib, n_req=a.requests)
        best = res["batched"]
        line = {"metric": METRIC, "value": best["reads_per_s"], "unit": "reads/s", "n_gpus": 1, "steps": 1,
                "warmup": 1, "ms_per_step": round(best["wall_s"] * 1e3, 2), "higher_is_better": True,
                "scaling": "weak", "vs_baseline": None, "dtype": "u8/int32", "data": "synthetic",
                "config": {"workload": "serving config 5: " + " ".join(config5.__doc__.split())},
                "p50_latency_s": best["p50_latency_s"], "p99_latency_s": best["p99_latency_s"], "serving": res}
        print(json.dumps(line), flush=True)
        return
    if a.config != "default":
        cpu = not a.no_cpu
        if a.config == "1":
            res = config1(lib, cpu)
        elif a.config == "3":
            res = config3(lib, cpu, a.reads_c3)
        else:
            res = config4(lib, cpu, a.reads_c4, skip_plain=a.no_plain)
        best = res["batched"]
        line = {"metric": METRIC, "value": best["reads_per_s"], "unit": "reads/s", "n_gpus": 1, "steps": 1,
                "warmup": 1, "ms_per_step": round(best["seconds"] * 1e3, 2), "higher_is_better": True,
                "scaling": "weak", "vs_baseline": None, "dtype": "u8/int32", "data": "synthetic",
                "config": {"workload": f"mapping config {a.config}: " + " ".join(
                    {"1": config1, "3": config3, "4": config4}[a.config].__doc__.split())},
                "mapping": res}
        print(json.dumps(line), flush=True)
        return
    L, W, d = 75, 8, 7
    t0 = time.perf_counter()
    g, reads, pos, stride = make_filter_workload(a.reads, L, W, d)
    log(f"workload: {a.reads} reads x 2 windows built in {time.perf_counter() - t0:.1f}s")
    n_cand = 2 * a.reads
    cells = n_cand * L * 2 * W
    ms, wall, scores, passf, cr, cp = gpu_filter(lib, g, reads, pos, stride, L, W, a.steps, a.warmup)
    gcups = cells / (ms * 1e-3) / 1e9
    log(f"filter: {ms:.3f} ms/launch (events), {wall * 1e3:.3f} ms/step (wall), {gcups:.1f} GCUPS, "
        f"pass rate {passf.mean():.3f}")
    # part of the CPU-baseline leg: a spot check against the CPU restatement (the full parity tests
    # live in tests/)
    spot = None
    orc = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
    if not a.no_cpu and os.path.exists(orc):
        from storm_amd import synth
        ol = C.CDLL(orc)
        ol.orc_filter_score.restype = C.c_int
        ol.orc_filter_score.argtypes = [C.c_void_p, C.c_long, C.c_void_p, C.c_int, C.c_int] + [C.c_int] * 4
        pad = 64
        gp = np.zeros(pad + (len(g) + 3) // 4 + pad, np.uint8)
        gp[pad:pad + (len(g) + 3) // 4] = synth.pack_codes(g)
        sel = np.linspace(0, n_cand - 1, 2000).astype(np.int64)
        want = np.array([min(255, ol.orc_filter_score(C.c_void_p(gp.ctypes.data + pad), int(cp[c]),
                                                      C.c_void_p(reads[int(cr[c])].ctypes.data), L, W, 5, -4, -9, -4))
                         for c in sel])
        spot = bool(np.array_equal(scores[sel].astype(np.int64), want))
        log(f"spot check vs oracle on 2000 candidates: {'identical' if spot else 'MISMATCH'}")
    # stage (d): the DP + traceback of the true-origin passes
    n_dp = min(a.dp_records, a.reads)
    dp_ms, dp_res, dp_tb, dp_rs, dp_diag = gpu_align_stage(lib, g, reads, pos, stride, L, 7, d, n_dp, max(1, a.steps // 4))
    dp_cells = n_dp * L * 15
    dp_gcups = dp_cells / (dp_ms * 1e-3) / 1e9
    n_tb = int((dp_res["tb_len"] > 0).sum())
    log(f"align stage: {n_dp} records in {dp_ms:.2f} ms = {dp_gcups:.1f} GCUPS ({n_dp / dp_ms / 1e3:.1f} M records/s), "
        f"{n_tb} tracebacks")
    pc_dt, _ = percall_filter(lib, g, reads, pos, stride, L, W, min(a.percall, a.reads))
    pc_gcups = 2 * min(a.percall, a.reads) * L * 2 * W / pc_dt / 1e9
    log(f"per-call drop-in row: {pc_dt:.2f}s for {min(a.percall, a.reads)} calls, {pc_gcups:.3f} GCUPS")
    mapping = None if a.no_map else config1(lib, cpu=not a.no_cpu)
    cpu = None
    if not a.no_cpu:
        cb = cpu_baseline(g, reads, pos, stride, L, d, n_sample=500_000)
        if cb and "sse2" in cb:
            nc, t, cores = cb["sse2"]
            cpu = {"value": round(nc * L * 2 * W / t / 1e9, 3), "unit": "GCUPS", "cores": cores, "kind": "reference",
                   "sample": f"reference SSE2 filter (alignment_sse2__align_pair, default Makefile flags) over "
                             f"{nc} candidate evaluations of the same workload (first 500k reads x 2 windows, "
                             f"repeated), OpenMP, {t:.1f}s"}
            if "avx2" in cb:
                nc2, t2, _ = cb["avx2"]
                cpu["avx2_gcups"] = round(nc2 * L * 2 * W / t2 / 1e9, 3)
            log(f"cpu baseline: {cpu['value']} GCUPS on {cores} threads")
    dp_cpu = None
    if not a.no_cpu:
        dp_cpu = cpu_align_baseline(g, reads, L, 7, dp_rs, dp_diag, dp_res, dp_tb)
        if dp_cpu:
            log(f"align stage cpu baseline: {dp_cpu['value']} GCUPS on 1 thread; GPU records identical on the sample: "
                f"{dp_cpu['gpu_records_identical']}")
    dp_achieved = dp_gcups * VALU_INSTR_PER_DP_CELL_I7_L75 / 1e3
    align_stage = {
        "workload": f"DP + traceback of {n_dp} filter passes (the true-origin windows of the filter workload), "
                    f"L={L}, -i 7 (15 band columns), -t 115, levels 3",
        "kernel": "k_align<7,false>", "records": n_dp, "tracebacks": n_tb, "ms": round(dp_ms, 3),
        "gcups": round(dp_gcups, 1), "records_per_s": round(n_dp / (dp_ms * 1e-3), 1),
        "roofline": {"bound": "valu", "achieved": round(dp_achieved, 3), "peak": VALU_PEAK_INT32,
                     "unit": "T VALU lane-instr/s", "frac": round(dp_achieved / VALU_PEAK_INT32, 4),
                     "work_per_cell": VALU_INSTR_PER_DP_CELL_I7_L75},
        "cpu_baseline": dp_cpu}
    achieved = gcups * VALU_INSTR_PER_CELL_W8_L75 / 1e3          # T lane-instructions/s
    algorithmic = gcups * OPS_PER_CELL / 1e3                       # T reference-ops/s
    line = {
        "metric": METRIC, "value": round(gcups, 2), "unit": "GCUPS", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(wall * 1e3, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u8", "data": "synthetic",
        "config": {"workload": f"banded u8 filter, {n_cand} candidates ({a.reads} reads x 2 windows: true origin, "
                               f"random decoy), L={L}, -i 7 (W={W}, {2 * W} diagonals), z=100, 1 Mbp i.i.d. genome",
                   "cells_per_step": cells, "kernel": "k_filter<8>"},
        "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": VALU_PEAK_INT32,
                     "unit": "T VALU lane-instr/s", "frac": round(achieved / VALU_PEAK_INT32, 4),
                     "traffic": int(HBM_BYTES_PER_CAND_W8_L75 * n_cand),
                     "work_per_cell": VALU_INSTR_PER_CELL_W8_L75,
                     "hbm": {"achieved": round(ALG_BYTES_PER_CAND_W8_L75 * n_cand / (ms * 1e-3) / 1e9, 1),
                             "peak": 8000.0, "unit": "GB/s",
                             "frac": round(ALG_BYTES_PER_CAND_W8_L75 * n_cand / (ms * 1e-3) / 8e12, 4),
                             "algorithmic_bytes_per_candidate": ALG_BYTES_PER_CAND_W8_L75,
                             "measured_bytes_per_candidate": HBM_BYTES_PER_CAND_W8_L75},
                     "algorithmic": {"ops_per_cell": OPS_PER_CELL, "achieved_tops": round(algorithmic, 2),
                                     "peak_packed16_tops": VALU_PEAK_PK16,
                                     "frac": round(algorithmic / VALU_PEAK_PK16, 4)}},
        "cpu_baseline": cpu,
        "kernel_ms": round(ms, 4),
        "percall_gcups": round(pc_gcups, 4),
        "batched_over_percall": round(gcups / pc_gcups, 1),
        "spot_check_identical": spot,
        "align_stage": align_stage,
        "mapping": mapping,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
