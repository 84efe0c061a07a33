"""Benchmark of the SToRM mapping hot path on one MI355X: every workload of the metric in one run.

Headline (`value`, one "step" = one launch of the banded filter over the whole batch): the filter
alone (10M candidate pairs = 5M reads of 75 nt x 2 windows, true origin - 7 and a uniform decoy,
on a 1 Mbp i.i.d. genome, -i 7: W = 8 lanes, 16 diagonals, u8 scores, z = 100), inputs resident in
HBM, filter GCUPS against the VALU issue roofline.  The same JSON line carries:
  * the DP stage alone (DP + traceback of 2M filter passes) against the reference's
    alignment__align / alignment__traceback on the host cores;
  * the mapping workloads 1 (100k x 50 nt, 1 Mbp), 3 (5M color reads of 35-75 bp, 50 Mbp) and 4
    (20M x 100 nt, 100 Mbp with repeats, -M 4): reads/s from packed reads in host memory to the
    hit map and tracebacks in host memory (storm_mapper_map_records, the records the SAM text is
    made of), batched (persistent fused kernel) and plain (staged kernels, one batch per read
    length), first call and warm, per-stage times of the timed calls, identical records on both
    paths; the SAM text (formatted on the device) timed separately; and the reference program's
    own mapping phase on the host cores over a sample (per read length for workload 3), with the
    GPU SAM compared to the reference's;
  * workload 5: Poisson arrivals of requests on the serving loop (resident indexes, admission by
    HBM, continuous batching, SAM records per request) and on the plain per-request server; p50 /
    p99 latency to the last SAM record, reads/s, identical results on both.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--config all|2|1|3|4|5] [--no-cpu]
"""
from __future__ import annotations

import argparse
import ctypes as C
import hashlib
import json
import os
import subprocess
import sys
import tempfile
import threading
import time
from types import SimpleNamespace

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
REF = os.path.join(ROOT, "oracle", "_ref")

METRIC = "mapped reads/s + filter GCUPS, batched vs plain GPU; p50/p99 request latency"
OPS_PER_CELL = 13            # compute chain of the reference's SIMD cell (alignment_simd.c:3810-3880)
VALU_PEAK_INT32 = 39.3       # T VALU lane-instructions/s: 256 CU x 64 lanes x 2.4 GHz
VALU_PEAK_PK16 = 78.6        # T 16-bit ops/s with packed VALU instructions (2 per lane-instruction)
HBM_PEAK_GBPS = 8000.0
# VALU lane-instructions k_filter<8> issues per cell at L=75 (SQ_INSTS_VALU x 64 / cells, rocprofv3 --pmc,
# profiles/k_filter_W8_L75_10Mcand.md); the kernel's own work measure for the VALU-issue roofline
VALU_INSTR_PER_CELL_W8_L75 = 4.755
# VALU lane-instructions k_align<7,false> issues per DP cell (L x 15 band columns) on the DP stage
# workload (one pass per record with directions, the walk of every accepted record), measured the
# same way (profiles/k_align_I7_L75_2M_records.md)
VALU_INSTR_PER_DP_CELL_I7_L75 = 32.53
# HBM bytes per candidate of the same dispatch: (FETCH_SIZE x 2.00 + WRITE_SIZE x 1.00) x 1024 /
# candidates, the factors calibrated on a device-to-device copy of known size in the same run
# (profiles/k_filter_W8_L75_10Mcand.md)
HBM_BYTES_PER_CAND_W8_L75 = 45.1
# HBM bytes fetched per index lookup by k_seed_count (FETCH_SIZE x 2.00 x 1024 / lookups, the factor
# calibrated on a device-to-device copy in the same run; profiles/k_seed_count_lookups.md): each
# random first[key] read takes a 128-byte line
LOOKUP_FETCH_BYTES = {"config1": 131.3, "config3": 131.7}
# algorithmic bytes per candidate: packed read (19 B at L = 75) + genome window (L + 2W - 1 = 90
# codes, 23 B) + the u8 score and the pass byte
ALG_BYTES_PER_CAND_W8_L75 = 44.0

SEEDS_N = "####--#-#--##--####;#####--###-#-###;###-##-#######"   # seed.h:29 (nucleotide defaults)
SEEDS_C = "####----##--##-####;####-###--#----####;####-####-####"   # seed.h:31 (color defaults)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# ------------------------------------------------------------------------------------------ host
def host_info() -> dict:
    """The host cores: nproc, lscpu's sockets / cores / threads per core, and the threads the CPU
    baselines use (the process's CPU set, capped by OMP_NUM_THREADS when the environment sets it:
    on a shared GPU box that is the CPU share of one GPU)."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    info = {"nproc": avail, "os_cpu_count": os.cpu_count()}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, check=True).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core", "CPU(s)"):
                info[k] = v
        if "Socket(s)" in info and "Core(s) per socket" in info:
            info["physical_cores"] = int(info["Socket(s)"]) * int(info["Core(s) per socket"])
    except (OSError, subprocess.CalledProcessError, ValueError):
        pass
    omp = os.environ.get("OMP_NUM_THREADS")
    used = avail
    if omp and omp.isdigit() and int(omp) > 0:
        used = min(avail, int(omp))
    info["threads_used"] = used
    info["threads_note"] = ("the process's CPU set" if used == avail else
                            f"OMP_NUM_THREADS={omp} (the CPU share of this process) of {avail} available")
    return info


HOST = None


def cores() -> int:
    return HOST["threads_used"]


# ------------------------------------------------------------------------------------------ filter
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
    so = lib.load()
    lib.check(so.storm_filter_plan_create(0, P(gp.ctypes.data), len(g), P(reads.ctypes.data), stride, P(rl.ctypes.data),
                                          n_reads, P(cr.ctypes.data), P(cp.ctypes.data), len(cr), W, 5, -4, -9, -4, 100,
                                          C.byref(h)))
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


def percall_filter(lib, g, reads, pos, L, n_calls):
    """The drop-in row: alignment_gfx950__init_pair + __align_pair, one call per read (run.c:660)."""
    from storm_amd import synth
    P = C.c_void_p
    so = lib.load()
    pad = 1024
    gp = np.zeros(pad + (len(g) + 3) // 4 + pad, np.uint8)
    gp[pad:pad + (len(g) + 3) // 4] = synth.pack_codes(g)
    lib.check(so.alignment_gfx950__set_genome(P(gp.ctypes.data), gp.nbytes, pad))
    so.alignment_gfx950__init_pair(5, 4, 9, 4, 100, L)
    t0 = time.perf_counter()
    for k in range(n_calls):
        win = np.ascontiguousarray(pos[k])
        so.alignment_gfx950__align_pair(P(gp.ctypes.data + pad), P(win.ctypes.data), P(reads[k].ctypes.data))
    dt = time.perf_counter() - t0
    so.alignment_gfx950__clean()
    return dt


def cpu_filter_baseline(g, reads, pos, L, d, n_sample, target_s=8.0):
    """The reference's SIMD filter (alignment_<isa>__align_pair of alignment_simd.c) built with the
    default Makefile flags (SSE2, Makefile:7) and with -mavx2, OpenMP over the reads on the host
    cores, on the first n_sample reads of the workload (2 windows each), repeated to ~target_s."""
    from storm_amd import synth
    pad = 1024
    gp = np.zeros(pad + (len(g) + 3) // 4 + pad, np.uint8)
    gp[pad:pad + (len(g) + 3) // 4] = synth.pack_codes(g)
    n = min(n_sample, reads.shape[0])
    rd = np.ascontiguousarray(reads[:n])
    seg = np.arange(0, 2 * n + 1, 2, dtype=np.int64)
    ps = np.ascontiguousarray(pos[:n].reshape(-1))
    bits = np.zeros(2 * n, np.uint8)
    P = C.c_void_p
    out = {}
    for isa, name, so_name in ((0, "sse2", "libref_filter_sse2.so"), (1, "avx2", "libref_filter_avx2.so")):
        path = os.path.join(REF, so_name)
        if not os.path.exists(path):
            continue
        lib = C.CDLL(path)
        lib.ref_filter_time.restype = C.c_double
        lib.ref_filter_time.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, P, C.c_long, P, P, P, C.c_long, P, C.c_int,
                                        C.c_int]
        t = lib.ref_filter_time(isa, d, L, 100, P(gp.ctypes.data + pad), n, P(seg.ctypes.data), P(ps.ctypes.data),
                                P(rd.ctypes.data), rd.shape[1], P(bits.ctypes.data), cores(), 1)
        if t <= 0:
            continue
        reps = max(1, int(target_s / max(t, 1e-3)))
        t = lib.ref_filter_time(isa, d, L, 100, P(gp.ctypes.data + pad), n, P(seg.ctypes.data), P(ps.ctypes.data),
                                P(rd.ctypes.data), rd.shape[1], P(bits.ctypes.data), cores(), reps)
        out[name] = (2 * n * reps, t)
    return out


def oracle_spot_check(g, reads, scores, cr, cp, L, W, n_check=2000):
    """The CPU restatement's u8 score of a sample of the filter workload's candidates."""
    from storm_amd import synth
    orc = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
    if not os.path.exists(orc):
        return None
    ol = C.CDLL(orc)
    ol.orc_filter_score.restype = C.c_int
    ol.orc_filter_score.argtypes = [C.c_void_p, C.c_long, C.c_void_p, C.c_int, C.c_int] + [C.c_int] * 4
    pad = 64
    gp = np.zeros(pad + (len(g) + 3) // 4 + pad, np.uint8)
    gp[pad:pad + (len(g) + 3) // 4] = synth.pack_codes(g)
    sel = np.linspace(0, len(cr) - 1, n_check).astype(np.int64)
    want = np.array([min(255, ol.orc_filter_score(C.c_void_p(gp.ctypes.data + pad), int(cp[c]),
                                                  C.c_void_p(reads[int(cr[c])].ctypes.data), L, W, 5, -4, -9, -4))
                     for c in sel])
    return bool(np.array_equal(scores[sel].astype(np.int64), want))


# ------------------------------------------------------------------------------------------ DP stage
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
    return ms.value, res, tb, diag


def cpu_align_baseline(g, reads, L, I, t, diag, res, tb, n_sample=200_000, target_s=4.0):
    """The reference's alignment__align + alignment__traceback (alignment.c:488-734, its SSE2 build)
    as process_read calls them for a filter pass, one AlignmentType per OpenMP thread on the host
    cores, over the first n_sample records of the DP workload; every record compared with the GPU's
    (score; for accepted records the last-row column, the traceback start and its codes)."""
    from storm_amd import synth
    path = os.path.join(REF, "libref_align.so")
    if not os.path.exists(path):
        return None
    lib = C.CDLL(path)
    P = C.c_void_p
    lib.ref_align_time.restype = C.c_double
    lib.ref_align_time.argtypes = [P, P, P, C.c_long, C.c_int, P, C.c_long] + [C.c_int] * 8 + [P] * 6 + [C.c_int]
    n = min(n_sample, len(diag))
    pad = 1024
    gbytes = (len(g) + 3) // 4
    gp = np.zeros(pad + gbytes + pad, np.uint8)
    gp[pad:pad + gbytes] = synth.pack_codes(g)
    mp = np.full(pad + gbytes + pad, 0x55, np.uint8)   # the reference's masked buffer: 0x55 padding
    mp[pad:pad + gbytes] = 0
    rd = np.ascontiguousarray(reads[:n])
    dg = np.ascontiguousarray(diag[:n], np.int32)
    sc, bj, bi0, bj0, tl = (np.zeros(n, np.int32) for _ in range(5))
    tbr = np.zeros((n, 144), np.uint8)
    args = lambda reps: (P(gp.ctypes.data + pad), P(mp.ctypes.data + pad), P(rd.ctypes.data), rd.shape[1], L,
                         P(dg.ctypes.data), n, I, t, 5, -4, -9, -4, cores(), reps, P(sc.ctypes.data), P(bj.ctypes.data),
                         P(bi0.ctypes.data), P(bj0.ctypes.data), P(tl.ctypes.data), P(tbr.ctypes.data), 144)
    t1 = lib.ref_align_time(*args(1))
    reps = max(1, int(target_s / max(t1, 1e-3)))
    dt = lib.ref_align_time(*args(reps))
    r = res[:n]
    acc = sc > t
    same = bool(np.array_equal(r["score"], sc))
    same = same and bool(np.array_equal(r["best_j"][acc], bj[acc]) and np.array_equal(r["best_i0"][acc], bi0[acc]) and
                         np.array_equal(r["best_j0"][acc], bj0[acc]) and np.array_equal(r["tb_len"][acc], tl[acc]))
    if same:   # the tb_len codes of each accepted record (the unused low nibble of an odd length's last byte aside)
        col = np.arange(144)[None, :]
        lens = tl[:, None]
        keep = np.where(col < lens // 2, 0xFF, np.where((col == lens // 2) & (lens % 2 == 1), 0xF0, 0)).astype(np.uint8)
        keep[~acc] = 0
        same = bool(np.array_equal(tb[:n] & keep, tbr & keep))
    return {"value": round(n * L * (2 * I + 1) / dt / 1e9, 3), "unit": "GCUPS", "cores": cores(), "kind": "reference",
            "sample": f"the reference's alignment__align + alignment__traceback (SSE2 build, libref_align.so) on the "
                      f"first {n} records, OpenMP, {dt:.2f}s per pass", "records_per_s": round(n / dt, 1),
            "gpu_records_identical": same}


# ------------------------------------------------------------------------------------------ mapping
def ref_harness(color: int):
    p = os.path.join(REF, "ref-map-harness-color" if color else "ref-map-harness-nucleotide")
    return p if os.path.exists(p) else None


def cli_args(kw) -> list:
    out = ["-i", str(kw["allowed_indels"]), "-z", str(kw["z"]), "-t", str(kw["t"])]
    if kw.get("max_hits"):
        out += ["-M", str(kw["max_hits"])]
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


def sam_records_body(path):
    with open(path) as f:
        return [l for l in f if not l.startswith("@")]


def cpu_mapping_baseline(lib, m, d, gpath, pr, kw, color, per_length: int, max_reads: int):
    """The reference program (its own sources, SSE2 flags, OpenMP on the host cores) on a sample of
    the workload through oracle/_ref/ref-map-harness-*: one run per read length (the reference
    maps one length per run, reads_data.c:156-163), the sequences and seed indexes built once,
    timed from the end of the index build to the start of the SAM output (its forward and reverse
    process_read passes, run.c:865-994).  reads/s = sample reads / summed mapping time.  The GPU
    maps each run's file (loaded and ordered by tag like the reference) and its SAM records are
    compared with the reference's."""
    exe = ref_harness(color)
    if exe is None:
        return None
    lens = np.asarray(pr.len)
    runs = []
    for Lv in sorted(np.unique(lens), reverse=True):   # longest first: the first run pads the sequence
        idx = np.nonzero(lens == Lv)[0][:per_length]
        if len(idx):
            runs.append((int(Lv), idx))
    total = sum(len(i) for _, i in runs)
    if total > max_reads:   # an even share of max_reads per length
        share = max(1, max_reads // len(runs))
        runs = [(Lv, idx[:share]) for Lv, idx in runs]
        total = sum(len(i) for _, i in runs)
    args = [exe, os.path.join(d, "ref_times.txt"), "-g", gpath] + cli_args(kw) + \
        ["-s", SEEDS_C if color else SEEDS_N, "-N", str(cores()), "-v", "0"]
    files = []
    for k, (Lv, idx) in enumerate(runs):
        rargs, f = write_read_files(d, pr, idx, color, f"ref_{k}")
        files.append(f)
        args += ["--"] + rargs + ["-o", os.path.join(d, f"ref_{k}.sam")]
    t0 = time.perf_counter()
    # the reference ORs color reads into an uninitialised buffer; a fixed perturbation byte makes it deterministic
    subprocess.run(args, check=True, env=dict(os.environ, MALLOC_PERTURB_="255"), stdout=subprocess.DEVNULL,
                   stderr=subprocess.DEVNULL)
    wall = time.perf_counter() - t0
    secs = [float(l.split()[3]) for l in open(os.path.join(d, "ref_times.txt"))]
    same = True
    for k, (rpath, qpath) in enumerate(files):
        rd = lib.load_reads(rpath, qpath, color=color)
        v = SimpleNamespace(codes=rd["codes"], qual=rd["qual"], len=rd["len"], first_base=rd["first_base"],
                            first_qual=rd["first_qual"])
        out = os.path.join(d, f"gpu_{k}.sam")
        m.map_sam(v, out, mode=lib.STORM_MODE_BATCHED, names=rd["names"])
        same = same and sam_records_body(out) == sam_records_body(os.path.join(d, f"ref_{k}.sam"))
    t_map = sum(secs)
    return {"value": round(total / t_map, 1), "unit": "reads/s", "cores": cores(), "kind": "reference",
            "sample": f"{total} reads in {len(runs)} run(s) of one read length each ({per_length if len(runs) > 1 else total}"
                      f" per length); the reference's mapping phase (process_read, forward and reverse) summed over the "
                      f"runs: {t_map:.2f}s; harness wall time with its one genome load and index build {wall:.1f}s",
            "runs": len(runs), "map_seconds": round(t_map, 3),
            "gpu_sam_identical_to_reference": bool(same)}


def map_config(lib, name, gpath, pr, kw, color, chunk, sam_reads, cpu=True, ref_per_length=0, ref_max=0):
    """Batched (persistent fused kernel in a hipGraph) and plain (staged kernels, one batch per read
    length) on one resident mapper: reads/s to the whole hit map and tracebacks in host memory
    (storm_mapper_map_records), in calls of up to `chunk` reads; first call (code objects, autotune,
    graph capture, buffer growth) and warm pass; per-stage times summed over the warm calls; the
    records of both paths compared; the SAM text formatted on the device timed apart."""
    from storm_amd import lib as L
    n = len(pr.len)
    M = kw["max_hits"]
    t0 = time.perf_counter()
    m = L.Mapper(gpath, color=color, seeds=SEEDS_C if color else SEEDS_N, **kw)
    res = {"reads": n, "genome_load_index_s": round(time.perf_counter() - t0, 2), "max_hits": M,
           "timed_region": "packed reads in host memory -> hit map records + tracebacks in host memory "
                           "(storm_mapper_map_records); SAM text timed apart"}
    chunks = [(lo, min(n, lo + chunk)) for lo in range(0, n, chunk)]
    views = [m.view(pr, lo, hi) for lo, hi in chunks]
    c0 = chunks[0][1] - chunks[0][0]
    bufs = {mode: (np.zeros((c0, 4), L.HIT_RECORD), np.zeros((c0, M, L.TB_BYTES), np.uint8)) for mode in (0, 1)}
    first = {}
    for mode in (1, 0):   # first calls, on the first chunk
        h, tb = bufs[mode]
        t = time.perf_counter()
        m.records_call(views[0][0], mode, h, tb)
        first[mode] = time.perf_counter() - t
    keys = ("gpu_ms_seed", "gpu_ms_lookup", "gpu_ms_filter", "gpu_ms_align", "gpu_ms_replay", "n_candidates",
            "n_filter_pass", "n_hits", "n_lookups", "n_mapped", "n_fallback")
    reps = 1
    if len(chunks) == 1 and max(first.values()) < 0.15:
        reps = 5   # a short workload: the warm figure is the mean of several calls
    wall = {0: 0.0, 1: 0.0}
    sums = {0: dict.fromkeys(keys, 0.0), 1: dict.fromkeys(keys, 0.0)}
    same = True
    for _ in range(reps):
        for (v, keep), (lo, hi) in zip(views, chunks):
            k = hi - lo
            for mode in (1, 0):
                h, tb = bufs[mode]
                t = time.perf_counter()
                st = m.records_call(v, mode, h[:k], tb[:k])
                wall[mode] += time.perf_counter() - t
                for q in keys:
                    sums[mode][q] += st[q]
            same = same and np.array_equal(bufs[0][0][:k], bufs[1][0][:k]) and np.array_equal(bufs[0][1][:k],
                                                                                               bufs[1][1][:k])
    for mode in (0, 1):
        for q in keys:
            sums[mode][q] /= reps
        wall[mode] /= reps
    tb_, tp = wall[1], wall[0]
    sb, sp = sums[1], sums[0]
    res["calls"] = {"chunk_reads": chunk, "calls_per_pass": len(chunks), "warm_passes": reps}
    res["batched"] = {
        "reads_per_s": round(n / tb_, 1), "seconds": round(tb_, 4), "mapped": int(sb["n_mapped"]),
        "first_call_reads_per_s": round(c0 / first[1], 1), "first_call_s": round(first[1], 4),
        "stage_ms": {"fused_kernel": round(sb["gpu_ms_filter"], 2),
                     "staged_fallback": round(sb["gpu_ms_seed"] + sb["gpu_ms_align"] + sb["gpu_ms_replay"], 2),
                     "host_and_copies": round(tb_ * 1e3 - sb["gpu_ms_filter"] - sb["gpu_ms_seed"] - sb["gpu_ms_align"] -
                                              sb["gpu_ms_replay"], 2)},
        "fallback_reads": int(sb["n_fallback"]), "candidates": int(sb["n_candidates"]),
        "filter_passes": int(sb["n_filter_pass"]), "lookups": int(sb["n_lookups"]), "raw_hits": int(sb["n_hits"])}
    gpu_p = sp["gpu_ms_seed"] + sp["gpu_ms_filter"] + sp["gpu_ms_align"] + sp["gpu_ms_replay"]
    res["plain"] = {
        "reads_per_s": round(n / tp, 1), "seconds": round(tp, 4), "mapped": int(sp["n_mapped"]),
        "first_call_reads_per_s": round(c0 / first[0], 1), "first_call_s": round(first[0], 4),
        "stage_ms": {"seeds_lookups_candidates": round(sp["gpu_ms_seed"], 2), "of_which_lookup": round(sp["gpu_ms_lookup"], 2),
                     "filter": round(sp["gpu_ms_filter"], 2), "dp_traceback": round(sp["gpu_ms_align"], 2),
                     "hit_map_replay": round(sp["gpu_ms_replay"], 2), "host_and_copies": round(tp * 1e3 - gpu_p, 2)},
        "index_lookup_GBps": round(8 * sp["n_lookups"] / (sp["gpu_ms_lookup"] * 1e-3) / 1e9, 1)
        if sp["gpu_ms_lookup"] > 0 else None}
    if res["plain"]["index_lookup_GBps"]:
        res["plain"]["index_lookup_hbm_frac"] = round(res["plain"]["index_lookup_GBps"] / HBM_PEAK_GBPS, 4)
        fb = LOOKUP_FETCH_BYTES.get(name)
        if fb:   # the bytes the lookups move: a cache line per random first[key] read
            gbps = fb * sp["n_lookups"] / (sp["gpu_ms_lookup"] * 1e-3) / 1e9
            res["plain"]["index_lookup_fetched"] = {"bytes_per_lookup": fb, "GBps": round(gbps, 1),
                                                    "hbm_frac": round(gbps / HBM_PEAK_GBPS, 3)}
    res["batched_over_plain"] = round(tp / tb_, 2)
    res["plain_records_equal_batched"] = bool(same)
    log(f"{name}: batched {n / tb_:.0f} reads/s ({tb_:.3f}s; first call {first[1]:.3f}s), plain {n / tp:.0f} reads/s "
        f"({tp:.3f}s) -> {tp / tb_:.2f}x; records identical: {same}")
    del bufs
    if sam_reads:
        k = min(n, sam_reads)
        st = m.map_sam(pr, os.devnull, mode=1, lo=0, hi=k)
        res["sam_text"] = {"reads": k, "seconds": round(st["seconds_sam"], 4),
                           "reads_per_s": round(k / max(st["seconds_sam"], 1e-9), 1),
                           "where": "records formatted on the device (k_sam.hip), text copied to the host and written"}
        log(f"{name}: SAM text {k / max(st['seconds_sam'], 1e-9):.0f} reads/s")
    if cpu and ref_max:
        cb = cpu_mapping_baseline(L, m, os.path.dirname(gpath), gpath, pr, kw, color, ref_per_length, ref_max)
        res["cpu_baseline"] = cb
        if cb:
            log(f"{name}: reference CPU {cb['value']} reads/s on {cb['cores']} threads; GPU SAM identical: "
                f"{cb['gpu_sam_identical_to_reference']}")
    m.close()
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
    return map_config(lib, "config1", gpath, pr, kw, 0, 100_000, 100_000, cpu, 100_000, 100_000)


def config3(lib, cpu=True, n_reads=5_000_000, g=None, gpath=None):
    """50 Mbp genome with GC drawn in 40-60 % per 10 kb window; 5M SOLiD color reads of 34-74
    colors after the primer (35-75 bp reads), 3 % color errors, 1 % SNPs, Q10-35; the three
    default weight-12 color seeds, -i 3 -z 85 -t 100 -M 1."""
    from storm_amd import synth
    t0 = time.perf_counter()
    pr = synth.fast_color_reads(g, n_reads, (34, 74), 303, snp=0.01, color_err=0.03, qlo=10, qhi=35)
    log(f"config3: {n_reads} reads generated in {time.perf_counter() - t0:.1f}s")
    kw = dict(allowed_indels=3, z=85, t=100, max_hits=1)
    return map_config(lib, "config3", gpath, pr, kw, 1, n_reads, 1_000_000, cpu, 2000, 82_000)


def config4(lib, cpu=True, n_reads=20_000_000):
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
    return map_config(lib, "config4", gpath, pr, kw, 0, min(n_reads, 4_000_000), 1_000_000, cpu, 40_000, 40_000)


# ------------------------------------------------------------------------------------------ serving
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
    per request collects its hits and SAM records.  Returns per-request latencies (submit -> last
    SAM record ready), the wall time from the first submission to the last result, and per request
    a digest of its hits and SAM text."""
    lat = [0.0] * len(reqs)
    info = [None] * len(reqs)
    digest = [None] * len(reqs)
    errs = []

    def waiter(k, rid):
        try:
            h, res, sam = server.wait_sam(rid)
            lat[k] = res["seconds_latency"]
            info[k] = res
            hh = hashlib.blake2b(digest_size=16)
            for f in ("score", "ref_id", "ref_start", "sense"):
                hh.update(h[f].tobytes())
            digest[k] = (hh.hexdigest(), hashlib.blake2b(sam, digest_size=16).hexdigest(), len(sam))
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
    return np.array(lat), wall, info, digest


def config5(lib, g, gpath, n_req=40, load=0.8, lanes=3, max_batch=1 << 18):
    """Serving: the config-3 genome (50 Mbp, GC 40-60 % per 10 kb window) with nucleotide and color
    indexes resident; Poisson arrivals of requests of 1k-1M reads (log-uniform), lengths 36-150,
    50/50 nucleotide/color, -i 3 or 7 per request, -M 1; admission by free HBM; continuous
    batching, SAM records per request; against the plain per-request path on the same arrivals."""
    from storm_amd import synth
    t0 = time.perf_counter()
    reqs = serve_trace(g, n_req)
    sizes = np.array([len(pr.len) for pr, _ in reqs])
    total = int(sizes.sum())
    log(f"config5: {n_req} requests, {total} reads generated in {time.perf_counter() - t0:.1f}s")
    res = {"requests": n_req, "reads": total, "request_reads_min_max": [int(sizes.min()), int(sizes.max())],
           "latency": "submit -> the request's last SAM record ready (records formatted on the device per batch)"}
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
            srv.wait_sam(rid)
    # capacity: every request submitted at once (closed loop)
    _, wall_c, _, _ = serve_run(srv, reqs, np.zeros(n_req))
    cap = total / wall_c
    log(f"config5: saturated batched capacity {cap:.0f} reads/s ({wall_c:.2f}s)")
    rate = load * cap / sizes.mean()
    arrivals = np.cumsum(synth.rng(55).exponential(1.0 / rate, size=n_req))
    arrivals -= arrivals[0]
    s0 = srv.stats()
    lat_b, wall_b, info_b, dig_b = serve_run(srv, reqs, arrivals)
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
                      "mapped": int(sum(i["n_mapped"] for i in info_b)),
                      "sam_bytes": int(sum(x[2] for x in dig_b))}
    log(f"config5: batched {total / wall_b:.0f} reads/s, p50 {np.percentile(lat_b, 50):.3f}s "
        f"p99 {np.percentile(lat_b, 99):.3f}s")
    # the same arrivals on a batched server that serves requests in admission order
    srv = lib.Server(gpath, lanes=lanes, max_batch_reads=max_batch, schedule=lib.STORM_SCHED_FIFO)
    for _ in range(lanes):
        for rid in [srv.submit(pr, **kw) for pr, kw in warm]:
            srv.wait_sam(rid)
    lat_f, wall_f, _, dig_f = serve_run(srv, reqs, arrivals)
    srv.close()
    res["batched"]["schedule"] = "deadline = submit + 4 x estimated remaining device time (STORM_SCHED_SIZE)"
    res["batched_fifo"] = {"reads_per_s": round(total / wall_f, 1), "wall_s": round(wall_f, 3),
                           "p50_latency_s": round(float(np.percentile(lat_f, 50)), 4),
                           "p99_latency_s": round(float(np.percentile(lat_f, 99)), 4),
                           "schedule": "admission order (STORM_SCHED_FIFO)",
                           "sam_equals_batched": all(a[1] == b[1] for a, b in zip(dig_f, dig_b))}
    log(f"config5: batched FIFO {total / wall_f:.0f} reads/s, p50 {np.percentile(lat_f, 50):.3f}s "
        f"p99 {np.percentile(lat_f, 99):.3f}s")
    t0 = time.perf_counter()
    srv = lib.Server(gpath, mode=lib.STORM_MODE_PLAIN)
    log(f"config5: plain server created in {time.perf_counter() - t0:.1f}s")
    for pr, kw in warm:
        srv.wait_sam(srv.submit(pr, **kw))
    lat_p, wall_p, info_p, dig_p = serve_run(srv, reqs, arrivals)
    # the plain server's capacity: every request submitted at once, as for the batched server
    _, wall_pc, _, _ = serve_run(srv, reqs, np.zeros(n_req))
    srv.close()
    cap_p = total / wall_pc
    res["plain"] = {"reads_per_s": round(total / wall_p, 1), "wall_s": round(wall_p, 3),
                    "p50_latency_s": round(float(np.percentile(lat_p, 50)), 4),
                    "p99_latency_s": round(float(np.percentile(lat_p, 99)), 4),
                    "saturated_reads_per_s": round(cap_p, 1),
                    "mapped": int(sum(i["n_mapped"] for i in info_p))}
    # on the same arrivals (the batched server keeps up with them, so its figure is the offered
    # load; the plain one does not), and capacity against capacity
    res["batched_over_plain"] = round(wall_p / wall_b, 2)
    res["saturated_batched_over_plain"] = round(cap / cap_p, 2)
    res["served_equals_plain_per_request"] = {
        "hits": all(a[0] == b[0] for a, b in zip(dig_b, dig_p)),
        "sam": all(a[1] == b[1] for a, b in zip(dig_b, dig_p))}
    log(f"config5: plain {total / wall_p:.0f} reads/s, p50 {np.percentile(lat_p, 50):.3f}s "
        f"p99 {np.percentile(lat_p, 99):.3f}s; batched/plain {wall_p / wall_b:.2f}x; saturated {cap_p:.0f} reads/s "
        f"({cap / cap_p:.2f}x); identical: {res['served_equals_plain_per_request']}")
    return res


# ------------------------------------------------------------------------------------------ main
def filter_and_dp(lib, a):
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
    spot = None if a.no_cpu else oracle_spot_check(g, reads, scores, cr, cp, L, W)
    n_dp = min(a.dp_records, a.reads)
    dp_ms, dp_res, dp_tb, dp_diag = gpu_align_stage(lib, g, reads, pos, stride, L, 7, d, n_dp, max(1, a.steps // 4))
    dp_gcups = n_dp * L * 15 / (dp_ms * 1e-3) / 1e9
    n_tb = int((dp_res["tb_len"] > 0).sum())
    log(f"align stage: {n_dp} records in {dp_ms:.2f} ms = {dp_gcups:.1f} GCUPS, {n_tb} tracebacks")
    pc_n = min(a.percall, a.reads)
    pc_dt = percall_filter(lib, g, reads, pos, L, pc_n)
    pc_gcups = 2 * pc_n * L * 2 * W / pc_dt / 1e9
    log(f"per-call drop-in row: {pc_dt:.2f}s for {pc_n} calls, {pc_gcups:.3f} GCUPS")
    cpu = None
    dp_cpu = None
    if not a.no_cpu:
        cb = cpu_filter_baseline(g, reads, pos, L, d, n_sample=500_000)
        if "sse2" in cb:
            nc, t = cb["sse2"]
            cpu = {"value": round(nc * L * 2 * W / t / 1e9, 3), "unit": "GCUPS", "cores": cores(), "kind": "reference",
                   "sample": f"the reference's SSE2 filter (alignment_sse2__align_pair, default Makefile flags, "
                             f"libref_filter_sse2.so) over {nc} candidate evaluations of the same workload (first "
                             f"500k reads x 2 windows, repeated), OpenMP on {cores()} threads, {t:.1f}s"}
            if "avx2" in cb:
                nc2, t2 = cb["avx2"]
                cpu["avx2_gcups"] = round(nc2 * L * 2 * W / t2 / 1e9, 3)
            log(f"cpu baseline: {cpu['value']} GCUPS (SSE2), {cpu.get('avx2_gcups')} (AVX2) on {cores()} threads")
        dp_cpu = cpu_align_baseline(g, reads, L, 7, 115, dp_diag, dp_res, dp_tb)
        if dp_cpu:
            log(f"align stage cpu baseline: {dp_cpu['value']} GCUPS on {cores()} threads; GPU records identical: "
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
                             "peak": HBM_PEAK_GBPS, "unit": "GB/s",
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
    }
    return line


def main():
    global HOST
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--reads", type=int, default=5_000_000, help="reads of the filter workload (2 windows each)")
    ap.add_argument("--no-cpu", action="store_true", help="leave out the reference CPU baselines")
    ap.add_argument("--percall", type=int, default=20000)
    ap.add_argument("--dp-records", type=int, default=2_000_000, help="records of the DP stage measurement")
    ap.add_argument("--config", default="all", choices=["all", "2", "1", "3", "4", "5"],
                    help="all (default): every workload in one line; 2: the filter headline alone; 1/3/4/5: one workload")
    ap.add_argument("--reads-c3", type=int, default=5_000_000)
    ap.add_argument("--reads-c4", type=int, default=20_000_000)
    ap.add_argument("--requests", type=int, default=40, help="config 5: requests in the arrival trace")
    ap.add_argument("--calib-copy", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device to device 3 times (calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {a.gpus}); the path shards reads across GPUs by design "
              f"but this build measures one GPU.", file=sys.stderr)
        sys.exit(2)
    import torch
    from storm_amd import lib, synth
    assert torch.cuda.is_available()
    torch.zeros(1, device="cuda:0")
    lib.load()
    HOST = host_info()
    log(f"host: {json.dumps(HOST)}")
    if a.calib_copy:
        src = torch.ones(a.calib_copy << 18, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration: copied {a.calib_copy} MiB device to device 3 times")
        del src, dst
    cpu = not a.no_cpu
    t_all = time.perf_counter()
    line = None
    if a.config in ("all", "2"):
        line = filter_and_dp(lib, a)
    cfgs = {}
    if a.config in ("all", "1"):
        cfgs["1"] = config1(lib, cpu)
    if a.config in ("all", "3", "5"):
        d = tempfile.mkdtemp(prefix="storm_c3_")
        g3 = synth.gc_window_genome(50_000_000, 3)
        g3path = os.path.join(d, "g.fa")
        synth.write_fasta_fast(g3path, "chr1", g3)
        if a.config in ("all", "3"):
            cfgs["3"] = config3(lib, cpu, a.reads_c3, g3, g3path)
        if a.config in ("all", "5"):
            cfgs["5"] = config5(lib, g3, g3path, n_req=a.requests)
    if a.config in ("all", "4"):
        cfgs["4"] = config4(lib, cpu, a.reads_c4)
    if line is None:   # one mapping workload alone: its batched reads/s is the line's value
        k = next(iter(cfgs))
        r = cfgs[k]
        best = r["batched"]
        line = {"metric": METRIC, "value": best["reads_per_s"], "unit": "reads/s", "n_gpus": 1, "steps": 1, "warmup": 1,
                "ms_per_step": round(best.get("seconds", best.get("wall_s", 0)) * 1e3, 2), "higher_is_better": True,
                "scaling": "weak", "vs_baseline": None, "dtype": "u8/int32", "data": "synthetic",
                "config": {"workload": f"mapping config {k}: " + " ".join(
                    {"1": config1, "3": config3, "4": config4, "5": config5}[k].__doc__.split())}}
    if cfgs:
        line["configs"] = cfgs
        line["batched_over_plain"] = {k: v.get("batched_over_plain") for k, v in cfgs.items()}
    line["host"] = HOST
    line["bench_seconds"] = round(time.perf_counter() - t_all, 1)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
