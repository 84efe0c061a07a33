"""ctypes binding of libstorm_hip.so (include/storm_gpu.h).

The HIP library is required: importing the binding fails loudly when it has not been built,
there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# STORM_LIB names another build of the library (the phase-timer measurement build:
# make -C storm_amd/csrc phases)
LIB_PATH = os.environ.get("STORM_LIB") or os.path.join(HERE, "_build", "libstorm_hip.so")

STORM_MODE_PLAIN = 0
STORM_MODE_BATCHED = 1
STORM_SCHED_FIFO = 0
STORM_SCHED_SIZE = 1


class StormError(RuntimeError):
    pass


class Params(C.Structure):
    """storm_params: the reference's -M mapping options (main.c:188-433)."""
    _fields_ = [(n, C.c_int) for n in (
        "color", "match", "mismatch", "gap_open", "gap_extend", "allowed_indels", "t", "z", "qual_base",
        "stdev_distance", "sense_mask", "max_hits", "uniqueness", "print_all", "ignore_lowercase",
        "strict_length")] + [("seeds", C.c_char_p), ("greedy", C.c_int), ("unmapped_path", C.c_char_p)]


class MapStats(C.Structure):
    _fields_ = [(n, C.c_int64) for n in ("n_reads", "n_mapped", "n_candidates", "n_filter_pass", "n_accepted",
                                          "n_lookups", "n_hits", "n_fallback")] + \
               [(n, C.c_double) for n in ("seconds_load", "seconds_index", "seconds_map", "seconds_sam",
                                          "gpu_ms_seed", "gpu_ms_filter", "gpu_ms_align", "gpu_ms_replay",
                                          "gpu_ms_lookup")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ReadsView(C.Structure):
    _fields_ = [("n_reads", C.c_int64), ("names", C.POINTER(C.c_char_p)), ("codes", C.c_void_p),
                ("code_stride", C.c_int64), ("qual", C.c_void_p), ("qual_stride", C.c_int64),
                ("len", C.c_void_p), ("first_base", C.c_void_p), ("first_qual", C.c_void_p),
                ("db_name", C.c_char_p), ("no_qual", C.c_void_p)]


class ServerConfig(C.Structure):
    """storm_server_config (include/storm_gpu.h)."""
    _fields_ = [("mode", C.c_int), ("lanes", C.c_int), ("cross_request", C.c_int), ("max_batch_reads", C.c_int64),
                ("hbm_budget_bytes", C.c_int64), ("hbm_reserve_bytes", C.c_int64), ("encodings", C.c_int),
                ("seeds_nucleotide", C.c_char_p), ("seeds_color", C.c_char_p), ("stdev_distance", C.c_int),
                ("ignore_lowercase", C.c_int), ("sam_records", C.c_int),
                ("schedule", C.c_int)]


class ServerStats(C.Structure):
    _fields_ = [(n, C.c_int64) for n in (
        "submitted", "completed", "failed", "reads_done", "batches", "batch_reads", "batch_requests",
        "max_requests_per_batch", "max_lanes_busy", "admission_waits", "arena_bytes", "arena_used_bytes",
        "arena_peak_bytes", "pending", "active", "n_fallback", "candidates")] + [("gpu_busy_seconds", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class RequestResult(C.Structure):
    _fields_ = [("n_reads", C.c_int64), ("n_mapped", C.c_int64), ("status", C.c_int), ("batches", C.c_int),
                ("seconds_queued", C.c_double), ("seconds_latency", C.c_double), ("t_submit", C.c_double),
                ("t_done", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise StormError(f"{LIB_PATH} is missing: run __graft_entry__.build() (or make -C storm_amd/csrc)")
    lib = C.CDLL(LIB_PATH)
    P, I, L, D = C.c_void_p, C.c_int, C.c_int64, C.c_double
    lib.storm_last_error.restype = C.c_char_p
    lib.storm_params_default.argtypes = [C.POINTER(Params), I]
    lib.storm_filter_plan_create.argtypes = [I, P, L, P, L, P, L, P, P, L, I, I, I, I, I, I, C.POINTER(P)]
    lib.storm_filter_run.argtypes = [P, I, C.POINTER(D)]
    lib.storm_filter_fetch.argtypes = [P, P, P]
    lib.storm_filter_plan_destroy.argtypes = [P]
    lib.storm_align_plan_create.argtypes = [I, P, P, L, P, L, P, L, P, P, L, P, P, L, C.POINTER(Params),
                                            C.POINTER(P)]
    lib.storm_align_run.argtypes = [P, I, C.POINTER(D)]
    lib.storm_align_fetch.argtypes = [P, P, P]
    lib.storm_align_plan_destroy.argtypes = [P]
    lib.storm_align_run_records_noalloc.argtypes = [P, P, P, L, P, P, P]
    lib.storm_map_files.argtypes = [I, C.c_char_p, C.c_char_p, C.c_char_p, C.POINTER(Params), I, C.c_char_p, I,
                                    P, C.POINTER(MapStats)]
    lib.storm_mapper_create.argtypes = [I, C.c_char_p, C.POINTER(Params), C.POINTER(P)]
    lib.storm_mapper_map.argtypes = [P, C.POINTER(ReadsView), I, C.c_char_p, I, P, C.POINTER(MapStats)]
    lib.storm_mapper_map_hits.argtypes = [P, C.POINTER(ReadsView), I, P, P, P, P, C.POINTER(MapStats)]
    lib.storm_mapper_map_records.argtypes = [P, C.POINTER(ReadsView), I, P, P, C.POINTER(MapStats)]
    lib.storm_mapper_free_bytes.argtypes = [P]
    lib.storm_mapper_free_bytes.restype = L
    lib.storm_mapper_destroy.argtypes = [P]
    lib.storm_server_config_default.argtypes = [C.POINTER(ServerConfig)]
    lib.storm_server_create.argtypes = [I, C.c_char_p, C.POINTER(ServerConfig), C.POINTER(P)]
    lib.storm_server_submit.argtypes = [P, C.POINTER(ReadsView), C.POINTER(Params), C.POINTER(L)]
    lib.storm_server_wait.argtypes = [P, L, P, P, P, P, C.POINTER(RequestResult)]
    lib.storm_server_wait_sam.argtypes = [P, L, C.POINTER(C.c_void_p), C.POINTER(L), P, P, P, P,
                                          C.POINTER(RequestResult)]
    lib.storm_free.argtypes = [P]
    lib.storm_server_get_stats.argtypes = [P, C.POINTER(ServerStats)]
    lib.storm_server_destroy.argtypes = [P]
    lib.storm_index_build.argtypes = [I, P, L, P, P, I, C.c_char_p, I, C.POINTER(P)]
    lib.storm_index_sizes.argtypes = [P, C.POINTER(L), C.POINTER(L), C.POINTER(L), C.POINTER(D)]
    lib.storm_index_fetch.argtypes = [P, P, P]
    lib.storm_index_destroy.argtypes = [P]
    lib.storm_reads_load.argtypes = [I, C.c_char_p, C.c_char_p, C.POINTER(Params), I, C.POINTER(P)]
    lib.storm_reads_get.argtypes = [P, C.POINTER(ReadsView), P, C.POINTER(I), C.POINTER(D)]
    lib.storm_reads_destroy.argtypes = [P]
    lib.alignment_gfx950__set_genome.argtypes = [P, C.c_size_t, C.c_size_t]
    for w in ("octa", "quad", "pair", "mono"):
        getattr(lib, f"alignment_gfx950__init_{w}").argtypes = [C.c_uint] * 6
        f = getattr(lib, f"alignment_gfx950__align_{w}")
        f.argtypes = [P, P, P]
        f.restype = C.c_uint
    _lib = lib
    return lib


def check(rc: int) -> None:
    if rc != 0:
        raise StormError(load().storm_last_error().decode())


def params(color: int = 0, **kw) -> Params:
    p = Params()
    load().storm_params_default(C.byref(p), color)
    for k, v in kw.items():
        if k in ("seeds", "unmapped_path"):
            v = v.encode() if isinstance(v, str) else v
        setattr(p, k, v)
    return p


def _ptr(a: np.ndarray) -> C.c_void_p:
    return C.c_void_p(a.ctypes.data)


class FilterPlan:
    """Batch filter over candidates with all inputs resident in HBM (storm_filter_plan_*)."""

    def __init__(self, genome_codes: np.ndarray, reads_codes, read_len: np.ndarray, cand_read: np.ndarray,
                 cand_pos: np.ndarray, W: int, match=5, mismatch=-4, gap_open=-9, gap_extend=-4, z=100, device=0):
        from .synth import pack_codes
        lib = load()
        g = pack_codes(np.ascontiguousarray(genome_codes, dtype=np.uint8))
        read_len = np.ascontiguousarray(read_len, dtype=np.int32)
        stride = (int(read_len.max()) + 3) // 4 + 8
        rp = np.zeros(len(read_len) * stride + 64, np.uint8)
        for k in range(len(read_len)):
            rp[k * stride:k * stride + (int(read_len[k]) + 3) // 4] = pack_codes(np.asarray(reads_codes[k], np.uint8))
        self._keep = (g, rp)
        cand_read = np.ascontiguousarray(cand_read, dtype=np.int32)
        cand_pos = np.ascontiguousarray(cand_pos, dtype=np.int32)
        self.n = len(cand_read)
        h = C.c_void_p()
        check(lib.storm_filter_plan_create(device, _ptr(g), len(genome_codes), _ptr(rp), stride, _ptr(read_len),
                                           len(read_len), _ptr(cand_read), _ptr(cand_pos), self.n, W, match, mismatch,
                                           gap_open, gap_extend, z, C.byref(h)))
        self.h = h

    def run(self, reps: int = 1) -> float:
        ms = C.c_double()
        check(load().storm_filter_run(self.h, reps, C.byref(ms)))
        return ms.value

    def fetch(self):
        s = np.zeros(self.n, np.uint8)
        p = np.zeros(self.n, np.uint8)
        check(load().storm_filter_fetch(self.h, _ptr(s), _ptr(p)))
        return s, p

    def close(self):
        if self.h:
            load().storm_filter_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


ALIGN_RESULT = np.dtype([("score", "<i4"), ("bound", "<i4"), ("win", "<i4"), ("best_j", "<i2"),
                         ("best_i0", "<i2"), ("best_j0", "<i2"), ("tb_len", "<i2")])
TB_BYTES = 144
MIN_SCORE = -16383
# storm_hit (HitInfoType, hit_map.h:17-40)
HIT_RECORD = np.dtype([(f, "<i4") for f in ("score", "ref_id", "ref_start", "ref_end", "read_start", "read_end",
                                             "sense", "tb_len")])


class AlignPlan:
    """Batch DP + traceback over (read strand, diagonal) records with all inputs resident in HBM
    (storm_align_plan_*).  reads_codes: per read its forward codes; qual: per read its levels
    0..3, or None (every level 3); first_qual: color space, per read the primer-side level."""

    def __init__(self, genome_codes: np.ndarray, genome_mask, reads_codes, qual, read_len: np.ndarray,
                 rec_rs: np.ndarray, rec_diag: np.ndarray, p: Params, first_qual=None, device=0):
        from .synth import pack_codes
        lib = load()
        g = pack_codes(np.ascontiguousarray(genome_codes, dtype=np.uint8))
        gm = None if genome_mask is None else np.ascontiguousarray(genome_mask, dtype=np.uint8)
        read_len = np.ascontiguousarray(read_len, dtype=np.int32)
        n = len(read_len)
        stride = (int(read_len.max()) + 3) // 4 if n else 1
        rp = np.zeros(n * stride + 16, np.uint8)
        qp = None if qual is None else np.zeros(n * stride + 16, np.uint8)
        for k in range(n):
            nb = (int(read_len[k]) + 3) // 4
            rp[k * stride:k * stride + nb] = pack_codes(np.asarray(reads_codes[k], np.uint8))
            if qp is not None:
                qp[k * stride:k * stride + nb] = pack_codes(np.asarray(qual[k], np.uint8))
        fq = None if first_qual is None else np.ascontiguousarray(first_qual, dtype=np.uint8)
        rec_rs = np.ascontiguousarray(rec_rs, dtype=np.int32)
        rec_diag = np.ascontiguousarray(rec_diag, dtype=np.int32)
        self._keep = (g, gm, rp, qp, fq, read_len, rec_rs, rec_diag)
        self.n = len(rec_rs)
        h = C.c_void_p()
        nul = C.c_void_p()
        check(lib.storm_align_plan_create(device, _ptr(g), nul if gm is None else _ptr(gm), len(genome_codes),
                                          _ptr(rp), stride, nul if qp is None else _ptr(qp), stride,
                                          nul if fq is None else _ptr(fq), _ptr(read_len), n, _ptr(rec_rs),
                                          _ptr(rec_diag), self.n, C.byref(p), C.byref(h)))
        self.h = h

    def run(self, reps: int = 1) -> float:
        ms = C.c_double()
        check(load().storm_align_run(self.h, reps, C.byref(ms)))
        return ms.value

    def fetch(self):
        """(results, tb): the structured storm_align_result array and n x TB_BYTES traceback bytes."""
        res = np.zeros(self.n, ALIGN_RESULT)
        tb = np.zeros((self.n, TB_BYTES), np.uint8)
        check(load().storm_align_fetch(self.h, _ptr(res), _ptr(tb)))
        return res, tb

    def run_records_noalloc(self, d_rec_rs: int, d_rec_diag: int, n_rec: int, d_results: int, d_tb: int,
                            stream: int = 0) -> None:
        """storm_align_run_records_noalloc: records and outputs are device addresses (e.g. the
        data_ptr() of device tensors); stream 0 = the plan's stream, waited for."""
        check(load().storm_align_run_records_noalloc(self.h, C.c_void_p(d_rec_rs), C.c_void_p(d_rec_diag), n_rec,
                                                     C.c_void_p(d_results), C.c_void_p(d_tb),
                                                     C.c_void_p(stream) if stream else None))

    def close(self):
        if self.h:
            load().storm_align_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def build_index(packed: np.ndarray, size: int, region_start, region_end, seed: str, stdev_distance: int = 100,
                device: int = 0):
    """storm_index_*: the CSR index of one seed over one packed sequence, built on the device.

    Returns (first, map, n_erased, device_ms): first has 4^w + 1 entries, map first[4^w]
    (index.c:157-188 layout)."""
    lib = load()
    packed = np.ascontiguousarray(packed, dtype=np.uint8)
    rs = np.ascontiguousarray(region_start, dtype=np.int32)
    re_ = np.ascontiguousarray(region_end, dtype=np.int32)
    h = C.c_void_p()
    check(lib.storm_index_build(device, _ptr(packed), int(size), _ptr(rs), _ptr(re_), len(rs), seed.encode(),
                                int(stdev_distance), C.byref(h)))
    try:
        nk, nm, ne, ms = C.c_int64(), C.c_int64(), C.c_int64(), C.c_double()
        check(lib.storm_index_sizes(h, C.byref(nk), C.byref(nm), C.byref(ne), C.byref(ms)))
        first = np.zeros(nk.value + 1, np.int32)
        mp = np.zeros(max(nm.value, 1), np.int32)
        check(lib.storm_index_fetch(h, _ptr(first), _ptr(mp)))
        return first, mp[:nm.value], ne.value, ms.value
    finally:
        lib.storm_index_destroy(h)


def load_reads(path: str, qual: str | None = None, use_device: bool = True, device: int = 0, **kw) -> dict:
    """storm_reads_*: a reads file loaded into packed arrays ordered by tag (load_reads_db +
    sort_reads_db).  FASTQ files are parsed on the device when use_device is set."""
    lib = load()
    p = params(**kw)
    h = C.c_void_p()
    check(lib.storm_reads_load(device, path.encode(), qual.encode() if qual else None, C.byref(p), int(use_device),
                               C.byref(h)))
    try:
        v = ReadsView()
        n = None
        dev, ms = C.c_int(), C.c_double()
        check(lib.storm_reads_get(h, C.byref(v), None, C.byref(dev), C.byref(ms)))
        n = v.n_reads
        tags = np.zeros(max(n, 1), np.uint64)
        check(lib.storm_reads_get(h, C.byref(v), _ptr(tags), None, None))

        def arr(ptr, count, dtype):
            if not ptr or count == 0:
                return None
            ct = {np.uint8: C.c_uint8, np.int32: C.c_int32}[dtype]
            return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ct)), shape=(count,)).copy()
        out = {
            "n": n,
            "names": [v.names[i].decode() for i in range(n)],
            "codes": arr(v.codes, n * v.code_stride, np.uint8).reshape(n, v.code_stride),
            "qual": None if not v.qual else arr(v.qual, n * v.qual_stride, np.uint8).reshape(n, v.qual_stride),
            "len": arr(v.len, n, np.int32),
            "first_base": arr(v.first_base, n, np.uint8),
            "first_qual": arr(v.first_qual, n, np.uint8),
            "db_name": v.db_name.decode() if v.db_name else None,
            "no_qual": arr(v.no_qual, n, np.uint8),
            "tags": tags[:n],
            "on_device": bool(dev.value),
            "device_ms": ms.value,
        }
        return out
    finally:
        lib.storm_reads_destroy(h)


def map_files(genome: str, reads: str, sam: str, qual: str | None = None, mode: int = STORM_MODE_BATCHED,
              device: int = 0, argv=None, **kw) -> dict:
    """storm_map_files: the reference's whole -M run on the GPU, SAM written to `sam`."""
    lib = load()
    p = params(**kw)
    st = MapStats()
    argv = argv or ["storm-gpu"]
    arr = (C.c_char_p * len(argv))(*[a.encode() for a in argv])
    check(lib.storm_map_files(device, genome.encode(), reads.encode(), qual.encode() if qual else None, C.byref(p),
                              mode, sam.encode(), len(argv), C.cast(arr, C.c_void_p), C.byref(st)))
    return st.as_dict()


class Mapper:
    """storm_mapper_*: genome and seed indexes resident in HBM, reads mapped from host arrays.

    `reads` is any object with the fields of synth.PackedReads: codes (n, stride) u8 packed
    codes, qual (n, stride) u8 packed 2-bit levels or None, len (n,) i32, and for color reads
    first_base / first_qual (n,) u8.
    """

    def __init__(self, genome: str, device: int = 0, **kw):
        lib = load()
        self.params = params(**kw)
        h = C.c_void_p()
        check(lib.storm_mapper_create(device, genome.encode(), C.byref(self.params), C.byref(h)))
        self.h = h

    @staticmethod
    def view(reads, lo: int = 0, hi: int | None = None, names=None) -> tuple[ReadsView, tuple]:
        hi = len(reads.len) if hi is None else hi
        codes = np.ascontiguousarray(reads.codes[lo:hi])
        qual = None if reads.qual is None else np.ascontiguousarray(reads.qual[lo:hi])
        lens = np.ascontiguousarray(reads.len[lo:hi], dtype=np.int32)
        fb = None if reads.first_base is None else np.ascontiguousarray(reads.first_base[lo:hi])
        fq = None if reads.first_qual is None else np.ascontiguousarray(reads.first_qual[lo:hi])
        v = ReadsView()
        v.n_reads = hi - lo
        keep = [codes, qual, lens, fb, fq]
        if names is not None:
            arr = (C.c_char_p * (hi - lo))(*[n.encode() for n in names[lo:hi]])
            keep.append(arr)
            v.names = C.cast(arr, C.POINTER(C.c_char_p))
        v.codes = codes.ctypes.data
        v.code_stride = codes.shape[1]
        v.qual = None if qual is None else qual.ctypes.data
        v.qual_stride = 0 if qual is None else qual.shape[1]
        v.len = lens.ctypes.data
        v.first_base = None if fb is None else fb.ctypes.data
        v.first_qual = None if fq is None else fq.ctypes.data
        v.db_name = None
        return v, tuple(keep)

    def map_sam(self, reads, sam: str, mode: int = STORM_MODE_BATCHED, lo: int = 0, hi: int | None = None,
                names=None, argv=None) -> dict:
        v, keep = self.view(reads, lo, hi, names)
        st = MapStats()
        argv = argv or ["storm-gpu"]
        arr = (C.c_char_p * len(argv))(*[a.encode() for a in argv])
        check(load().storm_mapper_map(self.h, C.byref(v), mode, sam.encode(), len(argv), C.cast(arr, C.c_void_p),
                                      C.byref(st)))
        del keep
        return st.as_dict()

    def map_hits(self, reads, mode: int = STORM_MODE_BATCHED, lo: int = 0, hi: int | None = None):
        """Top-4 hits per read: dict of (n, 4) int32 arrays score/ref_id/ref_start/sense, and stats."""
        v, keep = self.view(reads, lo, hi)
        n = v.n_reads
        out = {k: np.zeros((n, 4), np.int32) for k in ("score", "ref_id", "ref_start", "sense")}
        st = MapStats()
        check(load().storm_mapper_map_hits(self.h, C.byref(v), mode, _ptr(out["score"]), _ptr(out["ref_id"]),
                                           _ptr(out["ref_start"]), _ptr(out["sense"]), C.byref(st)))
        del keep
        return out, st.as_dict()

    def map_records(self, reads, mode: int = STORM_MODE_BATCHED, lo: int = 0, hi: int | None = None,
                    with_tb: bool = True, out=None):
        """The whole hit map in host memory (storm_mapper_map_records): (hits, tb, stats) with hits an
        (n, 4) HIT_RECORD array and tb an (n, M, TB_BYTES) u8 array (M = -M, 4 in the ordered mode)
        or None.  `out` = (hits, tb) arrays to reuse (at least n rows)."""
        v, keep = self.view(reads, lo, hi)
        n = v.n_reads
        M = self.params.max_hits or 4
        if out is None:
            hits = np.zeros((n, 4), HIT_RECORD)
            tb = np.zeros((n, M, TB_BYTES), np.uint8) if with_tb else None
        else:
            hits, tb = out
            hits = hits[:n]
            tb = tb[:n] if (with_tb and tb is not None) else None
        st = MapStats()
        check(load().storm_mapper_map_records(self.h, C.byref(v), mode, _ptr(hits),
                                              None if tb is None else _ptr(tb), C.byref(st)))
        del keep
        return hits, tb, st.as_dict()

    def records_call(self, v: ReadsView, mode: int, hits: np.ndarray, tb) -> dict:
        """storm_mapper_map_records on a prepared view (Mapper.view) into the given arrays."""
        st = MapStats()
        check(load().storm_mapper_map_records(self.h, C.byref(v), mode, _ptr(hits), None if tb is None else _ptr(tb),
                                              C.byref(st)))
        return st.as_dict()

    def free_bytes(self) -> int:
        return int(load().storm_mapper_free_bytes(self.h))

    def close(self):
        if self.h:
            load().storm_mapper_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """storm_server_*: genome and indexes resident in HBM (both encodings), requests mapped by
    continuous batching (mode=STORM_MODE_BATCHED) or one at a time with the staged kernels
    (mode=STORM_MODE_PLAIN, the per-request baseline)."""

    def __init__(self, genome: str, device: int = 0, **cfg):
        lib = load()
        c = ServerConfig()
        lib.storm_server_config_default(C.byref(c))
        for k, v in cfg.items():
            if k in ("seeds_nucleotide", "seeds_color") and isinstance(v, str):
                v = v.encode()
            setattr(c, k, v)
        self.cfg = c
        h = C.c_void_p()
        check(lib.storm_server_create(device, genome.encode(), C.byref(c), C.byref(h)))
        self.h = h
        self._keep = {}

    def submit(self, reads, lo: int = 0, hi: int | None = None, **kw) -> int:
        """Queues reads[lo:hi] (synth.PackedReads fields) with the given storm_params fields."""
        v, keep = Mapper.view(reads, lo, hi)
        p = params(**kw)
        rid = C.c_int64()
        check(load().storm_server_submit(self.h, C.byref(v), C.byref(p), C.byref(rid)))
        self._keep[rid.value] = v.n_reads   # the server copied the reads; only the size is kept
        del keep
        return rid.value

    def wait(self, rid: int):
        """Blocks until request `rid` is done: (dict of (n, 4) int32 arrays, result dict)."""
        n = self._keep.pop(rid)
        out = {k: np.zeros((n, 4), np.int32) for k in ("score", "ref_id", "ref_start", "sense")}
        res = RequestResult()
        rc = load().storm_server_wait(self.h, rid, _ptr(out["score"]), _ptr(out["ref_id"]), _ptr(out["ref_start"]),
                                      _ptr(out["sense"]), C.byref(res))
        check(rc)
        return out, res.as_dict()

    def wait_sam(self, rid: int):
        """As wait(), plus the request's SAM records as bytes: (hits, result, sam)."""
        n = self._keep.pop(rid)
        out = {k: np.zeros((n, 4), np.int32) for k in ("score", "ref_id", "ref_start", "sense")}
        res = RequestResult()
        text = C.c_void_p()
        nb = C.c_int64()
        rc = load().storm_server_wait_sam(self.h, rid, C.byref(text), C.byref(nb), _ptr(out["score"]),
                                          _ptr(out["ref_id"]), _ptr(out["ref_start"]), _ptr(out["sense"]),
                                          C.byref(res))
        check(rc)
        sam = C.string_at(text, nb.value)
        load().storm_free(text)
        return out, res.as_dict(), sam

    def stats(self) -> dict:
        st = ServerStats()
        check(load().storm_server_get_stats(self.h, C.byref(st)))
        return st.as_dict()

    def close(self):
        if self.h:
            load().storm_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
