"""The host read loader (storm_reads_load with the host path) against the oracle's loader
(reads_data.c:59-67,351-566): names, packed codes and quality levels, color primer fields, tags,
read order after the tag sort and the read-group name."""
import ctypes as C
import os

import numpy as np
import pytest

from reads_files import cases, corrupt_cases


class OrcRead(C.Structure):
    _fields_ = [("info", C.c_char_p), ("first_base", C.c_uint8), ("first_qual", C.c_uint8),
                ("qual", C.POINTER(C.c_uint8)), ("seq", C.POINTER(C.c_uint8)), ("tag", C.c_ulong)]


class OrcReads(C.Structure):
    _fields_ = [("name", C.c_char_p), ("read_len", C.c_int), ("n", C.c_long), ("reads", C.POINTER(OrcRead))]


class OrcParams(C.Structure):
    _fields_ = [(f, C.c_int) for f in ("color", "match", "mismatch", "gap_open", "gap_extend", "allowed_indels",
                                       "simd_diags", "t", "z", "qual_base", "filter_repeats", "stdev_distance",
                                       "n_bytes", "sense_mask", "ignore_lowercase", "unordered_M", "uniqueness",
                                       "uniqueness_threshold", "print_all", "greedy")]


def oracle_reads(oracle_built, path, color):
    lib = C.CDLL(os.path.join(oracle_built, "liboracle.so"))
    lib.orc_params_default.argtypes = [C.POINTER(OrcParams), C.c_int]
    lib.orc_load_reads.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(OrcParams), C.POINTER(OrcReads)]
    lib.orc_load_reads.restype = C.c_long
    lib.orc_sort_reads.argtypes = [C.POINTER(OrcReads)]
    lib.orc_reads_free.argtypes = [C.POINTER(OrcReads)]
    p = OrcParams()
    lib.orc_params_default(C.byref(p), color)
    db = OrcReads()
    assert lib.orc_load_reads(path.encode(), None, C.byref(p), C.byref(db)) > 0
    lib.orc_sort_reads(C.byref(db))
    L = db.read_len
    nb = (L + 3) // 4
    out = {"names": [], "codes": [], "qual": [], "tags": [], "fb": [], "fq": [], "L": L, "db_name": db.name.decode()}
    for i in range(db.n):
        r = db.reads[i]
        out["names"].append(r.info.decode())
        out["codes"].append(bytes(r.seq[:nb]))
        out["qual"].append(bytes(r.qual[:nb]) if r.qual else None)
        out["tags"].append(int(r.tag))
        out["fb"].append(r.first_base)
        out["fq"].append(r.first_qual)
    lib.orc_reads_free(C.byref(db))
    return out


def test_host_loader_matches_oracle(tmp_path, oracle_built):
    from storm_amd import lib
    missing = 0
    for path, color in cases(str(tmp_path)):
        o = oracle_reads(oracle_built, path, color)
        h = lib.load_reads(path, use_device=False, color=color)
        assert not h["on_device"]
        n, L = h["n"], o["L"]
        nb = (L + 3) // 4
        assert n == len(o["names"]) and np.all(h["len"] == L), path
        assert h["names"] == o["names"], path
        assert h["db_name"] == o["db_name"], path
        assert [int(t) for t in h["tags"]] == o["tags"], path
        nq = h["no_qual"] if h["no_qual"] is not None else np.zeros(n, np.uint8)
        for i in range(n):
            assert bytes(h["codes"][i, :nb]) == o["codes"][i], (path, i)
            if o["qual"][i] is None:   # quality line read as a comment: no quality, levels 3
                assert nq[i] == 1 and np.all(h["qual"][i] == 0xFF), (path, i)
            else:
                assert nq[i] == 0 and bytes(h["qual"][i, :nb]) == o["qual"][i], (path, i)
        if color:
            assert list(h["first_base"]) == o["fb"], path
            assert all(h["first_qual"][i] == (3 if nq[i] else o["fq"][i]) for i in range(n)), path
        missing += int(nq.sum())
    assert missing > 0   # some quality lines start with '#' or are empty


def test_host_loader_rejects_corrupt_fastq(tmp_path):
    """Blank lines, empty quality lines and empty sequence lines make the reference's FASTQ reader
    stop with "Reads file seams corrupt" (reads_data.c:429-433); the loader returns an error."""
    from storm_amd import lib
    for path in corrupt_cases(str(tmp_path)):
        with pytest.raises(lib.StormError, match="corrupt"):
            lib.load_reads(path, use_device=False, color=0)
