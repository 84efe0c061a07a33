"""The CPU restatement against the reference's own outputs (golden fixtures written by the
compiled reference, tests/golden/make_golden.py): whole -M SAM files and per-candidate filter
scores."""
import ctypes as C
import glob
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, sam_body, sam_cases, unpack_case


@pytest.mark.parametrize("case", sam_cases())
def test_oracle_sam_matches_reference(case, tmp_path, oracle_built):
    meta, paths = unpack_case(case, str(tmp_path))
    binary = os.path.join(oracle_built, "oracle-nucleotide" if meta["mode"] == "n" else "oracle-color")
    if meta["mode"] == "n":
        rargs = ["-r", paths["reads.fq"]]
    else:
        rargs = ["-r", paths["reads.csfasta"], "-q", paths["reads.qual"]]
    out = str(tmp_path / "oracle.sam")
    subprocess.run([binary, "-g", paths["genome.fa"]] + rargs + meta["args"] + ["-o", out], check=True,
                   stderr=subprocess.DEVNULL)
    assert sam_body(out) == sam_body(paths["ref.sam"])


def _oracle_lib(path):
    lib = C.CDLL(os.path.join(path, "liboracle.so"))
    lib.orc_filter_score.restype = C.c_int
    lib.orc_filter_score.argtypes = [C.c_void_p, C.c_long, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                     C.c_int]
    return lib


@pytest.mark.parametrize("fx", sorted(glob.glob(os.path.join(GOLDEN, "filter_*.npz"))), ids=os.path.basename)
def test_oracle_filter_matches_reference_sse2(fx, oracle_built):
    z = np.load(fx)
    lib = _oracle_lib(oracle_built)
    gp, pad, rp, stride = z["genome_packed"], int(z["pad"]), z["reads_packed"], int(z["stride"])
    L, W = int(z["L"]), int(z["W"])
    got = np.array([lib.orc_filter_score(C.c_void_p(gp.ctypes.data + pad), int(p),
                                         C.c_void_p(rp.ctypes.data + int(r) * stride), L, W, 5, -4, -9, -4)
                    for p, r in zip(z["pos"], z["read_of"])])
    assert np.array_equal(np.minimum(got, 255), z["scores"].astype(np.int64))
