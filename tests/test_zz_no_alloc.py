"""storm_align_run_records_noalloc: the DP + traceback of device-resident records with the static
direction scratch, against storm_align_run + storm_align_fetch of the same plan (the entry point it
is a variant of), record by record and bit for bit, on two calls in a row (the second smaller),
plus records outside the plan and a caller stream."""
import ctypes as C

import numpy as np
import pytest

from storm_amd import lib
from test_gpu_align import _case

pytestmark = pytest.mark.gpu


def _expected(color, I, t, with_qual, seed):
    genome, mask, reads, quals, lens, fq, rs, diags = _case(seed, color, I, t, with_qual)
    p = lib.params(color, allowed_indels=I, t=t)
    plan = lib.AlignPlan(genome, mask, reads, quals, lens, rs, diags, p, first_qual=fq if color else None)
    plan.run(1)
    res, tb = plan.fetch()
    return plan, rs, diags, res, tb


class _Hip:
    """The HIP runtime the library itself runs on (its libamdhip64.so.7, already loaded), for the
    caller-side device buffers and stream of these tests."""

    def __init__(self):
        lib.load()
        self.h = C.CDLL("libamdhip64.so.7")
        P = C.c_void_p
        self.h.hipMalloc.argtypes = [C.POINTER(P), C.c_size_t]
        self.h.hipFree.argtypes = [P]
        self.h.hipMemcpy.argtypes = [P, P, C.c_size_t, C.c_int]
        self.h.hipStreamCreate.argtypes = [C.POINTER(P)]
        self.h.hipStreamSynchronize.argtypes = [P]
        self.h.hipStreamDestroy.argtypes = [P]

    def ok(self, rc):
        assert rc == 0, f"HIP error {rc}"

    def to_device(self, a: np.ndarray) -> int:
        p = C.c_void_p()
        self.ok(self.h.hipMalloc(C.byref(p), max(a.nbytes, 1)))
        self.ok(self.h.hipMemcpy(p, C.c_void_p(a.ctypes.data), a.nbytes, 1))   # host to device
        return p.value

    def to_host(self, d: int, out: np.ndarray) -> np.ndarray:
        self.ok(self.h.hipMemcpy(C.c_void_p(out.ctypes.data), C.c_void_p(d), out.nbytes, 2))   # device to host
        return out

    def free(self, *ds):
        for d in ds:
            self.ok(self.h.hipFree(C.c_void_p(d)))


def _noalloc(hip, plan, rs, diags, on_own_stream=False):
    n = len(rs)
    d_rs = hip.to_device(np.ascontiguousarray(rs, np.int32))
    d_dg = hip.to_device(np.ascontiguousarray(diags, np.int32))
    d_res = hip.to_device(np.zeros(n * lib.ALIGN_RESULT.itemsize, np.uint8))
    d_tb = hip.to_device(np.zeros(n * lib.TB_BYTES, np.uint8))
    if on_own_stream:   # the caller's stream: the call returns at once, the caller waits
        st = C.c_void_p()
        hip.ok(hip.h.hipStreamCreate(C.byref(st)))
        plan.run_records_noalloc(d_rs, d_dg, n, d_res, d_tb, st.value)
        hip.ok(hip.h.hipStreamSynchronize(st))
        hip.ok(hip.h.hipStreamDestroy(st))
    else:
        plan.run_records_noalloc(d_rs, d_dg, n, d_res, d_tb)
    res = hip.to_host(d_res, np.zeros(n * lib.ALIGN_RESULT.itemsize, np.uint8)).view(lib.ALIGN_RESULT)
    tb = hip.to_host(d_tb, np.zeros(n * lib.TB_BYTES, np.uint8)).reshape(n, lib.TB_BYTES)
    hip.free(d_rs, d_dg, d_res, d_tb)
    return res, tb


def _same(res_a, tb_a, res_b, tb_b, where):
    assert np.array_equal(res_a, res_b), where
    for k in range(len(res_a)):
        nb = (int(res_a[k]["tb_len"]) + 1) // 2
        assert np.array_equal(tb_a[k, :nb], tb_b[k, :nb]), (where, k)


@pytest.mark.parametrize("color,I,t", [(0, 31, 60), (1, 7, 80)], ids=lambda v: str(v))
def test_noalloc_equals_align_run_twice(color, I, t):
    hip = _Hip()
    plan, rs, diags, res, tb = _expected(color, I, t, True, 7000 + 10 * I + color)
    n = len(rs)
    # call 1: every record, repeated past one launch's share of the array (at -i 31 and 256-code
    # reads a launch holds 2048 records), in a shuffled order
    rng = np.random.default_rng(11)
    idx1 = rng.permutation(np.tile(np.arange(n), 5000 // n + 1))[:5000]
    print(f"noalloc color={color} -i {I}: call 1 with {len(idx1)} records", flush=True)
    r1, t1 = _noalloc(hip, plan, rs[idx1], diags[idx1])
    _same(r1, t1, res[idx1], tb[idx1], "call 1")
    assert (r1["tb_len"] > 0).any()
    # call 2: fewer records, other ones, on the caller's stream
    idx2 = rng.choice(n, size=min(n, 97), replace=False)
    print(f"noalloc color={color} -i {I}: call 2 with {len(idx2)} records", flush=True)
    r2, t2 = _noalloc(hip, plan, rs[idx2], diags[idx2], on_own_stream=True)
    _same(r2, t2, res[idx2], tb[idx2], "call 2")
    plan.close()


def test_noalloc_records_outside_the_plan():
    hip = _Hip()
    plan, rs, diags, res, tb = _expected(0, 3, 100, True, 7100)
    n_reads = int(rs.max()) // 2 + 1
    bad_rs = np.array([rs[0], 2 * n_reads, -1, rs[1]], np.int32)
    bad_dg = np.array([diags[0], diags[0], diags[0], 1 << 30], np.int32)
    r, _ = _noalloc(hip, plan, bad_rs, bad_dg)
    assert np.array_equal(r[:1], res[:1])
    for k in (1, 2, 3):
        assert (r[k]["score"], r[k]["bound"], r[k]["best_j"], r[k]["tb_len"]) == (lib.MIN_SCORE, lib.MIN_SCORE, -2, 0)
    # no records: nothing to do; null buffers: refused
    plan.run_records_noalloc(0, 0, 0, 0, 0)
    with pytest.raises(lib.StormError):
        plan.run_records_noalloc(0, 0, 5, 0, 0)
    plan.close()
