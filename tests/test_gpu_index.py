"""Device index build (storm_index_build, k_index.hip) against the oracle's index (index.c:19-195).

Same inputs on both sides: the packed sequence and masked regions as the oracle's FASTA loader
holds them (N runs, lowercase runs with and without -j, runs at the first and last code, runs
closer than a seed span, a sequence shorter than the seeds), nucleotide and color encodings,
-l 100 / 3 (repeat filter firing) / 0.  Integer arrays: first[] and map[] must be identical.
"""
import os

import numpy as np
import pytest

from storm_amd import lib, synth
from test_oracle_index import make_cases, oracle, oracle_index, oracle_sequence, write_masked_fasta

pytestmark = pytest.mark.gpu

NUC_SEEDS = ["####--#-#--##--####", "#####--###-#-###", "###-##-#######", "##-#", "#-##--#"]
COL_SEEDS = ["####----##--##-####", "####-###--#----####", "####-####-####"]


def check_same(lib_o, path, color, lower, ref_no, seed, l):
    packed, size, rs, re_ = oracle_sequence(lib_o, path, color, lower, ref_no)
    f_o, m_o = oracle_index(lib_o, path, color, lower, ref_no, seed, l)
    f_d, m_d, erased, ms = lib.build_index(packed, size, rs, re_, seed, l)
    print(f"ref {ref_no} size {size} seed {seed} -l {l}: {len(m_d)} positions, {erased} keys erased, {ms:.2f} ms",
          flush=True)
    assert np.array_equal(f_d, f_o), (ref_no, seed, l)
    assert np.array_equal(m_d, m_o), (ref_no, seed, l)
    return erased


@pytest.mark.parametrize("color", [0, 1])
def test_device_index_equals_oracle_masked_cases(tmp_path, oracle_built, color):
    lib_o = oracle(oracle_built)
    path, n_refs = make_cases(tmp_path)
    erased = 0
    for lower in (0, 1):
        for ref_no in range(n_refs):
            for seed in (COL_SEEDS if color else NUC_SEEDS):
                for l in (100, 3, 0):
                    erased += check_same(lib_o, path, color, lower, ref_no, seed, l)
    assert erased > 0


def test_device_index_equals_oracle_repeat_genome(tmp_path, oracle_built):
    """8 Mbp with segmental duplications and tandem repeats (keys with thousands of positions)."""
    lib_o = oracle(oracle_built)
    g = synth.repeat_genome(8_000_000, 11)
    path = os.path.join(str(tmp_path), "rep.fa")
    nn = np.zeros(len(g), bool)
    nn[3_000_000:3_000_500] = True
    write_masked_fasta(path, [("chr1", g, np.zeros(len(g), bool), nn)])
    for seed in NUC_SEEDS[:3]:
        check_same(lib_o, path, 0, 0, 0, seed, 100)
    check_same(lib_o, path, 0, 0, 0, NUC_SEEDS[0], 3)


def test_device_index_structure_at_100mbp():
    """100 Mbp i.i.d. genome (the size of the repeat mapping workload): first[] non-decreasing and
    ending at the number of indexed windows, every key's positions ascending and carrying that key."""
    n = 100_000_000
    g = synth.random_genome(n, 12)
    packed = synth.pack_codes(g)
    seed = NUC_SEEDS[0]
    f, m, erased, ms = lib.build_index(packed, n, [], [], seed, 100)
    print(f"100 Mbp index build: {ms:.1f} ms on the device, {erased} keys erased", flush=True)
    span = len(seed)
    assert f[0] == 0 and f[-1] == len(m) and np.all(np.diff(f) >= 0)
    assert erased == 0 and len(m) == n - span + 1
    assert np.array_equal(np.sort(m), np.arange(n - span + 1, dtype=np.int32))
    offs = np.array([k for k, c in enumerate(seed) if c == "#"])
    rng = np.random.default_rng(3)
    for k in rng.integers(0, 1 << 24, 2000):
        pos = m[f[k]:f[k + 1]]
        assert np.all(np.diff(pos) > 0)
        if len(pos):
            codes = g[pos[:, None] + offs[None, :]].astype(np.int64)
            keys = np.zeros(len(pos), np.int64)
            for j in range(len(offs)):
                keys = (keys << 2) | codes[:, j]
            assert np.all(keys == k)
