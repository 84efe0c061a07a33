"""The entry points with the reference's own signatures and struct layouts
(include/storm_ref_dropin.h) inside a program built from the reference's sources
(oracle/_ref/dropin-check-*): index_gfx950__build_reference against index__build_reference
(sizes, first[], every key's positions through the reference's index__get_extern_next_hit), with
masked N runs and the repeat filter firing, and alignment_gfx950__align_batch against
alignment__align + alignment__traceback record by record under three thresholds."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu
REF = os.path.join(ROOT, "oracle", "_ref")
SEEDS_N = "####--#-#--##--####;#####--###-#-###;###-##-#######"
SEEDS_C = "####----##--##-####;####-###--#----####;####-####-####"


def _genome(path, seed):
    from storm_amd import synth
    g = synth.repeat_genome(300_000, seed, 0.2, 0.15)
    letters = bytearray(synth.BASES[g].tobytes())
    rng = np.random.default_rng(seed)
    for _ in range(6):   # N runs: masked regions
        a = int(rng.integers(1000, len(letters) - 3000))
        b = a + int(rng.integers(20, 900))
        letters[a:b] = b"N" * (b - a)
    with open(path, "wb") as f:
        f.write(b">chr1\n")
        for k in range(0, len(letters), 60):
            f.write(bytes(letters[k:k + 60]) + b"\n")
    return g


@pytest.mark.parametrize("color,I,t,l", [(0, 3, 100, 100), (0, 7, 115, 3), (1, 3, 100, 100), (1, 7, 90, 3)],
                         ids=lambda v: str(v))
def test_dropins_equal_reference(color, I, t, l, tmp_path):
    from storm_amd import synth
    exe = os.path.join(REF, "dropin-check-color" if color else "dropin-check-nucleotide")
    if not os.path.exists(exe):
        pytest.skip("oracle/_ref/dropin-check-* not built")
    d = str(tmp_path)
    gpath = os.path.join(d, "g.fa")
    g = _genome(gpath, 11 + I + color)
    if color:
        pr = synth.fast_color_reads(g, 300, (49, 49), 5 + I, snp=0.01, color_err=0.03, qlo=10, qhi=35)
        sr = synth.packed_to_simreads(pr, np.arange(300))
        synth.write_csfasta(os.path.join(d, "r.csfasta"), os.path.join(d, "r.qual"), sr)
        rargs = [os.path.join(d, "r.csfasta"), os.path.join(d, "r.qual")]
    else:
        pr = synth.fast_nucleotide_reads(g, 300, 60, 5 + I, sub=0.02, indel=0.003, qlo=5, qhi=40)
        synth.write_fastq(os.path.join(d, "r.fq"), synth.packed_to_simreads(pr, np.arange(300)))
        rargs = [os.path.join(d, "r.fq"), "-"]
    print(f"drop-in check color={color} -i {I} -t {t} -l {l}", flush=True)
    out = subprocess.run([exe, gpath] + rargs + [SEEDS_C if color else SEEDS_N, str(I), str(t), str(l)],
                         capture_output=True, text=True, timeout=300)
    print(out.stdout, out.stderr[-2000:])
    assert out.returncode == 0
    f = out.stdout.split()
    stats = dict(zip(f[0::2], map(int, f[1::2])))
    assert stats["mismatches"] == 0 and stats["keys"] > 1000 and stats["records"] > 500 and stats["accepted"] > 100
