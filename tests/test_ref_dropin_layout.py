"""The boundary against the reference's own headers (CPU, no GPU call; skipped without the
reference sources):
  * the library's mirrors of ReferenceDBType, MaskedRegionType, SeedType, IndexType,
    AlignmentParamType and AlignmentType (storm_ref_abi_layout) have the offsets the reference's
    headers give, in both builds (storm-nucleotide with NUCLEOTIDES, storm-color without);
  * every C binding INTEGRATION.md marks as compiling with the reference's headers compiles with
    them (gcc -fsyntax-only, both builds), next to include/storm_gpu.h and
    include/storm_ref_dropin.h."""
import ctypes as C
import os
import re
import subprocess

import pytest

from conftest import ROOT

REF = "/root/reference"
pytestmark = pytest.mark.skipif(not os.path.exists(os.path.join(REF, "index.h")), reason="reference sources absent")

LAYOUT_C = r"""
#include <stddef.h>
#include <stdio.h>
#include "alignment.h"
#include "index.h"
#define P(x) printf("%ld\n", (long)(x))
int main(void) {
  P(sizeof(MaskedRegionType)); P(offsetof(MaskedRegionType, start)); P(offsetof(MaskedRegionType, end));
  P(offsetof(MaskedRegionType, next));
  P(sizeof(SeedType)); P(offsetof(SeedType, weight)); P(offsetof(SeedType, length)); P(offsetof(SeedType, seed));
  P(offsetof(SeedType, positions)); P(offsetof(SeedType, positions_count)); P(offsetof(SeedType, mask));
  P(sizeof(IndexType)); P(offsetof(IndexType, db)); P(offsetof(IndexType, seed)); P(offsetof(IndexType, first));
  P(offsetof(IndexType, map)); P(offsetof(IndexType, index_size)); P(offsetof(IndexType, map_size));
  P(sizeof(AlignmentType)); P(offsetof(AlignmentType, params)); P(offsetof(AlignmentType, read_len));
  P(offsetof(AlignmentType, ref_len)); P(offsetof(AlignmentType, read)); P(offsetof(AlignmentType, read_quality));
  P(offsetof(AlignmentType, reference)); P(offsetof(AlignmentType, reference_masked)); P(offsetof(AlignmentType, best_i));
  P(offsetof(AlignmentType, best_j)); P(offsetof(AlignmentType, best_i0)); P(offsetof(AlignmentType, best_j0));
  P(offsetof(AlignmentType, traceback)); P(offsetof(AlignmentType, traceback_seq_len));
  P(offsetof(AlignmentType, matrix)); P(offsetof(AlignmentType, range_first));
  P(sizeof(AlignmentCellType)); P(offsetof(AlignmentCellType, score));
  P(sizeof(AlignmentParamType)); P(offsetof(AlignmentParamType, gap)); P(offsetof(AlignmentParamType, pair_scores));
  P(offsetof(AlignmentParamType, allowed_indels));
  P(sizeof(ReferenceDBType)); P(offsetof(ReferenceDBType, name)); P(offsetof(ReferenceDBType, size));
#ifdef NUCLEOTIDES
  P(-1);
#else
  P(offsetof(ReferenceDBType, first_base));
#endif
  P(offsetof(ReferenceDBType, sequence)); P(offsetof(ReferenceDBType, sequence_masked));
  P(offsetof(ReferenceDBType, masked_regions));
  return 0;
}
"""


@pytest.mark.parametrize("color", [0, 1], ids=["nucleotide", "color"])
def test_struct_mirrors_match_reference_headers(color, tmp_path):
    from storm_amd import lib
    src = tmp_path / "layout.c"
    src.write_text(LAYOUT_C)
    exe = str(tmp_path / "layout")
    subprocess.run(["gcc", "-fopenmp", "-I", REF] + ([] if color else ["-DNUCLEOTIDES"]) + [str(src), "-o", exe],
                   check=True, capture_output=True)
    want = [int(x) for x in subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()]
    so = lib.load()
    so.storm_ref_abi_layout.argtypes = [C.c_int, C.POINTER(C.c_int64), C.c_int]
    out = (C.c_int64 * 64)()
    n = so.storm_ref_abi_layout(color, out, 64)
    assert list(out[:n]) == want


def _blocks():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    return re.findall(r"<!-- compiles with the reference's headers -->\s*```c\n(.*?)```", text, re.S)


def test_integration_has_reference_bindings():
    assert len(_blocks()) >= 3


@pytest.mark.parametrize("color", [0, 1], ids=["nucleotide", "color"])
def test_integration_bindings_compile_with_reference_headers(color, tmp_path):
    for k, body in enumerate(_blocks()):
        src = tmp_path / f"binding{k}.c"
        src.write_text('#include "run.h"\n#include "hit_map.h"\n#include "storm_gpu.h"\n#include "storm_ref_dropin.h"\n'
                       + body)
        r = subprocess.run(["gcc", "-fsyntax-only", "-Wall", "-Werror=implicit-function-declaration",
                            "-Werror=incompatible-pointer-types", "-Werror=int-conversion", "-DSTORM_GFX950", "-fopenmp",
                            "-msse2", "-I", REF, "-I", os.path.join(ROOT, "include")] +
                           ([] if color else ["-DNUCLEOTIDES"]) + [str(src)], capture_output=True, text=True)
        assert r.returncode == 0, (k, r.stderr[-3000:])
