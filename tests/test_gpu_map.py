"""Whole mapping runs on the GPU against the reference's SAM (golden fixtures): -M hit map output,
and the ordered (genome map) output of the cases without -M, contextual or greedy (-G)."""
import os

import pytest

from conftest import sam_body, sam_cases, unpack_case

pytestmark = pytest.mark.gpu


def _run(meta, paths, out, mode):
    from storm_amd import lib
    color = 1 if meta["mode"] == "c" else 0
    a = meta["args"]
    kw = dict(color=color, max_hits=0)
    i = 0
    while i < len(a):
        f, v = a[i], a[i + 1] if i + 1 < len(a) else None
        if f == "-s": kw["seeds"] = v
        elif f == "-i": kw["allowed_indels"] = int(v)
        elif f == "-M": kw["max_hits"] = int(v)
        elif f == "-z": kw["z"] = int(v)
        elif f == "-t": kw["t"] = int(v)
        elif f == "-l": kw["stdev_distance"] = int(v)
        elif f == "-u": kw["uniqueness"] = int(v)
        elif f == "-A": kw["print_all"] = 1; i -= 1
        elif f == "-G": kw["greedy"] = 1; i -= 1
        else: raise AssertionError(f)
        i += 2
    if color:
        return lib.map_files(paths["genome.fa"], paths["reads.csfasta"], out, qual=paths["reads.qual"], mode=mode, **kw)
    return lib.map_files(paths["genome.fa"], paths["reads.fq"], out, mode=mode, **kw)


@pytest.mark.parametrize("mode", [0, 1], ids=["plain", "batched"])
@pytest.mark.parametrize("case", sam_cases())
def test_gpu_sam_matches_reference(case, mode, tmp_path):
    meta, paths = unpack_case(case, str(tmp_path))
    out = str(tmp_path / "gpu.sam")
    _run(meta, paths, out, mode)
    got, want = sam_body(out), sam_body(paths["ref.sam"])
    assert len(got) == len(want)
    assert got == want
