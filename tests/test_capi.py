"""The C-ABI library loads and exports every function include/storm_gpu.h and
include/storm_ref_dropin.h declare (no GPU call)."""
import ctypes as C
import os
import re

from conftest import ROOT


def test_library_exports_every_declared_symbol():
    from storm_amd import lib
    so = lib.load()
    hdr = open(os.path.join(ROOT, "include", "storm_gpu.h")).read()
    names = set(re.findall(r"\b((?:storm|alignment_gfx950)_\w+)\s*\(", hdr))
    assert len(names) > 20
    for n in sorted(names):
        assert hasattr(so, n), n


def test_library_exports_reference_signature_dropins():
    """storm_ref_dropin.h maps each name to its _nt / _cs build by NUCLEOTIDES; both must exist."""
    from storm_amd import lib
    so = lib.load()
    hdr = open(os.path.join(ROOT, "include", "storm_ref_dropin.h")).read()
    declared = set(re.findall(r"\b((?:storm|index_gfx950|alignment_gfx950)_\w+)\s*\(", hdr))
    mapped = dict(re.findall(r"#define\s+(\w+)\s+(\w+_(?:nt|cs))\b", hdr))
    assert {"index_gfx950__build_reference", "index_gfx950__destroy", "alignment_gfx950__align_batch",
            "storm_ref_abi_layout"} <= declared
    exported = set()
    for n in declared:
        if n in mapped:
            exported |= {n + "_nt", n + "_cs"}
        else:
            exported.add(n)
    assert set(mapped.values()) <= exported
    for n in sorted(exported):
        assert hasattr(so, n), n


def test_default_params_mirror_reference():
    from storm_amd import lib
    p = lib.params(0)
    assert (p.match, p.mismatch, p.gap_open, p.gap_extend) == (5, -4, -9, -4)     # alignment.h:24-27
    assert (p.allowed_indels, p.t, p.z, p.qual_base) == (3, 115, 100, 33)       # alignment.h:28,51-52
    assert p.stdev_distance == 100 and p.sense_mask == 3


def test_errors_are_returned_not_exited(tmp_path):
    from storm_amd import lib
    import pytest
    with pytest.raises(lib.StormError):
        lib.map_files(str(tmp_path / "missing.fa"), str(tmp_path / "missing.fq"), str(tmp_path / "o.sam"),
                      max_hits=0)


def test_server_config_defaults_and_argument_errors(tmp_path):
    """storm_server_config_default and the argument checks of storm_server_create (no GPU call)."""
    import pytest
    from storm_amd import lib
    so = lib.load()
    c = lib.ServerConfig()
    so.storm_server_config_default(C.byref(c))
    assert (c.mode, c.lanes, c.cross_request, c.max_batch_reads, c.encodings) == (lib.STORM_MODE_BATCHED, 3, 1,
                                                                                  1 << 18, 3)
    assert c.hbm_reserve_bytes == 8 << 30 and c.stdev_distance == 100
    for bad in (dict(lanes=0), dict(lanes=17), dict(max_batch_reads=0), dict(encodings=0), dict(mode=7)):
        with pytest.raises(lib.StormError):
            lib.Server(str(tmp_path / "missing.fa"), **bad)


def test_index_build_argument_errors():
    """storm_index_build rejects bad region lists and seeds before any device call."""
    import numpy as np
    import pytest
    from storm_amd import lib
    packed = np.zeros(100, np.uint8)
    for starts, ends in (([5, 3], [9, 8]),        # not ascending
                         ([5], [5]),             # empty region
                         ([5, 8], [10, 12]),     # overlapping
                         ([390], [401])):        # past the sequence end
        with pytest.raises(lib.StormError, match="masked region"):
            lib.build_index(packed, 400, starts, ends, "##-#")
    with pytest.raises(lib.StormError):
        lib.build_index(packed, 400, [], [], "##-x")
    with pytest.raises(lib.StormError, match="exactly one seed"):
        lib.build_index(packed, 400, [], [], "##-#;#-#")


def test_align_plan_argument_errors():
    """storm_align_plan_create rejects bad reads, records and options before any device call."""
    import numpy as np
    import pytest
    from storm_amd import lib
    g = np.zeros(2000, np.uint8)
    reads = [np.zeros(50, np.uint8), np.zeros(60, np.uint8)]
    lens = np.array([50, 60], np.int32)
    ok_rs, ok_h = np.array([0, 3], np.int32), np.array([10, 500], np.int32)
    cases = [
        (dict(rs=np.array([4], np.int32), h=np.array([10], np.int32)), "read strand"),
        (dict(rs=np.array([-1], np.int32), h=np.array([10], np.int32)), "read strand"),
        (dict(rs=np.array([1], np.int32), h=np.array([2000 + 1024 - 100], np.int32)), "genome end"),
        (dict(p=lib.params(0, allowed_indels=32)), "-i"),
        (dict(p=lib.params(0, mismatch=3)), "scores"),
        (dict(lens=np.array([50, 257], np.int32), reads=[np.zeros(50, np.uint8), np.zeros(257, np.uint8)]),
         "length"),
        (dict(lens=np.array([0, 60], np.int32), reads=[np.zeros(0, np.uint8), np.zeros(60, np.uint8)]), "length"),
    ]
    for kw, msg in cases:
        with pytest.raises(lib.StormError, match=msg):
            lib.AlignPlan(g, None, kw.get("reads", reads), None, kw.get("lens", lens), kw.get("rs", ok_rs),
                          kw.get("h", ok_h), kw.get("p", lib.params(0)))
