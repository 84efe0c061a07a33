import gzip
import json
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")
ORACLE_BUILD = os.path.join(ROOT, "oracle", "_build")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def sam_cases():
    return sorted(d[4:] for d in os.listdir(GOLDEN) if d.startswith("sam_"))


def unpack_case(name, dest):
    """Decompress a golden SAM case into `dest`; returns (meta, paths)."""
    src = os.path.join(GOLDEN, "sam_" + name)
    meta = json.load(open(os.path.join(src, "args.json")))
    paths = {}
    for f in os.listdir(src):
        if f.endswith(".gz"):
            out = os.path.join(dest, f[:-3])
            with gzip.open(os.path.join(src, f), "rb") as a, open(out, "wb") as b:
                shutil.copyfileobj(a, b)
            paths[f[:-3]] = out
    return meta, paths


def sam_body(path):
    """SAM lines without the @PG line (its CL: echoes argv, hit_map.c:310-315)."""
    with open(path) as f:
        return [l for l in f if not l.startswith("@PG")]


@pytest.fixture(scope="session")
def oracle_built():
    if not os.path.exists(os.path.join(ORACLE_BUILD, "liboracle.so")):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "oracle"], check=True,
                       stdout=subprocess.DEVNULL)
    return ORACLE_BUILD
