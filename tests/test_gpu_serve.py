"""Serving loop (storm_server_*): requests of mixed sizes, read lengths (36-150), encodings
(nucleotide / color) and bands (-i 3 / 7) share the device by continuous batching.  Under -M a
read's hits depend only on the read, the genome and the parameters (hit_map.c:50-144), so every
request must get identical slots (all four ranks: score, reference, start, strand) alone, in the
mix, under memory-limited admission, through the plain per-request path, and -- for one
single-length request -- in the CPU restatement's SAM."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu
SEEDS_N = "####--#-#--##--####;#####--###-#-###;###-##-#######"
SEEDS_C = "####----##--##-####;####-###--#----####;####-####-####"
SIZES = [300, 5000, 1200, 9000, 50, 2500, 7000, 800, 3000, 400, 6000, 150]
CLASSES = [(0, 3), (1, 7), (0, 7), (1, 3)]   # (color, -i)


def class_params(color, i):
    return dict(color=color, allowed_indels=i, z=85 if color else 100, t=100, max_hits=4)


@pytest.fixture(scope="module")
def serve_case(tmp_path_factory):
    from storm_amd import synth
    d = tmp_path_factory.mktemp("serve")
    g = synth.gc_window_genome(2_000_000, 31)
    gpath = str(d / "g.fa")
    synth.write_fasta_fast(gpath, "chr1", g)
    reqs = []
    for k, n in enumerate(SIZES):
        color, i = CLASSES[k % len(CLASSES)]
        if color:
            pr = synth.fast_color_reads(g, n, (35, 149), 500 + k, snp=0.01, color_err=0.03, qlo=10, qhi=35)
        else:
            pr = synth.fast_nucleotide_reads(g, n, (36, 150), 500 + k, sub=0.015, indel=0.002, qlo=15, qhi=40)
        reqs.append((pr, class_params(color, i)))
    return d, gpath, g, reqs


def _run(server, reqs, together):
    out = [None] * len(reqs)
    if together:
        ids = [server.submit(pr, **kw) for pr, kw in reqs]
        for k, rid in enumerate(ids):
            out[k] = server.wait(rid)
    else:
        for k, (pr, kw) in enumerate(reqs):
            out[k] = server.wait(server.submit(pr, **kw))
    return out


def _same(a, b):
    return all(np.array_equal(a[f], b[f]) for f in ("score", "ref_id", "ref_start", "sense"))


@pytest.fixture(scope="module")
def alone(serve_case):
    from storm_amd import lib
    _, gpath, _, reqs = serve_case
    s = lib.Server(gpath, lanes=3, max_batch_reads=4096)
    print("server: requests one at a time", flush=True)
    res = _run(s, reqs, together=False)
    s.close()
    return res


def test_mix_equals_alone(serve_case, alone):
    from storm_amd import lib
    _, gpath, _, reqs = serve_case
    s = lib.Server(gpath, lanes=3, max_batch_reads=4096)
    print("server: all requests submitted together", flush=True)
    mix = _run(s, reqs, together=True)
    st = s.stats()
    s.close()
    print(st, flush=True)
    for k in range(len(reqs)):
        assert _same(mix[k][0], alone[k][0]), f"request {k}"
        assert mix[k][1]["status"] == 0
    assert st["max_requests_per_batch"] >= 2          # batches crossed request boundaries
    assert mix[3][1]["batches"] >= 3                   # and 9000 reads spread over batches of 4096
    assert st["completed"] == len(reqs) and st["failed"] == 0
    mapped = sum(int((m[0]["ref_id"][:, 0] >= 0).sum()) for m in mix)
    assert mapped > 0.6 * sum(SIZES)


def test_alone_equals_plain_mapper(serve_case, alone):
    """Each request's slots equal storm_mapper_map_hits on the staged (plain) kernels."""
    from storm_amd import lib
    _, gpath, _, reqs = serve_case
    for color, i in CLASSES:
        m = lib.Mapper(gpath, seeds=SEEDS_C if color else SEEDS_N, **class_params(color, i))
        for k, (pr, kw) in enumerate(reqs):
            if (kw["color"], kw["allowed_indels"]) != (color, i):
                continue
            print(f"mapper plain: request {k} ({len(pr.len)} reads, color={color}, -i {i})", flush=True)
            h, _ = m.map_hits(pr, mode=lib.STORM_MODE_PLAIN)
            assert _same(h, alone[k][0]), f"request {k}"
        m.close()


def test_admission_by_device_memory(serve_case, alone):
    """A budget that holds about one large request at a time: requests wait for memory, the
    arena never exceeds the budget, the results do not change."""
    from storm_amd import lib
    _, gpath, _, reqs = serve_case
    budget = 2 << 20
    s = lib.Server(gpath, lanes=2, max_batch_reads=8192, hbm_budget_bytes=budget)
    print("server: 2 MiB admission budget", flush=True)
    mix = _run(s, reqs, together=True)
    st = s.stats()
    s.close()
    print(st, flush=True)
    assert st["arena_bytes"] <= budget
    assert st["arena_peak_bytes"] <= st["arena_bytes"]
    assert st["admission_waits"] >= 1
    for k in range(len(reqs)):
        assert _same(mix[k][0], alone[k][0]), f"request {k}"


def test_request_larger_than_budget_is_rejected(serve_case):
    from storm_amd import lib
    _, gpath, _, reqs = serve_case
    s = lib.Server(gpath, lanes=1, hbm_budget_bytes=1 << 20)
    pr, kw = reqs[3]   # 9000 reads: more than 1 MiB of reads and results
    with pytest.raises(lib.StormError, match="budget"):
        s.submit(pr, **kw)
    s.close()


def test_plain_server_equals_batched(serve_case, alone):
    from storm_amd import lib
    _, gpath, _, reqs = serve_case
    s = lib.Server(gpath, mode=lib.STORM_MODE_PLAIN)
    print("server: plain per-request mode", flush=True)
    res = _run(s, reqs[:6], together=True)
    s.close()
    for k in range(6):
        assert _same(res[k][0], alone[k][0]), f"request {k}"


def test_served_request_matches_oracle_sam(serve_case, oracle_built):
    """One single-length nucleotide request, -i 3 -M 4: the served slots are the records of the
    CPU restatement's SAM (rank order; POS, strand flag, AS:i), and the served SAM records are
    its SAM lines."""
    from storm_amd import lib, synth
    d, gpath, g, _ = serve_case
    pr = synth.fast_nucleotide_reads(g, 400, 100, 77, sub=0.015, indel=0.002, qlo=15, qhi=40)
    s = lib.Server(gpath, lanes=2, max_batch_reads=256)
    h, _, sam = s.wait_sam(s.submit(pr, **class_params(0, 3)))
    s.close()
    rpath = str(d / "oracle_reads.fq")
    synth.write_fastq(rpath, synth.packed_to_simreads(pr, np.arange(len(pr.len))))
    out = str(d / "oracle.sam")
    subprocess.run([os.path.join(oracle_built, "oracle-nucleotide"), "-g", gpath, "-r", rpath, "-s", SEEDS_N,
                    "-i", "3", "-M", "4", "-z", "100", "-t", "100", "-o", out], check=True, stderr=subprocess.DEVNULL)
    recs = {}
    body = []
    for line in open(out):
        if line.startswith("@"):
            continue
        body.append(line)
        f = line.rstrip("\n").split("\t")
        score = int(next(x for x in f[11:] if x.startswith("AS:i:"))[5:])
        recs.setdefault(int(f[0][1:]), []).append((score, int(f[3]) - 1, -1 if int(f[1]) & 16 else 1))
    assert len(recs) > 0.6 * len(pr.len)
    for r in range(len(pr.len)):
        want = recs.get(r, [])
        got = [(int(h["score"][r, k]), int(h["ref_start"][r, k]), int(h["sense"][r, k]))
               for k in range(4) if h["ref_id"][r, k] >= 0]
        assert got == want, f"read {r}"
    # the reference writes reads in its tag order (reads_data.c:59-67), a request in the order given:
    # the same lines read by read
    def by_read(lines):
        out = {}
        for line in lines:
            out.setdefault(line.split("\t", 1)[0], []).append(line)
        return out
    assert by_read(sam.decode().splitlines(keepends=True)) == by_read(body)


def _sam_body(path):
    return [l for l in open(path) if not l.startswith("@")]


def test_served_sam_equals_mapper_sam(serve_case, tmp_path):
    """Every request's SAM records, served in the mix (batches across requests, lengths and
    encodings), are the lines storm_mapper_map writes for the same reads alone."""
    from storm_amd import lib
    _, gpath, _, reqs = serve_case
    s = lib.Server(gpath, lanes=3, max_batch_reads=4096)
    print("server: SAM records of every request in the mix", flush=True)
    ids = [s.submit(pr, **kw) for pr, kw in reqs]
    served = [s.wait_sam(rid) for rid in ids]
    st = s.stats()
    s.close()
    assert st["max_requests_per_batch"] >= 2
    mappers = {}
    for k, (pr, kw) in enumerate(reqs):
        key = (kw["color"], kw["allowed_indels"])
        if key not in mappers:
            mappers[key] = lib.Mapper(gpath, seeds=SEEDS_C if kw["color"] else SEEDS_N, **kw)
        out = str(tmp_path / f"req{k}.sam")
        mappers[key].map_sam(pr, out, mode=lib.STORM_MODE_PLAIN)
        want = _sam_body(out)
        got = served[k][2].decode().splitlines(keepends=True)
        assert len(got) == len(want) and got == want, f"request {k}"
        assert served[k][1]["n_mapped"] > 0
    for m in mappers.values():
        m.close()


def test_plain_server_sam_equals_batched(serve_case):
    from storm_amd import lib
    _, gpath, _, reqs = serve_case
    out = {}
    for mode in (lib.STORM_MODE_PLAIN, lib.STORM_MODE_BATCHED):
        s = lib.Server(gpath, mode=mode)
        ids = [s.submit(pr, **kw) for pr, kw in reqs[:4]]
        out[mode] = [s.wait_sam(rid)[2] for rid in ids]
        s.close()
    assert out[lib.STORM_MODE_PLAIN] == out[lib.STORM_MODE_BATCHED]


def test_failed_upload_fails_only_its_request(serve_case, alone, monkeypatch):
    """One admitted region's upload is made to fail (fault injection from the host, no device
    work): that request, and only it, ends with an error; the requests batched with it get the
    same slots as alone."""
    from storm_amd import lib
    _, gpath, _, reqs = serve_case
    monkeypatch.setenv("STORM_SERVER_FAIL_UPLOAD", "2")
    s = lib.Server(gpath, lanes=2, max_batch_reads=4096)
    monkeypatch.delenv("STORM_SERVER_FAIL_UPLOAD")
    print("server: the second region upload fails", flush=True)
    ids = [s.submit(pr, **kw) for pr, kw in reqs[:6]]
    failed = []
    for k, rid in enumerate(ids):
        try:
            h, res = s.wait(rid)
        except lib.StormError as e:
            assert "injected upload failure" in str(e)
            failed.append(k)
            continue
        assert res["status"] == 0
        assert _same(h, alone[k][0]), f"request {k}"
    st = s.stats()
    s.close()
    assert len(failed) == 1 and st["failed"] == 1 and st["completed"] == 5


@pytest.mark.parametrize("schedule", ["size", "fifo"])
def test_schedule_order(serve_case, alone, schedule):
    """A small request submitted just after a large one of the same class (one lane, 1000-read
    batches, the large one needs 9): under STORM_SCHED_SIZE it completes first, under
    STORM_SCHED_FIFO after the large one; either way every read gets the slots it gets alone."""
    from storm_amd import lib
    _, gpath, _, reqs = serve_case
    big, kw = reqs[3]
    assert len(big.len) == 9000
    sched = lib.STORM_SCHED_SIZE if schedule == "size" else lib.STORM_SCHED_FIFO
    s = lib.Server(gpath, lanes=1, max_batch_reads=1000, schedule=sched)
    print(f"server: schedule {schedule}, 9000-read request then a 100-read one", flush=True)
    ib = s.submit(big, **kw)
    ism = s.submit(big, 0, 100, **kw)
    hb, rb = s.wait(ib)
    hs, rs = s.wait(ism)
    s.close()
    assert rb["status"] == 0 and rs["status"] == 0
    assert _same(hb, alone[3][0])
    assert _same(hs, {f: alone[3][0][f][:100] for f in hs})
    if schedule == "size":
        assert rs["t_done"] < rb["t_done"], (rs, rb)
    else:
        assert rs["t_done"] >= rb["t_done"], (rs, rb)
