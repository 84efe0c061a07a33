"""FASTQ files for the read-loader tests: seeded, with the cases the reference's loader handles
(N and lowercase bases, CRLF line ends, a '# Title:' comment, names with blanks and '@',
quality characters at the -b base)."""
import os

import numpy as np

BASES = np.frombuffer(b"ACGT", np.uint8)


def fastq_records(n, L, seed, color=False, lower=0.0, ns=0.0, qlo=0, qhi=40, varlen=None):
    g = np.random.default_rng(seed)
    recs = []
    for i in range(n):
        ln = int(g.integers(varlen[0], varlen[1] + 1)) if varlen else L
        if color:
            s = b"T" + (48 + g.integers(0, 4, ln + 1)).astype(np.uint8).tobytes()   # primer base + colors
            s = s[:ln + 2]
        else:
            c = BASES[g.integers(0, 4, ln)].copy()
            low = g.random(ln) < lower
            c[low] = c[low] + 32
            c[g.random(ln) < ns] = ord("N")
            s = c.tobytes()
        qn = ln + 1 if color else ln
        q = (33 + g.integers(qlo, qhi + 1, qn)).astype(np.uint8).tobytes()
        if i % 89 == 7:
            q = b"#" + q[1:]           # quality line starting with '#': a comment to the reference
        name = f"r{i}" if i % 7 else f"r{i} extra field"
        if i % 11 == 3:
            name = f"r{i}@x"
        recs.append((name, s, q))
    return recs


def write_fastq(path, recs, crlf=False, title=None, blank_every=0):
    nl = b"\r\n" if crlf else b"\n"
    with open(path, "wb") as f:
        if title:
            f.write(b"# Title: " + title.encode() + nl)
            f.write(b"# another comment" + nl)
        for i, (name, s, q) in enumerate(recs):
            if blank_every and i and i % blank_every == 0:
                f.write(nl)
            f.write(b"@" + name.encode() + nl + s + nl + b"+" + nl + q + nl)


def cases(dirpath):
    """(path, color) for every loader case."""
    out = []
    p = os.path.join(dirpath, "n_plain.fq")
    write_fastq(p, fastq_records(3000, 50, 1))
    out.append((p, 0))
    p = os.path.join(dirpath, "n_mixed.fastq")
    write_fastq(p, fastq_records(2500, 75, 2, lower=0.05, ns=0.02, qlo=0), crlf=True, title="run 7")
    out.append((p, 0))
    p = os.path.join(dirpath, "c_plain.fq")
    write_fastq(p, fastq_records(2000, 34, 3, color=True, qlo=0, qhi=35))
    out.append((p, 1))
    p = os.path.join(dirpath, "c_title.fq")
    write_fastq(p, fastq_records(1500, 49, 4, color=True), crlf=True, title="solid")
    out.append((p, 1))
    return out


def varlen_case(dirpath, color):
    p = os.path.join(dirpath, f"var_{color}.fq")
    write_fastq(p, fastq_records(4000, 0, 5 + color, color=bool(color), lower=0.02, ns=0.01, varlen=(30, 150)))
    return p


def corrupt_cases(dirpath):
    """Files the reference's FASTQ reader rejects as corrupt (reads_data.c:429-433)."""
    out = []
    p = os.path.join(dirpath, "blank.fq")          # blank line between records
    write_fastq(p, fastq_records(800, 60, 9), blank_every=100)
    out.append(p)
    recs = fastq_records(300, 60, 10)
    recs[120] = (recs[120][0], recs[120][1], b"")  # empty quality line
    p = os.path.join(dirpath, "emptyq.fq")
    write_fastq(p, recs)
    out.append(p)
    recs = fastq_records(300, 60, 11)
    recs[7] = (recs[7][0], b"", recs[7][2])        # empty sequence line
    p = os.path.join(dirpath, "emptys.fq")
    write_fastq(p, recs)
    out.append(p)
    return out


def write_fastq_hash_quals(path, reads, every=13):
    """synth.SimReads as FASTQ (Phred+33) where the quality line of every `every`-th read starts
    with '#' (Q2): the reference types such a line as a comment and maps the read without quality."""
    with open(path, "wb") as f:
        for k, (s, q) in enumerate(zip(reads.seqs, reads.quals)):
            qq = (33 + np.asarray(q)).astype(np.uint8)
            if k % every == 3:
                qq[0] = ord("#")
            f.write(b"@r%d\n" % k + BASES[s].tobytes() + b"\n+\n" + qq.tobytes() + b"\n")
