"""The classify front-end's host code without a GPU: bin/k2g_classify's
parse, blocks and batches, reader and GPU-call threads, -Q masking, output
lines, -C/-U records and report, built with AddressSanitizer and
UndefinedBehaviorSanitizer over a test double of the library
(tests/emu/frontend_double.cc).  The double's results are a fixed function of
each fragment's bases, recomputed here from the reads, so a result handed to
the wrong fragment, a dropped or repeated fragment, or a changed base fails
the comparison; the sanitizers fail any out-of-bounds or undefined access.
The classification itself is checked against the oracle by the GPU tests of
the same front-end (test_gpu_cli_*)."""
import os
import random
import struct
import subprocess
import zlib

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
TOY = os.path.join(HERE, "golden", "toy")
EXE = os.path.join(HERE, "emu", "k2g_classify_double")
DB = ["-H", f"{TOY}/hash.k2d", "-t", f"{TOY}/taxo.k2d", "-o", f"{TOY}/opts.k2d"]
K = struct.unpack("<Q", open(f"{TOY}/opts.k2d", "rb").read()[:8])[0]
QUAL = [chr(33 + i) for i in range(41)]


@pytest.fixture(scope="module", autouse=True)
def built():
    r = subprocess.run(["make", "-C", os.path.join(HERE, "emu"), "k2g_classify_double"], capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stderr[-2000:]


def run(args, env=None, stdin=None):
    e = dict(os.environ, ASAN_OPTIONS="abort_on_error=0:halt_on_error=1:detect_leaks=1",
             UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", **(env or {}))
    r = subprocess.run([EXE, *DB, *args], capture_output=True, env=e, timeout=600,
                       input=stdin.encode() if stdin is not None else None)
    err = r.stderr.decode(errors="replace")
    assert r.returncode == 0 and "Sanitizer" not in err and "runtime error" not in err, err[-3000:]
    return r.stdout.decode()


def masked(seq, qual, q):
    return "".join("x" if q and ord(b) - 33 < q else s for s, b in zip(seq, qual)) if qual is not None else seq


def expected(frags, q=0):
    """frags: (id, mate 1, mate 2 or None, quals 1, quals 2) per fragment."""
    out = []
    for fid, m1, m2, q1, q2 in frags:
        a, b = masked(m1, q1, q), (masked(m2, q2, q) if m2 is not None else None)
        call = (zlib.crc32((a + (b or "")).encode()) >> 7) & 1
        e1 = max(0, len(a) - K + 1)
        runs = [f"0:{e1}"] if e1 else []
        if b is not None:
            if fid[-2:] in ("/1", "/2"):
                fid = fid[:-2]
            e2 = max(0, len(b) - K + 1)
            runs += ["|:|"] + ([f"0:{e2}"] if e2 else [])
        ln = f"{len(a)}|{len(b)}" if b is not None else f"{len(a)}"
        out.append(f"{'C' if call else 'U'}\t{fid}\t{call}\t{ln}\t{' '.join(runs) if runs else '0:0'}\n")
    return "".join(out)


def reads(rng, n, lo=0, hi=600, long_every=0):
    out = []
    for i in range(n):
        L = rng.randint(lo, hi) if not (long_every and i % long_every == 0) else rng.randint(10000, 60000)
        out.append("".join(rng.choices("ACGTACGTACGTN", k=L)))
    return out


def quals(rng, s):
    return "".join(rng.choices(QUAL, k=len(s)))


def fasta(ids, seqs, width=0, eol="\n"):
    parts = []
    for i, s in zip(ids, seqs):
        lines = [s[j:j + width] for j in range(0, len(s), width)] if width else [s]
        parts.append(f">{i} c{len(s) % 7}{eol}" + eol.join(lines) + eol)
    return "".join(parts)


def fastq(ids, seqs, qs):
    return "".join(f"@{i} c{len(s) % 7}\n{s}\n+\n{q}\n" for i, s, q in zip(ids, seqs, qs))


def record(fid, cmt, seq, qual, call):
    suffix = " kraken:taxid|1" if call else ""
    if qual is None:
        return f">{fid}{suffix} {cmt}\n{seq}\n"
    return f"@{fid}{suffix} {cmt}\n{seq}\n+\n{qual}\n"


@pytest.mark.parametrize("batch,threads", [(None, "4"), ("200000", "3"), ("4000000", "1")])
def test_fasta_single_reads(tmp_path, batch, threads):
    rng = random.Random(11)
    seqs = reads(rng, 12000, 0, 500, long_every=997)
    ids = [f"r{i}" for i in range(len(seqs))]
    p = tmp_path / "s.fa"
    p.write_text(fasta(ids, seqs))
    got = run(["-p", threads, str(p)], env={"K2G_BATCH_BASES": batch} if batch else None)
    assert got == expected([(i, s, None, None, None) for i, s in zip(ids, seqs)])


def test_multiline_crlf_fasta_several_files_and_stdin(tmp_path):
    rng = random.Random(12)
    seqs = reads(rng, 6000, 0, 3000)
    ids = [f"m{i}" for i in range(len(seqs))]
    a, b, c = tmp_path / "a.fa", tmp_path / "b.fa", tmp_path / "empty.fa"
    a.write_text(fasta(ids[:3000], seqs[:3000], width=70))
    b.write_text(fasta(ids[3000:], seqs[3000:], width=61, eol="\r\n"))
    c.write_text("")
    # an empty record's lone "\r" line is one base, as the reference reads it
    crlf = [s if s or j < 3000 else "\r" for j, s in enumerate(seqs)]
    want = expected([(i, s, None, None, None) for i, s in zip(ids, crlf)])
    assert run(["-p", "4", str(a), str(c), str(b)], env={"K2G_BATCH_BASES": "300000"}) == want
    want = expected([(i, s, None, None, None) for i, s in zip(ids, seqs)])
    assert run(["-p", "2"], stdin=fasta(ids, seqs, width=50)) == want


@pytest.mark.parametrize("q", [0, 20])
def test_fastq_single_reads_with_cu_records(tmp_path, q):
    rng = random.Random(13)
    seqs = reads(rng, 9000, 0, 400, long_every=1500)
    qs = [quals(rng, s) for s in seqs]
    ids = [f"q{i}" for i in range(len(seqs))]
    p = tmp_path / "s.fq"
    p.write_text(fastq(ids, seqs, qs))
    args = ["-p", "4", "-C", str(tmp_path / "c.fq"), "-U", str(tmp_path / "u.fq")] + (["-Q", str(q)] if q else [])
    got = run(args + [str(p)], env={"K2G_BATCH_BASES": "500000"})
    frags = [(i, s, None, qq, None) for i, s, qq in zip(ids, seqs, qs)]
    assert got == expected(frags, q)
    calls = [ln[0] == "C" for ln in got.splitlines()]
    for name, want_c in (("c.fq", True), ("u.fq", False)):
        want = "".join(record(i, f"c{len(s) % 7}", masked(s, qq, q), qq, c)
                       for (i, s, _, qq, _), c in zip(frags, calls) if c == want_c)
        assert (tmp_path / name).read_text() == want


@pytest.mark.parametrize("kind", ["fasta", "fastq", "fastq-Q", "fastq-CU"])
def test_two_file_pairs(tmp_path, kind):
    """Uneven files over several 10 000-pair blocks: the longer file's extra
    records are dropped with the last block."""
    rng = random.Random(14)
    m1, m2 = reads(rng, 25003, 0, 300), reads(rng, 25000, 0, 300)
    q1, q2 = [quals(rng, s) for s in m1], [quals(rng, s) for s in m2]
    ids1, ids2 = [f"p{i}/1" for i in range(len(m1))], [f"p{i}/2" for i in range(len(m2))]
    f1, f2 = tmp_path / "x_1", tmp_path / "x_2"
    fq = kind != "fasta"
    f1.write_text(fastq(ids1, m1, q1) if fq else fasta(ids1, m1))
    f2.write_text(fastq(ids2, m2, q2) if fq else fasta(ids2, m2))
    q = 20 if kind == "fastq-Q" else 0
    args = ["-p", "4", "-P"] + (["-Q", "20"] if q else [])
    if kind == "fastq-CU":
        args += ["-C", str(tmp_path / "c#.fq"), "-U", str(tmp_path / "u#.fq")]
    got = run(args + [str(f1), str(f2)], env={"K2G_BATCH_BASES": "3000000"})
    frags = [(ids1[i], m1[i], m2[i], q1[i] if fq else None, q2[i] if fq else None) for i in range(len(m2))]
    assert got == expected(frags, q)
    if kind == "fastq-CU":
        calls = [ln[0] == "C" for ln in got.splitlines()]
        for kind_c, want_c in (("c", True), ("u", False)):
            for mate in (0, 1):
                ids, seqs, qs = (ids1, m1, q1) if mate == 0 else (ids2, m2, q2)
                want = "".join(record(ids[i], f"c{len(seqs[i]) % 7}", seqs[i], qs[i], c)
                               for i, c in enumerate(calls) if c == want_c)
                assert (tmp_path / f"{kind_c}_{mate + 1}.fq").read_text() == want


def test_interleaved_pairs_with_an_odd_record_and_report(tmp_path):
    rng = random.Random(15)
    seqs = reads(rng, 4001, 0, 400)
    ids = [f"i{j // 2}/{j % 2 + 1}" for j in range(len(seqs))]
    p = tmp_path / "i.fa"
    p.write_text(fasta(ids, seqs))
    rep = tmp_path / "r.kreport"
    got = run(["-p", "4", "-S", "-R", str(rep), "-K", str(p)], env={"K2G_BATCH_BASES": "100000"})
    frags = [(ids[2 * j], seqs[2 * j], seqs[2 * j + 1], None, None) for j in range(2000)]
    assert got == expected(frags)
    n_c = sum(ln[0] == "C" for ln in got.splitlines())
    lines = rep.read_text().splitlines()
    assert lines[0].split("\t")[1:3] == [str(2000 - n_c)] * 2 and lines[0].endswith("unclassified")
    assert lines[1].split("\t")[1] == str(n_c) and lines[1].endswith("root")
