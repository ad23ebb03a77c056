"""The classify front-end (k2g_classify) on inputs its parse and batching
must survive, each against the oracle's command line on the same toy
database: an empty file, standard input, several files in one run, a record
longer than a parse chunk, multi-line and CRLF FASTA at scale, interleaved
pairs with an odd record, and batches small enough that records carry from
one batch to the next."""
import os
import random
import subprocess

import pytest

from helpers import ORACLE_BIN

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
TOY = os.path.join(HERE, "golden", "toy")
INPUTS = os.path.join(HERE, "golden", "toy_inputs")
CLI = os.path.join(ROOT, "kraken2_amd", "bin", "k2g_classify")
ORACLE_CLI = os.path.join(ORACLE_BIN, "k2cpu_classify")


def _genomes():
    seqs = []
    for name in sorted(os.listdir(INPUTS)):
        if name.endswith(".fa"):
            cur = []
            for line in open(os.path.join(INPUTS, name)):
                if line.startswith(">"):
                    if cur:
                        seqs.append("".join(cur))
                    cur = []
                else:
                    cur.append(line.strip())
            if cur:
                seqs.append("".join(cur))
    return [s for s in seqs if len(s) > 1000]


def _reads(rng, genomes, n, lo, hi):
    out = []
    for _ in range(n):
        g = genomes[rng.randrange(len(genomes))]
        L = rng.randint(lo, min(hi, len(g)))
        p = rng.randrange(len(g) - L + 1)
        s = list(g[p:p + L])
        for j in range(len(s)):
            if rng.random() < 0.01:
                s[j] = rng.choice("ACGTN")
        out.append("".join(s))
    return out


def _fasta(reads, width=0, eol="\n", prefix="r"):
    parts = []
    for i, s in enumerate(reads):
        lines = [s[j:j + width] for j in range(0, len(s), width)] if width else [s]
        parts.append(f">{prefix}{i} c{i % 7}{eol}" + eol.join(lines or [""]) + eol)
    return "".join(parts)


def _both(args, tmp_path, env=None, stdin=None):
    base = ["-H", f"{TOY}/hash.k2d", "-t", f"{TOY}/taxo.k2d", "-o", f"{TOY}/opts.k2d", "-p", "4", "-g", "2"]
    outs = []
    for exe in (ORACLE_CLI, CLI):
        e = dict(os.environ, **(env or {})) if exe == CLI else None
        r = subprocess.run([exe, *base, *args], capture_output=True, env=e,
                           input=stdin.encode() if stdin is not None else None, timeout=300)
        assert r.returncode == 0, (exe, r.stderr[-500:])
        outs.append(r.stdout)
    return outs


def test_empty_file_and_stdin(tmp_path):
    empty = tmp_path / "e.fa"
    empty.write_text("")
    ref, got = _both([str(empty)], tmp_path)
    assert ref == got == b""
    rng = random.Random(1)
    text = _fasta(_reads(rng, _genomes(), 500, 100, 3000), width=60)
    ref, got = _both([], tmp_path, stdin=text)
    assert got == ref and got.count(b"\n") == 500


def test_several_files_one_run(tmp_path):
    rng = random.Random(2)
    g = _genomes()
    paths = []
    for k in range(3):
        p = tmp_path / f"f{k}.fa"
        p.write_text(_fasta(_reads(rng, g, 300 + 100 * k, 50, 2000), prefix=f"f{k}_"))
        paths.append(str(p))
    ref, got = _both(paths, tmp_path)
    assert got == ref and got.count(b"\n") == 1200


def test_record_longer_than_a_parse_chunk(tmp_path):
    """A 70 MB record (the parse reads 64 MiB chunks) between short ones."""
    rng = random.Random(3)
    g = _genomes()
    short = _reads(rng, g, 200, 100, 500)
    big = "".join(rng.choice(g)[:60000] for _ in range(1200))[:70_000_000]
    p = tmp_path / "big.fa"
    p.write_text(_fasta(short[:100]) + f">big one\n{big}\n" + _fasta(short[100:], prefix="s"))
    ref, got = _both([str(p)], tmp_path)
    assert got == ref and got.count(b"\n") == 201


@pytest.mark.parametrize("batch_bases", [None, "300000"])
def test_multiline_crlf_fasta_and_small_batches(tmp_path, batch_bases):
    """Multi-line records (joined in place), CRLF lines, and batches of a
    few hundred kb so records carry from batch to batch."""
    rng = random.Random(4)
    g = _genomes()
    reads = _reads(rng, g, 6000, 30, 6000)
    p = tmp_path / "ml.fa"
    p.write_text(_fasta(reads[:3000], width=70) + _fasta(reads[3000:], width=61, eol="\r\n", prefix="w"))
    env = {"K2G_BATCH_BASES": batch_bases} if batch_bases else None
    ref, got = _both([str(p)], tmp_path, env=env)
    assert got == ref and got.count(b"\n") == 6000


def test_interleaved_pairs_with_an_odd_record(tmp_path):
    rng = random.Random(5)
    reads = _reads(rng, _genomes(), 2001, 100, 300)
    p = tmp_path / "i.fa"
    p.write_text("".join(f">p{i // 2}/{i % 2 + 1}\n{s}\n" for i, s in enumerate(reads)))
    ref, got = _both(["-S", str(p)], tmp_path, env={"K2G_BATCH_BASES": "100000"})
    assert got == ref and got.count(b"\n") == 1000


def _fastq(reads, rng, prefix):
    return "".join(f"@{prefix}{i} q{i % 5}\n{s}\n+\n{''.join(chr(33 + rng.randrange(41)) for _ in s)}\n"
                   for i, s in enumerate(reads))


@pytest.mark.parametrize("quals", ["dropped", "masked", "written"])
def test_two_file_fastq_pairs(tmp_path, quals):
    """Two-file FASTQ pairs over several 10 000-pair blocks, with uneven
    files (the longer one's extra records are dropped) and batches small
    enough to hold a few blocks: qualities dropped by the reader (no -Q, no
    -C/-U), used for -Q masking, and written back in the -C/-U records."""
    rng = random.Random(6)
    g = _genomes()
    m1 = _reads(rng, g, 25003, 60, 250)
    m2 = _reads(rng, g, 25000, 60, 250)
    f1, f2 = tmp_path / "a_1.fq", tmp_path / "a_2.fq"
    f1.write_text(_fastq(m1, rng, "p"))
    f2.write_text(_fastq(m2, rng, "p"))
    args = ["-P", str(f1), str(f2)]
    if quals == "masked":
        args = ["-Q", "20"] + args
    outs = []
    for exe in (ORACLE_CLI, CLI):
        extra = []
        if quals == "written":
            extra = ["-C", str(tmp_path / f"{os.path.basename(exe)}_c#.fq"),
                     "-U", str(tmp_path / f"{os.path.basename(exe)}_u#.fq")]
        base = ["-H", f"{TOY}/hash.k2d", "-t", f"{TOY}/taxo.k2d", "-o", f"{TOY}/opts.k2d", "-p", "4", "-g", "2"]
        env = dict(os.environ, K2G_BATCH_BASES="4000000") if exe == CLI else None
        r = subprocess.run([exe, *base, *extra, *args], capture_output=True, env=env, timeout=300)
        assert r.returncode == 0, (exe, r.stderr[-500:])
        outs.append(r.stdout)
    assert outs[1] == outs[0] and outs[1].count(b"\n") == 25000
    if quals == "written":
        for kind in ("c", "u"):
            for mate in ("1", "2"):
                ref = (tmp_path / f"k2cpu_classify_{kind}_{mate}.fq").read_bytes()
                got = (tmp_path / f"k2g_classify_{kind}_{mate}.fq").read_bytes()
                assert got == ref
