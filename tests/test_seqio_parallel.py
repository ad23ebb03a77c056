"""The front-ends' chunked record readers (Reader::next_records,
Reader::next_flat and Reader::next_inplace: plain FASTA cut at record
starts and parsed on several
threads, everything else read serially) return exactly the records of the
serial reader (Reader::next),
on the reference's edge cases: multi-line and CRLF records, comments, empty
records, FASTQ, and lines starting with '@' or '+' inside FASTA."""
import os
import random
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
TOY = os.path.join(HERE, "golden", "toy")


@pytest.fixture(scope="module")
def checker():
    r = subprocess.run(["make", "-C", os.path.join(HERE, "emu"), "reader_check"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return os.path.join(HERE, "emu", "reader_check")


def both(checker, path, chunk, threads):
    """(serial records, chunked records); the flat and in-place parses'
    records must equal the serial ones too."""
    r = subprocess.run([checker, path, str(chunk), str(threads)], capture_output=True)
    assert r.returncode == 0
    serial, chunked, flat, inplace = r.stdout.split(b"----\n")
    assert flat == serial
    assert inplace == serial
    return serial, chunked


def random_fasta(rng, n, marker_at=None):
    out = []
    for i in range(n):
        head = f">r{i}" + (f" comment {i}" if rng.random() < 0.3 else "")
        seq = "".join(rng.choice("ACGTNacgt") for _ in range(rng.randrange(0, 400)))
        lines = [seq[j:j + 70] for j in range(0, len(seq), 70)] or [""]
        eol = "\r\n" if rng.random() < 0.1 else "\n"
        if i == marker_at:
            lines.insert(1, "+not a quality line")
        out.append(head + eol + eol.join(lines) + eol)
    return "".join(out)


@pytest.mark.parametrize("chunk,threads", [(1 << 26, 8), (3000, 4), (200, 3), (100000, 16)])
def test_chunked_reader_matches_serial_on_fasta(checker, chunk, threads, tmp_path):
    rng = random.Random(chunk)
    p = tmp_path / "r.fa"
    p.write_text(random_fasta(rng, 3000))
    serial, chunked = both(checker, str(p), chunk, threads)
    assert serial.count(b"\n") == 3000
    assert chunked == serial


@pytest.mark.parametrize("name", ["edge.fa", "reads_1.fq"])
def test_chunked_reader_matches_serial_on_golden_inputs(checker, name):
    for chunk, threads in ((1 << 20, 8), (500, 4)):
        serial, chunked = both(checker, os.path.join(TOY, name), chunk, threads)
        assert chunked == serial


def test_chunked_reader_falls_back_at_marker_lines(checker, tmp_path):
    """A '+' line inside FASTA switches the record rules to quality reading;
    the chunked reader must hand over to the serial one there."""
    rng = random.Random(3)
    p = tmp_path / "m.fa"
    p.write_text(random_fasta(rng, 2000, marker_at=1500) + "@q1\nACGT\n+\nIIII\n>last\nAC\n")
    for chunk, threads in ((4000, 4), (1 << 20, 8)):
        serial, chunked = both(checker, str(p), chunk, threads)
        assert chunked == serial


def test_chunked_reader_parallel_reads_of_a_large_file(checker, tmp_path):
    """Chunks of a few MB are copied in by several preads side by side."""
    rng = random.Random(11)
    recs = []
    for i in range(30000):
        L = rng.randrange(50, 500)
        recs.append(f">big{i}\n" + "".join(rng.choice("ACGT") for _ in range(L)) + "\n")
    p = tmp_path / "big.fa"
    p.write_text("".join(recs))
    assert p.stat().st_size > 8_000_000
    for chunk, threads in ((4 << 20, 8), (3 << 20, 3)):
        serial, chunked = both(checker, str(p), chunk, threads)
        assert serial.count(b"\n") == 30000
        assert chunked == serial


@pytest.mark.parametrize("case", [b">a\nAC\n>", b">a\nAC\n>b", b">\nACGT\n>x y\tz\r\nA\r\n\r\nC\r\n",
                                  b">a b\n\r\n>b\n\rA\n", b">a\x00b c\nA\x00C\n>c\n", b">a\nAC\r",
                                  b">a\r\nAC\r\n>b\r\n\r\n>c"])
def test_readers_agree_on_record_edges(checker, case, tmp_path):
    """A lone '>' or an id at the end of input, empty ids, tabs and NULs in
    headers, lone '\\r' lines and a last line without '\\n'."""
    p = tmp_path / "e.fa"
    p.write_bytes(case)
    for chunk, threads in ((3, 2), (1 << 20, 4)):
        serial, chunked = both(checker, str(p), chunk, threads)
        assert chunked == serial


def test_work_pool_runs_each_item_once():
    """WorkPool::run back to back: every item once, by its own run's function."""
    r = subprocess.run(["make", "-C", os.path.join(HERE, "emu"), "pool_check"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    for threads in ("2", "8", "16"):
        r = subprocess.run([os.path.join(HERE, "emu", "pool_check"), threads, "5000"], capture_output=True,
                           text=True, timeout=120)
        assert r.returncode == 0, r.stdout


def random_fastq(rng, n, eol="\n", weird_at=None):
    """4-line FASTQ whose quality lines often start with '@' or '+' (what a
    parallel cut must not mistake for a record start)."""
    out = []
    for i in range(n):
        seq = "".join(rng.choice("ACGTN") for _ in range(rng.randrange(1, 300)))
        qual = "".join(rng.choice("@+!#5?IJ") for _ in range(len(seq)))
        head = f"@q{i}" + (f" note {i}" if rng.random() < 0.3 else "")
        plus = "+" + (f"q{i}" if rng.random() < 0.2 else "")
        e = eol if rng.random() < 0.9 else "\r\n"
        rec = head + e + seq + e + plus + e + qual + e
        if i == weird_at:  # a two-line sequence: next() alone reads it
            rec = head + e + seq + e + seq + e + plus + e + qual + qual + e
        out.append(rec)
    return "".join(out)


@pytest.mark.parametrize("chunk,threads", [(1 << 26, 8), (3000, 4), (700, 3), (100000, 16)])
def test_chunked_readers_match_serial_on_fastq(checker, chunk, threads, tmp_path):
    rng = random.Random(chunk + 7)
    p = tmp_path / "r.fq"
    p.write_text(random_fastq(rng, 3000))
    serial, chunked = both(checker, str(p), chunk, threads)
    assert serial.count(b"\n") == 3000
    assert chunked == serial


def test_fastq_readers_hand_over_at_a_multiline_record(checker, tmp_path):
    rng = random.Random(9)
    p = tmp_path / "m.fq"
    p.write_text(random_fastq(rng, 2000, weird_at=1500) + ">fa\nACGT\n@z\nAC\n+\n@@\n")
    for chunk, threads in ((4000, 4), (1 << 20, 8)):
        serial, chunked = both(checker, str(p), chunk, threads)
        assert chunked == serial


@pytest.mark.parametrize("case", [b"@a\nACGT\n+\nIIII", b"@a\nACGT\n+\nIII\n@b\nAC\n+\nII\n",
                                  b"@a x\r\nAC\r\n+\r\n@@\r\n@b\nA\n+\n+\n", b"@a\nACGT\n+\nIIIII\n@b\nA\n+\nI\n",
                                  b"@a\n\n+\n\n@b\nA\n+\nI\n", b"@a\nA\n+\n", b"@\nAC\n+\nII\n@"])
def test_fastq_record_edges(checker, case, tmp_path):
    """Quality line without '\\n' at the end, a short or long quality line
    (next() ends the input there), CRLF, an empty sequence, a record cut
    after its '+' line, empty ids and a lone '@' at the end."""
    p = tmp_path / "e.fq"
    p.write_bytes(case)
    for chunk, threads in ((3, 2), (1 << 20, 4)):
        serial, chunked = both(checker, str(p), chunk, threads)
        assert chunked == serial
