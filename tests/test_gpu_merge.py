"""k2g_merge on the GPU (the reference's merge tool, src/merge.cc).

The front-end (kraken2_amd/bin/k2g_merge, merging through libk2g's
k2g_merge entry points) against the compiled reference's outputs
(tests/golden/merge) and against the oracle (oracle/bin/k2cpu_merge) on
random hit lists; the entry points through ctypes on empty and ragged
batches, checked against the oracle restatement.
"""
import hashlib
import json
import os
import subprocess
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
GOLD = os.path.join(HERE, "golden")
sys.path.insert(0, GOLD)
sys.path.insert(0, HERE)
sys.path.insert(0, ROOT)
from make_golden import MERGE_CASES  # noqa: E402
from helpers import RAGGED_MERGE_LINES, random_kraken_pairs, taxonomy_external_ids  # noqa: E402

import kraken2_amd as k2  # noqa: E402

pytestmark = pytest.mark.gpu

EXE = os.path.join(ROOT, "kraken2_amd", "bin", "k2g_merge")
ORACLE = os.path.join(ROOT, "oracle", "bin", "k2cpu_merge")
EXPECTED = json.load(open(os.path.join(GOLD, "merge", "expected.json")))
TOY_TAXO = os.path.join(GOLD, "toy", "taxo.k2d")


def run_merge(exe, taxo, in1, in2, opts, d, threads=4, batch=0):
    out, hdr, rep = (os.path.join(d, x) for x in ("out.kraken", "headers.txt", "report.txt"))
    print(f"merge: {os.path.basename(exe)} {' '.join(opts)} {os.path.basename(in1)} {os.path.basename(in2)}"
          f" batch {batch}", flush=True)
    env = dict(os.environ, K2G_MERGE_BATCH_LINES=str(batch)) if batch else None
    r = subprocess.run([exe, "-T", str(threads), "-t", taxo, "-o", out, "-c", hdr, "-r", rep, *opts, in1, in2],
                       capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stderr
    return {k: open(p, "rb").read() for k, p in (("output", out), ("headers", hdr), ("report", rep))}


@pytest.mark.parametrize("case", sorted(MERGE_CASES))
def test_gpu_merge_matches_reference_golden(case, tmp_path):
    tax, in1, in2, opts = MERGE_CASES[case]
    got = run_merge(EXE, os.path.join(GOLD, tax, "taxo.k2d"), os.path.join(GOLD, in1 + ".kraken"),
                    os.path.join(GOLD, in2 + ".kraken"), opts, str(tmp_path))
    for key, exp in EXPECTED[case].items():
        assert got[key].count(b"\n") == exp["lines"], (case, key)
        assert hashlib.sha256(got[key]).hexdigest() == exp["sha256"], (case, key)


@pytest.mark.parametrize("seed,opts,batch", [(11, [], 0), (12, ["-i", "0.2", "-n", "-z"], 3000), (13, ["-i", "0.6"], 7)])
def test_gpu_merge_random_vs_oracle(seed, opts, batch, tmp_path):
    a, b = random_kraken_pairs(taxonomy_external_ids(TOY_TAXO), 40000, seed)
    pa, pb = tmp_path / "a.kraken", tmp_path / "b.kraken"
    pa.write_text(a)
    pb.write_text(b)
    (tmp_path / "gpu").mkdir()
    (tmp_path / "cpu").mkdir()
    got = run_merge(EXE, TOY_TAXO, str(pa), str(pb), opts, str(tmp_path / "gpu"), batch=batch)
    exp = run_merge(ORACLE, TOY_TAXO, str(pa), str(pb), opts, str(tmp_path / "cpu"))
    for key in ("output", "headers", "report"):
        assert got[key] == exp[key], key


def test_gpu_merge_ragged_lines_vs_oracle(tmp_path):
    """Hit lists whose totals differ between the inputs (the second list
    shorter or longer than the first), count-0 and unknown taxa."""
    lines_a, lines_b = RAGGED_MERGE_LINES
    pa, pb = tmp_path / "a.kraken", tmp_path / "b.kraken"
    pa.write_text("".join(lines_a))
    pb.write_text("".join(lines_b))
    for opts in ([], ["-i", "0.5"]):
        d1, d2 = tmp_path / f"g{len(opts)}", tmp_path / f"c{len(opts)}"
        d1.mkdir()
        d2.mkdir()
        got = run_merge(EXE, TOY_TAXO, str(pa), str(pb), opts, str(d1))
        exp = run_merge(ORACLE, TOY_TAXO, str(pa), str(pb), opts, str(d2))
        assert got == exp


def test_merge_entry_points_empty_and_errors():
    m = k2.Merger(TOY_TAXO)
    try:
        z = np.zeros(1, np.uint64)
        out = m.merge(np.zeros(0, np.int32), z, np.zeros(0, np.int32), z)
        assert len(out.call) == 0 and len(out.tokens) == 0
        # a read with no items in the first list prints nothing
        out = m.merge(np.zeros(0, np.int32), np.zeros(2, np.uint64), np.array([211044, 5], np.int32),
                      np.array([0, 1], np.uint64))
        assert out.token_off.tolist() == [0, 0] and out.call.tolist() == [0]
        with pytest.raises(k2.K2GError):  # different read counts
            m.merge(np.zeros(0, np.int32), np.zeros(2, np.uint64), np.zeros(0, np.int32), z)
        with pytest.raises(k2.K2GError):  # offsets not starting at 0
            m.merge(np.array([1, 1], np.int32), np.array([1, 1], np.uint64), np.array([1, 1], np.int32),
                    np.array([1, 1], np.uint64))
        with pytest.raises(k2.K2GError):  # decreasing offsets
            m.merge(np.array([1, 1], np.int32), np.array([0, 1, 0], np.uint64), np.array([1, 1], np.int32),
                    np.array([0, 1, 1], np.uint64))
    finally:
        m.close()


def _printed(v):  # int_to_string of src/merge.cc:251-326 for the values classify writes
    return {0x7FFFFFFF: "A", 0x7FFFFFFE: "|", 0x7FFFFFFD: "-"}.get(int(v), str(int(v)))


def test_merge_text_mode_equals_pairs_mode():
    """The entry point's two result forms: pairs (token_off, tokens, joined)
    printed as write_hit_list prints them equal the text form, on random
    lists with markers and unknown taxids, at two confidences."""
    ext = [e for e in taxonomy_external_ids(TOY_TAXO)[1:]]
    rng = np.random.default_rng(5)
    n = 3000

    def lists():
        off = np.zeros(n + 1, np.uint64)
        items = []
        for i in range(n):
            k = int(rng.integers(1, 12))
            for _ in range(k):
                r = rng.random()
                t = (0x7FFFFFFF if r < 0.05 else 0 if r < 0.3 else 10 ** 9 if r < 0.32
                     else int(rng.choice(ext)))
                items.append((t, int(rng.integers(1, 9))))
            off[i + 1] = off[i] + k
        return np.array(items, np.int32).reshape(-1, 2), off

    (ia, oa), (ib, ob) = lists(), lists()
    m = k2.Merger(TOY_TAXO)
    try:
        for conf in (0.0, 0.4):
            pairs = m.merge(ia, oa, ib, ob, confidence=conf)
            text = m.merge(ia, oa, ib, ob, confidence=conf, text=True)
            assert np.array_equal(pairs.call, text.call)
            for i in range(n):
                s, e = int(pairs.token_off[i]), int(pairs.token_off[i + 1])
                want = "".join(_printed(pairs.tokens[k, 0]) + ":" + _printed(pairs.tokens[k, 1]) +
                               ("" if pairs.joined[k] else " ") for k in range(s, e))
                assert text.text[i] == want.encode(), i
    finally:
        m.close()


def test_merge_text_entry_point_checks_item_counts(tmp_path):
    """k2g_merge_text tokenizes on the GPU; its hit lists equal the oracle's
    merged output, and slot counts that are not the fields' ':' counts are
    rejected."""
    lines_a = [l.encode() for l in RAGGED_MERGE_LINES[0]]
    lines_b = [l.encode() for l in RAGGED_MERGE_LINES[1]]
    pa, pb = tmp_path / "a.kraken", tmp_path / "b.kraken"
    pa.write_bytes(b"".join(lines_a))
    pb.write_bytes(b"".join(lines_b))
    exp = run_merge(ORACLE, TOY_TAXO, str(pa), str(pb), [], str(tmp_path))["output"].decode().splitlines()
    m = k2.Merger(TOY_TAXO)
    try:
        out = m.merge_text(lines_a, lines_b)
        assert [t.decode() for t in out.text] == [l.split("\t")[4] for l in exp]
        with pytest.raises(k2.K2GError):
            m.merge_text(lines_a, lines_b, item_counts=[5, 1, 1, 3, 3])
        # item offsets counted on the GPU for one side or both, and empty input
        want = [l.split("\t")[4] for l in exp]
        for sides in ((True, True), (True, False), (False, True)):
            out = m.merge_text(lines_a, lines_b, device_counts=sides)
            assert [t.decode() for t in out.text] == want, sides
        for sides in ((True, True), (False, True)):
            out = m.merge_text([], [], device_counts=sides)
            assert len(out.text) == 0 and len(out.call) == 0
    finally:
        m.close()
