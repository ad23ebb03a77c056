"""Multi-database merge (the reference's merge tool, src/merge.cc) on the CPU.

The oracle restatement (oracle/bin/k2cpu_merge) and the k2g_merge front-end
with its merge kernel executed on the host (tests/emu/k2g_merge_emu: the
front-end's own host code, the kernel's device code through the SIMT shim)
are checked against the compiled reference's outputs (tests/golden/merge,
written by tests/golden/make_golden.py) and against each other on random
hit lists; where the reference binary is built (oracle/_ref/merge) the
random cases are checked against it too.
"""
import hashlib
import json
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
GOLD = os.path.join(HERE, "golden")
sys.path.insert(0, GOLD)
sys.path.insert(0, HERE)
from make_golden import MERGE_CASES  # noqa: E402
from helpers import RAGGED_MERGE_LINES, random_kraken_pairs, taxonomy_external_ids  # noqa: E402

ORACLE = os.path.join(ROOT, "oracle", "bin", "k2cpu_merge")
REF = os.path.join(ROOT, "oracle", "_ref", "merge")
EXPECTED = json.load(open(os.path.join(GOLD, "merge", "expected.json")))


@pytest.fixture(scope="module")
def emu_merge():
    r = subprocess.run(["make", "-C", os.path.join(HERE, "emu"), "k2g_merge_emu"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return os.path.join(HERE, "emu", "k2g_merge_emu")


# the emulated front-end is built with AddressSanitizer; the OpenMP runtime
# keeps one allocation of its own until exit, which is not a leak of ours
SAN_ENV = {"ASAN_OPTIONS": "detect_leaks=0"}


def run_merge(exe, taxo, in1, in2, opts, d, threads=3, env=None):
    out, hdr, rep = (os.path.join(d, x) for x in ("out.kraken", "headers.txt", "report.txt"))
    r = subprocess.run([exe, "-T", str(threads), "-t", taxo, "-o", out, "-c", hdr, "-r", rep, *opts, in1, in2],
                       capture_output=True, text=True, env=dict(env or os.environ, **SAN_ENV))
    assert r.returncode == 0, r.stderr
    return {k: open(p, "rb").read() for k, p in (("output", out), ("headers", hdr), ("report", rep))}


def golden_case(case):
    tax, in1, in2, opts = MERGE_CASES[case]
    return (os.path.join(GOLD, tax, "taxo.k2d"), os.path.join(GOLD, in1 + ".kraken"),
            os.path.join(GOLD, in2 + ".kraken"), opts)


def check_digests(got, case):
    for key, exp in EXPECTED[case].items():
        assert got[key].count(b"\n") == exp["lines"], (case, key)
        assert hashlib.sha256(got[key]).hexdigest() == exp["sha256"], (case, key)


@pytest.mark.parametrize("case", sorted(MERGE_CASES))
def test_oracle_merge_matches_reference_golden(case, tmp_path):
    check_digests(run_merge(ORACLE, *golden_case(case), str(tmp_path)), case)


@pytest.mark.parametrize("search", [False, True])
@pytest.mark.parametrize("case", sorted(MERGE_CASES))
def test_emulated_merge_front_end_matches_reference_golden(emu_merge, case, search, tmp_path):
    """search: taxids looked up by binary search over the sorted map (the
    library's path for external ids of 2^26 and above) instead of the
    direct-mapped table."""
    env = dict(os.environ, K2G_EMU_MERGE_SEARCH="1" if search else "0")
    check_digests(run_merge(emu_merge, *golden_case(case), str(tmp_path), threads=4, env=env), case)


@pytest.mark.parametrize("seed,opts,batch", [(1, [], 0), (2, ["-i", "0.15"], 7), (3, ["-i", "0.5", "-n", "-z"], 100),
                                             (4, ["-i", "1.0"], 1)])
def test_random_hit_lists_emulated_vs_oracle(emu_merge, seed, opts, batch, tmp_path):
    """batch: lines per front-end batch (0 = default), so the pipeline of
    reader, two merge slots and printing runs over many batches."""
    taxo = os.path.join(GOLD, "toy", "taxo.k2d")
    a, b = random_kraken_pairs(taxonomy_external_ids(taxo), 1500, seed)
    pa, pb = tmp_path / "a.kraken", tmp_path / "b.kraken"
    pa.write_text(a)
    pb.write_text(b)
    runs = {}
    env = dict(os.environ, K2G_MERGE_BATCH_LINES=str(batch)) if batch else None
    for name, exe in (("oracle", ORACLE), ("emulated", emu_merge)) + ((("reference", REF),)
                                                                      if os.path.exists(REF) else ()):
        d = tmp_path / name
        d.mkdir()
        runs[name] = run_merge(exe, taxo, str(pa), str(pb), opts, str(d), env=env)
    for name, got in runs.items():
        for key in ("output", "headers", "report"):
            assert got[key] == runs["oracle"][key], (name, key)


def test_ragged_hit_lists_emulated_vs_oracle(emu_merge, tmp_path):
    taxo = os.path.join(GOLD, "toy", "taxo.k2d")
    pa, pb = tmp_path / "a.kraken", tmp_path / "b.kraken"
    pa.write_text("".join(RAGGED_MERGE_LINES[0]))
    pb.write_text("".join(RAGGED_MERGE_LINES[1]))
    for i, opts in enumerate(([], ["-i", "0.5"])):
        d1, d2 = tmp_path / f"e{i}", tmp_path / f"o{i}"
        d1.mkdir()
        d2.mkdir()
        assert run_merge(emu_merge, taxo, str(pa), str(pb), opts, str(d1)) == \
            run_merge(ORACLE, taxo, str(pa), str(pb), opts, str(d2))


def test_merge_command_line_errors(emu_merge, tmp_path):
    taxo = os.path.join(GOLD, "toy", "taxo.k2d")
    a = os.path.join(GOLD, "merge", "a_edge.kraken")
    out = str(tmp_path / "o")

    def code(*args):
        return subprocess.run([emu_merge, *args], capture_output=True, env=dict(os.environ, **SAN_ENV)).returncode

    assert code("-t", taxo, a, a) == 1            # no -o (src/merge.cc:1034-1036)
    assert code("-o", out, a, a) == 2              # no -t
    assert code("-o", out, "-t", taxo) == 3        # no input
    assert code("-o", out, "-t", taxo, a) == 4     # one input
    assert code("-o", out, "-t", taxo, a, a, a) == 5
    assert code("-o", out, "-t", taxo, "-m", a, a) == 64  # mpa-style report not provided
    short = tmp_path / "short.kraken"
    short.write_text("".join(open(a).readlines()[:10]))
    assert code("-o", out, "-t", taxo, a, str(short)) == 65  # line counts differ
