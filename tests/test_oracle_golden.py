"""Pins the CPU oracle (oracle/k2cpu) to the reference.

Every expected value here was written by the compiled reference or by a
harness linking its sources (tests/golden/make_golden.py); the oracle must
reproduce each one exactly.  The brute-force scanner test pins the window
definition independently of both implementations."""
import hashlib
import json
import os
import random
import shutil
import subprocess

import numpy as np
import pytest

from helpers import ORACLE_BIN, Oracle, run

HERE = os.path.dirname(os.path.abspath(__file__))
G = os.path.join(HERE, "golden")
TOY = os.path.join(G, "toy")
PROT = os.path.join(G, "prot")


@pytest.fixture(scope="module")
def orc():
    return Oracle()


def _scan(orc, k, l, mask, toggle, dna, start, finish, s):
    b = s.encode()
    cap = len(b) + 8
    amb = np.zeros(cap, np.uint8)
    mins = np.zeros(cap, np.uint64)
    n = orc.lib.k2cpu_scan(k, l, mask, toggle, dna, start, finish, b, len(b), amb.ctypes.data,
                           mins.ctypes.data, cap)
    assert n >= 0
    return list(zip(amb[:n].tolist(), mins[:n].tolist()))


def test_scanner_matches_reference_vectors(orc):
    cases = open(os.path.join(G, "scan_cases.txt")).read().splitlines()
    exp_lines = open(os.path.join(G, "scan_expected.txt")).read().splitlines()
    pos = 0
    for line in cases:
        k, l, mask, tog, dna, a, b, s = line.split(" ")
        s = "" if s == "-" else s
        got = _scan(orc, int(k), int(l), int(mask, 16), int(tog, 16), int(dna), int(a), int(b), s)
        assert exp_lines[pos].startswith("# ")
        n = int(exp_lines[pos][2:])
        exp = [(int(x.split()[0]), int(x.split()[1], 16)) for x in exp_lines[pos + 1:pos + 1 + n]]
        pos += 1 + n
        assert got == exp, line


def test_scanner_revcom_version0_vectors(orc):
    """Old databases' reverse complement (revcom_version 0) against the
    reference scanner."""
    cases = open(os.path.join(G, "scan_rc0_cases.txt")).read().splitlines()
    exp_lines = open(os.path.join(G, "scan_rc0_expected.txt")).read().splitlines()
    pos = 0
    for line in cases:
        k, l, mask, tog, dna, a, b, s = line.split(" ")
        s = "" if s == "-" else s
        bs = s.encode()
        amb = np.zeros(len(bs) + 1, np.uint8)
        mins = np.zeros(len(bs) + 1, np.uint64)
        n = orc.lib.k2cpu_scan_rc(int(k), int(l), int(mask, 16), int(tog, 16), int(dna), 0, int(a), int(b), bs,
                                  len(bs), amb.ctypes.data, mins.ctypes.data, len(bs) + 1)
        got = [(int(amb[i]), int(mins[i])) for i in range(n)]
        assert exp_lines[pos].startswith("# ")
        m = int(exp_lines[pos][2:])
        exp = [(int(x.split()[0]), int(x.split()[1], 16)) for x in exp_lines[pos + 1:pos + 1 + m]]
        pos += 1 + m
        assert got == exp, line
    assert pos == len(exp_lines)


def _brute(k, l, mask, toggle, s):
    """Definition of the emitted minimizers of a scan from position 0:
    one output per k-mer end e >= k-1; ambiguous iff a non-ACGT base lies in
    [e-k+2, e] (for k == l: in [e-l+1, e]); otherwise the minimum over the
    window's l-mers that contain no non-ACGT base of
    (canonical(lmer) & mask) ^ toggle, then ^ toggle."""
    code = {"A": 0, "C": 1, "G": 2, "T": 3}
    L = len(s)
    lm = (1 << (2 * l)) - 1
    tog = toggle & lm

    def cand(st):
        sub = s[st:st + l].upper()
        if any(c not in code for c in sub):
            return None
        x = 0
        for c in sub:
            x = (x << 2) | code[c]
        rc = 0
        for c in reversed(sub):
            rc = (rc << 2) | (3 - code[c])
        v = min(x, rc)
        if mask:
            v &= mask
        return v ^ tog

    out = []
    for e in range(k - 1, L):
        lo = e - k + 2 if k > l else e - l + 1
        amb = any(s[i].upper() not in code for i in range(max(lo, 0), e + 1))
        if amb:
            out.append((1, None))
            continue
        vals = [cand(st) for st in range(e - k + 1, e - l + 2)]
        vals = [v for v in vals if v is not None]
        out.append((0, min(vals) ^ tog))
    return out


@pytest.mark.parametrize("k,l,masked", [(35, 31, True), (35, 31, False), (12, 8, False), (9, 9, False),
                                        (20, 5, True)])
def test_scanner_brute_force_definition(orc, k, l, masked):
    rng = random.Random(k * 100 + l)
    mask = 0
    if masked:
        for i in range(l):
            mask |= (3 if (i % 3) else 0) << (2 * i)
    for t in range(150):
        L = rng.randrange(0, 160)
        s = "".join(rng.choice("ACGTACGTACGTNacgtR") for _ in range(L))
        got = _scan(orc, k, l, mask, 0xe37e28c4271b5a2d, 1, 0, -1, s)
        exp = _brute(k, l, mask, 0xe37e28c4271b5a2d, s)
        assert len(got) == len(exp)
        for (ga, gm), (ea, em) in zip(got, exp):
            assert ga == ea
            if not ea:
                assert gm == em


def test_hll_matches_reference_vectors(orc):
    script = open(os.path.join(G, "hll_cases.txt")).read()
    exp = [list(map(int, x.split())) for x in open(os.path.join(G, "hll_expected.txt")).read().splitlines()]
    out = np.zeros(2 * len(exp) + 8, np.uint64)
    n = orc.lib.k2cpu_hll_script(script.encode(), out.ctypes.data, out.shape[0])
    got = out[:n].reshape(-1, 2).tolist()
    assert got == exp


def test_translation_matches_reference_vectors(orc):
    cases = open(os.path.join(G, "translate_cases.txt")).read().splitlines()
    exp = open(os.path.join(G, "translate_expected.txt")).read().splitlines()
    for i, s in enumerate(cases):
        s = "" if s == "-" else s
        buf = np.zeros(6 * (len(s) // 3 + 2) + 16, np.uint8)
        n = orc.lib.k2cpu_translate(s.encode(), len(s), buf.ctypes.data, buf.shape[0])
        frames = bytes(buf[:n]).split(b"\0")[:6]
        frames = [f.decode() or "-" for f in frames]
        assert frames == exp[6 * i:6 * i + 6], s


def _opts_fields(path):
    b = open(path, "rb").read()
    k, l, mask, tog = np.frombuffer(b[:32], np.uint64).tolist()
    dna = b[32]
    minh = int(np.frombuffer(b[40:48], np.uint64)[0])
    rc, dbv, dbt = np.frombuffer(b[48:60], np.int32).tolist()
    return (k, l, mask, tog, dna, minh, rc, dbv, dbt)


def test_build_matches_reference_bytes(tmp_path):
    """The oracle's deterministic build writes the reference's hash.k2d and
    taxo.k2d byte for byte (opts.k2d field by field: its padding is
    uninitialised in the reference)."""
    inp = os.path.join(G, "toy_inputs")
    lib = tmp_path / "lib.fna"
    with open(lib, "w") as f:
        for name in sorted(os.listdir(inp)):
            if name.endswith(".fa"):
                f.write(open(os.path.join(inp, name)).read())
    tmpl = "1" * 17 + "01" * 7
    with open(lib, "rb") as f:
        est = int(run([os.path.join(ORACLE_BIN, "k2cpu_estimate_capacity"), "-k", "35", "-l", "31",
                       "-S", tmpl], stdin=f).stdout)
    cap = int((est + 8192) / 0.7)
    assert 32 + 4 * cap == os.path.getsize(os.path.join(TOY, "hash.k2d"))
    with open(lib, "rb") as f:
        run([os.path.join(ORACLE_BIN, "k2cpu_build_db"), "-k", "35", "-l", "31", "-S", tmpl,
             "-H", str(tmp_path / "hash.k2d"), "-t", str(tmp_path / "taxo.k2d"), "-o", str(tmp_path / "opts.k2d"),
             "-n", inp, "-m", os.path.join(inp, "seqid2taxid.map"), "-c", str(cap), "-B", "16384",
             "-b", "2048", "-r", "0"], stdin=f)
    for fn in ("hash.k2d", "taxo.k2d"):
        assert open(tmp_path / fn, "rb").read() == open(os.path.join(TOY, fn), "rb").read(), fn
    assert _opts_fields(tmp_path / "opts.k2d") == _opts_fields(os.path.join(TOY, "opts.k2d"))


def _interleave(r1, r2, out):
    with open(r1) as a, open(r2) as b, open(out, "w") as o:
        la, lb = a.readlines(), b.readlines()
        for i in range(0, len(la), 4):
            o.writelines(la[i:i + 4] + lb[i:i + 4])


TOY_RUNS = {
    "edge_g0": ["-T", "0", "-g", "0", "-R", "{rep}", "-K", "{d}/edge.fa"],
    "edge_g2_names": ["-T", "0", "-g", "2", "-n", "{d}/edge.fa"],
    "edge_conf05": ["-T", "0.5", "-g", "0", "-R", "{rep}", "-z", "{d}/edge.fa"],
    "edge_quick": ["-q", "-g", "2", "{d}/edge.fa"],
    "paired": ["-P", "-T", "0.1", "-g", "2", "-R", "{rep}", "-K", "{d}/reads_1.fq", "{d}/reads_2.fq"],
    "paired_q20": ["-P", "-Q", "20", "-g", "2", "{d}/reads_1.fq", "{d}/reads_2.fq"],
    "interleaved": ["-S", "-g", "2", "-R", "{rep}", "{tmp}/interleaved.fq"],
    "single_cu": ["-g", "2", "-C", "{tmp}/single_cu.C.fq", "-U", "{tmp}/single_cu.U.fq", "{d}/reads_1.fq"],
}


def run_golden_case(binary, name, tmp, threads="3", extra=()):
    _interleave(os.path.join(TOY, "reads_1.fq"), os.path.join(TOY, "reads_2.fq"), os.path.join(tmp, "interleaved.fq"))
    rep = os.path.join(tmp, f"{name}.kreport")
    args = [a.format(d=TOY, rep=rep, tmp=tmp) for a in TOY_RUNS[name]]
    out = run([binary, "-H", f"{TOY}/hash.k2d", "-t", f"{TOY}/taxo.k2d", "-o", f"{TOY}/opts.k2d",
               "-p", threads, *extra, *args]).stdout
    assert out == open(os.path.join(TOY, f"{name}.kraken")).read(), name
    if "{rep}" in TOY_RUNS[name]:
        assert open(rep).read() == open(os.path.join(TOY, f"{name}.kreport")).read(), name
    if name == "single_cu":
        digests = json.load(open(os.path.join(TOY, "sequence_outputs.json")))
        for fn, h in digests.items():
            assert hashlib.sha256(open(os.path.join(tmp, fn), "rb").read()).hexdigest() == h, fn


@pytest.mark.parametrize("name", sorted(TOY_RUNS))
def test_oracle_classify_matches_reference(name, tmp_path):
    run_golden_case(os.path.join(ORACLE_BIN, "k2cpu_classify"), name, str(tmp_path))


@pytest.mark.parametrize("name", ["single", "paired"])
def test_oracle_translated_matches_reference(name, tmp_path):
    args = {"single": ["-g", "0", "-R", str(tmp_path / "r.kreport"), "-K", f"{PROT}/reads_1.fa"],
            "paired": ["-P", "-T", "0.2", "-g", "2", f"{PROT}/reads_1.fa", f"{PROT}/reads_2.fa"]}[name]
    out = run([os.path.join(ORACLE_BIN, "k2cpu_classify"), "-H", f"{PROT}/hash.k2d", "-t", f"{PROT}/taxo.k2d",
               "-o", f"{PROT}/opts.k2d", *args]).stdout
    assert out == open(os.path.join(PROT, f"{name}.kraken")).read()
    if name == "single":
        assert open(tmp_path / "r.kreport").read() == open(os.path.join(PROT, "single.kreport")).read()


def test_table_invariant_every_library_minimizer_is_found(orc):
    """fmix64, the modulo and the probe are pinned by the table itself: every
    non-ambiguous minimizer of the library a table was built from must be
    found, with a value that is an ancestor-or-self of its genome's taxon."""
    inp = os.path.join(G, "toy_inputs")
    idmap = dict(l.split("\t") for l in open(os.path.join(inp, "seqid2taxid.map")).read().splitlines())
    h = orc.open(TOY)
    mask = 0
    tmpl = "1" * 17 + "01" * 7
    for i, c in enumerate(reversed(tmpl)):
        if c == "1":
            mask |= 3 << (2 * i)
    found = 0
    for name in sorted(os.listdir(inp)):
        if not name.endswith(".fa"):
            continue
        txt = open(os.path.join(inp, name)).read().split(">")[1:]
        for rec in txt[:2]:
            lines = rec.splitlines()
            seq = "".join(lines[1:])[:3000]
            got = _scan(orc, 35, 31, mask, 0xe37e28c4271b5a2d, 1, 0, -1, seq)
            keys = np.array([m for a, m in got if not a], dtype=np.uint64)
            vals = orc.get_batch(h, keys)
            assert (vals != 0).all()
            found += len(keys)
    assert found > 1000
