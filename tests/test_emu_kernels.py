"""The classify kernels' device code, executed on the CPU by the SIMT
emulation (tests/emu: 64 host threads per wave, barrier-synchronised
ballots and shuffles), against the CPU oracle.  Covers the kernel logic
without a GPU; the same code runs on the MI355X in tests/test_gpu_*.py."""
import os
import random
import subprocess

import numpy as np
import pytest

from helpers import Oracle, chimeric_reads, edge_reads, read_fasta

HERE = os.path.dirname(os.path.abspath(__file__))
G = os.path.join(HERE, "golden")


@pytest.fixture(scope="module")
def emu_lib():
    r = subprocess.run(["make", "-C", os.path.join(HERE, "emu")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    import sys
    sys.path.insert(0, os.path.join(HERE, "emu"))
    import emu
    return emu, emu.load()


def _check(emu, lib, d, seq, off, pflags, *, conf=0.0, g=0, quick=False, tab=256, mode=1, hll=False,
           cap_pct=None, unit_elements=None):
    import kraken2_amd as k2
    db = emu.read_db(d)
    out = emu.classify(lib, db, seq, off, pflags, conf=conf, min_groups=g, quick=quick, tab_cap=tab, mode=mode,
                       hll=hll, kmer_counts=hll, cap_pct=cap_pct, unit_elements=unit_elements)
    orc = Oracle()
    h = orc.open(d)
    n = len(pflags)
    paired_all = bool(pflags.all())
    assert pflags.all() or not pflags.any()
    ref = orc.classify(h, seq, off, paired=paired_all, quick=quick, confidence=conf, min_hit_groups=g)
    for i in range(n):
        ro = int(out["run_off"][i])
        runs = out["runs"][2 * ro:2 * (ro + int(out["n_runs"][i]))].reshape(-1, 2)
        if out["status"][i] & 2:
            # out of run space: runs truncated (so the call, taken from the
            # written runs, is not final: the library classifies the fragment
            # again into exact space); the counters are counted in full
            assert int(out["n_runs"][i]) > int(out["run_off"][i + 1] - out["run_off"][i])
            assert (int(out["groups"][i]), int(out["total"][i])) == (int(ref["groups"][i]), int(ref["total"][i]))
            continue
        got = (int(out["call"][i]), int(out["groups"][i]))
        exp = (int(ref["call"][i]), int(ref["groups"][i]))
        assert got == exp, i
        if not quick:
            assert int(out["total"][i]) == int(ref["total"][i])
            got_text = k2.format_hitlist(runs, db["ext"])
            assert got_text == ref["text"][i], (i, int(out["status"][i]), got_text[:400], ref["text"][i][:400])
    return out



@pytest.mark.parametrize("mode,units", [(0, None), (1, None), (1, 300)])
def test_emulated_single_end(emu_lib, synth_db, mode, units):
    """Plain, persistent with one fragment per claim, and persistent with
    work units of about 300 elements (several short reads per claim)."""
    import kraken2_amd as k2
    emu, lib = emu_lib
    _, genomes = read_fasta(os.path.join(synth_db, "library.fna"))
    reads = edge_reads(genomes, 60, seed=31 + mode)
    seq, off = k2.pack_fragments(reads)
    out = _check(emu, lib, synth_db, seq, off, np.zeros(len(reads), np.uint8), g=2, mode=mode,
                 unit_elements=units)
    if units:
        assert 1 < len(out["units"]) - 1 < len(reads)


def test_emulated_paired_confidence(emu_lib, synth_db):
    import kraken2_amd as k2
    emu, lib = emu_lib
    _, genomes = read_fasta(os.path.join(synth_db, "library.fna"))
    m1 = edge_reads(genomes, 40, seed=41)
    m2 = edge_reads(genomes, 40, seed=42)
    seq, off = k2.pack_fragments(m1, m2)
    _check(emu, lib, synth_db, seq, off, np.ones(40, np.uint8), conf=0.3, g=2)


def test_emulated_quick_and_dense_table(emu_lib, synth_db):
    import kraken2_amd as k2
    emu, lib = emu_lib
    _, genomes = read_fasta(os.path.join(synth_db, "library.fna"))
    reads = edge_reads(genomes, 40, seed=51, min_len=100, max_len=300)
    seq, off = k2.pack_fragments(reads)
    _check(emu, lib, synth_db, seq, off, np.zeros(40, np.uint8), g=2, quick=True)
    out = _check(emu, lib, synth_db, seq, off, np.zeros(40, np.uint8), g=0, tab=2)
    # a 2-slot table overflows for reads with 3+ distinct taxa: flagged for the dense pass
    assert (out["status"] & 1).any()


def test_emulated_translated(emu_lib):
    import kraken2_amd as k2
    emu, lib = emu_lib
    d = os.path.join(G, "prot")
    names, seqs = [], []
    for line in open(os.path.join(d, "reads_1.fa")):
        if not line.startswith(">"):
            seqs.append(line.strip())
    seqs = seqs[:40]
    seq, off = k2.pack_fragments(seqs)
    _check(emu, lib, d, seq, off, np.zeros(len(seqs), np.uint8), g=0)


def test_emulated_run_space_overflow(emu_lib, synth_db):
    """With little run space reserved, fragments with more runs are flagged
    (status bit 1) while their counters stay exact; the rest match the oracle
    in full."""
    import kraken2_amd as k2
    emu, lib = emu_lib
    _, genomes = read_fasta(os.path.join(synth_db, "library.fna"))
    reads = chimeric_reads(genomes, 12, seed=61) + edge_reads(genomes, 12, seed=62)
    seq, off = k2.pack_fragments(reads)
    out = _check(emu, lib, synth_db, seq, off, np.zeros(len(reads), np.uint8), g=2, cap_pct=1)
    assert (out["status"] & 2).any() and not (out["status"] & 2).all()


def test_emulated_revcom_version0(emu_lib, synth_db_rc0):
    import kraken2_amd as k2
    emu, lib = emu_lib
    _, genomes = read_fasta(os.path.join(synth_db_rc0, "library.fna"))
    reads = edge_reads(genomes, 40, seed=71)
    seq, off = k2.pack_fragments(reads)
    _check(emu, lib, synth_db_rc0, seq, off, np.zeros(len(reads), np.uint8), g=0)


@pytest.mark.parametrize("name", ["one_empty_102", "full_128", "small_37", "sparse_1001"])
def test_emulated_probe_edge_tables(emu_lib, edge_tables, name):
    """Random tables with 8-bit compact keys: chains that wrap past the last
    cell into the padding copies (lean probe), a table with no empty cell and
    one under 64 cells (bounded probe)."""
    import kraken2_amd as k2
    emu, lib = emu_lib
    d = edge_tables[name]
    _, genomes = read_fasta(os.path.join(d, "library.fna"))
    reads = edge_reads(genomes, 16, seed=81, min_len=40, max_len=160)
    seq, off = k2.pack_fragments(reads)
    out = _check(emu, lib, d, seq, off, np.zeros(len(reads), np.uint8), g=0)
    assert (out["groups"] > 0).any()


@pytest.mark.parametrize("extras", [False, True])
def test_emulated_split_long_fragments(emu_lib, synth_db, extras):
    """Long fragments split over waves (segments of 2048 k-mers, cut points
    moved past ambiguous bases, mates split separately), stitched: every
    fragment's call, counters, hitlist, HLL records and the per-taxon k-mer
    counts equal the unsplit scan's, and calls and hitlists the oracle's."""
    emu, lib = emu_lib
    import kraken2_amd as k2
    names, genomes = read_fasta(os.path.join(synth_db, "library.fna"))
    rng = random.Random(31)
    m1 = chimeric_reads(genomes, 10, seed=32, seg=700, length=9000)
    m2 = chimeric_reads(genomes, 10, seed=33, seg=300, length=6000)
    # ambiguous stretches near and at the would-be cut points, lowercase, short mates
    for r in range(len(m1)):
        s = list(m1[r])
        for c in (2048 + 34 - 3, 4096 + 20, 6000):
            for j in range(c, min(len(s), c + rng.choice([1, 5, 40]))):
                s[j] = rng.choice("NRYn")
        m1[r] = "".join(s).lower() if r % 3 == 0 else "".join(s)
    m2[4] = m2[4][:20]
    m2[5] = "N" * 5000
    m1 += edge_reads(genomes, 6, seed=34)
    m2 += edge_reads(genomes, 6, seed=35)
    flags = np.array([i % 2 for i in range(len(m1))], np.uint8)
    m2 = [b if f else "" for b, f in zip(m2, flags)]
    seq, off = k2.pack_fragments(m1, m2)
    db = emu.read_db(synth_db)
    whole = emu.classify(lib, db, seq, off, flags, mode=1, hll=extras, kmer_counts=extras)
    split, segs = emu.classify_split(lib, db, seq, off, flags, hll=extras, kmer_counts=extras)
    assert segs > 20
    n = len(flags)
    for i in range(n):
        for f in ("call", "groups", "total", "n_runs", "status", "n_recs"):
            assert int(split[f][i]) == int(whole[f][i]), (i, f)
        ro = int(whole["run_off"][i])
        nr = int(whole["n_runs"][i])
        assert np.array_equal(split["runs"][2 * ro:2 * (ro + nr)], whole["runs"][2 * ro:2 * (ro + nr)]), i
        if extras:
            nrec = int(whole["n_recs"][i])
            assert np.array_equal(split["recs"][2 * ro:2 * (ro + nrec)], whole["recs"][2 * ro:2 * (ro + nrec)]), i
    if extras:
        assert np.array_equal(split["kmer_counts"], whole["kmer_counts"])
    orc = Oracle()
    h = orc.open(synth_db)
    for pf in (0, 1):
        idx = np.flatnonzero(flags == pf)
        parts1 = [m1[i] for i in idx]
        parts2 = [m2[i] for i in idx] if pf else None
        s2, o2 = k2.pack_fragments(parts1, parts2)
        ref = orc.classify(h, s2, o2, paired=bool(pf))
        for j, i in enumerate(idx):
            ro = int(split["run_off"][i])
            runs = split["runs"][2 * ro:2 * (ro + int(split["n_runs"][i]))].reshape(-1, 2)
            assert (int(split["call"][i]), k2.format_hitlist(runs, db["ext"])) == \
                (int(ref["call"][j]), ref["text"][j]), i
