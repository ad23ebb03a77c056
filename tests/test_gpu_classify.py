"""GPU parity: the HIP classify core against the CPU oracle on the same DB.

Bit-exact comparison of call, minimizer hit groups, confidence denominator
and the hitlist text of every fragment (src/classify.cc:773-991)."""
import os
import random

import numpy as np
import pytest

from helpers import Oracle, chimeric_reads, edge_reads, read_fasta

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env(synth_db):
    import kraken2_amd as k2
    db = k2.Db.from_dir(synth_db)
    batch = k2.Batch(db)
    orc = Oracle()
    h = orc.open(synth_db)
    names, genomes = read_fasta(os.path.join(synth_db, "library.fna"))
    yield k2, db, batch, orc, h, genomes
    batch.close()
    db.close()


def _compare(k2, db, out, ref, n):
    ext = db.external_ids
    for i in range(n):
        runs = out.runs[out.run_off[i]:out.run_off[i + 1]]
        got = k2.format_hitlist(runs, ext)
        assert (int(out.call[i]), int(out.hit_groups[i]), int(out.total_kmers[i]), got) == \
               (int(ref["call"][i]), int(ref["groups"][i]), int(ref["total"][i]), ref["text"][i]), \
               f"fragment {i}"


def test_get_batch_all_cells_and_misses(env, synth_db):
    k2, db, batch, orc, h, genomes = env
    rng = np.random.default_rng(5)
    keys = rng.integers(0, 2**62, size=200000, dtype=np.uint64)
    assert np.array_equal(db.get_batch(keys), orc.get_batch(h, keys))


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("paired", [False, True])
def test_classify_edge_reads(env, mode, paired):
    k2, db, batch, orc, h, genomes = env
    n = 3000
    m1 = edge_reads(genomes, n, seed=100 + mode + 2 * paired)
    m2 = edge_reads(genomes, n, seed=200 + mode) if paired else None
    seq, off = k2.pack_fragments(m1, m2)
    out = batch.classify(seq, off, paired=paired, mode=mode, min_hit_groups=2)
    ref = orc.classify(h, seq, off, paired=paired, min_hit_groups=2)
    _compare(k2, db, out, ref, n)


@pytest.mark.parametrize("conf,g,quick", [(0.0, 0, False), (0.3, 2, False), (0.75, 1, False),
                                          (1.0, 3, False), (0.0, 2, True), (0.5, 0, True)])
def test_classify_confidence_and_quick(env, conf, g, quick):
    k2, db, batch, orc, h, genomes = env
    n = 2000
    m1 = edge_reads(genomes, n, seed=300, min_len=100, max_len=300)
    seq, off = k2.pack_fragments(m1)
    out = batch.classify(seq, off, confidence=conf, min_hit_groups=g, quick=quick, mode=1)
    ref = orc.classify(h, seq, off, confidence=conf, min_hit_groups=g, quick=quick)
    if quick:
        for i in range(n):
            assert (int(out.call[i]), int(out.hit_groups[i])) == (int(ref["call"][i]), int(ref["groups"][i]))
    else:
        _compare(k2, db, out, ref, n)


def test_long_reads_and_dense_table(env):
    """Long reads cross many chunks; a 2-slot LDS table forces the dense path."""
    k2, db, batch, orc, h, genomes = env
    rng = random.Random(7)
    m1 = []
    for i in range(60):
        g = genomes[rng.randrange(len(genomes))]
        L = rng.randint(500, min(len(g), 15000))
        p = rng.randrange(len(g) - L + 1)
        s = list(g[p:p + L])
        for j in range(len(s)):
            if rng.random() < 0.03:
                s[j] = rng.choice("ACGTN")
        m1.append("".join(s))
    seq, off = k2.pack_fragments(m1)
    ref = orc.classify(h, seq, off, confidence=0.1, min_hit_groups=2)
    for slots, mode in ((256, 0), (2, 1), (4, 0)):
        out = batch.classify(seq, off, confidence=0.1, min_hit_groups=2, mode=mode, table_slots=slots)
        _compare(k2, db, out, ref, len(m1))
        if slots == 2:
            assert out.redone > 0


def test_mixed_paired_flags(env):
    k2, db, batch, orc, h, genomes = env
    n = 1000
    m1 = edge_reads(genomes, n, seed=400, min_len=50, max_len=250)
    m2 = edge_reads(genomes, n, seed=401, min_len=50, max_len=250)
    flags = np.array([i % 3 == 0 for i in range(n)], dtype=np.uint8)
    m2x = [m2[i] if flags[i] else "" for i in range(n)]
    seq, off = k2.pack_fragments(m1, m2x)
    out = batch.classify(seq, off, paired_flags=flags, mode=1)
    for pflag in (0, 1):
        idx = [i for i in range(n) if flags[i] == pflag]
        sub1 = [m1[i] for i in idx]
        sub2 = [m2x[i] for i in idx] if pflag else None
        s2, o2 = k2.pack_fragments(sub1, sub2)
        ref = orc.classify(h, s2, o2, paired=bool(pflag))
        for j, i in enumerate(idx):
            runs = out.runs[out.run_off[i]:out.run_off[i + 1]]
            assert (int(out.call[i]), k2.format_hitlist(runs, db.external_ids)) == \
                   (int(ref["call"][j]), ref["text"][j])


@pytest.mark.parametrize("mode,slots,hll", [(1, 256, False), (0, 256, False), (1, 2, False), (1, 256, True)])
def test_run_space_overflow_redo(env, mode, slots, hll):
    """With 1% of the element count reserved for runs, fragments that need
    more are classified again into exact space; results equal the oracle and
    the exact-space run, HLL records included."""
    k2, db, batch, orc, h, genomes = env
    reads = chimeric_reads(genomes, 150, seed=500) + edge_reads(genomes, 300, seed=501)
    seq, off = k2.pack_fragments(reads)
    n = len(reads)
    ref = orc.classify(h, seq, off, min_hit_groups=2)
    exact = batch.classify(seq, off, min_hit_groups=2, mode=mode, table_slots=slots, hll=hll,
                           kmer_counts=hll, run_cap_pct=100)
    assert batch.redo_counts()[0] == 0
    d_exact = batch.digests(n)
    out = batch.classify(seq, off, min_hit_groups=2, mode=mode, table_slots=slots, hll=hll,
                         kmer_counts=hll, run_cap_pct=1)
    space, dense = batch.redo_counts()
    assert space > 0
    if slots == 2:
        assert dense > 0
    _compare(k2, db, out, ref, n)
    assert np.array_equal(batch.digests(n), d_exact)
    if hll:
        assert np.array_equal(out.kmer_counts, exact.kmer_counts)
        for i in range(n):
            a = out.hll[out.hll_off[i]:out.hll_off[i + 1]]
            b = exact.hll[exact.hll_off[i]:exact.hll_off[i + 1]]
            assert np.array_equal(a, b), i


def test_digests_plain_vs_persistent(env):
    k2, db, batch, orc, h, genomes = env
    reads = edge_reads(genomes, 2000, seed=600)
    seq, off = k2.pack_fragments(reads)
    batch.classify(seq, off, min_hit_groups=2, mode=0)
    d0 = batch.digests(len(reads))
    batch.classify(seq, off, min_hit_groups=2, mode=1)
    d1 = batch.digests(len(reads))
    assert np.array_equal(d0, d1)
    # a digest depends on the fragment's results: reads with different calls differ
    out = batch.classify(seq, off, min_hit_groups=2, mode=1)
    calls = out.call
    i, j = int(np.argmax(calls != calls[0])), 0
    assert calls[i] != calls[j] and d1[i] != d1[j]


@pytest.mark.parametrize("unit_elements", [1, 200, 4096, 1 << 30])
def test_work_unit_sizes(env, unit_elements):
    """Persistent work units from one fragment each up to the whole batch."""
    k2, db, batch, orc, h, genomes = env
    reads = edge_reads(genomes, 400, seed=700) + chimeric_reads(genomes, 20, seed=701, length=5000)
    seq, off = k2.pack_fragments(reads)
    out = batch.classify(seq, off, min_hit_groups=2, mode=1, unit_elements=unit_elements)
    ref = orc.classify(h, seq, off, min_hit_groups=2)
    _compare(k2, db, out, ref, len(reads))


def test_classify_revcom_version0(synth_db_rc0):
    """A database marked revcom_version 0 (old reverse complement)."""
    import kraken2_amd as k2
    db = k2.Db.from_dir(synth_db_rc0)
    batch = k2.Batch(db)
    orc = Oracle()
    h = orc.open(synth_db_rc0)
    _, genomes = read_fasta(os.path.join(synth_db_rc0, "library.fna"))
    reads = edge_reads(genomes, 1500, seed=800)
    seq, off = k2.pack_fragments(reads)
    for mode in (0, 1):
        out = batch.classify(seq, off, mode=mode)
        ref = orc.classify(h, seq, off)
        _compare(k2, db, out, ref, len(reads))
    batch.close()
    db.close()


@pytest.mark.parametrize("name", ["one_empty_102", "full_128", "small_37", "sparse_1001"])
@pytest.mark.parametrize("mode", [0, 1])
def test_probe_edge_tables(edge_tables, name, mode):
    """Random tables with 8-bit compact keys (frequent key matches along long
    chains): probes that wrap past the last cell (lean probe over the padding
    copies), a table with no empty cell and one under 64 cells (bounded
    probe), through the library's load path; plus Get on random keys."""
    import kraken2_amd as k2
    d = edge_tables[name]
    db = k2.Db.from_dir(d)
    batch = k2.Batch(db)
    orc = Oracle()
    h = orc.open(d)
    try:
        _, genomes = read_fasta(os.path.join(d, "library.fna"))
        reads = edge_reads(genomes, 300, seed=91 + mode, min_len=40, max_len=300)
        seq, off = k2.pack_fragments(reads)
        out = batch.classify(seq, off, mode=mode)
        ref = orc.classify(h, seq, off)
        _compare(k2, db, out, ref, len(reads))
        assert (out.hit_groups > 0).any()
        keys = np.random.default_rng(7).integers(0, 2**62, size=5000, dtype=np.uint64)
        assert np.array_equal(db.get_batch(keys), orc.get_batch(h, keys))
    finally:
        batch.close()
        db.close()


def test_read_of_2_31_bases_rejected(env):
    """Reads of 2^31 or more bases are refused before any kernel reads the
    sequence (positions within a read are 32-bit): the offsets alone claim
    such a read here, over a 64-byte sequence buffer that is never read."""
    import ctypes as C
    from kraken2_amd.synth import Hip
    k2, db, batch, orc, h, genomes = env
    hip = Hip()
    d_seq = hip.upload(np.frombuffer(b"ACGT" * 16, np.uint8))
    d_off = hip.upload(np.array([0, 1 << 31, 1 << 31], np.uint64))
    try:
        lib = k2.load_library()
        p = k2._lib.ClassifyParams(0.0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0)
        rc = lib.k2g_classify_device(batch._h, C.byref(p), d_seq, d_off, None, 1, None)
        assert rc == 64 and b"2^31" in lib.k2g_last_error()
    finally:
        hip.free(d_seq)
        hip.free(d_off)


def test_mode_chosen_per_batch(env):
    """Mode 2 picks the persistent kernel for uneven lengths and the plain one
    for similar lengths; both batches equal the oracle."""
    k2, db, batch, orc, h, genomes = env
    for reads in (edge_reads(genomes, 500, seed=301, min_len=140, max_len=160),
                  chimeric_reads(genomes, 4, seed=302, length=6000) + edge_reads(genomes, 400, seed=303)):
        seq, off = k2.pack_fragments(reads)
        out = batch.classify(seq, off, mode=2, min_hit_groups=2)
        ref = orc.classify(h, seq, off, min_hit_groups=2)
        _compare(k2, db, out, ref, len(reads))


def _gapped_singles(reads, genomes, rng):
    """Single reads laid out as a FASTA file read in place: the bytes after
    each read (mate 2's range, which a single read does not read) hold a
    newline, the next header and a stretch of genome that would hit the
    table if it were scanned."""
    parts, off, pos = [b">r0 first\n"], [], 10
    for i, s in enumerate(reads):
        s = s.encode()
        off += [pos, pos + len(s)]
        parts.append(s)
        pos += len(s)
        g = genomes[rng.randrange(len(genomes))]
        p = rng.randrange(len(g) - 200)
        gap = b"\n>r%d comment\n" % (i + 1) + g[p:p + rng.randint(0, 200)].encode() + b"\n"
        parts.append(gap)
        pos += len(gap)
    off.append(pos)
    return b"".join(parts), np.array(off, dtype=np.uint64)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_single_reads_ignore_the_mate2_range(env, mode):
    """Bytes between single reads change nothing: calls, hitlists, k-mer
    counts and HLL records (per hit, and per block and taxon) equal those of
    the same reads packed back to back, short and long reads alike."""
    k2, db, batch, orc, h, genomes = env
    rng = random.Random(31 + mode)
    reads = edge_reads(genomes, 1500, seed=500 + mode, min_len=30, max_len=400)
    for _ in range(12):  # long reads take the segment path
        g = genomes[rng.randrange(len(genomes))]
        L = rng.randint(5000, min(len(g), 40000))
        p = rng.randrange(len(g) - L + 1)
        reads.insert(rng.randrange(len(reads)), g[p:p + L])
    seq0, off0 = k2.pack_fragments(reads)
    seq1, off1 = _gapped_singles(reads, genomes, rng)
    n = len(reads)
    for hll in (1, 2):
        a = batch.classify(seq0, off0, mode=mode, min_hit_groups=2, confidence=0.1, kmer_counts=True, hll=hll)
        b = batch.classify(seq1, off1, mode=mode, min_hit_groups=2, confidence=0.1, kmer_counts=True, hll=hll)
        for f in ("call", "hit_groups", "total_kmers", "run_off", "runs", "kmer_counts"):
            assert np.array_equal(getattr(a, f), getattr(b, f)), (f, hll)
        if hll == 1:
            assert np.array_equal(a.hll_off, b.hll_off) and np.array_equal(a.hll, b.hll)
        else:
            for key in a.sets:
                assert np.array_equal(a.sets[key], b.sets[key]), key
    ref = orc.classify(h, seq0, off0, confidence=0.1, min_hit_groups=2)
    _compare(k2, db, b, ref, n)
