"""GPU parity at the benchmark configurations' own shapes.

* Probe statistics (lookups and distinct 128-byte table lines per fragment,
  the numerator of the gather roofline) equal the oracle's count from
  walking the same table.
* Config 1 whole: 64 genomes x 200 kb built into a table by the GPU build
  front-ends, 100k x 150 bp reads (1% substitutions, 10% random), -T 0 -g 2:
  every fragment's call, hit groups, total k-mers and hitlist against the
  oracle reading the same hash.k2d.
* Configs 2 and 3 at full size: the 4000-genome table (~2.7e9 cells, built on
  the GPU), 50M pairs 2x150 with 20% off-target reads and 10M uneven-length
  reads (150 bp singles, 2x250 pairs, long reads up to 50 kb with indels):
  the plain and persistent kernels agree on every fragment (per-fragment
  result digests) and a stratified sample of thousands of fragments, long
  reads included, equals the oracle, probe statistics included.
Reference semantics: src/classify.cc:773-991 (ClassifySequence, ResolveTree,
hitlist text), src/compact_hash.h:354-392 (GetBatch).
"""
import os
import shutil

import numpy as np
import pytest

from helpers import (K2SYNTH, Oracle, compare_with_oracle, edge_reads, large_scratch_dir, make_synthetic_db,
                     read_fasta, run)

pytestmark = pytest.mark.gpu

RUN_SPACE_LIMIT = int(1.2e10)  # hitlist runs reserved per launch (8 bytes each)


def _log(msg):
    print(f"[configs] {msg}", flush=True)


# ---------------------------------------------------------------- probe statistics

@pytest.mark.parametrize("mode", [0, 1])
def test_probe_statistics_match_oracle(synth_db, mode):
    """Per-fragment lookups and distinct probe lines (want_probe_stats)
    equal the oracle's walk of the same table (HashTable::probe_lines)."""
    import kraken2_amd as k2
    db = k2.Db.from_dir(synth_db)
    batch = k2.Batch(db)
    orc = Oracle()
    h = orc.open(synth_db)
    try:
        _, genomes = read_fasta(os.path.join(synth_db, "library.fna"))
        reads = edge_reads(genomes, 1500, seed=910 + mode, max_len=3000)
        seq, off = k2.pack_fragments(reads)
        out = batch.classify(seq, off, mode=mode, min_hit_groups=2, probe_stats=True)
        assert int(out.probe_lines.sum()) > int(out.lookups.sum()) > 0
        compare_with_oracle(k2, db, out, orc, h, seq, off, np.zeros(len(reads), np.uint8), min_hit_groups=2,
                            lines=True)
        # pairs: both mates' lookups count
        m2 = edge_reads(genomes, 1500, seed=920 + mode, max_len=400)
        seq, off = k2.pack_fragments(reads, m2)
        out = batch.classify(seq, off, paired=True, mode=mode, probe_stats=True)
        compare_with_oracle(k2, db, out, orc, h, seq, off, np.ones(len(reads), np.uint8), lines=True)
    finally:
        orc.lib.k2cpu_db_close(h)
        batch.close()
        db.close()


@pytest.mark.parametrize("name", ["one_empty_102", "full_128", "small_37", "sparse_1001"])
def test_probe_statistics_edge_tables(edge_tables, name):
    """Probe runs that wrap past the last cell, tables with no empty cell
    (a miss visits every cell) and tables under one 128-byte line."""
    import kraken2_amd as k2
    d = edge_tables[name]
    db = k2.Db.from_dir(d)
    batch = k2.Batch(db)
    orc = Oracle()
    h = orc.open(d)
    try:
        _, genomes = read_fasta(os.path.join(d, "library.fna"))
        reads = edge_reads(genomes, 300, seed=930, min_len=40, max_len=300)
        seq, off = k2.pack_fragments(reads)
        for mode in (0, 1):
            out = batch.classify(seq, off, mode=mode, probe_stats=True)
            compare_with_oracle(k2, db, out, orc, h, seq, off, np.zeros(len(reads), np.uint8), lines=True)
    finally:
        orc.lib.k2cpu_db_close(h)
        batch.close()
        db.close()


# ---------------------------------------------------------------- config 1

@pytest.fixture(scope="module")
def c1(tmp_path_factory):
    """Config 1: seed 1, root -> 4 phyla -> 16 genera -> 64 species, 200 kb
    genomes (species = genus root + 5% substitutions), k=35 l=31, load 0.7,
    built by the GPU build front-ends; 100k x 150 bp reads."""
    import kraken2_amd as k2
    d = str(tmp_path_factory.mktemp("c1"))
    _log("config 1: building the database with k2g_estimate_capacity / k2g_build_db")
    make_synthetic_db(d, seed=1, levels="4,4,4", genome_len=200000,
                      tools=(os.path.join(k2.BIN_DIR, "k2g_estimate_capacity"),
                             os.path.join(k2.BIN_DIR, "k2g_build_db")))
    reads = os.path.join(d, "reads.fa")
    run([K2SYNTH, "reads", "--lib", os.path.join(d, "library.fna"), "--out", reads, "--seed", "1", "--n",
         "100000", "--len", "150", "--sub", "0.01", "--random-frac", "0.1"])
    _, seqs = read_fasta(reads)
    return d, seqs


def test_c1_every_read_matches_oracle(c1):
    import kraken2_amd as k2
    d, seqs = c1
    assert len(seqs) == 100000
    db = k2.Db.from_dir(d)
    assert db.info.capacity > 5_000_000
    batch = k2.Batch(db)
    orc = Oracle()
    h = orc.open(d)
    try:
        seq, off = k2.pack_fragments(seqs)
        digests = []
        for mode in (0, 1, 2):
            _log(f"config 1: 100000 reads on the GPU, mode {mode}")
            out = batch.classify(seq, off, confidence=0.0, min_hit_groups=2, mode=mode, probe_stats=True)
            digests.append(batch.digests(len(seqs)))
            if mode == 1:
                _log("config 1: the oracle on every read")
                n = compare_with_oracle(k2, db, out, orc, h, seq, off, np.zeros(len(seqs), np.uint8),
                                        min_hit_groups=2, lines=True)
                assert n == len(seqs)
                classified = int((out.call != 0).sum())
                assert 0.5 * len(seqs) < classified < len(seqs)  # 10% are random reads
        assert np.array_equal(digests[0], digests[1]) and np.array_equal(digests[1], digests[2])
    finally:
        orc.lib.k2cpu_db_close(h)
        batch.close()
        db.close()


# ---------------------------------------------------------------- configs 2 and 3

@pytest.fixture(scope="module")
def c2_table():
    """The config-2 table: 4000 genomes (lognormal, median 2.5 Mb, siblings
    sharing ~92%), estimated and built on the GPU, written as hash.k2d."""
    from kraken2_amd.synth import Workload
    tmp = large_scratch_dir("k2g_test_c2_")
    try:
        _log(f"config 2 table: 4000 genomes into {tmp}")
        wl = Workload(tmp, seed=2, n_genomes=4000, log=_log)
        yield wl, tmp
    finally:
        wl_ = locals().get("wl")
        if wl_ is not None and wl_.d_genomes is not None:
            wl_.free_genomes()
        shutil.rmtree(tmp, ignore_errors=True)


def _chunks(off, paired, k, limit=RUN_SPACE_LIMIT, pct=25):
    """Contiguous fragment ranges whose reserved hitlist space stays under limit."""
    n = len(paired)
    l1 = (off[1:2 * n:2] - off[0:2 * n:2]).astype(np.int64)
    l2 = (off[2::2] - off[1:2 * n:2]).astype(np.int64)
    e = np.maximum(l1 - k + 1, 0) + np.where(paired == 1, np.maximum(l2 - k + 1, 0) + 1, 0)
    space = np.minimum(e, e * pct // 100 + 64)
    csum = np.cumsum(space)
    out, s, base = [], 0, 0
    while s < n:
        e_ = max(int(np.searchsorted(csum, base + limit, side="right")), s + 1)
        out.append((s, e_))
        base = int(csum[e_ - 1])
        s = e_
    return out


def _device_digests(k2, batch, rs, mode, confidence, min_hit_groups):
    """Classifies the device-resident read set in run-space chunks; returns
    per-fragment digests."""
    import ctypes as C
    lib = k2.load_library()
    k = int(batch.db.info.k)
    dig = np.zeros(rs.n, np.uint64)
    for s, e in _chunks(rs.off, rs.paired, k):
        p = k2._lib.ClassifyParams(confidence, min_hit_groups, 0, 0, mode, 0, 0, 0, 0, 0, 0)
        rc = lib.k2g_classify_device(batch._h, C.byref(p), rs.d_seq, C.c_void_p(rs.d_off.value + 16 * s),
                                     C.c_void_p(rs.d_paired.value + s), e - s, None)
        if rc:
            raise k2.K2GError(rc, lib.k2g_last_error().decode())
        dig[s:e] = batch.digests(e - s)
    return dig


def _stratified_sample(kind, per_kind, seed):
    rng = np.random.default_rng(seed)
    idx = []
    for kd in np.unique(kind):
        where = np.flatnonzero(kind == kd)
        idx.append(rng.choice(where, size=min(per_kind, len(where)), replace=False))
    return np.sort(np.concatenate(idx))


def _full_size_check(k2, wl, tmp, n, *, f_single, f_paired, paired_len, off_frac, seed, confidence, per_kind):
    db = k2.Db.from_dir(tmp)
    batch = k2.Batch(db)
    rs = None
    orc = Oracle()
    h = None
    try:
        _log(f"generating {n} reads (seed {seed}) on the GPU")
        rs = wl.make_reads(n, f_single=f_single, f_paired=f_paired, paired_len=paired_len, off_frac=off_frac,
                           seed=seed, group_pairs=True)
        _log(f"{rs.n} fragments, {rs.bases / 1e9:.2f} Gbp; plain kernel")
        d_plain = _device_digests(k2, batch, rs, 0, confidence, 2)
        _log("persistent kernel")
        d_pers = _device_digests(k2, batch, rs, 1, confidence, 2)
        mism = int((d_plain != d_pers).sum())
        assert mism == 0, f"{mism} fragments differ between the plain and persistent kernels"
        sample = _stratified_sample(rs.kind, per_kind, seed + 100)
        seq, off, pflags = rs.host_fragments(sample)
        out = batch.classify(seq, off, paired_flags=pflags, confidence=confidence, min_hit_groups=2, mode=1,
                             probe_stats=True)
        # the sample classified on its own gives the full run's results
        assert np.array_equal(batch.digests(len(sample)), d_pers[sample])
        _log(f"oracle on {len(sample)} sampled fragments (opening the table)")
        h = orc.open(tmp)
        checked = compare_with_oracle(k2, db, out, orc, h, seq, off, pflags, confidence=confidence,
                                      min_hit_groups=2, lines=True)
        assert checked == len(sample)
        _log(f"{checked} fragments equal the oracle, probe statistics included")
        return rs.kind[sample]
    finally:
        if h:
            orc.lib.k2cpu_db_close(h)
        if rs is not None:
            rs.free()
        batch.close()
        db.close()


def test_c3_full_size(c2_table):
    """Config 3: 10M reads, 40% single 150 bp, 30% pairs 2x250, 30% long
    (lognormal median 8 kb, cap 50 kb, 2.5% substitutions + 2.5% indels),
    -T 0.1 -g 2, against the config-2 table."""
    import kraken2_amd as k2
    wl, tmp = c2_table
    kinds = _full_size_check(k2, wl, tmp, 10_000_000, f_single=0.4, f_paired=0.3, paired_len=250, off_frac=0.0,
                             seed=3, confidence=0.1, per_kind=1500)
    assert set(np.unique(kinds).tolist()) == {0, 1, 2}  # singles, pairs and long reads all sampled


def test_c2_full_size(c2_table):
    """Config 2: 50M pairs 2x150, insert N(350, 50), 1% substitutions, 20%
    off-target reads with GC 0.41, -T 0.1 -g 2 -P."""
    import kraken2_amd as k2
    wl, tmp = c2_table
    _full_size_check(k2, wl, tmp, 50_000_000, f_single=0.0, f_paired=1.0, paired_len=150, off_frac=0.2, seed=2,
                     confidence=0.1, per_kind=4000)
