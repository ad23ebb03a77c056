"""GPU parity of the graph-replayed pipeline (k2g_pipeline_*): every replay
of one recorded graph, at batch sizes from empty to its capacity, gives the
results k2g_classify gives on the same fragments (bit-exact checksums and
per-fragment digests), and those equal the oracle's on a sample; the
autotuner's configurations all agree."""
import os
import random

import numpy as np
import pytest

from helpers import Oracle, chimeric_reads, compare_with_oracle, edge_reads, read_fasta

pytestmark = pytest.mark.gpu

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _log(msg):
    print(f"[pipeline] {msg}", flush=True)


@pytest.fixture(scope="module")
def env(synth_db):
    import kraken2_amd as k2
    db = k2.Db.from_dir(synth_db)
    batch = k2.Batch(db)
    names, genomes = read_fasta(os.path.join(synth_db, "library.fna"))
    yield k2, db, batch, genomes, synth_db
    batch.close()
    db.close()


def _mixed_set(genomes, n, seed):
    """Single reads and pairs interleaved (per-fragment flags), edge cases
    and long chimeric reads mixed in."""
    rng = np.random.default_rng(seed)
    m1 = edge_reads(genomes, n, seed=seed, max_len=600)
    m2 = edge_reads(genomes, n, seed=seed + 1, max_len=300)
    long_ = chimeric_reads(genomes, max(1, n // 50), seed=seed + 2, length=5000)
    for j, r in enumerate(long_):
        m1[(j * 37) % n] = r
    flags = (rng.random(n) < 0.4).astype(np.uint8)
    m2 = [b if f else "" for b, f in zip(m2, flags)]
    return m1, m2, flags


def _pack(k2, m1, m2):
    seq, off = k2.pack_fragments(m1, m2)
    return seq, off


@pytest.mark.parametrize("extras", [False, True])
def test_host_pipeline_replays_match_classify(env, extras):
    """One graph, replayed on batches of several sizes (empty, one fragment,
    capacity): each replay's results equal k2g_classify's."""
    k2, db, batch, genomes, path = env
    kw = dict(confidence=0.1, min_hit_groups=2, hll=extras, kmer_counts=extras)
    m1, m2, flags = _mixed_set(genomes, 6000, seed=31)
    seq, off = _pack(k2, m1, m2)
    pl = k2.Pipeline(db, 6000, len(seq), host_io=True, **kw)
    try:
        for lo, hi in [(0, 6000), (10, 11), (0, 0), (100, 2600), (0, 6000), (5000, 6000)]:
            s_m1, s_m2, s_fl = m1[lo:hi], m2[lo:hi], flags[lo:hi]
            sseq, soff = _pack(k2, s_m1, s_m2)
            _log(f"host pipeline extras={extras} fragments [{lo}, {hi})")
            got = pl.run(sseq, soff, paired_flags=s_fl, checksum_only=True)
            want = batch.classify_checksum(sseq, soff, paired_flags=s_fl, mode=1, **kw)
            assert got == want, f"fragments [{lo}, {hi})"
        # one replay's full results, field by field
        out = pl.run(seq, off, paired_flags=flags)
        ref = batch.classify(seq, off, paired_flags=flags, mode=1, **kw)
        for f in ("call", "hit_groups", "total_kmers", "run_off", "runs"):
            assert np.array_equal(getattr(out, f), getattr(ref, f)), f
        if extras:
            assert np.array_equal(out.hll_off, ref.hll_off) and np.array_equal(out.hll, ref.hll)
            assert np.array_equal(out.kmer_counts, ref.kmer_counts)
        assert out.redone == ref.redone
    finally:
        pl.close()


def test_host_pipeline_matches_oracle(env):
    k2, db, batch, genomes, path = env
    orc = Oracle()
    h = orc.open(path)
    m1, m2, flags = _mixed_set(genomes, 3000, seed=41)
    seq, off = _pack(k2, m1, m2)
    pl = k2.Pipeline(db, 4000, len(seq) + 1000, host_io=True, confidence=0.3, min_hit_groups=2)
    try:
        out = pl.run(seq, off, paired_flags=flags)
        compare_with_oracle(k2, db, out, orc, h, seq, off, flags, confidence=0.3, min_hit_groups=2)
    finally:
        pl.close()


def test_dense_redo_inside_the_graph(env):
    """A 2-slot hit table sends most fragments down the dense redo, whose
    count the graph reads from device memory."""
    k2, db, batch, genomes, path = env
    rng = random.Random(9)
    m1 = chimeric_reads(genomes, 300, seed=51, length=3000) + \
        ["".join(rng.choice("ACGT") for _ in range(rng.randint(0, 200))) for _ in range(200)]
    seq, off = k2.pack_fragments(m1)
    kw = dict(table_slots=2, hll=True, kmer_counts=True)
    pl = k2.Pipeline(db, 600, len(seq), host_io=True, **kw)
    try:
        for n in (500, 120, 500):
            sseq, soff = k2.pack_fragments(m1[:n])
            got = pl.run(sseq, soff)
            ref = batch.classify(sseq, soff, mode=1, **kw)
            assert got.redone == ref.redone and got.redone > 0
            for f in ("call", "hit_groups", "total_kmers", "run_off", "runs", "hll_off", "hll", "kmer_counts"):
                assert np.array_equal(getattr(got, f), getattr(ref, f)), (n, f)
    finally:
        pl.close()


def test_host_pipeline_capacity_errors(env):
    k2, db, batch, genomes, path = env
    m1 = edge_reads(genomes, 100, seed=61, min_len=100, max_len=200)
    seq, off = k2.pack_fragments(m1)
    pl = k2.Pipeline(db, 50, len(seq), host_io=True)
    try:
        with pytest.raises(k2.K2GError, match="fragments"):
            pl.run(seq, off)
        s2, o2 = k2.pack_fragments(m1[:50])
        assert pl.run(s2, o2, checksum_only=True) == batch.classify_checksum(s2, o2, mode=1)
    finally:
        pl.close()
    pl = k2.Pipeline(db, 100, 1000, host_io=True)
    try:
        with pytest.raises(k2.K2GError, match="bases"):
            pl.run(seq, off)
    finally:
        pl.close()
    with pytest.raises(k2.K2GError):
        k2.Pipeline(db, 10, 100, host_io=True, tuning={"bucket_shift": 3})
    with pytest.raises(k2.K2GError):
        k2.Pipeline(db, 10, 100, host_io=True, tuning={"probe_depth": 9})


def _device_set(k2, genomes, n, seed):
    from kraken2_amd.synth import Hip
    m1, m2, flags = _mixed_set(genomes, n, seed)
    seq, off = _pack(k2, m1, m2)
    hip = Hip()
    d_seq = hip.upload(np.frombuffer(seq, np.uint8))
    d_off = hip.upload(off)
    d_fl = hip.upload(flags)
    return hip, seq, off, flags, d_seq, d_off, d_fl


def _addr(p, delta):
    import ctypes as C
    return C.c_void_p(p.value + delta)


def test_device_pipeline_replays_match_classify(env):
    """Device-resident ranges (offsets not starting at zero) through one
    graph: per-fragment digests equal k2g_classify's."""
    k2, db, batch, genomes, path = env
    n = 8000
    hip, seq, off, flags, d_seq, d_off, d_fl = _device_set(k2, genomes, n, seed=71)
    pl = k2.Pipeline(db, n, len(seq), host_io=False, min_hit_groups=2)
    try:
        for lo, hi in [(0, n), (3, 4), (0, 0), (1234, 5000), (0, n)]:
            _log(f"device pipeline fragments [{lo}, {hi})")
            ms = pl.run_device(d_seq, _addr(d_off, 16 * lo), _addr(d_fl, lo), hi - lo)
            assert ms >= 0
            got = pl.digests(hi - lo)
            soff = off[2 * lo:2 * hi + 1]
            batch.classify(seq, soff, paired_flags=flags[lo:hi], mode=1, min_hit_groups=2)
            want = batch.digests(hi - lo)
            assert np.array_equal(got, want), f"fragments [{lo}, {hi})"
    finally:
        pl.close()
        for p in (d_seq, d_off, d_fl):
            hip.free(p)


def test_device_pipeline_overflow_pool(env):
    """A 1% hitlist reservation: long chimeric fragments outgrow it, take
    exact space from the overflow pool and are classified again inside the
    same replay (some of them also through the dense redo)."""
    k2, db, batch, genomes, path = env
    n = 3000
    hip, seq, off, flags, d_seq, d_off, d_fl = _device_set(k2, genomes, n, seed=101)
    kw = dict(min_hit_groups=2, run_cap_pct=1)
    for slots in (0, 4):
        pl = k2.Pipeline(db, n, len(seq), host_io=False, table_slots=slots, **kw)
        try:
            for lo, hi in [(0, n), (700, 2100), (0, n)]:
                _log(f"overflow pool table_slots={slots} fragments [{lo}, {hi})")
                pl.run_device(d_seq, _addr(d_off, 16 * lo), _addr(d_fl, lo), hi - lo)
                space, dense = pl.redo_counts()
                got = pl.digests(hi - lo)
                batch.classify(seq, off[2 * lo:2 * hi + 1], paired_flags=flags[lo:hi], mode=1, table_slots=slots,
                               **kw)
                want = batch.digests(hi - lo)
                assert np.array_equal(got, want), f"fragments [{lo}, {hi})"
                assert (space, dense) == batch.redo_counts()
                assert space > 0 and (dense > 0) == (slots == 4)
        finally:
            pl.close()
    for p in (d_seq, d_off, d_fl):
        hip.free(p)


def test_device_pipeline_over_capacity_is_refused(env):
    """A device batch whose hitlist space exceeds the pipeline's bases
    capacity is caught on the device (nothing classified) and reported."""
    k2, db, batch, genomes, path = env
    n = 400
    hip, seq, off, flags, d_seq, d_off, d_fl = _device_set(k2, genomes, n, seed=81)
    pl = k2.Pipeline(db, n, 64, host_io=False)
    try:
        with pytest.raises(k2.K2GError, match="capacity"):
            pl.run_device(d_seq, d_off, d_fl, n)
        # the same graph still serves a batch that fits
        small = int(np.argmin(off[2:2 * n + 1:2] - off[0:2 * n:2]))
        assert int(off[2 * small + 2] - off[2 * small]) + 1 <= 64 + n
        pl.run_device(d_seq, _addr(d_off, 16 * small), _addr(d_fl, small), 1)
        batch.classify(seq, off[2 * small:2 * small + 3], paired_flags=flags[small:small + 1], mode=1)
        assert np.array_equal(pl.digests(1), batch.digests(1))
    finally:
        pl.close()
        for p in (d_seq, d_off, d_fl):
            hip.free(p)


def test_autotuner_configurations_agree(env):
    k2, db, batch, genomes, path = env
    n = 6000
    hip, seq, off, flags, d_seq, d_off, d_fl = _device_set(k2, genomes, n, seed=91)
    try:
        _log("tuning")
        best, rep = k2.tune_pipeline(db, d_seq, d_off, d_fl, n, stride=2, min_hit_groups=2)
        _log(f"tuned {best} {rep}")
        assert rep["mismatched"] == 0
        assert rep["evaluated"] >= 10
        assert rep["sample_fragments"] == n // 2
        assert 0 < rep["best_ms"] <= rep["default_ms"]
        assert best["unit_elements"] in (1024, 2048, 4096, 8192, 16384, 32768)
        assert 0 <= best["bucket_shift"] <= 2 and 1 <= best["probe_depth"] <= 8
        pl = k2.Pipeline(db, n, len(seq), host_io=False, tuning=best, min_hit_groups=2)
        try:
            tuning, grid = pl.params()
            assert tuning == best and grid >= 8
            pl.run_device(d_seq, d_off, d_fl, n)
            batch.classify(seq, off, paired_flags=flags, mode=1, min_hit_groups=2)
            assert np.array_equal(pl.digests(n), batch.digests(n))
        finally:
            pl.close()
    finally:
        for p in (d_seq, d_off, d_fl):
            hip.free(p)


def test_translated_pipeline_matches_classify():
    """The translated variant (six frames per mate) on the reference-built
    protein database."""
    import kraken2_amd as k2
    prot = os.path.join(G, "prot")
    db = k2.Db.from_dir(prot)
    batch = k2.Batch(db)
    _, r1 = read_fasta(os.path.join(prot, "reads_1.fa"))
    _, r2 = read_fasta(os.path.join(prot, "reads_2.fa"))
    try:
        for paired in (False, True):
            seq, off = k2.pack_fragments(r1, r2 if paired else None)
            kw = dict(paired=paired, min_hit_groups=2, hll=True, kmer_counts=True)
            pl = k2.Pipeline(db, len(r1), len(seq), host_io=True, **kw)
            try:
                for n in (len(r1), len(r1) // 3, len(r1)):
                    sseq, soff = k2.pack_fragments(r1[:n], r2[:n] if paired else None)
                    got = pl.run(sseq, soff, checksum_only=True)
                    assert got == batch.classify_checksum(sseq, soff, mode=1, **kw), (paired, n)
            finally:
                pl.close()
    finally:
        batch.close()
        db.close()


def test_host_pipeline_runs_from_its_staging(env):
    """A batch written straight into the pipeline's pinned staging runs
    without a copy and gives k2g_classify's results."""
    k2, db, batch, genomes, path = env
    m1, m2, flags = _mixed_set(genomes, 2000, seed=111)
    seq, off = _pack(k2, m1, m2)
    pl = k2.Pipeline(db, 2000, len(seq), host_io=True, min_hit_groups=2, hll=True, kmer_counts=True)
    try:
        st_seq, st_off, st_pf = pl.staging(2000, len(seq))
        for lo, hi in [(0, 2000), (500, 900)]:
            sseq, soff = _pack(k2, m1[lo:hi], m2[lo:hi])
            n = hi - lo
            st_seq[:len(sseq)] = np.frombuffer(sseq, np.uint8)
            st_off[:2 * n + 1] = soff
            st_pf[:n] = flags[lo:hi]
            got = pl.run(st_seq[:max(1, len(sseq))], st_off[:2 * n + 1], paired_flags=st_pf[:n], checksum_only=True)
            want = batch.classify_checksum(sseq, soff, paired_flags=flags[lo:hi], mode=1, min_hit_groups=2,
                                           hll=True, kmer_counts=True)
            assert got == want, (lo, hi)
    finally:
        pl.close()
    dev = k2.Pipeline(db, 10, 1000, host_io=False)
    try:
        with pytest.raises(k2.K2GError, match="staging"):
            dev.staging(10, 1000)
    finally:
        dev.close()


@pytest.mark.parametrize("host_io,extras", [(False, False), (True, True)])
def test_pipeline_splits_long_reads(env, host_io, extras):
    """Batches whose longest reads outweigh a wave's share are split over
    waves inside the graph (segments, then the stitch): results equal
    k2g_classify's, replay after replay, and a batch of short reads alone is
    not split."""
    k2, db, batch, genomes, path = env
    kw = dict(confidence=0.1, min_hit_groups=2, hll=extras, kmer_counts=extras)
    rng = random.Random(91)
    m1 = edge_reads(genomes, 1500, seed=92, min_len=50, max_len=300)
    for _ in range(12):
        g = genomes[rng.randrange(len(genomes))]
        L = rng.randint(12000, min(len(g), 40000))
        p = rng.randrange(len(g) - L + 1)
        m1.insert(rng.randrange(len(m1)), g[p:p + L])
    seq, off = _pack(k2, m1, None)
    n = len(m1)
    want = batch.classify(seq, off, mode=1, **kw)
    pl = k2.Pipeline(db, n, len(seq), host_io=host_io, **kw)
    try:
        if host_io:
            for rep in range(2):
                _log(f"host pipeline with long reads, replay {rep}")
                got = pl.run(seq, off)
                for f in ("call", "hit_groups", "total_kmers", "run_off", "runs"):
                    assert np.array_equal(getattr(got, f), getattr(want, f)), f
                if extras:
                    assert np.array_equal(got.hll_off, want.hll_off) and np.array_equal(got.hll, want.hll)
                    assert np.array_equal(got.kmer_counts, want.kmer_counts)
                frags, segs = pl.split_counts()
                assert frags >= 12 and segs > frags, (frags, segs)
            short = [r for r in m1 if len(r) < 1000]
            s2, o2 = _pack(k2, short, None)
            _log("host pipeline, short reads only")
            got = pl.run(s2, o2)
            ref = batch.classify(s2, o2, mode=1, **kw)
            assert np.array_equal(got.call, ref.call) and np.array_equal(got.runs, ref.runs)
            assert pl.split_counts() == (0, 0)
        else:
            from kraken2_amd.synth import Hip
            hip = Hip()
            d_seq, d_off = hip.upload(np.frombuffer(seq, np.uint8)), hip.upload(off)
            try:
                for rep in range(2):
                    _log(f"device pipeline with long reads, replay {rep}")
                    pl.run_device(d_seq, d_off, None, n)
                    batch.classify(seq, off, mode=1, **kw)
                    assert np.array_equal(pl.digests(n), batch.digests(n))
                    frags, segs = pl.split_counts()
                    assert frags >= 12 and segs > frags, (frags, segs)
            finally:
                hip.free(d_seq)
                hip.free(d_off)
    finally:
        pl.close()
