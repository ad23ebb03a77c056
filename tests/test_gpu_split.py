"""GPU parity of long fragments split over waves (persistent mode): batches
whose longest fragments outweigh a wave's share classify those fragments as
segments of about 2048 k-mers and stitch them.  Every array of the results
equals the plain kernel's (which never splits) and the oracle's, with HLL
records and k-mer counts, with reserved and exact hitlist space (the
out-of-space and dense-table redos of split fragments included), and in the
request server."""
import os
import random

import numpy as np
import pytest

from helpers import Oracle, chimeric_reads, compare_with_oracle, edge_reads, read_fasta

pytestmark = pytest.mark.gpu

FIELDS = ("call", "hit_groups", "total_kmers", "run_off", "runs", "hll_off", "hll", "kmer_counts")


def _log(msg):
    print(f"[split] {msg}", flush=True)


@pytest.fixture(scope="module")
def env(synth_db):
    import kraken2_amd as k2
    db = k2.Db.from_dir(synth_db)
    batch = k2.Batch(db)
    names, genomes = read_fasta(os.path.join(synth_db, "library.fna"))
    yield k2, db, batch, genomes, synth_db
    batch.close()
    db.close()


def long_batch(k2, genomes, seed, n_short=400):
    """Long chimeric reads and pairs (5-40 kb, ambiguous stretches and
    lowercase near the cut points, a mate shorter than k, an all-N mate)
    among short edge-case reads; pair flags per fragment."""
    rng = random.Random(seed)
    long1 = [chimeric_reads(genomes, 1, seed=seed + i, seg=rng.choice([300, 800, 3000]),
                            length=rng.choice([5000, 9000, 17000, 40000]))[0] for i in range(12)]
    long2 = [chimeric_reads(genomes, 1, seed=seed + 100 + i, seg=500, length=rng.choice([34, 3000, 8000]))[0]
             for i in range(12)]
    for r in range(len(long1)):
        s = list(long1[r])
        for c in (2048 + 34 - 1, 2048 + 34 + 3, 4130, rng.randrange(len(s))):
            for j in range(c, min(len(s), c + rng.choice([1, 3, 34, 35, 200]))):
                s[j] = rng.choice("NRYnx")
        long1[r] = "".join(s).lower() if r % 4 == 0 else "".join(s)
    long2[3] = long2[3][:20]
    long2[7] = "N" * 6000
    m1 = long1 + edge_reads(genomes, n_short, seed=seed + 7)
    m2 = long2 + edge_reads(genomes, n_short, seed=seed + 8, min_len=50, max_len=250)
    flags = np.array([(i % 3 != 0) for i in range(len(m1))], np.uint8)
    order = list(range(len(m1)))
    rng.shuffle(order)
    m1 = [m1[i] for i in order]
    m2 = [m2[i] if flags[i] else "" for i in order]
    flags = flags[order]
    seq, off = k2.pack_fragments(m1, m2)
    return seq, off, flags


def _same(a, b, what):
    for f in FIELDS:
        x, y = getattr(a, f), getattr(b, f)
        assert (x is None) == (y is None) and (x is None or np.array_equal(x, y)), (what, f)


@pytest.mark.parametrize("extras,pct,slots", [(False, 0, 0), (True, 0, 0), (True, 100, 0), (True, 1, 0),
                                              (True, 0, 4), (False, 1, 4)])
def test_split_equals_plain(env, extras, pct, slots):
    k2, db, batch, genomes, path = env
    seq, off, flags = long_batch(k2, genomes, seed=11 + pct + slots)
    kw = dict(paired_flags=flags, confidence=0.2, min_hit_groups=2, hll=extras, kmer_counts=extras,
              run_cap_pct=pct, table_slots=slots)
    out = batch.classify(seq, off, mode=1, **kw)
    fr, segs = batch.split_counts()
    redo = batch.redo_counts()
    _log(f"extras={extras} pct={pct} slots={slots}: {fr} fragments split into {segs} segments, redone {redo}")
    assert fr >= 10 and segs > 3 * fr
    if pct == 1:
        assert redo[0] > 0
    if slots:
        assert redo[1] > 0
    plain = batch.classify(seq, off, mode=0, **kw)
    assert batch.split_counts() == (0, 0)
    _same(out, plain, "split vs plain")


def test_split_matches_oracle(env):
    k2, db, batch, genomes, path = env
    orc = Oracle()
    h = orc.open(path)
    seq, off, flags = long_batch(k2, genomes, seed=41)
    out = batch.classify(seq, off, mode=1, paired_flags=flags, confidence=0.3, min_hit_groups=2)
    assert batch.split_counts()[0] > 0
    compare_with_oracle(k2, db, out, orc, h, seq, off, flags, confidence=0.3, min_hit_groups=2)


def test_split_requests_in_server(env):
    """Served requests with long reads (split in their own launches) equal
    the plain kernel on the same fragments."""
    k2, db, batch, genomes, path = env
    reqs = [long_batch(k2, genomes, seed=60 + i, n_short=40 * i) for i in range(6)]
    kw = dict(confidence=0.1, min_hit_groups=2, kmer_counts=True, hll=True)
    srv = k2.Server(db, 2 << 30, slots=3)
    try:
        tickets = [srv.submit(s, o, paired_flags=f, **kw) for s, o, f in reqs]
        outs = [srv.wait(t)[0] for t in tickets]
    finally:
        srv.close()
    for i, ((s, o, f), out) in enumerate(zip(reqs, outs)):
        _same(out, batch.classify(s, o, paired_flags=f, mode=0, **kw), f"request {i}")
