"""GPU parity of the per-(input block, taxon) HLL sets reduced on the device
(want_hll = 2): the sets equal the ones the per-hit records give when cut in
the reference's input blocks (src/classify.cc:523-543), the register-switch
flag follows src/hyperloglogplus.cc:471-509 at and around its 1024-value
edge, and kreports folded from the sets equal kreports folded from the
records and the oracle's classify driver on the same reads."""
import os

import numpy as np
import pytest

from helpers import Oracle, chimeric_reads, edge_reads, read_fasta

pytestmark = pytest.mark.gpu


def _log(msg):
    print(f"[hll sets] {msg}", flush=True)


@pytest.fixture(scope="module")
def env(synth_db):
    import kraken2_amd as k2
    db = k2.Db.from_dir(synth_db)
    batch = k2.Batch(db)
    names, genomes = read_fasta(os.path.join(synth_db, "library.fna"))
    yield k2, db, batch, genomes, synth_db
    batch.close()
    db.close()


def input_blocks(off, flags):
    """Block starts: runs of pairs 10 000 at a time, runs of single reads
    until 3 MiB of mate-1 bases (src/classify.cc:523-543)."""
    n = len(flags)
    starts, i = [], 0
    while i < n:
        starts.append(i)
        s = i
        if flags[i]:
            while i < n and flags[i] and i - s < 10000:
                i += 1
        else:
            bases = 0
            while i < n and not flags[i] and bases < 3 * 1024 * 1024:
                bases += int(off[2 * i + 1] - off[2 * i])
                i += 1
    return starts + [n]


def sets_from_records(out, off, flags):
    """(block, taxon, dense, n, values) per group, from per-hit records."""
    starts = input_blocks(off, flags)
    rows = []
    for b in range(len(starts) - 1):
        recs = out.hll[out.hll_off[starts[b]]:out.hll_off[starts[b + 1]]]
        for t in np.unique(recs[:, 0]):
            v = recs[recs[:, 0] == t, 1]
            u = np.unique(v)
            before_last = len(u) - (1 if int((v == v[-1]).sum()) == 1 else 0)
            rows.append((b, int(t), int(before_last >= 1024), len(v), u))
    return rows


def assert_sets_equal(sets, rows):
    assert len(sets["block"]) == len(rows)
    for g, (b, t, dense, n, u) in enumerate(rows):
        got = (int(sets["block"][g]), int(sets["taxon"][g]), int(sets["dense"][g]), int(sets["n"][g]))
        assert got == (b, t, dense, n), (g, got, (b, t, dense, n))
        assert np.array_equal(sets["vals"][sets["off"][g]:sets["off"][g + 1]], u), g


def test_sets_over_many_blocks_match_records(env, tmp_path):
    """Runs of long single reads (3 MiB blocks) and of pairs (10 000-pair
    blocks) in one call: sets, flags and kreports as from the records."""
    k2, db, batch, genomes, path = env
    parts = [(chimeric_reads(genomes, 3500, seed=1, seg=600, length=2500), None),
             (edge_reads(genomes, 12000, seed=2, min_len=80, max_len=150),
              edge_reads(genomes, 12000, seed=3, min_len=80, max_len=150)),
             (edge_reads(genomes, 700, seed=4), None),
             (edge_reads(genomes, 90, seed=5), edge_reads(genomes, 90, seed=6))]
    m1, m2, flags = [], [], []
    for a, b in parts:
        m1 += a
        m2 += b if b is not None else [""] * len(a)
        flags += [1 if b is not None else 0] * len(a)
    flags = np.array(flags, np.uint8)
    seq, off = k2.pack_fragments(m1, m2)
    kw = dict(mode=1, min_hit_groups=2, confidence=0.1, kmer_counts=True, paired_flags=flags)
    rec = batch.classify(seq, off, hll=True, **kw)
    out = batch.classify(seq, off, hll=2, **kw)
    assert out.hll_off is None and out.sets is not None
    for f in ("call", "hit_groups", "total_kmers", "run_off", "runs", "kmer_counts"):
        assert np.array_equal(getattr(out, f), getattr(rec, f)), f
    rows = sets_from_records(rec, off, flags)
    assert len(input_blocks(off, flags)) - 1 >= 6
    assert any(r[2] for r in rows) and not all(r[2] for r in rows)
    _log(f"{len(rows)} groups over {len(input_blocks(off, flags)) - 1} blocks, "
         f"{sum(r[2] for r in rows)} switched to registers")
    assert_sets_equal(out.sets, rows)
    assert int(out.sets["n"].sum()) == int(rec.hll_off[-1]) == int(rec.kmer_counts.sum())
    for name, o in (("sets", out), ("records", rec)):
        k2.write_report(db, off, o, str(tmp_path / f"{name}.kreport"), paired_flags=flags, kmer_data=True)
    assert (tmp_path / "sets.kreport").read_text() == (tmp_path / "records.kreport").read_text()
    # the checksum covers the sets
    a = batch.classify_checksum(seq, off, hll=2, **kw)
    assert a == batch.classify_checksum(seq, off, hll=2, **kw)
    assert a != batch.classify_checksum(seq, off, hll=True, **kw)


def _edge_reads_1024(k2, batch, genome, k):
    """One-k-mer reads of one genome and the first prefix whose most frequent
    taxon reaches exactly 1024 distinct hashes, by a new hash."""
    rng = np.random.default_rng(17)
    pos = rng.permutation(len(genome) - k)[:8000]
    reads = [genome[p:p + k] for p in pos]
    seq, off = k2.pack_fragments(reads)
    rec = batch.classify(seq, off, mode=1, hll=True)
    taxa, counts = np.unique(rec.hll[:, 0], return_counts=True)
    t = int(taxa[np.argmax(counts)])
    seen, last = set(), None
    for f in range(len(reads)):
        r = rec.hll[rec.hll_off[f]:rec.hll_off[f + 1]]
        for tax, e in r.tolist():
            if tax == t:
                seen.add(e)
                if len(seen) == 1024 and last is None:
                    last = f
        if last is not None:
            break
    assert last is not None
    # an earlier read whose only record is a hash of taxon t
    rep = next(f for f in range(last) if rec.hll_off[f + 1] - rec.hll_off[f] == 1
               and rec.hll[rec.hll_off[f], 0] == t)
    return reads[:last + 1], reads[rep], t


@pytest.mark.parametrize("repeat_last", [False, True])
def test_register_switch_at_1024(env, tmp_path, repeat_last):
    """1024 distinct hashes of one taxon in one block: the sketch switched
    iff the block's last record of that taxon repeats an earlier hash (then
    1024 were held when it arrived).  The kreport's estimate, which differs
    between the two states, equals the oracle's."""
    k2, db, batch, genomes, path = env
    k = int(db.info.k)
    reads, rep, t = _edge_reads_1024(k2, batch, genomes[0], k)
    if repeat_last:
        reads = reads + [rep]
    seq, off = k2.pack_fragments(reads)
    flags = np.zeros(len(reads), np.uint8)
    rec = batch.classify(seq, off, mode=1, hll=True, kmer_counts=True)
    out = batch.classify(seq, off, mode=1, hll=2, kmer_counts=True)
    rows = sets_from_records(rec, off, flags)
    g = [i for i, r in enumerate(rows) if r[1] == t]
    assert len(g) == 1 and len(rows[g[0]][4]) == 1024 and rows[g[0]][2] == int(repeat_last)
    assert_sets_equal(out.sets, rows)
    # kreports: from the sets, and from the oracle's driver on the reads as a file
    orc = Oracle()
    h = orc.open(path)
    try:
        fa = tmp_path / "r.fa"
        fa.write_text("".join(f">r{i}\n{r}\n" for i, r in enumerate(reads)))
        orc.run_files(h, str(fa), kraken_out=str(tmp_path / "orc.kraken"), report=str(tmp_path / "orc.kreport"),
                      kmer_data=True, threads=4)
    finally:
        orc.lib.k2cpu_db_close(h)
    k2.write_report(db, off, out, str(tmp_path / "gpu.kreport"), kmer_data=True)
    assert (tmp_path / "gpu.kreport").read_text() == (tmp_path / "orc.kreport").read_text()


def test_server_sets_equal_alone_and_refusals(env):
    """Served requests with want_hll 2 equal the same fragments alone; the
    paths without host offsets refuse it."""
    k2, db, batch, genomes, path = env
    rng = np.random.default_rng(23)
    reqs = []
    for i in range(12):
        n = int(rng.integers(1, 3000))
        paired = bool(i % 2)
        m1 = edge_reads(genomes, n, seed=100 + i, min_len=50, max_len=400)
        m2 = edge_reads(genomes, n, seed=200 + i, min_len=50, max_len=200) if paired else None
        seq, off = k2.pack_fragments(m1, m2)
        reqs.append((seq, off, paired))
    kw = dict(confidence=0.1, min_hit_groups=2, kmer_counts=True, hll=2)
    srv = k2.Server(db, 1 << 30, slots=4)
    try:
        tickets = [srv.submit(s, o, paired=p, **kw) for s, o, p in reqs]
        outs = [srv.wait(tk)[0] for tk in tickets]
    finally:
        srv.close()
    for i, ((s, o, p), out) in enumerate(zip(reqs, outs)):
        alone = batch.classify(s, o, paired=p, mode=2, **kw)
        for f in ("call", "hit_groups", "total_kmers", "run_off", "runs", "kmer_counts"):
            assert np.array_equal(getattr(out, f), getattr(alone, f)), (i, f)
        for f in ("block", "taxon", "dense", "n", "off", "vals"):
            assert np.array_equal(out.sets[f], alone.sets[f]), (i, f)
    import ctypes as C
    lib = k2.load_library()
    p = k2._lib.ClassifyParams(0.0, 0, 0, 0, 1, 0, 2, 0, 0, 0, 0)
    s, o, _ = reqs[0]
    from kraken2_amd.synth import Hip
    hip = Hip()
    d_seq = hip.upload(np.frombuffer(s, np.uint8))
    d_off = hip.upload(o)
    try:
        rc = lib.k2g_classify_device(batch._h, C.byref(p), d_seq, d_off, None, (len(o) - 1) // 2, None)
        assert rc != 0 and "host fragments" in lib.k2g_last_error().decode()
    finally:
        hip.free(d_seq)
        hip.free(d_off)
    with pytest.raises(k2.K2GError, match="want_hll"):
        k2.Pipeline(db, 100, 10000, host_io=True, hll=2)
