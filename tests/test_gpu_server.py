"""The request server: FIFO admission against a device-memory budget,
concurrent requests on separate streams, the table resident.  Every
request's results must equal the same fragments classified alone
(src/classify.cc:373-430 runs each request as a whole classify() call)."""
import os
import random
import threading
import time

import numpy as np
import pytest

from helpers import Oracle, chimeric_reads, edge_reads, read_fasta

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env(synth_db):
    import kraken2_amd as k2
    db = k2.Db.from_dir(synth_db)
    batch = k2.Batch(db)
    _, genomes = read_fasta(os.path.join(synth_db, "library.fna"))
    yield k2, db, batch, genomes
    batch.close()
    db.close()


def _requests(k2, genomes, count, seed):
    rng = random.Random(seed)
    reqs = []
    for r in range(count):
        size = int(round(10 ** rng.uniform(0, 3.3)))  # 1 .. ~2000 fragments
        paired = rng.random() < 0.3
        pick = rng.random()
        if pick < 0.2:
            m1 = chimeric_reads(genomes, max(1, size // 20), seed=seed * 1000 + r, length=rng.choice([600, 3000]))
        else:
            m1 = edge_reads(genomes, size, seed=seed * 1000 + r, min_len=50, max_len=400)
        m2 = edge_reads(genomes, len(m1), seed=seed * 1000 + r + 500, min_len=50, max_len=300) if paired else None
        seq, off = k2.pack_fragments(m1, m2)
        params = dict(paired=paired, confidence=rng.choice([0.0, 0.1, 0.5]), min_hit_groups=rng.choice([0, 2]),
                      hll=rng.random() < 0.5, table_slots=rng.choice([0, 0, 0, 4]))
        params["kmer_counts"] = params["hll"]
        reqs.append((seq, off, params))
    return reqs


def _same(a, b):
    for f in ("call", "hit_groups", "total_kmers", "run_off", "runs"):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f
    if a.hll is not None or b.hll is not None:
        assert np.array_equal(a.hll_off, b.hll_off) and np.array_equal(a.hll, b.hll)
        assert np.array_equal(a.kmer_counts, b.kmer_counts)


def test_server_mix_equals_alone(env):
    """Requests submitted together (queued behind a small budget, several on
    the GPU at once) give exactly the results of each one classified alone."""
    k2, db, batch, genomes = env
    reqs = _requests(k2, genomes, 40, seed=7)
    alone = [batch.classify(seq, off, **p) for seq, off, p in reqs]
    budget = 48 << 20
    srv = k2.Server(db, budget, slots=4)
    try:
        tickets = [srv.submit(seq, off, **p) for seq, off, p in reqs]
        for t, ref in zip(tickets, alone):
            out, times = srv.wait(t)
            _same(out, ref)
            assert times["total_ms"] >= times["gpu_ms"] >= 0
        st = srv.stats()
    finally:
        srv.close()
    assert st["peak_bytes"] <= st["budget"] == budget
    assert st["max_concurrent"] > 1


def test_server_threads_and_oracle(env, synth_db):
    """Submitters on several threads; results checked against the CPU oracle."""
    k2, db, batch, genomes = env
    reqs = _requests(k2, genomes, 24, seed=11)
    srv = k2.Server(db, 256 << 20, slots=3)
    outs = [None] * len(reqs)
    errs = []

    def client(idx):
        try:
            for i in idx:
                seq, off, p = reqs[i]
                t = srv.submit(seq, off, **p)
                time.sleep(0.001)
                outs[i] = srv.wait(t)[0]
        except Exception as e:  # reported below
            errs.append(e)

    th = [threading.Thread(target=client, args=(list(range(j, len(reqs), 4)),)) for j in range(4)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    srv.close()
    assert not errs, errs
    orc = Oracle()
    h = orc.open(synth_db)
    for (seq, off, p), out in zip(reqs, outs):
        ref = orc.classify(h, seq, off, paired=p["paired"], confidence=p["confidence"],
                           min_hit_groups=p["min_hit_groups"])
        n = len(out.call)
        for i in range(n):
            runs = out.runs[out.run_off[i]:out.run_off[i + 1]]
            assert (int(out.call[i]), k2.format_hitlist(runs, db.external_ids)) == \
                   (int(ref["call"][i]), ref["text"][i])


def test_server_rejects_request_over_budget(env):
    k2, db, batch, genomes = env
    srv = k2.Server(db, 1 << 20, slots=1)
    try:
        m1 = edge_reads(genomes, 20000, seed=3, min_len=300, max_len=400)
        seq, off = k2.pack_fragments(m1)
        with pytest.raises(k2.K2GError) as ei:
            srv.submit(seq, off)
        assert ei.value.code == 64
        t = srv.submit(*k2.pack_fragments(m1[:3]))  # a small one still runs
        out, _ = srv.wait(t)
        assert len(out.call) == 3
    finally:
        srv.close()


def test_server_pinned_inputs_and_checksums(env):
    """Requests whose reads sit in pinned host memory; results compared by
    checksum with the same fragments classified alone."""
    import kraken2_amd as k2
    from kraken2_amd._lib import pinned_array
    k2m, db, batch, genomes = env
    reqs = _requests(k2, genomes, 16, seed=13)
    srv = k2.Server(db, 64 << 20, slots=3)
    try:
        tickets, alone = [], []
        for seq, off, p in reqs:
            buf = pinned_array(len(seq))
            buf[:] = np.frombuffer(seq, np.uint8)
            tickets.append(srv.submit(buf, off, **p))
            alone.append(batch.classify_checksum(seq, off, **p))
        for t, ref in zip(tickets, alone):
            h, _ = srv.wait(t, checksum_only=True)
            assert h == ref
    finally:
        srv.close()


def test_server_checksum_without_hll_records(env):
    """A request whose fragments give no HLL record (one random pair, no
    hits) has an empty record array on both paths: equal checksums."""
    k2m, db, batch, genomes = env
    import kraken2_amd as k2
    rng = np.random.default_rng(17)
    m1 = ["".join(rng.choice(list("ACGT"), 150)) for _ in range(1)]
    m2 = ["".join(rng.choice(list("ACGT"), 150)) for _ in range(1)]
    seq, off = k2.pack_fragments(m1, m2)
    p = dict(paired=True, confidence=0.1, min_hit_groups=2, hll=True, kmer_counts=True)
    out = batch.classify(seq, off, **p)
    assert len(out.hll) == 0
    srv = k2.Server(db, 64 << 20, slots=2)
    try:
        h, _ = srv.wait(srv.submit(seq, off, **p), checksum_only=True)
    finally:
        srv.close()
    assert h == batch.classify_checksum(seq, off, **p)
