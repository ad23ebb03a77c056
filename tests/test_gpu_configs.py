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
    built by the GPU build front-ends; 100k x 150 bp reads.  The same inputs
    also built by the oracle's deterministic build (the reference's)."""
    import kraken2_amd as k2
    d = str(tmp_path_factory.mktemp("c1"))
    _log("config 1: building the database with k2g_estimate_capacity / k2g_build_db")
    make_synthetic_db(d, seed=1, levels="4,4,4", genome_len=200000,
                      tools=(os.path.join(k2.BIN_DIR, "k2g_estimate_capacity"),
                             os.path.join(k2.BIN_DIR, "k2g_build_db")))
    d_orc = str(tmp_path_factory.mktemp("c1_oracle_build"))
    _log("config 1: the same inputs built by the oracle")
    make_synthetic_db(d_orc, seed=1, levels="4,4,4", genome_len=200000)
    reads = os.path.join(d, "reads.fa")
    run([K2SYNTH, "reads", "--lib", os.path.join(d, "library.fna"), "--out", reads, "--seed", "1", "--n",
         "100000", "--len", "150", "--sub", "0.01", "--random-frac", "0.1"])
    _, seqs = read_fasta(reads)
    return d, seqs, d_orc


def test_c1_table_builds(c1):
    """Every minimizer the build inserts is found in the GPU-built and the
    oracle-built config-1 tables; each cell holds the LCA of the taxa of the
    minimizers whose lookup ends there (so a value differs from a
    minimizer's own LCA only where it shares its cell, a compacted-key
    collision); minimizers alone in their cell in both tables have equal
    values.  k2g_build_verify on a GPU build reports the counts this host
    check computes from that build's cells."""
    import kraken2_amd as k2
    from kraken2_amd.synth import Hip
    from helpers import DNA_SEED_TEMPLATE, taxonomy_external_ids
    from helpers import buildcheck as bc
    d, _, d_orc = c1
    opts = np.fromfile(os.path.join(d, "opts.k2d"), np.uint64, 4)
    k, l, mask, toggle = (int(x) for x in opts)
    assert mask == k2.expand_spaced_seed_mask(DNA_SEED_TEMPLATE, 2)
    assert open(os.path.join(d, "taxo.k2d"), "rb").read() == open(os.path.join(d_orc, "taxo.k2d"), "rb").read()
    raw = open(os.path.join(d, "taxo.k2d"), "rb").read()
    nn = int(np.frombuffer(raw, np.uint64, 1, 8)[0])
    parent = np.frombuffer(raw, np.uint64, 7 * nn, 32).reshape(nn, 7)[:, 0].astype(np.int64)
    ext = taxonomy_external_ids(os.path.join(d, "taxo.k2d"))
    e2i = {e: i for i, e in enumerate(ext)}
    names, genomes = read_fasta(os.path.join(d, "library.fna"))
    idmap = dict(line.split("\t") for line in open(os.path.join(d, "seqid2taxid.map")).read().split("\n") if line)
    taxa = [e2i[int(idmap[nm])] for nm in names]
    orc = Oracle()
    _log("config 1: inserted pairs from the oracle scanner")
    keys, tax = bc.inserted_pairs(orc, genomes, taxa, k, l, mask, toggle, 16384, 2048)
    uk, own = bc.key_lcas(keys, tax, parent)
    res = {}
    for name, dd in (("gpu", d), ("oracle", d_orc)):
        cells, kb, vb, size = bc.read_table(os.path.join(dd, "hash.k2d"))
        r = bc.check_table(cells, kb, vb, size, parent, uk, own)
        _log(f"config 1 {name} table: " + str({x: r[x] for x in r if x not in ("values", "shared_key_mask")}))
        for must_be_zero in ("missing", "not_covering", "unexplained", "orphan_cells"):
            assert r[must_be_zero] == 0, (name, must_be_zero, r[must_be_zero])
        assert r["differ"] <= r["shared_keys"]
        res[name] = r
    alone = ~res["gpu"]["shared_key_mask"] & ~res["oracle"]["shared_key_mask"]
    assert np.array_equal(res["gpu"]["values"][alone], res["oracle"]["values"][alone])
    assert np.array_equal(res["gpu"]["values"][alone].astype(np.int64), own[alone])

    # k2g_build_verify on an in-process GPU build of the same inputs
    cap = int(np.fromfile(os.path.join(d, "hash.k2d"), np.uint64, 1)[0])
    seq, goff = k2.pack_fragments(genomes)
    gstart = goff[0::2]
    hip = Hip()
    d_seq = hip.upload(np.frombuffer(seq, np.uint8))
    bld = k2.Builder(os.path.join(d, "taxo.k2d"), cap, k=k, l=l, spaced_seed_mask=mask, toggle_mask=toggle,
                     block_size=16384, subblock_size=2048)
    try:
        tx = np.array(taxa, np.uint32)
        bld.insert(seq, gstart, tx)
        chk = bld.verify_device(d_seq, gstart, tx)
        info = bld.info()
        host = bc.check_table(bld.cells(), info["key_bits"], info["value_bits"], info["size"], parent, uk, own)
    finally:
        bld.close()
        hip.free(d_seq)
    _log(f"config 1 k2g_build_verify: {chk}")
    assert chk["occurrences"] == len(keys)
    for f in ("keys", "cells", "missing", "not_covering", "shared_cells", "shared_keys", "differ",
              "unexplained", "orphan_cells"):
        assert chk[f] == host[f], (f, chk[f], host[f])


def test_c1_every_read_matches_oracle(c1):
    import kraken2_amd as k2
    d, seqs, _ = c1
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
        wl = Workload(tmp, seed=2, n_genomes=4000, verify=True, log=_log)
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


def _full_size_check(k2, wl, tmp, n, *, seed, confidence, per_kind, **read_kw):
    db = k2.Db.from_dir(tmp)
    batch = k2.Batch(db)
    rs = None
    orc = Oracle()
    h = None
    try:
        _log(f"generating {n} reads (seed {seed}) on the GPU")
        rs = wl.make_reads(n, seed=seed, **read_kw)
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


def test_c2_table_verified(c2_table):
    """The config-2 table built on the GPU: every inserted minimizer is
    found, its value covers each of its taxa, and every cell holds the LCA
    of the taxa of the minimizers ending there; minimizers whose value is not
    their own LCA all share a cell with another minimizer of the same
    compacted key (src/compact_hash.h:432-447)."""
    wl, _ = c2_table
    chk = wl.build_check
    for f in ("missing", "not_covering", "unexplained", "orphan_cells"):
        assert chk[f] == 0, (f, chk)
    assert chk["keys"] > 1.5e9 and chk["cells"] > 1.5e9
    assert chk["differ"] <= chk["shared_keys"] and chk["shared_cells"] * 2 <= chk["shared_keys"]


def test_c3_full_size(c2_table):
    """Config 3: 10M reads, 40% single 150 bp, 30% pairs 2x250, 30% long
    (lognormal median 8 kb, cap 50 kb, 2.5% substitutions + 2.5% indels),
    -T 0.1 -g 2, against the config-2 table."""
    import kraken2_amd as k2
    wl, tmp = c2_table
    kinds = _full_size_check(k2, wl, tmp, 10_000_000, f_single=0.4, f_paired=0.3, paired_len=250, off_frac=0.0,
                             seed=3, confidence=0.1, per_kind=1500, group_pairs=True)
    assert set(np.unique(kinds).tolist()) == {0, 1, 2}  # singles, pairs and long reads all sampled


def test_c2_full_size(c2_table):
    """Config 2: 50M pairs 2x150, insert N(350, 50), 1% substitutions, 20%
    off-target reads with GC 0.41, -T 0.1 -g 2 -P."""
    import kraken2_amd as k2
    wl, tmp = c2_table
    _full_size_check(k2, wl, tmp, 50_000_000, f_single=0.0, f_paired=1.0, paired_len=150, off_frac=0.2, seed=2,
                     confidence=0.1, per_kind=4000, group_pairs=True)


# ---------------------------------------------------------------- config 4

@pytest.fixture(scope="module")
def c4_table():
    """The config-4 protein table: 2000 taxa x 3000 proteins (lognormal
    lengths, median 300 aa, background amino-acid frequencies, siblings
    sharing ~80% of residues), k=15 l=12, built and verified on the GPU."""
    from kraken2_amd.synth import ProteinWorkload
    tmp = large_scratch_dir("k2g_test_c4_", need_gb=20)
    wl = None
    try:
        _log(f"config 4 table: 2000 x 3000 proteins into {tmp}")
        wl = ProteinWorkload(tmp, seed=4, verify=True, log=_log)
        yield wl, tmp
    finally:
        if wl is not None and wl.d_prot is not None:
            wl.free_proteomes()
        shutil.rmtree(tmp, ignore_errors=True)


def test_c4_table_verified(c4_table):
    wl, _ = c4_table
    chk = wl.build_check
    for f in ("missing", "not_covering", "unexplained", "orphan_cells"):
        assert chk[f] == 0, (f, chk)
    assert wl.residues > 1.8e9 and chk["keys"] > 5e8
    assert chk["differ"] <= chk["shared_keys"]


def test_c4_full_size(c4_table):
    """Config 4: 5M nucleotide reads of 150 bp back-translated from the
    proteomes (random synonymous codons, 1% substitutions, either strand),
    six-frame translated search (src/aa_translate.cc:41-100), -g 2."""
    import kraken2_amd as k2
    wl, tmp = c4_table
    _full_size_check(k2, wl, tmp, 5_000_000, seed=4, confidence=0.0, per_kind=4000)


# ---------------------------------------------------------------- config 5

def test_c5_requests_reports_match_oracle(c2_table, tmp_path):
    """Config-5-shaped serving against the resident config-2 table: requests
    of log-uniform size (1 to 20 000 fragments) sliced from the config-2 and
    config-3 read mixes, -T 0.1 -g 2 -R -K, submitted together to a server
    with four slots.  Every request's results equal the same fragments
    classified alone; its kraken output lines (call and hitlist) and its
    -R -K kreport (read counts, k-mer counts, distinct-minimizer estimates
    folded in the reference's input blocks) equal the oracle's classify
    driver run over the request's reads as one classify call
    (src/classify.cc:373-430 runs each request that way)."""
    import kraken2_amd as k2
    wl, tmp = c2_table
    db = k2.Db.from_dir(tmp)
    batch = k2.Batch(db)
    orc = Oracle()
    h = None
    pools = {}
    try:
        for name, kw in (("pairs", dict(f_single=0.0, f_paired=1.0, paired_len=150, off_frac=0.2, seed=52)),
                         ("uneven", dict(f_single=0.4, f_paired=0.3, paired_len=250, seed=53, group_pairs=True))):
            rs = wl.make_reads(40000, **kw)
            seq = rs.hip.download(rs.d_seq, rs.bases)
            pools[name] = (seq, rs.off.copy(), rs.paired.copy())
            rs.free()
        rng = np.random.default_rng(55)
        reqs = []
        for i in range(24):
            name = ("pairs", "uneven")[i % 2]
            seq, off, pf = pools[name]
            n_pool = len(pf)
            size = int(np.exp(rng.uniform(0, np.log(20000))))
            want_pair = bool(i % 4 == 1) if name == "uneven" else True
            idx = np.flatnonzero(pf == int(want_pair))
            s0 = int(idx[0]) + int(rng.integers(0, max(1, len(idx) - size)))
            e0 = min(s0 + size, int(idx[-1]) + 1, n_pool)
            a0, a2 = int(off[2 * s0]), int(off[2 * e0])
            reqs.append((seq[a0:a2], (off[2 * s0:2 * e0 + 1] - off[2 * s0]).astype(np.uint64), want_pair))
        kw = dict(confidence=0.1, min_hit_groups=2, kmer_counts=True, hll=True)
        srv = k2.Server(db, 8 << 30, slots=4)
        try:
            tickets = [srv.submit(s, o, paired=p, **kw) for s, o, p in reqs]
            outs = [srv.wait(t)[0] for t in tickets]
            st = srv.stats()
        finally:
            srv.close()
        assert st["max_concurrent"] > 1
        _log(f"config 5: {len(reqs)} requests served ({sum(len(o) // 2 for _, o, _ in reqs)} fragments); "
             "the oracle on each")
        h = orc.open(tmp)
        for i, ((s, o, p), out) in enumerate(zip(reqs, outs)):
            n = (len(o) - 1) // 2
            alone = batch.classify(s, o, paired=p, mode=2, **kw)
            for f in ("call", "hit_groups", "total_kmers", "run_off", "runs", "hll_off", "hll", "kmer_counts"):
                assert np.array_equal(getattr(out, f), getattr(alone, f)), (i, f)
            # the request's reads as input files, run through the oracle's driver
            d = tmp_path / f"r{i}"
            d.mkdir()
            m1, m2 = [], []
            for j in range(n):
                a0, a1, a2 = (int(x) for x in o[2 * j:2 * j + 3])
                m1.append(b">r%d%s\n%s\n" % (j, b"/1" if p else b"", s[a0:a1]))
                if p:
                    m2.append(b">r%d/2\n%s\n" % (j, s[a1:a2]))
            (d / "1.fa").write_bytes(b"".join(m1))
            if p:
                (d / "2.fa").write_bytes(b"".join(m2))
            orc.run_files(h, str(d / "1.fa"), str(d / "2.fa") if p else None, kraken_out=str(d / "orc.kraken"),
                          report=str(d / "orc.kreport"), kmer_data=True, confidence=0.1, min_hit_groups=2,
                          threads=16)
            lines = (d / "orc.kraken").read_text().splitlines()
            assert len(lines) == n
            for j, line in enumerate(lines):
                f = line.split("\t")
                runs = out.runs[out.run_off[j]:out.run_off[j + 1]]
                assert (f[0], int(f[2]), f[4]) == ("C" if out.call[j] else "U", int(db.external_ids[out.call[j]]),
                                                   k2.format_hitlist(runs, db.external_ids)), (i, j)
            k2.write_report(db, o, out, str(d / "gpu.kreport"), paired=p, kmer_data=True)
            assert (d / "gpu.kreport").read_text() == (d / "orc.kreport").read_text(), f"request {i} kreport"
    finally:
        if h:
            orc.lib.k2cpu_db_close(h)
        batch.close()
        db.close()
