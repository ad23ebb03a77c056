"""k2g_classify_device_noalloc (classify with every scratch buffer carved
from a static device array) against k2g_classify on the same fragments:
calls, hit groups, total k-mers, every hitlist run, HLL records and
per-taxon k-mer counts equal, over two calls in a row (the second smaller,
reusing the same scratch), the overflow path forced with a 2-slot hit
table, and the size limits rejected with K2G_EUSAGE."""
import ctypes as C
import os

import numpy as np
import pytest

from helpers import chimeric_reads, edge_reads, read_fasta

pytestmark = pytest.mark.gpu


class DeviceResults(C.Structure):
    _fields_ = [("call", C.c_void_p), ("hit_groups", C.c_void_p), ("total_kmers", C.c_void_p),
                ("n_runs", C.c_void_p), ("run_off", C.c_void_p), ("runs", C.c_void_p),
                ("run_capacity", C.c_uint64), ("hll", C.c_void_p), ("n_hll", C.c_void_p),
                ("kmer_counts", C.c_void_p)]


def _lib(k2):
    lib = k2.load_library()
    lib.k2g_classify_device_noalloc.argtypes = [C.c_void_p, C.POINTER(k2._lib.ClassifyParams), C.c_void_p,
                                                C.c_void_p, C.c_void_p, C.c_uint64, C.POINTER(DeviceResults),
                                                C.c_void_p]
    return lib


def _noalloc(k2, db, seq, off, *, paired=False, confidence=0.0, min_hit_groups=0, mode=1, hll=False,
             kmer_counts=False, table_slots=0, run_capacity=None):
    from kraken2_amd.synth import Hip
    lib = _lib(k2)
    hip = Hip()
    n = (len(off) - 1) // 2
    cap = int(off[-1]) + n + 16 if run_capacity is None else run_capacity
    nodes = int(db.info.node_count)
    bufs = {}
    try:
        bufs["seq"] = hip.upload(np.frombuffer(seq, np.uint8) if len(seq) else np.zeros(1, np.uint8))
        bufs["off"] = hip.upload(np.ascontiguousarray(off, np.uint64))
        for name, nbytes in (("call", 4 * n), ("groups", 4 * n), ("total", 4 * n), ("n_runs", 4 * n),
                             ("run_off", 8 * (n + 1)), ("runs", 8 * max(cap, 1)), ("hll", 8 * max(cap, 1)),
                             ("n_hll", 4 * n), ("kc", 8 * nodes)):
            bufs[name] = hip.malloc(nbytes)
        r = DeviceResults(bufs["call"], bufs["groups"], bufs["total"], bufs["n_runs"], bufs["run_off"],
                          bufs["runs"], cap, bufs["hll"] if hll else None, bufs["n_hll"] if hll else None,
                          bufs["kc"] if kmer_counts else None)
        p = k2._lib.ClassifyParams(confidence, min_hit_groups, 0, int(paired), mode, int(kmer_counts), int(hll),
                                   table_slots, 0, 0, 0)
        rc = lib.k2g_classify_device_noalloc(db._h, C.byref(p), bufs["seq"], bufs["off"], None, n, C.byref(r),
                                             None)
        if rc:
            return rc, lib.k2g_last_error().decode()
        assert hip.lib.hipDeviceSynchronize() == 0
        if n == 0:
            return 0, {"call": np.zeros(0, np.uint32)}

        def get(name, dtype, count):
            return np.frombuffer(hip.download(bufs[name], count * np.dtype(dtype).itemsize), dtype)

        res = {"call": get("call", np.uint32, n), "hit_groups": get("groups", np.uint32, n),
               "total_kmers": get("total", np.uint32, n), "n_runs": get("n_runs", np.uint32, n),
               "run_off": get("run_off", np.uint64, n + 1)}
        used = int(res["run_off"][-1])
        res["runs"] = get("runs", np.uint32, 2 * used).reshape(-1, 2)
        if hll:
            res["n_hll"] = get("n_hll", np.uint32, n)
            res["hll"] = get("hll", np.uint32, 2 * used).reshape(-1, 2)
        if kmer_counts:
            res["kmer_counts"] = get("kc", np.uint64, nodes)
        return 0, res
    finally:
        for v in bufs.values():
            hip.free(v)


def _same(out, res, hll, kmer_counts):
    n = len(out.call)
    assert np.array_equal(out.call, res["call"])
    assert np.array_equal(out.hit_groups, res["hit_groups"])
    assert np.array_equal(out.total_kmers, res["total_kmers"])
    for i in range(n):
        a = out.runs[out.run_off[i]:out.run_off[i + 1]]
        o = int(res["run_off"][i])
        b = res["runs"][o:o + int(res["n_runs"][i])]
        assert np.array_equal(a, b), f"runs of fragment {i}"
        if hll:
            a = out.hll[out.hll_off[i]:out.hll_off[i + 1]]
            b = res["hll"][o:o + int(res["n_hll"][i])]
            assert np.array_equal(a, b), f"HLL records of fragment {i}"
    if kmer_counts:
        assert np.array_equal(out.kmer_counts, res["kmer_counts"])


@pytest.fixture(scope="module")
def env(synth_db):
    import kraken2_amd as k2
    db = k2.Db.from_dir(synth_db)
    batch = k2.Batch(db)
    _, genomes = read_fasta(os.path.join(synth_db, "library.fna"))
    yield k2, db, batch, genomes
    batch.close()
    db.close()


def test_noalloc_equals_classify_two_calls(env):
    k2, db, batch, genomes = env
    # call 1: uneven lengths, persistent kernel, HLL records and counts, a
    # 2-slot hit table so the dense overflow path runs
    reads = edge_reads(genomes, 3000, seed=1201) + chimeric_reads(genomes, 40, seed=1202, length=5000)
    seq, off = k2.pack_fragments(reads)
    kw = dict(min_hit_groups=2, mode=1, hll=True, kmer_counts=True, table_slots=2)
    out = batch.classify(seq, off, **kw)
    assert out.redone > 0
    rc, res = _noalloc(k2, db, seq, off, **kw)
    assert rc == 0, res
    _same(out, res, True, True)
    # call 2, smaller, on the same scratch: pairs, plain kernel, confidence
    m1 = edge_reads(genomes, 400, seed=1203, min_len=50, max_len=300)
    m2 = edge_reads(genomes, 400, seed=1204, min_len=50, max_len=300)
    seq, off = k2.pack_fragments(m1, m2)
    kw = dict(paired=True, confidence=0.3, min_hit_groups=1, mode=0)
    out = batch.classify(seq, off, **kw)
    rc, res = _noalloc(k2, db, seq, off, **kw)
    assert rc == 0, res
    _same(out, res, False, False)
    # and mode 2 on a third, smaller batch
    seq, off = k2.pack_fragments(edge_reads(genomes, 100, seed=1205))
    out = batch.classify(seq, off, mode=2)
    rc, res = _noalloc(k2, db, seq, off, mode=2)
    assert rc == 0, res
    _same(out, res, False, False)


def test_noalloc_limits(env):
    k2, db, batch, genomes = env
    lib = _lib(k2)
    p = k2._lib.ClassifyParams(0.0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0)
    r = DeviceResults()
    # more fragments than the scratch is sized for: rejected before any device access
    rc = lib.k2g_classify_device_noalloc(db._h, C.byref(p), None, None, None, (1 << 19) + 1, C.byref(r), None)
    assert rc == 64 and b"at most 524288" in lib.k2g_last_error()
    # run capacity below the batch's element count
    seq, off = k2.pack_fragments(edge_reads(genomes, 50, seed=1206, min_len=200, max_len=300))
    rc, msg = _noalloc(k2, db, seq, off, run_capacity=10)
    assert rc == 64 and "run capacity" in msg
    # an empty batch is fine
    rc, res = _noalloc(k2, db, b"", np.zeros(1, np.uint64))
    assert rc == 0 and len(res["call"]) == 0
