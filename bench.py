#!/usr/bin/env python3
"""Benchmark: reads/s of the batched persistent GPU classify against the plain
one-launch-per-batch GPU classify on the same device-resident reads, both
checked bit-exact (against each other on every read, and against the CPU
oracle on a sample), with the reference CPU classify timed on this host's
cores beside them.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c3]

Workloads (all synthetic and seeded):
  c3  batched uneven-length classify: 4000 genomes (lognormal, median 2.5 Mb)
      built into a ~2.5e9-cell table on the GPU, 10M reads mixed 40% single
      150 bp / 30% pairs 2x250 / 30% long (lognormal median 8 kb, cap 50 kb,
      2.5% substitutions + 2.5% 1-3 bp indels), shuffled.
  c2  plain classify at real shape: the same table, 50M pairs 2x150 with
      20% off-target reads.
  c1  the small CPU-runnable case: 64 genomes x 200 kb, 100k x 150 bp.

A step is one classify pass over the whole read set already resident in HBM.
Prints one JSON line.
"""
import argparse
import ctypes as C
import hashlib
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0
METRIC = "reads/sec batched vs plain GPU classify (bit-exact) + p50/p99 request latency"
BLOCK_BASES = 3 * 1024 * 1024       # reference block for single reads (src/classify.cc:527)
BLOCK_PAIRS = 10000                 # reference block for pairs (src/classify.cc:31, 531)
DNA_SEED = "1" * 17 + "01" * 7      # kraken2-build's spaced seed for l = 31 (7 spaced positions)

WORKLOADS = {
    "c1": dict(kind="host", desc="c1: plain classify, 64 genomes x 200 kb, 100k x 150 bp single-end, -T 0 -g 2",
               levels="4,4,4", genome_len=200000, n_reads=100000, read_len=150, seed=1, confidence=0.0,
               min_hit_groups=2),
    "c3": dict(kind="gpu", desc="c3: batched uneven-length classify, 10M reads (40% 150 bp single, 30% 2x250 "
               "pairs, 30% long lognormal median 8 kb cap 50 kb with 5% errors) against the c2 table "
               "(4000 genomes, ~2.5e9 cells), -T 0.1 -g 2",
               seed=3, n_reads=10_000_000, f_single=0.4, f_paired=0.3, paired_len=250, off_frac=0.0,
               confidence=0.1, min_hit_groups=2),
    "c4": dict(kind="protein", desc="c4: translated classify, 5M x 150 bp reads back-translated (random synonymous "
               "codons, 1% substitutions, either strand) from 2000 synthetic proteomes x 3000 proteins (lognormal "
               "median 300 aa, siblings ~80% identical), protein table k=15 l=12, six-frame search, -T 0 -g 2",
               seed=4, n_reads=5_000_000, confidence=0.0, min_hit_groups=2),
    "c2": dict(kind="gpu", desc="c2: plain classify at real shape, 50M pairs 2x150 (20% off-target GC 0.41) "
               "against 4000 genomes / ~2.5e9 cells, -T 0.1 -g 2 -P",
               seed=2, n_reads=50_000_000, f_single=0.0, f_paired=1.0, paired_len=150, off_frac=0.2,
               confidence=0.1, min_hit_groups=2),
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def scratch_dir(prefix):
    for base in ("/dev/shm", None):
        try:
            if base and shutil.disk_usage(base).free < (40 << 30):
                continue
            return tempfile.mkdtemp(prefix=prefix, dir=base)
        except OSError:
            continue
    return tempfile.mkdtemp(prefix=prefix)


def run_space(off, paired, k, pct=25):
    """Hitlist space the library reserves per fragment: min(e, e * pct / 100 + 64)
    of its exact k-mer-element count e (nucleotide databases)."""
    l1 = (off[1::2][:len(paired)] - off[0::2][:len(paired)]).astype(np.int64)
    l2 = (off[2::2] - off[1::2][:len(paired)]).astype(np.int64)
    e = np.maximum(l1 - k + 1, 0) + np.where(paired == 1, np.maximum(l2 - k + 1, 0) + 1, 0)
    return np.minimum(e, e * pct // 100 + 64).astype(np.int64)


def split_chunks(bounds, limit):
    """Contiguous fragment ranges whose reserved hitlist space stays under limit."""
    csum = np.cumsum(bounds)
    chunks, s, base = [], 0, 0
    n = len(bounds)
    while s < n:
        e = int(np.searchsorted(csum, base + limit, side="right"))
        e = max(e, s + 1)
        chunks.append((s, e))
        base = int(csum[e - 1])
        s = e
    return chunks


def reference_blocks(off, paired):
    """The reference's input blocks over the fragment stream (src/classify.cc:
    523-543): unpaired reads until 3 MiB of bases, pairs 10 000 at a time; a
    block never mixes the two (they come from separate classify runs)."""
    n = len(paired)
    lens = (off[2::2] - off[0::2][:n]).astype(np.int64)
    blocks, s = [], 0
    while s < n:
        e = s
        if paired[s]:
            while e < n and paired[e] and e - s < BLOCK_PAIRS:
                e += 1
        else:
            bases = 0
            while e < n and not paired[e] and bases < BLOCK_BASES:
                bases += int(lens[e])
                e += 1
        blocks.append((s, e))
        s = e
    return blocks


FRONTEND_BATCH_BASES = 64 << 20    # the classify front-end's GPU batch (csrc/host/classify_main.cc)
FRONTEND_BATCH_BLOCKS = 4096


def frontend_batches(blocks, off, paired):
    """The classify front-end's GPU batches over the reference blocks: whole
    blocks until 64 MiB of bases or 4096 blocks, one launch each; single
    reads and pairs (separate inputs of the front-end) never share a batch."""
    out, s = [], 0
    while s < len(blocks):
        e, bases = s, 0
        kind = paired[blocks[s][0]]
        while e < len(blocks) and bases < FRONTEND_BATCH_BASES and e - s < FRONTEND_BATCH_BLOCKS and \
                paired[blocks[e][0]] == kind:
            bases += int(off[2 * blocks[e][1]]) - int(off[2 * blocks[e][0]])
            e += 1
        out.append((blocks[s][0], blocks[e - 1][1]))
        s = e
    return out


def cpu_model():
    """The host CPU's model name and logical CPU count (from /proc/cpuinfo)."""
    try:
        names = [l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("model name")]
        return f"{names[0]} ({len(names)} logical CPUs on the host)" if names else None
    except OSError:
        return None


class Runner:
    """Drives k2g_classify_device over ranges of device-resident fragments on one batch."""

    def __init__(self, k2, db, d_seq, d_off, d_paired, work):
        self.k2 = k2
        self.lib = k2.load_library()
        self.batch = k2.Batch(db)
        self.d_seq, self.d_off, self.d_paired = d_seq, d_off, d_paired
        self.work = work

    def run(self, s, e, mode, stats=False):
        lib = self.lib
        p = self.k2._lib.ClassifyParams(self.work["confidence"], self.work["min_hit_groups"], 0, 0, mode, 0, 0, 0,
                                        int(stats), 0)
        pp = C.c_void_p(self.d_paired.value + s) if self.d_paired is not None else None
        rc = lib.k2g_classify_device(self.batch._h, C.byref(p), self.d_seq,
                                     C.c_void_p(self.d_off.value + 16 * s), pp, e - s, None)
        ms = C.c_double()
        rc = rc or lib.k2g_batch_sync(self.batch._h, C.byref(ms))
        if rc:
            raise self.k2.K2GError(rc, lib.k2g_last_error().decode())
        return ms.value

    def close(self):
        self.batch.close()


def run_graph_pipeline(k2, db, runner, range_sets, off, paired, hip, batched_digests, a, work):
    """The persistent pass as graph replays (k2g_pipeline): autotuned on an
    even sample of the read set, then, for each set of fragment ranges, W
    warmup and K timed steps of one replay per range, beside the direct
    persistent path (k2g_classify_device) over the same ranges.  The
    replays' digests are compared with the batched pass's."""
    n = (len(off) - 1) // 2
    kw = dict(confidence=work["confidence"], min_hit_groups=work["min_hit_groups"])
    stride = max(1, n // max(1, a.tune_sample))
    t0 = time.perf_counter()
    best, rep = k2.tune_pipeline(db, runner.d_seq, runner.d_off, runner.d_paired, n, stride=stride, **kw)
    tune_s = time.perf_counter() - t0
    log(f"# graph pipeline: tuned on {rep['sample_fragments']} fragments in {tune_s:.1f}s "
        f"({rep['evaluated']} configurations, {rep['mismatched']} mismatched): {best}, "
        f"{rep['default_ms']:.2f} -> {rep['best_ms']:.2f} ms")

    def dev(s, e):
        return (runner.d_seq, C.c_void_p(runner.d_off.value + 16 * s),
                C.c_void_p(runner.d_paired.value + s) if runner.d_paired is not None else None, e - s)

    def timed(name, step):
        for _ in range(a.warmup):
            step()
        t = time.perf_counter()
        dms = []
        for _ in range(a.steps):
            dms.append(step())
            log(f"# {name}: step {len(dms)}/{a.steps} device {dms[-1]:.1f} ms")
        wall = (time.perf_counter() - t) / a.steps
        log(f"# {name}: {wall * 1e3:.1f} ms/step, device {np.mean(dms):.1f} ms")
        return wall, float(np.mean(dms))

    out = {"tuning": best,
           "tune": {"sample_fragments": rep["sample_fragments"], "stride": stride,
                    "configurations": rep["evaluated"], "mismatched": rep["mismatched"],
                    "default_ms": round(rep["default_ms"], 3), "best_ms": round(rep["best_ms"], 3),
                    "seconds": round(tune_s, 1)},
           "chain": "element bounds, reserved hitlist layout, length-bucket work units, persistent classify, "
                    "out-of-space redo into an overflow pool, dense redo: one HIP graph replay per range, no "
                    "host round trip"}
    all_equal = True
    for set_name, ranges in range_sets.items():
        cap_n = max(e - s for s, e in ranges)
        cap_b = max(int(off[2 * e] - off[2 * s]) for s, e in ranges)
        pl = k2.Pipeline(db, cap_n, cap_b, host_io=False, tuning=best, **kw)
        _, grid = pl.params()
        g_wall, g_ms = timed(f"graph over {set_name}", lambda: sum(pl.run_device(*dev(s, e)) for s, e in ranges))
        d_wall, d_ms = timed(f"direct over {set_name}", lambda: sum(runner.run(s, e, 1) for s, e in ranges))
        equal = None
        if batched_digests is not None:
            dig = np.zeros(n, np.uint64)
            for s, e in ranges:
                pl.run_device(*dev(s, e))
                dig[s:e] = pl.digests(e - s)
            equal = bool(np.array_equal(dig, batched_digests))
            all_equal = all_equal and equal
            log(f"# graph over {set_name}: digests equal the batched pass's: {equal}")
        pl.close()
        out[set_name] = {"replays_per_step": len(ranges), "grid": grid,
                         "graph": {"reads_per_s": round(n / g_wall, 1), "ms_per_step": round(g_wall * 1e3, 3),
                                   "device_ms_per_step": round(g_ms, 3)},
                         "direct": {"reads_per_s": round(n / d_wall, 1), "ms_per_step": round(d_wall * 1e3, 3),
                                    "kernel_ms_per_step": round(d_ms, 3)},
                         "graph_vs_direct": round(d_wall / g_wall, 3), "digests_equal_batched": equal}
    out["digests_equal_batched"] = all_equal if batched_digests is not None else None
    out["host_batches"] = run_host_pipeline(k2, db, runner, range_sets["frontend_batches"][:a.host_batches], off,
                                            paired, hip, best, a, work)
    return out


def run_host_pipeline(k2, db, runner, batches, off, paired, hip, tuning, a, work):
    """Upload and writeback included: front-end batches written into the host
    pipeline's pinned staging (as the classify front-end's reader writes its
    batches), each run by one graph replay that reads the batch over the bus
    and writes calls, counters and hitlists into pinned host memory.  The
    timed region is the run of each batch (k2g_pipeline_run); filling the
    staging is the reader's work and is not in it.  Every batch's results are
    checked against k2g_classify on the same fragments."""
    kw = dict(confidence=work["confidence"], min_hit_groups=work["min_hit_groups"])
    cap_n = max(e - s for s, e in batches)
    cap_b = max(int(off[2 * e] - off[2 * s]) for s, e in batches)
    pl = k2.Pipeline(db, cap_n, cap_b, host_io=True, tuning=tuning, **kw)
    lib = k2.load_library()
    st_seq, st_off, st_pf = pl.staging(cap_n, cap_b)
    host = [host_range(hip, runner.d_seq, off, paired, s, e) for s, e in batches]
    def run(j):
        sq, so, pf = host[j]
        n = len(pf)
        st_seq[:len(sq)] = np.frombuffer(sq, np.uint8)
        st_off[:2 * n + 1] = so
        st_pf[:n] = pf
        fr = k2._lib.Fragments(st_seq.ctypes.data, st_off.ctypes.data, st_pf.ctypes.data, n)
        res = k2._lib.Results()
        t = time.perf_counter()
        rc = lib.k2g_pipeline_run(pl._h, C.byref(fr), C.byref(res))
        dt = time.perf_counter() - t
        if rc:
            raise k2.K2GError(rc, lib.k2g_last_error().decode())
        return dt, res

    for j in range(min(len(host), a.warmup)):
        run(j)
    total = 0.0
    for step in range(a.steps):
        total += sum(run(j)[0] for j in range(len(host)))
        log(f"# host pipeline: step {step + 1}/{a.steps}")
    wall = total / a.steps
    equal = True
    for j, (sq, so, pf) in enumerate(host):
        _, res = run(j)
        h = C.c_uint64()
        lib.k2g_results_checksum(C.byref(res), C.byref(h))
        equal = equal and h.value == runner.batch.classify_checksum(sq, so, paired_flags=pf, mode=1, **kw)
    pl.close()
    n = sum(e - s for s, e in batches)
    bases = sum(len(x[0]) for x in host)
    log(f"# host pipeline: {len(batches)} front-end batches, {n} fragments, {bases / 1e9:.2f} Gbp in {wall * 1e3:.1f} ms "
        f"per pass ({n / wall / 1e6:.2f} M reads/s, {bases / wall / 1e9:.1f} GB/s up); results equal k2g_classify's: "
        f"{equal}")
    return {"batches": len(batches), "fragments": n, "bases": bases, "ms_per_pass": round(wall * 1e3, 3),
            "reads_per_s": round(n / wall, 1), "upload_gb_per_s": round(bases / wall / 1e9, 2),
            "results_equal_classify": equal,
            "timed": "k2g_pipeline_run per batch: upload over the bus, the classify chain, calls, counters and "
                     "hitlists written into pinned host memory; the batches lie in the pipeline's pinned staging"}


def run_host_workload(a, work, k2):
    """c1: generated on the host with k2synth, built by the GPU build front-ends."""
    from helpers import K2SYNTH, make_synthetic_db, run
    from kraken2_amd.synth import Hip
    tmp = a.keep or scratch_dir("k2g_c1_")
    os.makedirs(tmp, exist_ok=True)
    make_synthetic_db(tmp, seed=work["seed"], levels=work["levels"], genome_len=work["genome_len"],
                      tools=(os.path.join(k2.BIN_DIR, "k2g_estimate_capacity"),
                             os.path.join(k2.BIN_DIR, "k2g_build_db")))
    reads = os.path.join(tmp, "reads.fa")
    run([K2SYNTH, "reads", "--lib", os.path.join(tmp, "library.fna"), "--out", reads, "--seed",
         str(work["seed"]), "--n", str(work["n_reads"]), "--len", str(work["read_len"]), "--sub", "0.01",
         "--random-frac", "0.1"])
    seqs = [l.rstrip(b"\n") for l in open(reads, "rb") if not l.startswith(b">")]
    n = len(seqs)
    lens = np.array([len(s) for s in seqs], np.uint64)
    ends = np.cumsum(lens, dtype=np.uint64)
    off = np.zeros(2 * n + 1, np.uint64)
    off[0:2 * n:2] = np.concatenate([np.zeros(1, np.uint64), ends[:-1]])
    off[1:2 * n:2] = ends
    off[2 * n] = ends[-1]
    hip = Hip()
    seq = b"".join(seqs)
    d_seq = hip.upload(np.frombuffer(seq, np.uint8))
    d_off = hip.upload(off)
    paired = np.zeros(n, np.uint8)
    host = dict(seq=seq, off=off, reads_fa=reads, pairs_fa=None)
    return tmp, hip, d_seq, d_off, None, off, paired, host


def run_gpu_workload(a, work, k2):
    from kraken2_amd.synth import Workload
    tmp = a.keep or scratch_dir(f"k2g_{a.workload}_")
    wl = Workload(tmp, seed=2, n_genomes=max(5, int(4000 * a.scale)), verify=not a.no_verify, log=log)
    t0 = time.time()
    rs = wl.make_reads(max(100, int(work["n_reads"] * a.scale)), f_single=work["f_single"], f_paired=work["f_paired"],
                       paired_len=work["paired_len"], off_frac=work["off_frac"], seed=work["seed"],
                       group_pairs=True)
    log(f"# reads: {rs.n} fragments, {rs.bases / 1e9:.2f} Gbp in {time.time() - t0:.1f}s")
    pools = None if a.no_serve else make_pools(wl, a)
    wl.free_genomes()
    return tmp, wl.hip, rs.d_seq, rs.d_off, rs.d_paired, rs.off, rs.paired, dict(readset=rs, build_s=wl.build_s,
                                                                               capacity=wl.capacity,
                                                                               estimate=wl.estimate, pools=pools,
                                                                               build_check=wl.build_check)


def run_protein_workload(a, work, k2):
    """c4: proteomes, protein table and back-translated reads, all on the GPU."""
    from kraken2_amd.synth import ProteinWorkload
    tmp = a.keep or scratch_dir("k2g_c4_")
    wl = ProteinWorkload(tmp, seed=work["seed"], n_taxa=max(5, int(2000 * a.scale)), verify=not a.no_verify,
                         log=log)
    t0 = time.time()
    rs = wl.make_reads(max(100, int(work["n_reads"] * a.scale)), seed=work["seed"])
    log(f"# reads: {rs.n} fragments, {rs.bases / 1e9:.2f} Gbp in {time.time() - t0:.1f}s")
    wl.free_proteomes()
    return tmp, wl.hip, rs.d_seq, rs.d_off, rs.d_paired, rs.off, rs.paired, dict(
        readset=rs, build_s=wl.build_s, capacity=wl.capacity, estimate=wl.estimate, build_check=wl.build_check)


def make_pools(wl, a):
    """Host read pools that serving requests are sliced from: the paired
    2x150 mix with 20% off-target reads, and the uneven-length mix (single,
    pairs and long reads interleaved), both generated on the GPU."""
    pools = {}
    m = max(2000, int(a.serve_pool * min(1.0, a.scale * 20)))
    for name, key, seed in (("pairs 2x150", "c2", 52), ("uneven mix", "c3", 53)):
        w = WORKLOADS[key]
        rs = wl.make_reads(m, f_single=w["f_single"], f_paired=w["f_paired"], paired_len=w["paired_len"],
                           off_frac=w["off_frac"], seed=seed, group_pairs=True)
        from kraken2_amd._lib import pinned_array
        seq = pinned_array(rs.bases)  # requests upload from pinned memory
        seq[:] = np.frombuffer(rs.hip.download(rs.d_seq, rs.bases), np.uint8)
        pools[name] = dict(seq=seq, off=rs.off.copy(), paired=rs.paired.copy(), n=rs.n)
        rs.free()
    log(f"# serving pools: " + ", ".join(f"{k} {v['n']} fragments {len(v['seq']) / 1e6:.0f} Mbp"
                                          for k, v in pools.items()))
    return pools


def serve_requests(pools, n_req, max_reads, rng):
    """Request sizes log-uniform in [1, max_reads] fragments, each a random
    slice of one of the pools, within its single reads or within its pairs
    (a request is one classify call, which reads either)."""
    names = sorted(pools)
    reqs = []
    for _ in range(n_req):
        name = names[int(rng.integers(len(names)))]
        pool = pools[name]
        pf = pool["paired"]
        n_single = int((pf == 0).sum())  # the pools hold their single reads first
        lo, hi = (0, n_single) if rng.uniform() < n_single / len(pf) else (n_single, len(pf))
        size = min(hi - lo, int(np.exp(rng.uniform(0.0, np.log(max_reads)))))
        s0 = lo + int(rng.integers(0, hi - lo - size + 1))
        off = pool["off"][2 * s0:2 * (s0 + size) + 1]
        a0, a2 = int(off[0]), int(off[-1])
        reqs.append(dict(seq=pool["seq"][a0:a2], off=(off - off[0]).astype(np.uint64),
                         paired=pool["paired"][s0:s0 + size], pool=name, size=size, bases=a2 - a0))
    return reqs


def serve_at_rate(srv, reqs, rate, work, rng, hll, slots):
    """Poisson arrivals at `rate` requests/s (0: all at once) into a running
    server; returns per-request (results checksum, times), the server-clock
    span and the server's statistics."""
    arrivals = np.cumsum(rng.exponential(1.0 / rate, len(reqs))) if rate else np.zeros(len(reqs))
    tickets = [None] * len(reqs)
    results = [None] * len(reqs)
    submitted = [0.0] * len(reqs)
    ready = threading.Semaphore(0)
    errors = []
    turn = iter(range(len(reqs)))
    turn_lock = threading.Lock()

    def waiter():  # one per slot: results are checksummed while other requests run
        try:
            while True:
                with turn_lock:
                    i = next(turn, None)
                    if i is None:
                        return
                    ready.acquire()
                if tickets[i] is None:  # the submitter stopped
                    return
                results[i] = srv.wait(tickets[i], checksum_only=True)
        except Exception as e:  # surfaced after join
            errors.append(e)

    th = [threading.Thread(target=waiter) for _ in range(max(8, slots))]
    for t in th:
        t.start()
    t0 = time.perf_counter()
    try:
        for i, r in enumerate(reqs):
            delay = t0 + arrivals[i] - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            submitted[i] = time.perf_counter() - t0
            tickets[i] = srv.submit(r["seq"], r["off"], paired_flags=r["paired"], confidence=work["confidence"],
                                    min_hit_groups=work["min_hit_groups"], kmer_counts=True, hll=hll)
            ready.release()
    finally:
        for _ in range(len(reqs) + len(th)):  # wake any waiter still waiting for a ticket
            ready.release()
        for t in th:
            t.join()
    stats = srv.stats()
    if errors:
        raise errors[0]
    # the server's own clock: from the first arrival to the last request done
    # (the waiter threads' checksums run beside it, outside this span)
    wall = max(submitted[i] + results[i][1]["total_ms"] / 1e3 for i in range(len(reqs)))
    return results, wall, stats


def run_serving(k2, db, tmp, pools, a, work):
    """Config 5: the resident table serves requests of log-uniform size with
    -R -K (per-taxon counts and HLL records) under a device-memory budget.
    One request set is served closed-loop by one slot and by all slots (the
    saturation rates), then with Poisson arrivals at 0.5x, 1x and 2x the
    one-slot rate.  Every request must equal itself classified alone, and a
    spread of requests is checked against the oracle's classify driver:
    output lines and the -R -K kreport."""
    rng = np.random.default_rng(5)
    max_reads = max(10, int(200_000 * min(1.0, a.scale * 20)))
    budget = int(a.serve_budget_gb * (1 << 30))
    reqs = serve_requests(pools, a.serve_requests, max_reads, rng)
    reads = sum(r["size"] for r in reqs)
    bases = sum(r["bases"] for r in reqs)
    legs, digests = [], []
    # one server per slot count at a time, warmed by one untimed pass over
    # the request set (a resident service has its result memory laid out)
    servers = {}

    def server(slots):
        if slots not in servers:
            for old in list(servers):
                servers.pop(old).close()
            servers[slots] = k2.Server(db, budget, slots)
            serve_at_rate(servers[slots], reqs, 0, work, rng, a.serve_hll, slots)
            log(f"# serving: {slots}-slot server warmed")
        return servers[slots]

    def leg(name, rate, slots):
        res, wall, st = serve_at_rate(server(slots), reqs, rate, work, rng, a.serve_hll, slots)
        lat = np.array([t["total_ms"] for _, t in res])
        row = {"leg": name, "slots": slots, "offered_req_per_s": round(rate, 2) if rate else None,
               "completed_req_per_s": round(len(reqs) / wall, 2), "reads_per_s": round(reads / wall, 1),
               "mbp_per_s": round(bases / wall / 1e6, 1),
               "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
               "service_ms_mean": round(float(np.mean([t["gpu_ms"] for _, t in res])), 3),
               "max_concurrent": st["max_concurrent"], "peak_bytes": st["peak_bytes"]}
        log(f"# serving {name}: {row}")
        legs.append(row)
        digests.append([d for d, _ in res])
        return row

    one = leg("closed loop, one slot", 0, 1)
    leg(f"closed loop, {a.serve_slots} slots", 0, a.serve_slots)
    mu = one["completed_req_per_s"]
    for mult in (0.5, 1.0, 2.0):
        leg(f"Poisson {mult}x the one-slot rate", mult * mu, a.serve_slots)
    for srv in servers.values():
        srv.close()
    # the same saturation with the other -K form (per-hit records or the
    # GPU's per-(block, taxon) sets), on its own warmed server
    other = 3 - a.serve_hll
    srv = k2.Server(db, budget, a.serve_slots)
    try:
        serve_at_rate(srv, reqs, 0, work, rng, other, a.serve_slots)
        res, wall, st = serve_at_rate(srv, reqs, 0, work, rng, other, a.serve_slots)
    finally:
        srv.close()
    other_leg = {"hll": "per-(block, taxon) sets on the GPU" if other == 2 else "per-hit records",
                 "completed_req_per_s": round(len(reqs) / wall, 2), "mbp_per_s": round(bases / wall / 1e6, 1)}
    log(f"# serving closed loop, {a.serve_slots} slots, the other -K form: {other_leg}")
    # every request equals itself served alone: the one-slot closed loop runs
    # each request by itself, the other legs run them side by side
    alone = digests[0]
    mismatched = sorted({i for dig in digests[1:] for i, d in enumerate(dig) if d != alone[i]})
    kw = dict(confidence=work["confidence"], min_hit_groups=work["min_hit_groups"], kmer_counts=True,
              hll=a.serve_hll)
    # the oracle's classify driver on a spread of request sizes (up to
    # serve_oracle_max fragments, so the CPU check stays short)
    n_check = 0 if a.no_cpu else min(len(reqs), a.serve_oracle)
    by_size = sorted((i for i in range(len(reqs)) if reqs[i]["size"] <= a.serve_oracle_max),
                     key=lambda i: reqs[i]["size"])
    pick = sorted({by_size[int(j)] for j in np.linspace(0, len(by_size) - 1, n_check)}) if n_check and by_size else []
    checked = mism = rep_mism = 0
    if pick:
        from helpers import Oracle
        orc = Oracle()
        h = orc.open(tmp)
        srv = k2.Server(db, budget, 1)
        try:
            for i in pick:
                r = reqs[i]
                out, _ = srv.wait(srv.submit(r["seq"], r["off"], paired_flags=r["paired"], **kw))
                m, rm = oracle_request_check(k2, db, orc, h, tmp, r, out, work)
                mism += m
                rep_mism += rm
                checked += r["size"]
                log(f"# serving oracle check: request of {r['size']} fragments: {m} line and {rm} kreport "
                    "mismatches")
        finally:
            srv.close()
            orc.lib.k2cpu_db_close(h)
    log(f"# serving verify: {len(mismatched)} requests differ from themselves alone; oracle: {mism} fragment "
        f"and {rep_mism} kreport mismatches over {len(pick)} requests ({checked} fragments)")
    sat = legs[1]["completed_req_per_s"] / mu
    two = legs[-1]["completed_req_per_s"] / mu
    hll_form = ("per-(input block, taxon) distinct sets reduced on the GPU" if a.serve_hll == 2
                else "per-hit HLL records")
    return {"workload": "c5: %d requests of log-uniform size in [1, %d] fragments sliced from the pairs 2x150 "
                        "and uneven-length mixes (single reads or pairs), -R -K (counts + %s), against "
                        "the resident table" % (len(reqs), max_reads, hll_form),
            "requests": len(reqs), "reads": reads, "bases": bases,
            "budget_bytes": budget, "slots": a.serve_slots, "legs": legs,
            "saturation_vs_one_slot": round(sat, 3), "completed_at_2x_vs_one_slot": round(two, 3),
            "other_hll_form_saturation": other_leg,
            "bit_exact_vs_alone": not mismatched, "mismatched_requests": len(mismatched),
            "oracle_requests": len(pick), "oracle_fragments": checked, "oracle_mismatches": int(mism),
            "oracle_kreport_mismatches": int(rep_mism)}


def oracle_request_check(k2, db, orc, h, tmp, r, out, work):
    """One served request against the oracle's classify driver run over the
    request's reads as input files (one classify call, src/classify.cc:
    373-430): (fragments whose output line differs, 1 if the -R -K kreport
    differs)."""
    d = tempfile.mkdtemp(prefix="req_", dir=tmp)
    try:
        paired = bool(r["paired"][0]) if r["size"] else False
        seq, off = r["seq"], r["off"]
        m1, m2 = [], []
        for j in range(r["size"]):
            a0, a1, a2 = (int(x) for x in off[2 * j:2 * j + 3])
            m1.append(b">r%d%s\n%s\n" % (j, b"/1" if paired else b"", seq[a0:a1].tobytes()))
            if paired:
                m2.append(b">r%d/2\n%s\n" % (j, seq[a1:a2].tobytes()))
        f1, f2 = os.path.join(d, "1.fa"), os.path.join(d, "2.fa")
        open(f1, "wb").write(b"".join(m1))
        if paired:
            open(f2, "wb").write(b"".join(m2))
        orc.run_files(h, f1, f2 if paired else None, kraken_out=os.path.join(d, "o.kraken"),
                      report=os.path.join(d, "o.kreport"), kmer_data=True, confidence=work["confidence"],
                      min_hit_groups=work["min_hit_groups"], threads=os.cpu_count() or 1)
        mism = 0
        lines = open(os.path.join(d, "o.kraken")).read().splitlines()
        if len(lines) != r["size"]:
            return r["size"], 1
        for j, line in enumerate(lines):
            f = line.split("\t")
            runs = out.runs[out.run_off[j]:out.run_off[j + 1]]
            mism += (f[0], int(f[2]), f[4]) != ("C" if out.call[j] else "U", int(db.external_ids[out.call[j]]),
                                                 k2.format_hitlist(runs, db.external_ids))
        k2.write_report(db, off, out, os.path.join(d, "g.kreport"), paired=paired, kmer_data=True)
        rep = open(os.path.join(d, "g.kreport")).read() != open(os.path.join(d, "o.kreport")).read()
        return mism, int(rep)
    finally:
        shutil.rmtree(d, ignore_errors=True)


def write_sample_fasta(path1, path2, frags):
    """frags: list of (seq1 bytes, seq2 bytes or None, name)."""
    with open(path1, "wb") as f1:
        f2 = open(path2, "wb") if path2 else None
        for s1, s2, name in frags:
            if s2 is None:
                f1.write(b">%s\n%s\n" % (name.encode(), s1))
            else:
                f1.write(b">%s/1\n%s\n" % (name.encode(), s1))
                f2.write(b">%s/2\n%s\n" % (name.encode(), s2))
        if f2:
            f2.close()


def cpu_baseline(work, d, files, threads, exe=None):
    """The reference classify (compiled from its sources into oracle/_ref) when
    present, else the oracle restatement, on all host threads.  Single reads
    and pairs run as two invocations (one run cannot mix them); times add.
    exe: time that command line program instead (the GPU front-end)."""
    ref = os.path.join(ROOT, "oracle", "_ref", "classify")
    port = os.path.join(ROOT, "oracle", "bin", "k2cpu_classify")
    if exe:
        kind = "gpu front-end"
    else:
        exe, kind = (ref, "reference") if os.path.exists(ref) else (port, "port")
    total_n, total_s, total_bp = 0, 0.0, 0.0
    for args in files:
        cmd = [exe, "-H", f"{d}/hash.k2d", "-t", f"{d}/taxo.k2d", "-o", f"{d}/opts.k2d", "-p", str(threads),
               "-T", str(work["confidence"]), "-g", str(work["min_hit_groups"]), "-O", "/dev/null", *args]
        env = dict(os.environ, K2G_TIMING="1") if kind == "gpu front-end" else None
        r = subprocess.run(cmd, capture_output=True, text=True, env=env)
        if r.returncode != 0:
            log(f"# {kind} run failed: {r.stderr[-300:]}")
            return None
        for tl in r.stderr.splitlines():
            if "timing:" in tl:
                log(f"# {kind}{tl.strip().replace('timing:', ' phases:')}")
        m = re.search(r"(\d+) sequences \(([\d.]+) Mbp\) processed in ([\d.]+)s", r.stderr)
        total_n += int(m.group(1))
        total_bp += float(m.group(2))
        total_s += float(m.group(3))
    return {"value": round(total_n / total_s, 1), "unit": "reads/s", "cores": threads, "kind": kind,
            "cpu_model": cpu_model(), "binary": os.path.relpath(exe, ROOT),
            "sample": f"{total_n} fragments ({total_bp:.1f} Mbp) drawn evenly from the workload; the binary's own "
                      f"timed interval (after DB load, through parse, classify and output), {total_s:.2f} s"}


def outputs_equal(work, d, files, threads, gpu_cli):
    """The GPU front-end's kraken output on the sample files, byte for byte
    against the CPU baseline binary's (the reference's compiled classify when
    present, else the oracle restatement), outside the timed runs."""
    ref = os.path.join(ROOT, "oracle", "_ref", "classify")
    cpu_exe = ref if os.path.exists(ref) else os.path.join(ROOT, "oracle", "bin", "k2cpu_classify")
    for k, args in enumerate(files):
        digests = []
        for exe in (cpu_exe, gpu_cli):
            path = os.path.join(d, f"eq_{k}.kraken")
            cmd = [exe, "-H", f"{d}/hash.k2d", "-t", f"{d}/taxo.k2d", "-o", f"{d}/opts.k2d", "-p", str(threads),
                   "-T", str(work["confidence"]), "-g", str(work["min_hit_groups"]), "-O", path, *args]
            if subprocess.run(cmd, capture_output=True).returncode != 0:
                return False
            h = hashlib.sha256()
            with open(path, "rb") as f:
                for blk in iter(lambda: f.read(1 << 24), b""):
                    h.update(blk)
            digests.append(h.hexdigest())
            os.remove(path)
        if digests[0] != digests[1]:
            return False
    return True


def host_range(hip, d_seq, off, paired, s, e):
    """(seq bytes, rebased 2m+1 offsets, paired flags) of fragments [s, e)."""
    a0, a2 = int(off[2 * s]), int(off[2 * e])
    seq = hip.download(d_seq, a2 - a0, a0)
    return seq, (off[2 * s:2 * e + 1] - off[2 * s]).astype(np.uint64), paired[s:e]


def fragment_text(seq, off, j):
    a0, a1, a2 = (int(x) for x in off[2 * j:2 * j + 3])
    return seq[a0:a1], seq[a1:a2]


def oracle_check(k2, db, tmp, seq, off, pflags, work):
    """Classifies the fragments with the library (one batch) and with the CPU
    oracle; returns (mismatches, library digests)."""
    from helpers import Oracle
    n = len(pflags)
    batch = k2.Batch(db)
    batch.classify(seq, off, paired_flags=pflags, confidence=work["confidence"],
                   min_hit_groups=work["min_hit_groups"], mode=1)
    dig = batch.digests(n)
    out = batch.classify(seq, off, paired_flags=pflags, confidence=work["confidence"],
                         min_hit_groups=work["min_hit_groups"], mode=1)
    batch.close()
    orc = Oracle()
    h = orc.open(tmp)
    mism = 0
    for pflag in (0, 1):
        sel = [j for j in range(n) if pflags[j] == pflag]
        if not sel:
            continue
        parts, o = [], np.zeros(2 * len(sel) + 1, np.uint64)
        pos = 0
        for jj, j in enumerate(sel):
            m1, m2 = fragment_text(seq, off, j)
            parts += [m1, m2]
            o[2 * jj] = pos
            o[2 * jj + 1] = pos + len(m1)
            pos += len(m1) + len(m2)
        o[-1] = pos
        ref = orc.classify(h, b"".join(parts), o, paired=bool(pflag), confidence=work["confidence"],
                           min_hit_groups=work["min_hit_groups"])
        for jj, j in enumerate(sel):
            runs = out.runs[out.run_off[j]:out.run_off[j + 1]]
            got = (int(out.call[j]), int(out.hit_groups[j]), int(out.total_kmers[j]),
                   k2.format_hitlist(runs, db.external_ids))
            exp = (int(ref["call"][jj]), int(ref["groups"][jj]), int(ref["total"][jj]), ref["text"][jj])
            mism += got != exp
    orc.lib.k2cpu_db_close(h)
    return mism, dig


def profiled_traffic(workload):
    """HBM bytes per launch of the persistent kernel measured by counters for
    this workload (profiles/traffic.txt), or None.  Raw FETCH_SIZE is scaled
    by the random-line calibration recorded there (k_gather_ceiling's raw
    FETCH_SIZE per random 128-byte line it loads, against 128 bytes);
    WRITE_SIZE is bytes already."""
    path = os.path.join(ROOT, "profiles", "traffic.txt")
    if not os.path.exists(path):
        return None
    kern = calib = None
    for line in open(path):
        f = line.split()
        if len(f) != 4 or line.startswith("#"):
            continue
        if f[0] == workload:
            kern = (int(f[2]) * 1024, int(f[3]) * 1024)
        elif f[0] == "gather":
            calib = 128.0 / (int(f[2]) * 1024 / int(f[3]))
    if kern is None or calib is None:
        return None
    return int(kern[0] * calib + kern[1])


def copy_calibration():
    """1 GiB device-to-device copies through torch's copy kernel: the read and
    write byte counters of these dispatches, divided by 1 GiB, give the
    factor that scales the same counters of the classify kernel to bytes."""
    import torch
    n = 1 << 30
    src = torch.ones(n, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    log(f"# copy calibration: 3 x {n} bytes")


def run_merge_bench(a):
    """Multi-database merge at scale (the reference's merge tool, src/merge.cc):
    a synthetic library (k2synth) split into two halves by genome, one
    database per half built on the GPU (k2g_estimate_capacity, k2g_build_db),
    read pairs classified against each (k2g_classify -P), then the two kraken
    outputs merged under the taxonomy of the whole library, file to file, by
    k2g_merge and by the reference merge binary (the CPU baseline; the oracle
    restatement where it is absent).  Outputs must be byte-identical."""
    import kraken2_amd as k2
    from kraken2_amd.taxonomy import write_k2_taxonomy
    tmp = a.keep or scratch_dir("k2g_merge_")
    os.makedirs(tmp, exist_ok=True)
    synth = os.path.join(k2.BIN_DIR, "k2synth")
    bins = {n: os.path.join(k2.BIN_DIR, n) for n in ("k2g_estimate_capacity", "k2g_build_db", "k2g_classify",
                                                     "k2g_merge")}
    threads = min(16, os.cpu_count() or 1)

    def sh(cmd, **kw):
        r = subprocess.run(cmd, capture_output=True, **kw)
        if r.returncode != 0:
            raise RuntimeError(f"{cmd[0]} failed ({r.returncode}): {r.stderr[-400:]!r}")
        return r

    t0 = time.time()
    sh([synth, "db-inputs", "--out", tmp, "--seed", "5", "--levels", a.merge_levels, "--genome-len",
        str(a.merge_genome_len), "--sub", "0.05"])
    recs = open(os.path.join(tmp, "library.fna"), "rb").read().split(b">")[1:]
    idmap = dict(l.split("\t") for l in open(os.path.join(tmp, "seqid2taxid.map")).read().split("\n") if l)
    tax_dir = os.path.join(tmp, "taxonomy")
    for part in (0, 1):
        d = os.path.join(tmp, f"db{part}")
        os.makedirs(d, exist_ok=True)
        lib = b"".join(b">" + r for i, r in enumerate(recs) if i % 2 == part)
        names = [r.split(None, 1)[0].decode() for i, r in enumerate(recs) if i % 2 == part]
        with open(os.path.join(d, "seqid2taxid.map"), "w") as f:
            f.write("".join(f"{n}\t{idmap[n]}\n" for n in names))
        est = int(sh([bins["k2g_estimate_capacity"], "-k", "35", "-l", "31", "-S", DNA_SEED], input=lib).stdout)
        sh([bins["k2g_build_db"], "-k", "35", "-l", "31", "-S", DNA_SEED, "-H", f"{d}/hash.k2d", "-t",
            f"{d}/taxo.k2d", "-o", f"{d}/opts.k2d", "-n", tax_dir, "-m", f"{d}/seqid2taxid.map", "-c",
            str(int((est + 8192) / 0.7)), "-p", str(threads)], input=lib)
    taxo = os.path.join(tmp, "merged_taxo.k2d")
    write_k2_taxonomy(os.path.join(tax_dir, "nodes.dmp"), os.path.join(tax_dir, "names.dmp"),
                      [int(v) for v in idmap.values()], taxo)
    r1, r2 = os.path.join(tmp, "r1.fa"), os.path.join(tmp, "r2.fa")
    sh([synth, "reads", "--lib", os.path.join(tmp, "library.fna"), "--out", r1, "--out2", r2, "--seed", "7",
        "--n", str(a.merge_reads), "--len", "150", "--sub", "0.01", "--random-frac", "0.1"])
    outs = []
    for part in (0, 1):
        d = os.path.join(tmp, f"db{part}")
        o = os.path.join(tmp, f"k{part}.kraken")
        sh([bins["k2g_classify"], "-H", f"{d}/hash.k2d", "-t", f"{d}/taxo.k2d", "-o", f"{d}/opts.k2d", "-p",
            str(threads), "-P", "-g", "2", "-O", o, r1, r2])
        outs.append(o)
    in_bytes = sum(os.path.getsize(o) for o in outs)
    log(f"# merge inputs: {a.merge_reads} pairs classified against two half databases, "
        f"{in_bytes / 1e6:.0f} MB of kraken output; setup {time.time() - t0:.1f}s")

    def merge(exe, tag, env=None):
        d = os.path.join(tmp, tag)
        os.makedirs(d, exist_ok=True)
        cmd = [exe, "-T", str(threads), "-t", taxo, "-o", f"{d}/out", "-c", f"{d}/hdr", "-r", f"{d}/rep",
               "-i", str(a.merge_confidence), *outs]
        t = time.time()
        r = sh(cmd, env=env, text=True)
        secs = time.time() - t
        return secs, r.stderr, {k: open(f"{d}/{k}", "rb").read() for k in ("out", "hdr", "rep")}

    ref = os.path.join(ROOT, "oracle", "_ref", "merge")
    port = os.path.join(ROOT, "oracle", "bin", "k2cpu_merge")
    cpu_exe, kind = (ref, "reference") if os.path.exists(ref) else (port, "port")
    gpu_runs = []
    for i in range(a.steps + a.warmup):
        secs, err, got = merge(bins["k2g_merge"], "gpu", env=dict(os.environ, K2G_TIMING="1"))
        if i >= a.warmup:
            gpu_runs.append((secs, err))
    secs_gpu = min(s for s, _ in gpu_runs)
    timing = [e for s, e in gpu_runs if s == secs_gpu][0].strip().splitlines()[-1]
    m = re.search(r"kernels ([\d.]+)s", timing)
    kernel_s = float(m.group(1)) if m else None
    log(f"# k2g_merge: best {secs_gpu:.3f}s of {[round(x, 3) for x, _ in gpu_runs]} ({timing})")
    cpu_runs = [merge(cpu_exe, "cpu") for _ in range(max(1, a.steps))]
    secs_cpu, exp = min(r[0] for r in cpu_runs), cpu_runs[0][2]
    log(f"# {kind} merge -T {threads}: best {secs_cpu:.3f}s of {[round(r[0], 3) for r in cpu_runs]}")
    equal = got == exp
    log(f"# byte-identical output, headers and report: {equal}")
    n = a.merge_reads
    items = 0
    for o in outs:
        with open(o, "rb") as f:
            for line in f:
                items += line.count(b":")
    line = {"metric": "merged reads/s (multi-database merge of two kraken outputs, file to file)",
            "value": round(n / secs_gpu, 1), "unit": "reads/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
            "ms_per_step": round(secs_gpu * 1e3, 3), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "int32", "data": "synthetic",
            "config": {"workload": f"merge: {n} read pairs 2x150 classified (-P -g 2) against two databases built "
                                   f"from the two halves of a {a.merge_levels} synthetic library "
                                   f"({a.merge_genome_len} bp genomes), -i {a.merge_confidence}, -T {threads}",
                       "input_bytes": in_bytes, "hit_list_items": items},
            "kernel_ms": round(kernel_s * 1e3, 3) if kernel_s is not None else None,
            "kernel_reads_per_s": round(n / kernel_s, 1) if kernel_s else None,
            "cpu_baseline": {"value": round(n / secs_cpu, 1), "unit": "reads/s", "cores": threads, "kind": kind,
                             "sample": f"the whole workload ({n} reads), best of {len(cpu_runs)} runs, "
                                       f"{secs_cpu:.2f} s wall clock"},
            "vs_cpu": round(secs_cpu / secs_gpu, 2), "bit_exact_vs_cpu": bool(equal)}
    print(json.dumps(line), flush=True)
    if not a.keep:
        shutil.rmtree(tmp, ignore_errors=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c3", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-verify", action="store_true")
    ap.add_argument("--keep", default=None)
    ap.add_argument("--chunk-space", type=float, default=1.2e10,
                    help="hitlist runs reserved per persistent launch (8 bytes each)")
    ap.add_argument("--cpu-sample", type=int, default=None, help="fragments timed on the CPU")
    ap.add_argument("--gather-probes", type=float, default=2 ** 30)
    ap.add_argument("--plain-steps", type=int, default=1, help="timed steps of the plain per-block pass")
    ap.add_argument("--skip-plain", action="store_true",
                    help="time only the persistent and one-launch passes (profiling runs; no headline ratio)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device copy (calibrates HBM byte counters) and exit")
    ap.add_argument("--full-verify", action="store_true", help="digest every plain block, not an even spread")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving measurement")
    ap.add_argument("--no-pipeline", action="store_true", help="skip the graph-pipeline leg")
    ap.add_argument("--tune-sample", type=int, default=500_000, help="fragments the autotuner times on")
    ap.add_argument("--host-batches", type=int, default=8, help="front-end batches of the host-pipeline leg")
    ap.add_argument("--serve-requests", type=int, default=300, help="requests in the serving set")
    ap.add_argument("--serve-oracle", type=int, default=24, help="served requests checked against the oracle")
    ap.add_argument("--serve-oracle-max", type=int, default=50000,
                    help="largest request (fragments) checked against the oracle")
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--serve-hll", type=int, default=1, choices=(1, 2),
                    help="-K data of served requests: 1 per-hit records, 2 per-(block, taxon) sets from the GPU")
    ap.add_argument("--serve-budget-gb", type=float, default=64.0)
    ap.add_argument("--serve-pool", type=int, default=300_000, help="fragments per serving pool")
    ap.add_argument("--scale", type=float, default=1.0,
                    help="scale genomes and reads down for a quick functional run (not a bench number)")
    ap.add_argument("--merge", action="store_true",
                    help="measure the multi-database merge (k2g_merge) instead of classify")
    ap.add_argument("--merge-reads", type=int, default=2_000_000)
    ap.add_argument("--merge-levels", default="8,8,8")
    ap.add_argument("--merge-genome-len", type=int, default=100000)
    ap.add_argument("--merge-confidence", type=float, default=0.0)
    a = ap.parse_args()
    if a.gpus != 1:
        log(f"bench.py measures one GPU: --gpus {a.gpus} is not supported (reads would shard across GPUs "
            "as one process per GPU with a replicated table; this build runs on one GPU)")
        sys.exit(2)
    if a.copy_calibration:
        copy_calibration()
        return
    if a.merge:
        run_merge_bench(a)
        return
    import kraken2_amd as k2

    work = WORKLOADS[a.workload]
    t_start = time.time()
    if work["kind"] == "host":
        tmp, hip, d_seq, d_off, d_paired, off, paired, host = run_host_workload(a, work, k2)
    elif work["kind"] == "protein":
        tmp, hip, d_seq, d_off, d_paired, off, paired, host = run_protein_workload(a, work, k2)
    else:
        tmp, hip, d_seq, d_off, d_paired, off, paired, host = run_gpu_workload(a, work, k2)
    n = len(paired)
    total_bases = int(off[-1])
    db = k2.Db.from_dir(tmp)
    k = int(db.info.k)
    log(f"# DB resident: {db.info.capacity} cells, {db.info.device_bytes / 1e9:.2f} GB; "
        f"{n} fragments, {total_bases / 1e9:.2f} Gbp; setup {time.time() - t_start:.1f}s")

    chunks = split_chunks(run_space(off, paired, k), int(a.chunk_space))
    blocks = reference_blocks(off, paired)
    fbatches = frontend_batches(blocks, off, paired)
    runner = Runner(k2, db, d_seq, d_off, d_paired, work)

    passes = {
        "plain": (fbatches, 0),           # one plain launch per front-end batch (64 MiB of bases)
        "plain_blocks": (blocks, 0),      # one plain launch per reference input block (3 MiB / 10 000 pairs)
        "plain_one_launch": (chunks, 0),  # the plain kernel over the whole read set at once
        "batched": (chunks, 1),           # persistent, longest first
    }
    if a.skip_plain:
        del passes["plain"], passes["plain_blocks"]

    def run_pass(name):
        ranges, mode = passes[name]
        return sum(runner.run(s, e, mode) for s, e in ranges)

    # The persistent pass is the headline: W warmup steps, then K timed steps.
    # The plain per-block pass (thousands of launches at full size) is timed
    # over min(K, plain_steps) steps after a warmup over its first blocks.
    timing = {}
    for name in passes:
        steps = a.steps if name != "plain_blocks" else max(1, min(a.steps, a.plain_steps))
        for _ in range(a.warmup):
            if name == "plain_blocks":
                for s0, e0 in blocks[:16]:
                    runner.run(s0, e0, 0)
            else:
                run_pass(name)
        t = time.perf_counter()
        kms = []
        for _ in range(steps):
            kms.append(run_pass(name))
            log(f"# {name}: step {len(kms)}/{steps} kernel {kms[-1]:.1f} ms")
        wall = (time.perf_counter() - t) / steps
        timing[name] = (wall, float(np.mean(kms)), steps)
        log(f"# {name}: {wall * 1e3:.1f} ms/step ({len(passes[name][0])} launches), kernel {np.mean(kms):.1f} ms")
    batched_s, batched_kms, _ = timing["batched"]
    plain_s, plain_kms, plain_steps = timing.get("plain", (float("nan"), float("nan"), 0))
    pblk_s, pblk_kms, pblk_steps = timing.get("plain_blocks", (float("nan"), float("nan"), 0))

    # ---- verification and measurement passes (outside the timed region)
    digests = {}
    stats_pass_differs = [0]
    lookups = lines = n_runs_total = 0
    space_redone = dense_redone = 0
    checked = {}
    for name in passes:
        if a.no_verify and name != "batched":
            continue
        ranges, mode = passes[name]
        if name == "plain_blocks" and not a.full_verify and len(ranges) > 256:
            ranges = ranges[::max(1, len(ranges) // 256)]  # an even spread of the blocks
        dig = np.zeros(n, np.uint64)
        mask = np.zeros(n, bool)
        for s, e in ranges:
            mask[s:e] = True
            runner.run(s, e, mode)
            dig[s:e] = runner.batch.digests(e - s)
            if name == "batched":
                # probe statistics come from the kernels with per-fragment
                # counters (the wave scan for every read): their results must
                # equal the timed pass's
                runner.run(s, e, mode, stats=True)
                if not np.array_equal(runner.batch.digests(e - s), dig[s:e]):
                    stats_pass_differs[0] += 1
                c = runner.batch.fetch_counts(e - s, probe_stats=True)
                lookups += int(c["lookups"].sum())
                lines += int(c["probe_lines"].sum())
                n_runs_total += int(c["n_runs"].astype(np.int64).sum())
                sr, dr = runner.batch.redo_counts()
                space_redone += sr
                dense_redone += dr
        digests[name] = dig
        checked[name] = mask
        log(f"# digests of {name} done ({int(mask.sum())} fragments)")
    if a.no_verify:
        modes_agree, plain_checked = True, 0
    else:
        pm = checked.get("plain_blocks", np.zeros(n, bool))
        plain_checked = int(pm.sum())
        modes_agree = stats_pass_differs[0] == 0 and \
            (np.array_equal(digests["plain_blocks"][pm], digests["batched"][pm])
             if "plain_blocks" in digests else True) and \
            np.array_equal(digests["plain_one_launch"], digests["batched"]) and \
            (np.array_equal(digests["plain"], digests["batched"]) if "plain" in digests else True)

    graph = None if a.no_pipeline else run_graph_pipeline(
        k2, db, runner, {"chunks": chunks, "frontend_batches": fbatches}, off, paired, hip, digests.get("batched"), a,
        work)

    # CPU leg, part 1: the oracle checks an even sample of fragments (the
    # reference binary is timed further down); skipped with --no-cpu
    rng = np.random.default_rng(7)
    sample = np.sort(rng.choice(n, size=min(n, 4000), replace=False))
    if a.no_cpu:
        sample = sample[:0]
    if not len(sample):
        s_seq, so = b"", np.zeros(1, np.uint64)
    elif work["kind"] == "host":
        seq_all = host["seq"]
        parts, so = [], np.zeros(2 * len(sample) + 1, np.uint64)
        pos = 0
        for j, i in enumerate(sample.tolist()):
            a0, a2 = int(off[2 * i]), int(off[2 * i + 2])
            parts.append(seq_all[a0:a2])
            so[2 * j] = pos
            so[2 * j + 1] = pos + int(off[2 * i + 1]) - a0
            pos += a2 - a0
        so[-1] = pos
        s_seq = b"".join(parts)
    else:
        s_seq, so, _ = host["readset"].host_fragments(sample)
    mism, s_dig = oracle_check(k2, db, tmp, s_seq, so, paired[sample], work) if len(sample) else (0, np.zeros(0))
    sample_digest_agree = bool(np.array_equal(s_dig, digests["batched"][sample]))
    log(f"# verify: modes agree ({plain_checked} per-block / {n} front-end batch and one-launch fragments): "
        f"{modes_agree}; oracle sample mismatches: "
        f"{mism}/{len(sample)}; sample digests equal the bench run's: {sample_digest_agree}")

    # roofline of the persistent classify kernel: algorithmic bytes per launch
    #   = bases + per fragment (16 B offsets + 1 B pair flag + 4 B order + 16 B
    #     run offset/space + 24 B scalar outputs) + 128 B per distinct table
    #     line a probe touches + 8 B per hitlist run
    alg_bytes = total_bases + 61 * n + 128 * lines + 8 * n_runs_total
    kernel_s = batched_kms / 1e3
    achieved = alg_bytes / kernel_s / 1e9
    ceil_rate, ceil_ms = db.gather_ceiling(int(a.gather_probes))
    probe_rate = lines / kernel_s
    log(f"# gather ceiling {ceil_rate / 1e9:.2f} G lines/s ({ceil_ms:.1f} ms); classify "
        f"{probe_rate / 1e9:.2f} G lines/s")

    cpu = end_to_end = None
    if not a.no_cpu:
        if work["kind"] == "host":
            files = [[host["reads_fa"]]]
        else:
            m = min(n, a.cpu_sample or 1_000_000)
            w = min(1000, m)
            starts = np.linspace(0, n - w, max(1, m // w)).astype(np.int64)
            singles, pairs = [], []
            for s0 in starts.tolist():
                sq, so2, pf = host_range(hip, d_seq, off, paired, s0, s0 + w)
                for j in range(w):
                    m1, m2 = fragment_text(sq, so2, j)
                    (pairs if pf[j] else singles).append((m1, m2 if pf[j] else None, f"f{s0 + j}"))
            files = []
            if singles:
                write_sample_fasta(os.path.join(tmp, "cpu_single.fa"), None, singles)
                files.append([os.path.join(tmp, "cpu_single.fa")])
            if pairs:
                write_sample_fasta(os.path.join(tmp, "cpu_1.fa"), os.path.join(tmp, "cpu_2.fa"), pairs)
                files.append(["-P", os.path.join(tmp, "cpu_1.fa"), os.path.join(tmp, "cpu_2.fa")])
            log(f"# cpu sample: {len(singles)} unpaired + {len(pairs)} pairs in {len(starts)} windows")
        cpu = cpu_baseline(work, tmp, files, os.cpu_count() or 1)
        log(f"# cpu baseline: {cpu}")
        # the same files through the GPU front-end over the same timed interval
        # (parse, classify, format, write); the table is loaded before it
        gpu_cli = os.path.join(ROOT, "kraken2_amd", "bin", "k2g_classify")
        if os.path.exists(gpu_cli):
            end_to_end = cpu_baseline(work, tmp, files, os.cpu_count() or 1, exe=gpu_cli)
            log(f"# gpu end to end: {end_to_end}")
            if end_to_end and cpu:
                end_to_end["output_equal_" + cpu["kind"]] = outputs_equal(work, tmp, files, os.cpu_count() or 1,
                                                                          gpu_cli)
                log(f"# gpu end to end: kraken output equal to the {cpu['kind']} binary's: "
                    f"{end_to_end['output_equal_' + cpu['kind']]}")

    serving = None
    if work["kind"] == "gpu" and host.get("pools"):
        runner.close()
        host["readset"].free()
        serving = run_serving(k2, db, tmp, host["pools"], a, work)

    value = n / batched_s
    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "reads/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(batched_s * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64/u32 integer",
        "data": "synthetic (seeded generators; table built on the GPU)" if work["kind"] != "host"
                else "synthetic (seeded generator; table built by the GPU build front-ends, k2g_estimate_capacity and k2g_build_db)",
        "config": {"workload": work["desc"] + ("" if a.scale == 1.0 else f" (scaled x{a.scale})"), "fragments": n,
                   "bases": total_bases, "table_cells": int(db.info.capacity), "parallelism": "single GPU"},
        "mbp_per_s": round(total_bases / batched_s / 1e6, 1),
        "kernel_ms_per_step": round(batched_kms, 3),
        "persistent_launches_per_step": len(chunks),
        "plain": None if a.skip_plain else {
            "reads_per_s": round(n / plain_s, 1), "ms_per_step": round(plain_s * 1e3, 3),
            "kernel_ms_per_step": round(plain_kms, 3), "launches_per_step": len(fbatches),
            "timed_steps": plain_steps,
            "batches": "the classify front-end's GPU batches: whole reference blocks up to 64 MiB of bases "
                       "or 4096 blocks, one plain launch each"},
        "plain_blocks": None if a.skip_plain else {
            "reads_per_s": round(n / pblk_s, 1), "ms_per_step": round(pblk_s * 1e3, 3),
            "kernel_ms_per_step": round(pblk_kms, 3), "launches_per_step": len(blocks),
            "timed_steps": pblk_steps,
            "batches": "reference input blocks: 3 MiB of unpaired bases or 10 000 pairs, one plain launch each"},
        "plain_one_launch": {"reads_per_s": round(n / timing["plain_one_launch"][0], 1),
                             "ms_per_step": round(timing["plain_one_launch"][0] * 1e3, 3),
                             "kernel_ms_per_step": round(timing["plain_one_launch"][1], 3),
                             "launches_per_step": len(chunks)},
        "batched_vs_plain": None if a.skip_plain else round(plain_s / batched_s, 3),
        "batched_vs_plain_blocks": None if a.skip_plain else round(pblk_s / batched_s, 3),
        "batched_vs_plain_one_launch": round(timing["plain_one_launch"][0] / batched_s, 3),
        "batched_graph": graph,
        "bit_exact": bool(modes_agree and mism == 0 and sample_digest_agree and
                          (graph is None or graph["digests_equal_batched"] is not False)),
        "verify": {"modes_equal": bool(modes_agree), "plain_fragments_compared": plain_checked,
                   "one_launch_fragments_compared": n, "oracle_sample": int(len(sample)),
                   "oracle_mismatches": int(mism), "sample_digests_match_bench_run": sample_digest_agree},
        "lookups": lookups,
        "probe_lines": lines,
        "hitlist_runs": n_runs_total,
        "redone": {"run_space": space_redone, "dense_table": dense_redone},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": profiled_traffic(a.workload) if a.scale == 1.0 else None,
                     "algorithmic_bytes": int(alg_bytes)},
        "gather": {"classify_lines_per_s": round(probe_rate, 1), "ceiling_lines_per_s": round(ceil_rate, 1),
                   "frac_of_ceiling": round(probe_rate / ceil_rate, 4) if ceil_rate else None,
                   "ceiling": "independent 16-byte loads at random 128-byte lines of the resident table"},
        "cpu_baseline": cpu,
    }
    if serving:
        mid = serving["legs"][3]  # Poisson arrivals at the one-slot rate
        line["latency_ms"] = {"p50": mid["p50_ms"], "p99": mid["p99_ms"],
                              "offered_req_per_s": mid["offered_req_per_s"], "slots": mid["slots"]}
        line["serving"] = serving
    if cpu:
        line["vs_cpu"] = round(value / cpu["value"], 2)
    if end_to_end:  # PCIe, parsing and formatting included: never the headline value
        line["gpu_end_to_end"] = end_to_end
        line["end_to_end_vs_cpu"] = round(end_to_end["value"] / cpu["value"], 2)
    if work["kind"] in ("gpu", "protein"):
        line["config"].update({"table_estimate": host["estimate"], "gpu_build_s": round(host["build_s"], 2)})
        if host.get("build_check"):
            line["build_verify"] = host["build_check"]
    print(json.dumps(line), flush=True)
    if not serving:
        runner.close()
    db.close()
    if not a.keep:
        shutil.rmtree(tmp, ignore_errors=True)


if __name__ == "__main__":
    main()
