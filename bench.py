#!/usr/bin/env python3
"""Benchmark: reads/s of the batched persistent GPU classify against the plain
one-launch-per-batch GPU classify on the same device-resident reads, both
checked bit-exact against the CPU oracle, with the reference CPU classify
timed on this host beside them.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c1]

A step is one classify pass over the whole read set already resident in
HBM.  Prints one JSON line (rank 0, N=1).
"""
import argparse
import ctypes as C
import json
import os
import re
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # single-end 150 bp reads against a 64-genome, 86-node DB (k=35, l=31)
    "c1": dict(desc="plain classify, 64 genomes x 200 kb, 100k x 150 bp single-end, -T 0 -g 2",
               levels="4,4,4", genome_len=200000, n_reads=100000, read_len=150, seed=1,
               confidence=0.0, min_hit_groups=2, paired=False),
}


class Hip:
    """Minimal HIP runtime access for device-resident inputs (plumbing only)."""

    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so")
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.lib.hipFree.argtypes = [C.c_void_p]
        self.lib.hipDeviceSynchronize.argtypes = []

    def upload(self, arr: np.ndarray) -> C.c_void_p:
        p = C.c_void_p()
        n = max(1, arr.nbytes)
        if self.lib.hipMalloc(C.byref(p), n) != 0:
            raise RuntimeError("hipMalloc failed")
        if arr.nbytes and self.lib.hipMemcpy(p, arr.ctypes.data, arr.nbytes, 1) != 0:
            raise RuntimeError("hipMemcpy failed")
        return p

    def free(self, p):
        self.lib.hipFree(p)


def prepare(work: dict, d: str):
    from helpers import K2SYNTH, make_synthetic_db, run
    make_synthetic_db(d, seed=work["seed"], levels=work["levels"], genome_len=work["genome_len"])
    reads = os.path.join(d, "reads.fa")
    run([K2SYNTH, "reads", "--lib", os.path.join(d, "library.fna"), "--out", reads, "--seed",
         str(work["seed"]), "--n", str(work["n_reads"]), "--len", str(work["read_len"]), "--sub", "0.01",
         "--random-frac", "0.1"])
    return reads


def load_reads(path):
    seqs = []
    with open(path, "rb") as f:
        for line in f:
            if not line.startswith(b">"):
                seqs.append(line.rstrip(b"\n"))
    n = len(seqs)
    lens = np.array([len(s) for s in seqs], np.uint64)
    ends = np.cumsum(lens, dtype=np.uint64)
    off = np.zeros(2 * n + 1, np.uint64)
    off[0:2 * n:2] = np.concatenate([np.zeros(1, np.uint64), ends[:-1]])
    off[1:2 * n:2] = ends
    off[2 * n] = ends[-1]
    return b"".join(seqs), off, seqs


def cpu_baseline(work, d, reads, threads):
    """The reference classify (compiled from its sources into oracle/_ref) when
    present, else the oracle restatement, on all host threads."""
    ref = os.path.join(ROOT, "oracle", "_ref", "classify")
    port = os.path.join(ROOT, "oracle", "bin", "k2cpu_classify")
    exe, kind = (ref, "reference") if os.path.exists(ref) else (port, "port")
    cmd = [exe, "-H", f"{d}/hash.k2d", "-t", f"{d}/taxo.k2d", "-o", f"{d}/opts.k2d", "-p", str(threads),
           "-T", str(work["confidence"]), "-g", str(work["min_hit_groups"]), "-O", "/dev/null", reads]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        return None
    m = re.search(r"(\d+) sequences \(([\d.]+) Mbp\) processed in ([\d.]+)s", r.stderr)
    n, secs = int(m.group(1)), float(m.group(3))
    return {"value": n / secs, "unit": "reads/s", "cores": threads, "kind": kind,
            "sample": f"all {n} reads of the workload, timed interval of the binary itself "
                      f"(after DB load, through parse, classify and output), {secs:.3f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c1", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--keep", default=None, help="directory for the generated workload")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {a.gpus} is not supported (this build shards "
              "reads across GPUs only by running one process per GPU)", file=sys.stderr)
        sys.exit(2)
    work = WORKLOADS[a.workload]
    import kraken2_amd as k2
    from helpers import Oracle

    tmp = a.keep or tempfile.mkdtemp(prefix="k2g_bench_")
    os.makedirs(tmp, exist_ok=True)
    t0 = time.time()
    reads = prepare(work, tmp)
    seq, off, seqs = load_reads(reads)
    n = len(seqs)
    print(f"# workload {a.workload}: {n} reads ready in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)

    db = k2.Db.from_dir(tmp)
    lib = k2.load_library()
    hip = Hip()
    d_seq = hip.upload(np.frombuffer(seq, np.uint8))
    d_off = hip.upload(off)
    batch_p = k2.Batch(db)
    batch_b = k2.Batch(db)
    P = k2._lib.ClassifyParams
    params_b = P(work["confidence"], work["min_hit_groups"], 0, int(work["paired"]), 1, 0, 0, 0)
    params_p = P(work["confidence"], work["min_hit_groups"], 0, int(work["paired"]), 0, 0, 0, 0)

    # the reference's blocks: 3 MiB of bases per block for single reads
    blocks = []
    start, bases = 0, 0
    lens = np.diff(off)[0::2]
    for i in range(n):
        bases += int(lens[i])
        if bases >= 3 * 1024 * 1024 or i == n - 1:
            blocks.append((start, i + 1))
            start, bases = i + 1, 0
    off_ptr = d_off.value

    def step_batched():
        rc = lib.k2g_classify_device(batch_b._h, C.byref(params_b), d_seq, d_off, None, n, None)
        ms = C.c_double()
        rc |= lib.k2g_batch_sync(batch_b._h, C.byref(ms))
        if rc:
            raise k2.K2GError(rc, lib.k2g_last_error().decode())
        return ms.value

    def step_plain():
        tot = 0.0
        for (s, e) in blocks:
            rc = lib.k2g_classify_device(batch_p._h, C.byref(params_p), d_seq, C.c_void_p(off_ptr + 16 * s),
                                         None, e - s, None)
            ms = C.c_double()
            rc |= lib.k2g_batch_sync(batch_p._h, C.byref(ms))
            if rc:
                raise k2.K2GError(rc, lib.k2g_last_error().decode())
            tot += ms.value
        return tot

    for _ in range(a.warmup):
        step_plain()
        step_batched()
    hip.lib.hipDeviceSynchronize()
    t = time.perf_counter()
    kms_p = [step_plain() for _ in range(a.steps)]
    hip.lib.hipDeviceSynchronize()
    plain_s = (time.perf_counter() - t) / a.steps
    t = time.perf_counter()
    kms_b = [step_batched() for _ in range(a.steps)]
    hip.lib.hipDeviceSynchronize()
    batched_s = (time.perf_counter() - t) / a.steps
    print(f"# plain {plain_s * 1e3:.2f} ms/step, batched {batched_s * 1e3:.2f} ms/step", file=sys.stderr,
          flush=True)

    # parity of the last batched pass against the oracle, every read
    res = k2._lib.Results()
    rc = lib.k2g_batch_fetch(batch_b._h, C.byref(res))
    if rc:
        raise k2.K2GError(rc, lib.k2g_last_error().decode())
    out = k2._lib._copy_results(res, False, False)
    orc = Oracle()
    h = orc.open(tmp)
    ref = orc.classify(h, seq, off, confidence=work["confidence"], min_hit_groups=work["min_hit_groups"],
                       lines=True)
    mismatches = 0
    for i in range(n):
        runs = out.runs[out.run_off[i]:out.run_off[i + 1]]
        if int(out.call[i]) != int(ref["call"][i]) or k2.format_hitlist(runs, db.external_ids) != ref["text"][i]:
            mismatches += 1

    # roofline of the persistent kernel: algorithmic bytes per launch =
    # input bases + 8 B offsets per read + 128 B per distinct probed line +
    # outputs (call, groups, total, n_runs, status, n_recs = 24 B per read,
    # 8 B per hitlist run)
    lines = int(ref["lines"].sum())
    lookups = int(ref["lookups"].sum())
    n_runs = int(out.run_off[-1])
    alg_bytes = len(seq) + 16 * n + 128 * lines + 24 * n + 8 * n_runs
    kernel_s = float(np.mean(kms_b)) / 1e3
    achieved = alg_bytes / kernel_s / 1e9

    cpu = None if a.no_cpu else cpu_baseline(work, tmp, reads, os.cpu_count() or 1)
    value = n / batched_s
    line = {
        "metric": "reads/sec batched vs plain GPU classify (bit-exact) + p50/p99 request latency",
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
        "data": "synthetic (seeded generator; DB built by the deterministic build)",
        "config": {"workload": a.workload + ": " + work["desc"], "reads": n, "bases": len(seq),
                   "table_cells": int(db.info.capacity), "parallelism": "single GPU"},
        "plain": {"reads_per_s": round(n / plain_s, 1), "ms_per_step": round(plain_s * 1e3, 3),
                  "kernel_ms_per_step": round(float(np.mean(kms_p)), 3), "launches_per_step": len(blocks)},
        "batched_vs_plain": round(plain_s / batched_s, 3),
        "kernel_ms_per_step": round(float(np.mean(kms_b)), 3),
        "bit_exact": mismatches == 0,
        "mismatches": mismatches,
        "lookups": lookups,
        "probe_lines": lines,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    hip.free(d_seq)
    hip.free(d_off)
    batch_b.close()
    batch_p.close()
    db.close()


if __name__ == "__main__":
    main()
