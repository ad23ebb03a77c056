"""Python side of the SIMT emulation: runs the classify kernels' device code
on the CPU for a DB directory and a fragment batch (tests only)."""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def load():
    lib = C.CDLL(os.path.join(HERE, "libk2g_emu.so"))
    vp = C.c_void_p
    u32, u64, i32, dbl = C.c_uint32, C.c_uint64, C.c_int32, C.c_double
    lib.emu_classify.argtypes = [vp, u64, u32, u32, vp, u32, u32, u32, u32, u64, u64, u64, u32, dbl, i32, u32, u32,
                                 vp, vp, vp, u32, u64, vp, C.c_int, C.c_uint, vp, vp, vp, vp] + [vp] * 9
    lib.emu_classify_split.argtypes = [vp, u64, u32, u32, vp, u32, u32, u32, u64, u64, u64, u32, dbl, i32, u32,
                                       vp, vp, vp, u64, vp, C.c_uint, vp, vp, vp, u32] + [vp] * 9
    lib.emu_collectives.restype = C.c_ulonglong
    return lib


def read_db(d):
    raw = np.fromfile(os.path.join(d, "hash.k2d"), dtype=np.uint8)
    cap, size, kb, vb = np.frombuffer(raw[:32].tobytes(), np.uint64).tolist()
    cells = np.zeros((cap + 4 + 63) // 64 * 64, np.uint32)
    cells[:cap] = np.frombuffer(raw[32:].tobytes(), np.uint32)
    if cap >= 64:  # padding holds copies of the first cells, as the library loads it
        cells[cap:] = cells[np.arange(len(cells) - cap) % cap]
    tx = open(os.path.join(d, "taxo.k2d"), "rb").read()
    nn = int(np.frombuffer(tx[8:16], np.uint64)[0])
    nodes = np.frombuffer(tx[32:32 + 56 * nn], np.uint64).reshape(nn, 7)
    parent = nodes[:, 0].astype(np.uint32)
    ext = nodes[:, 5].copy()
    o = open(os.path.join(d, "opts.k2d"), "rb").read()
    k, l, mask, tog = np.frombuffer(o[:32], np.uint64).tolist()
    dna = o[32]
    minh = int(np.frombuffer(o[40:48], np.uint64)[0])
    rc = int(np.frombuffer(o[48:52], np.int32)[0]) if len(o) >= 52 else 0
    return dict(cells=cells, capacity=cap, kb=kb, vb=vb, parent=parent, ext=ext, k=k, l=l, mask=mask,
                toggle=tog, dna=dna, min_hash=minh, revcom=rc)


def bounds(db, off, paired_flags):
    n = (len(off) - 1) // 2
    k = db["k"]
    b = np.zeros(n, np.uint64)
    for i in range(n):
        e = 0
        for m in range(2 if paired_flags[i] else 1):
            L = int(off[2 * i + m + 1] - off[2 * i + m])
            if db["dna"]:
                e += max(0, L - k + 1)
            else:
                for fr in range(6):
                    if L < 3:
                        fl = 1
                    else:
                        first = 3 if fr % 3 == 0 else (4 if fr % 3 == 1 else 2)
                        fl = (L - 1 - first) // 3 + 1 if L - 1 >= first else 0
                    e += max(0, fl - k + 1)
                e += 5
        if paired_flags[i]:
            e += 1
        b[i] = e
    return b


def classify(lib, db, seq, off, paired_flags, *, conf=0.0, min_groups=0, quick=False, tab_cap=256, mode=1,
             grid=4, hll=False, kmer_counts=False, cap_pct=None, unit_elements=None):
    """cap_pct None reserves each fragment's exact element count for its runs;
    otherwise min(e, e * cap_pct // 100 + 64) as the library does."""
    n = (len(off) - 1) // 2
    off = np.ascontiguousarray(off, np.uint64)
    pf = np.ascontiguousarray(paired_flags, np.uint8)
    b = bounds(db, off, pf)
    caps = b.copy() if cap_pct is None else np.minimum(b, b * np.uint64(cap_pct) // np.uint64(100) + np.uint64(64))
    run_off = np.zeros(n + 1, np.uint64)
    run_off[1:] = np.cumsum(caps)
    tot = int(run_off[-1])
    order = np.argsort(-b.astype(np.int64), kind="stable").astype(np.uint32)
    units = n_units = None
    if unit_elements:  # windows of unit_elements elements over the longest-first order
        pre = np.concatenate([[0], np.cumsum(np.minimum(b[order], 0xFFFFFFFF).astype(np.uint64))[:-1]])
        w = pre // np.uint64(unit_elements)
        starts = np.flatnonzero(np.concatenate([[True], w[1:] != w[:-1]])) if n else np.zeros(0, np.int64)
        units = np.concatenate([starts, [n]]).astype(np.uint32)
        n_units = np.array([len(starts)], np.uint32)
    outs = {k: np.zeros(n, np.uint32) for k in ("call", "groups", "total", "n_runs", "status", "n_recs")}
    runs = np.zeros(2 * max(tot, 1), np.uint32)
    recs = np.zeros(2 * max(tot, 1), np.uint32) if hll else None
    kc = np.zeros(len(db["parent"]), np.uint64) if kmer_counts else None
    sb = np.frombuffer(seq if len(seq) else b"\0", np.uint8)
    lib.emu_classify(db["cells"].ctypes.data, db["capacity"], db["kb"], db["vb"], db["parent"].ctypes.data,
                     len(db["parent"]), db["k"], db["l"], db["dna"], db["mask"], db["toggle"], db["min_hash"],
                     db["revcom"], conf, min_groups, int(quick), tab_cap, sb.ctypes.data, off.ctypes.data, pf.ctypes.data, 0,
                     n, order.ctypes.data, mode, grid, run_off.ctypes.data, caps.ctypes.data,
                     units.ctypes.data if units is not None else None,
                     n_units.ctypes.data if n_units is not None else None,
                     *[outs[k].ctypes.data for k in ("call", "groups", "total", "n_runs", "status", "n_recs")],
                     runs.ctypes.data, recs.ctypes.data if recs is not None else None,
                     kc.ctypes.data if kc is not None else None)
    outs["run_off"] = run_off
    outs["units"] = units
    outs["runs"] = runs
    outs["recs"] = recs
    outs["kmer_counts"] = kc
    return outs


def classify_split(lib, db, seq, off, paired_flags, *, conf=0.0, min_groups=0, tab_cap=256, grid=4, hll=False,
                   kmer_counts=False, seg_pct=100):
    """The split path (nucleotide, persistent, exact fragment space): long
    fragments classified by segments and stitched.  Returns (outputs, segments)."""
    n = (len(off) - 1) // 2
    off = np.ascontiguousarray(off, np.uint64)
    pf = np.ascontiguousarray(paired_flags, np.uint8)
    b = bounds(db, off, pf)
    caps = b.copy()
    run_off = np.zeros(n + 1, np.uint64)
    run_off[1:] = np.cumsum(caps)
    tot = int(run_off[-1])
    order = np.argsort(-b.astype(np.int64), kind="stable").astype(np.uint32)
    outs = {k: np.zeros(n, np.uint32) for k in ("call", "groups", "total", "n_runs", "status", "n_recs")}
    runs = np.zeros(2 * max(tot, 1), np.uint32)
    recs = np.zeros(2 * max(tot, 1), np.uint32) if hll else None
    kc = np.zeros(len(db["parent"]), np.uint64) if kmer_counts else None
    sb = np.frombuffer(seq if len(seq) else b"\0", np.uint8)
    segs = lib.emu_classify_split(db["cells"].ctypes.data, db["capacity"], db["kb"], db["vb"],
                                  db["parent"].ctypes.data, len(db["parent"]), db["k"], db["l"], db["mask"],
                                  db["toggle"], db["min_hash"], db["revcom"], conf, min_groups, tab_cap,
                                  sb.ctypes.data, off.ctypes.data, pf.ctypes.data, n, order.ctypes.data, grid,
                                  run_off.ctypes.data, caps.ctypes.data, b.ctypes.data, seg_pct,
                                  *[outs[k].ctypes.data for k in ("call", "groups", "total", "n_runs", "status",
                                                                  "n_recs")],
                                  runs.ctypes.data, recs.ctypes.data if recs is not None else None,
                                  kc.ctypes.data if kc is not None else None)
    outs["run_off"] = run_off
    outs["runs"] = runs
    outs["recs"] = recs
    outs["kmer_counts"] = kc
    return outs, segs
