"""Host-side (numpy) verification of a built table, the checker of
k2g_build_verify on small databases.

The pairs a build inserts are enumerated with the oracle's scanner over the
deterministic build's windows (src/build_db.h:216-326: blocks of
block_size, each extended by k-1, cut into subblocks of subblock_size
extended by k-1; a window that does not start the sequence reports from
its (k-2)-th base, src/mmscanner.cc:258-259).  Within a window a
non-ambiguous minimizer is inserted when it differs from the previous
non-ambiguous one.  A table is then checked by simulating the linear probe
(src/compact_hash.h:354-374) of every distinct minimizer on its cells.
"""
import ctypes as C

import numpy as np

M1 = np.uint64(0xFF51AFD7ED558CCD)
M2 = np.uint64(0xC4CEB9FE1A85EC53)


def fmix64(x):
    x = x.astype(np.uint64, copy=True)
    with np.errstate(over="ignore"):
        x ^= x >> np.uint64(33)
        x *= M1
        x ^= x >> np.uint64(33)
        x *= M2
        x ^= x >> np.uint64(33)
    return x


def read_table(path):
    raw = np.fromfile(path, dtype=np.uint8)
    cap, size, kb, vb = (int(v) for v in raw[:32].view(np.uint64))
    cells = raw[32:32 + 4 * cap].view(np.uint32)
    return cells, kb, vb, size


def inserted_pairs(orc, genomes, taxa, k, l, mask, toggle, block, subblock, dna=True):
    """(minimizers, taxa) of every insertion the build makes, in window order.
    Protein sequences get the '*' the build appends (src/build_db.h:113-114)."""
    keys, tax = [], []
    for g, t in zip(genomes, taxa):
        raw = g.encode() if isinstance(g, str) else g
        if not dna and not raw.endswith(b"*"):
            raw += b"*"
        L = len(raw)
        cap = 2 * subblock + 2 * k + 16
        amb = np.zeros(cap, np.uint8)
        mins = np.zeros(cap, np.uint64)
        for j in range(0, L, block):
            bfin = min(j + block + k - 1, L)
            for i in range(j, bfin, subblock):
                fin = min(i + subblock + k - 1, bfin)
                n = orc.lib.k2cpu_scan(k, l, mask, toggle, int(dna), i, fin, raw, L, amb.ctypes.data,
                                       mins.ctypes.data, cap)
                assert n >= 0
                na = mins[:n][amb[:n] == 0]
                if not len(na):
                    continue
                keep = np.ones(len(na), bool)
                keep[1:] = na[1:] != na[:-1]
                ins = na[keep]
                keys.append(ins)
                tax.append(np.full(len(ins), t, np.uint32))
    return np.concatenate(keys), np.concatenate(tax)


def lca_vec(parent, a, b):
    """Elementwise LCA of internal ids (parent < child; 0 is no taxon)."""
    a = a.astype(np.int64, copy=True)
    b = b.astype(np.int64, copy=True)
    za, zb = a == 0, b == 0
    while True:
        ne = a != b
        if not ne.any():
            break
        up_a = ne & (a > b)
        up_b = ne & (b > a)
        a[up_a] = parent[a[up_a]]
        b[up_b] = parent[b[up_b]]
    out = a
    out[za] = b[za]
    out[zb] = a[zb]
    return out


def key_lcas(keys, tax, parent):
    """Distinct keys and the LCA of each key's taxa."""
    pairs = np.unique(np.stack([keys, tax.astype(np.uint64)], axis=1), axis=0)
    uk, start = np.unique(pairs[:, 0], return_index=True)
    end = np.append(start[1:], len(pairs))
    acc = pairs[start, 1].astype(np.int64)
    for j in range(1, int((end - start).max())):
        has = start + j < end
        acc[has] = lca_vec(parent, acc[has], pairs[start[has] + j, 1].astype(np.int64))
    return uk, acc


def probe(cells, kb, vb, keys):
    """(value, terminal cell or capacity) of each key's linear probe."""
    cap = len(cells)
    hc = fmix64(keys)
    ck = (hc >> np.uint64(64 - kb)).astype(np.uint32)
    idx = (hc % np.uint64(cap)).astype(np.int64)
    vmask = np.uint32((1 << vb) - 1)
    val = np.zeros(len(keys), np.uint32)
    where = np.full(len(keys), cap, np.int64)
    live = np.arange(len(keys))
    left = cap
    while len(live) and left:
        c = cells[idx[live]]
        empty = (c & vmask) == 0
        hit = ~empty & ((c >> np.uint32(vb)) == ck[live])
        val[live[hit]] = c[hit] & vmask
        where[live[hit]] = idx[live[hit]]
        live = live[~(empty | hit)]
        idx[live] = (idx[live] + 1) % cap
        left -= 1
    return val, where


def check_table(cells, kb, vb, size, parent, keys, own):
    """The counters k2g_build_verify reports, for one table."""
    val, where = probe(cells, kb, vb, keys)
    found = val != 0
    out = {"keys": len(keys), "missing": int((~found).sum())}
    # value an ancestor-or-self of the key's own LCA <=> of each of its taxa
    out["not_covering"] = int((lca_vec(parent, val[found], own[found]) != val[found]).sum())
    out["differ"] = int((val[found] != own[found]).sum())
    w = where[found]
    ow = own[found]
    order = np.argsort(w, kind="stable")
    w, ow = w[order], ow[order]
    uc, start, cnt = np.unique(w, return_index=True, return_counts=True)
    acc = ow[start].astype(np.int64)
    for j in range(1, int(cnt.max()) if len(cnt) else 1):
        has = cnt > j
        acc[has] = lca_vec(parent, acc[has], ow[start[has] + j])
    vmask = (1 << vb) - 1
    out["cells"] = len(uc)
    out["shared_cells"] = int((cnt > 1).sum())
    out["shared_keys"] = int(cnt[cnt > 1].sum())
    out["unexplained"] = int(((cells[uc] & vmask).astype(np.int64) != acc).sum())
    out["orphan_cells"] = size - len(uc)
    out["values"] = val
    out["shared_key_mask"] = np.isin(where, uc[cnt > 1])
    return out
