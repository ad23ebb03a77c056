"""CPU-side checks of the boundary: the library loads and exports exactly the
entry points include/k2g.h declares (no GPU call is made)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "k2g.h")).read()
    return sorted(set(re.findall(r"\b(k2g_[a-z_]+)\s*\(", text)))


def test_library_exports_header_symbols():
    import kraken2_amd as k2
    lib = ctypes.CDLL(k2.LIB_PATH)
    names = declared_functions()
    assert names, "no declarations parsed"
    for n in names:
        assert hasattr(lib, n), f"libk2g.so does not export {n}"
    assert sorted(k2.EXPORTED_SYMBOLS) == names


def test_results_checksum_is_host_only_and_sensitive():
    """k2g_results_checksum runs on the host: equal result sets hash equal,
    any changed element (a call, a run count, an HLL record) changes it."""
    import numpy as np
    import kraken2_amd._lib as L
    lib = L.load_library()
    rng = np.random.default_rng(3)
    n = 50
    arrs = dict(call=rng.integers(0, 9, n).astype(np.uint32), groups=rng.integers(0, 9, n).astype(np.uint32),
                total=rng.integers(0, 200, n).astype(np.uint32))
    run_off = np.concatenate([[0], np.cumsum(rng.integers(0, 4, n))]).astype(np.uint64)
    runs = rng.integers(0, 9, 2 * int(run_off[-1])).astype(np.uint32)
    hll_off = run_off.copy()
    hll = rng.integers(0, 2**32, 2 * int(hll_off[-1]), dtype=np.uint64).astype(np.uint32)
    kc = rng.integers(0, 9, 12).astype(np.uint64)

    def ptr(a):
        t = ctypes.c_uint64 if a.dtype == np.uint64 else ctypes.c_uint32
        return a.ctypes.data_as(ctypes.POINTER(t))

    def checksum():
        r = L.Results()
        r.n = n
        r.call, r.hit_groups, r.total_kmers = (ptr(arrs[k]) for k in ("call", "groups", "total"))
        r.run_off, r.runs = ptr(run_off), ptr(runs)
        r.hll_off, r.hll = ptr(hll_off), ptr(hll)
        r.kmer_counts, r.node_count = ptr(kc), 12
        h = ctypes.c_uint64()
        assert lib.k2g_results_checksum(ctypes.byref(r), ctypes.byref(h)) == 0
        return h.value

    base = checksum()
    assert checksum() == base
    for arr, i in ((arrs["call"], 7), (runs, 3), (hll, 1), (kc, 11)):
        arr[i] ^= 1
        assert checksum() != base
        arr[i] ^= 1
    assert checksum() == base

    # an empty array hashes the same whether its data pointer is null or not
    # (a server result of a request without HLL records has no record data)
    def empty_hll(null_data):
        r = L.Results()
        r.n = 1
        one = np.zeros(1, np.uint32)
        r.call, r.hit_groups, r.total_kmers = ptr(one), ptr(one), ptr(one)
        ro = np.array([0, 1], np.uint64)
        rr = np.array([0, 5], np.uint32)
        ho = np.zeros(2, np.uint64)
        hd = np.zeros(2, np.uint32)
        r.run_off, r.runs = ptr(ro), ptr(rr)
        r.hll_off = ptr(ho)
        if not null_data:
            r.hll = ptr(hd)
        h = ctypes.c_uint64()
        assert lib.k2g_results_checksum(ctypes.byref(r), ctypes.byref(h)) == 0
        return h.value

    assert empty_hll(True) == empty_hll(False)
