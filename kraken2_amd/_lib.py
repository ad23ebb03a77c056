"""ctypes bindings of libk2g.so (declared in include/k2g.h).

The shared library is built in-tree (kraken2_amd/lib/libk2g.so); nothing here
falls back to a CPU path: if the library or a GPU is missing, the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libk2g.so")
BIN_DIR = os.path.join(PKG_DIR, "bin")

CODE_MATE = 0xFFFFFFFF
CODE_FRAME = 0xFFFFFFFE
CODE_AMBIG = 0xFFFFFFFD

EXPORTED_SYMBOLS = (
    "k2g_db_open",
    "k2g_db_close",
    "k2g_db_get_info",
    "k2g_db_external_ids",
    "k2g_last_error",
    "k2g_get_batch",
    "k2g_get_batch_device",
    "k2g_batch_create",
    "k2g_batch_destroy",
    "k2g_classify",
    "k2g_classify_device",
    "k2g_batch_sync",
    "k2g_batch_fetch",
    "k2g_batch_stream",
    "k2g_batch_fetch_counts",
    "k2g_batch_redo_counts",
    "k2g_batch_split_counts",
    "k2g_batch_digests",
    "k2g_classify_device_noalloc",
    "k2g_db_gather_ceiling",
    "k2g_host_alloc",
    "k2g_host_free",
    "k2g_results_checksum",
    "k2g_report_write",
    "k2g_server_create",
    "k2g_server_submit",
    "k2g_server_wait",
    "k2g_server_release",
    "k2g_server_stats",
    "k2g_server_destroy",
    "k2g_build_create",
    "k2g_build_insert",
    "k2g_build_insert_device",
    "k2g_build_get_info",
    "k2g_build_verify",
    "k2g_build_download",
    "k2g_build_write",
    "k2g_build_destroy",
    "k2g_estimate_capacity",
    "k2g_estimate_capacity_device",
    "k2g_pipeline_create",
    "k2g_pipeline_run",
    "k2g_pipeline_run_device",
    "k2g_pipeline_staging",
    "k2g_pipeline_digests",
    "k2g_pipeline_get_params",
    "k2g_pipeline_redo_counts",
    "k2g_pipeline_split_counts",
    "k2g_pipeline_destroy",
    "k2g_pipeline_tune",
    "k2g_merge_open",
    "k2g_merge",
    "k2g_merge_close",
    "k2g_merge_text",
)

MERGE_AMBIG = 0x7FFFFFFF
MERGE_MATE = 0x7FFFFFFE
MERGE_FRAME = 0x7FFFFFFD


class K2GError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"k2g error {code}: {msg}")
        self.code = code


class DbInfo(C.Structure):
    _fields_ = [
        ("capacity", C.c_uint64),
        ("size", C.c_uint64),
        ("key_bits", C.c_uint64),
        ("value_bits", C.c_uint64),
        ("node_count", C.c_uint64),
        ("k", C.c_uint64),
        ("l", C.c_uint64),
        ("spaced_seed_mask", C.c_uint64),
        ("toggle_mask", C.c_uint64),
        ("minimum_acceptable_hash_value", C.c_uint64),
        ("dna_db", C.c_int32),
        ("revcom_version", C.c_int32),
        ("device_bytes", C.c_uint64),
    ]


class ClassifyParams(C.Structure):
    _fields_ = [
        ("confidence", C.c_double),
        ("min_hit_groups", C.c_int32),
        ("quick", C.c_int32),
        ("paired", C.c_int32),
        ("mode", C.c_int32),
        ("want_kmer_counts", C.c_int32),
        ("want_hll", C.c_int32),
        ("table_slots", C.c_int32),
        ("want_probe_stats", C.c_int32),
        ("run_cap_pct", C.c_int32),
        ("unit_elements", C.c_int32),
    ]


class ServerParams(C.Structure):
    _fields_ = [
        ("hbm_budget", C.c_uint64),
        ("slots", C.c_int32),
    ]


class Fragments(C.Structure):
    _fields_ = [
        ("seq", C.c_void_p),
        ("off", C.c_void_p),
        ("paired", C.c_void_p),
        ("n", C.c_uint64),
    ]


class Results(C.Structure):
    _fields_ = [
        ("n", C.c_uint64),
        ("call", C.POINTER(C.c_uint32)),
        ("hit_groups", C.POINTER(C.c_uint32)),
        ("total_kmers", C.POINTER(C.c_uint32)),
        ("run_off", C.POINTER(C.c_uint64)),
        ("runs", C.POINTER(C.c_uint32)),
        ("hll_off", C.POINTER(C.c_uint64)),
        ("hll", C.POINTER(C.c_uint32)),
        ("kmer_counts", C.POINTER(C.c_uint64)),
        ("lookups", C.POINTER(C.c_uint32)),
        ("probe_lines", C.POINTER(C.c_uint64)),
        ("node_count", C.c_uint64),
        ("redone", C.c_uint64),
        ("kernel_ms", C.c_double),
        ("n_sketches", C.c_uint64),
        ("sketch_block", C.POINTER(C.c_uint32)),
        ("sketch_taxon", C.POINTER(C.c_uint32)),
        ("sketch_dense", C.POINTER(C.c_uint32)),
        ("sketch_n", C.POINTER(C.c_uint64)),
        ("sketch_off", C.POINTER(C.c_uint64)),
        ("sketch_vals", C.POINTER(C.c_uint32)),
    ]


class HitLists(C.Structure):
    _fields_ = [("items", C.c_void_p), ("off", C.c_void_p), ("n", C.c_uint64)]


class HitListText(C.Structure):
    _fields_ = [("text", C.c_void_p), ("text_bytes", C.c_uint64), ("field", C.c_void_p), ("len", C.c_void_p),
                ("off", C.c_void_p), ("n", C.c_uint64)]


class MergeParams(C.Structure):
    _fields_ = [("confidence", C.c_float), ("text", C.c_int32)]


class MergeResults(C.Structure):
    _fields_ = [
        ("n", C.c_uint64),
        ("call", C.POINTER(C.c_int32)),
        ("token_off", C.POINTER(C.c_uint64)),
        ("tokens", C.POINTER(C.c_int32)),
        ("joined", C.POINTER(C.c_uint8)),
        ("text", C.POINTER(C.c_char)),
        ("text_off", C.POINTER(C.c_uint64)),
        ("kernel_ms", C.c_double),
    ]


_lib = None


def load_library(path: str = LIB_PATH):
    """Loads libk2g.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise K2GError(70, f"{path} is not built (run __graft_entry__.build())")
    lib = C.CDLL(path)
    vp, u64, i32 = C.c_void_p, C.c_uint64, C.c_int
    lib.k2g_db_open.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.POINTER(vp)]
    lib.k2g_db_close.argtypes = [vp]
    lib.k2g_db_get_info.argtypes = [vp, C.POINTER(DbInfo)]
    lib.k2g_db_external_ids.argtypes = [vp, vp, u64]
    lib.k2g_last_error.restype = C.c_char_p
    lib.k2g_get_batch.argtypes = [vp, vp, vp, C.c_size_t]
    lib.k2g_get_batch_device.argtypes = [vp, vp, vp, C.c_size_t, vp]
    lib.k2g_batch_create.argtypes = [vp, C.POINTER(vp)]
    lib.k2g_batch_destroy.argtypes = [vp]
    lib.k2g_classify.argtypes = [vp, C.POINTER(ClassifyParams), C.POINTER(Fragments), C.POINTER(Results)]
    lib.k2g_classify_device.argtypes = [vp, C.POINTER(ClassifyParams), vp, vp, vp, u64, vp]
    lib.k2g_batch_sync.argtypes = [vp, C.POINTER(C.c_double)]
    lib.k2g_batch_fetch.argtypes = [vp, C.POINTER(Results)]
    lib.k2g_batch_stream.argtypes = [vp]
    lib.k2g_batch_fetch_counts.argtypes = [vp] + [vp] * 6 + [u64]
    lib.k2g_batch_redo_counts.argtypes = [vp, C.POINTER(u64), C.POINTER(u64)]
    lib.k2g_batch_digests.argtypes = [vp, vp, u64]
    lib.k2g_batch_split_counts.argtypes = [vp, C.POINTER(u64), C.POINTER(u64)]
    lib.k2g_server_create.argtypes = [vp, C.POINTER(ServerParams), C.POINTER(vp)]
    lib.k2g_server_submit.argtypes = [vp, C.POINTER(ClassifyParams), C.POINTER(Fragments), C.POINTER(u64)]
    lib.k2g_server_wait.argtypes = [vp, u64, C.POINTER(Results), C.POINTER(C.c_double)]
    lib.k2g_server_release.argtypes = [vp, u64]
    lib.k2g_server_stats.argtypes = [vp, C.POINTER(u64), C.POINTER(C.c_int32), C.POINTER(u64)]
    lib.k2g_server_destroy.argtypes = [vp]
    lib.k2g_host_alloc.argtypes = [C.c_size_t, C.POINTER(vp)]
    lib.k2g_host_free.argtypes = [vp]
    lib.k2g_results_checksum.argtypes = [C.POINTER(Results), C.POINTER(u64)]
    lib.k2g_db_gather_ceiling.argtypes = [vp, u64, u64, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    lib.k2g_pipeline_create.argtypes = [vp, C.POINTER(ClassifyParams), C.POINTER(PipelineParams), C.POINTER(vp)]
    lib.k2g_pipeline_run.argtypes = [vp, C.POINTER(Fragments), C.POINTER(Results)]
    lib.k2g_pipeline_run_device.argtypes = [vp, vp, vp, vp, u64, C.POINTER(C.c_double)]
    lib.k2g_pipeline_staging.argtypes = [vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp)]
    lib.k2g_pipeline_digests.argtypes = [vp, vp, u64]
    lib.k2g_pipeline_get_params.argtypes = [vp, C.POINTER(PipelineParams), C.POINTER(C.c_int32)]
    lib.k2g_pipeline_redo_counts.argtypes = [vp, C.POINTER(u64), C.POINTER(u64)]
    lib.k2g_pipeline_split_counts.argtypes = [vp, C.POINTER(u64), C.POINTER(u64)]
    lib.k2g_pipeline_destroy.argtypes = [vp]
    lib.k2g_pipeline_tune.argtypes = [vp, C.POINTER(ClassifyParams), vp, vp, vp, u64, u64, C.POINTER(PipelineParams),
                                      C.POINTER(TuneReport)]
    lib.k2g_merge_open.argtypes = [C.c_char_p, C.POINTER(vp)]
    lib.k2g_merge.argtypes = [vp, C.POINTER(MergeParams), C.POINTER(HitLists), C.POINTER(HitLists),
                              C.POINTER(MergeResults)]
    lib.k2g_merge_close.argtypes = [vp]
    lib.k2g_merge_text.argtypes = [vp, C.POINTER(MergeParams), C.POINTER(HitListText), C.POINTER(HitListText),
                                   C.POINTER(MergeResults)]
    lib.k2g_batch_stream.restype = vp
    for name in EXPORTED_SYMBOLS:
        if name not in ("k2g_last_error", "k2g_batch_stream"):
            getattr(lib, name).restype = i32
    _lib = lib
    return lib


def _check(rc: int):
    if rc != 0:
        raise K2GError(rc, _lib.k2g_last_error().decode())


@dataclass
class ClassifyOutput:
    call: np.ndarray          # internal taxid per fragment
    hit_groups: np.ndarray
    total_kmers: np.ndarray
    run_off: np.ndarray       # n+1
    runs: np.ndarray          # (m, 2) uint32: code, count
    hll_off: np.ndarray | None
    hll: np.ndarray | None    # (r, 2) uint32: taxon, encoded hash
    kmer_counts: np.ndarray | None
    lookups: np.ndarray | None
    probe_lines: np.ndarray | None
    redone: int
    kernel_ms: float
    # hll=2: per-(input block, taxon) distinct sets reduced on the device:
    # arrays block, taxon, dense, n, off (groups + 1), vals
    sets: dict | None = None


class Db:
    """A resident database (hash.k2d + taxo.k2d + opts.k2d) on the current GPU."""

    def __init__(self, hash_k2d: str, taxo_k2d: str, opts_k2d: str):
        lib = load_library()
        h = C.c_void_p()
        _check(lib.k2g_db_open(hash_k2d.encode(), taxo_k2d.encode(), opts_k2d.encode(), C.byref(h)))
        self._h = h
        info = DbInfo()
        _check(lib.k2g_db_get_info(h, C.byref(info)))
        self.info = info
        ext = np.zeros(info.node_count, dtype=np.uint64)
        _check(lib.k2g_db_external_ids(h, ext.ctypes.data, info.node_count))
        self.external_ids = ext

    @classmethod
    def from_dir(cls, d: str) -> "Db":
        return cls(os.path.join(d, "hash.k2d"), os.path.join(d, "taxo.k2d"), os.path.join(d, "opts.k2d"))

    def close(self):
        if self._h:
            _lib.k2g_db_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def gather_ceiling(self, n_probes: int, seed: int = 1) -> tuple[float, float]:
        """(lines/s, kernel ms) of random 16-byte loads at 128-byte-aligned table offsets."""
        rate, ms = C.c_double(), C.c_double()
        _check(_lib.k2g_db_gather_ceiling(self._h, n_probes, seed, C.byref(rate), C.byref(ms)))
        return rate.value, ms.value

    def get_batch(self, keys: np.ndarray) -> np.ndarray:
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        out = np.zeros(keys.shape[0], dtype=np.uint32)
        _check(_lib.k2g_get_batch(self._h, keys.ctypes.data, out.ctypes.data, keys.shape[0]))
        return out


def pack_fragments(mates1, mates2=None):
    """Concatenates reads into (seq bytes, 2n+1 offsets).  mates2 None = single-end."""
    n = len(mates1)
    parts = []
    off = np.zeros(2 * n + 1, dtype=np.uint64)
    pos = 0
    for i in range(n):
        a = mates1[i].encode() if isinstance(mates1[i], str) else mates1[i]
        off[2 * i] = pos
        parts.append(a)
        pos += len(a)
        off[2 * i + 1] = pos
        if mates2 is not None:
            b = mates2[i].encode() if isinstance(mates2[i], str) else mates2[i]
            parts.append(b)
            pos += len(b)
    off[2 * n] = pos
    return b"".join(parts), off


class Batch:
    def __init__(self, db: Db):
        self.db = db
        h = C.c_void_p()
        _check(_lib.k2g_batch_create(db._h, C.byref(h)))
        self._h = h

    def close(self):
        if self._h:
            _lib.k2g_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def classify(self, seq: bytes, off: np.ndarray, *, paired=False, paired_flags=None,
                 confidence=0.0, min_hit_groups=0, quick=False, mode=0, kmer_counts=False,
                 hll=False, table_slots=0, probe_stats=False, run_cap_pct=0, unit_elements=0) -> ClassifyOutput:
        off = np.ascontiguousarray(off, dtype=np.uint64)
        n = (off.shape[0] - 1) // 2
        p = ClassifyParams(confidence, min_hit_groups, int(quick), int(paired), mode,
                           int(kmer_counts), int(hll), table_slots, int(probe_stats), run_cap_pct,
                           unit_elements)
        sbuf = np.frombuffer(seq, dtype=np.uint8) if len(seq) else np.zeros(1, np.uint8)
        flags = None
        if paired_flags is not None:
            flags = np.ascontiguousarray(paired_flags, dtype=np.uint8)
        fr = Fragments(sbuf.ctypes.data, off.ctypes.data, flags.ctypes.data if flags is not None else None, n)
        res = Results()
        _check(_lib.k2g_classify(self._h, C.byref(p), C.byref(fr), C.byref(res)))
        return _copy_results(res, hll, kmer_counts)

    def classify_checksum(self, seq, off: np.ndarray, **kw) -> int:
        """Like classify, but returns only the results' 64-bit checksum."""
        off = np.ascontiguousarray(off, dtype=np.uint64)
        n = (off.shape[0] - 1) // 2
        p = ClassifyParams(kw.get("confidence", 0.0), kw.get("min_hit_groups", 0), int(kw.get("quick", False)),
                           int(kw.get("paired", False)), kw.get("mode", 0), int(kw.get("kmer_counts", False)),
                           int(kw.get("hll", False)), kw.get("table_slots", 0), 0, kw.get("run_cap_pct", 0), 0)
        if isinstance(seq, np.ndarray):
            sbuf = np.ascontiguousarray(seq, dtype=np.uint8) if len(seq) else np.zeros(1, np.uint8)
        else:
            sbuf = np.frombuffer(seq, dtype=np.uint8) if len(seq) else np.zeros(1, np.uint8)
        pf = kw.get("paired_flags")
        flags = None if pf is None else np.ascontiguousarray(pf, dtype=np.uint8)
        fr = Fragments(sbuf.ctypes.data, off.ctypes.data, flags.ctypes.data if flags is not None else None, n)
        res = Results()
        _check(_lib.k2g_classify(self._h, C.byref(p), C.byref(fr), C.byref(res)))
        h = C.c_uint64()
        _check(_lib.k2g_results_checksum(C.byref(res), C.byref(h)))
        return h.value

    def digests(self, n: int) -> np.ndarray:
        """Per-fragment 64-bit digests of the last classify call's results."""
        out = np.zeros(max(1, n), dtype=np.uint64)
        _check(_lib.k2g_batch_digests(self._h, out.ctypes.data, n))
        return out[:n]

    def fetch_counts(self, n: int, probe_stats: bool = False) -> dict:
        """Per-fragment call, hit_groups, total_kmers, n_runs (and lookups,
        probe_lines) of the last call, without hitlists."""
        out = {k: np.zeros(max(1, n), np.uint32) for k in ("call", "hit_groups", "total_kmers", "n_runs")}
        lk = np.zeros(max(1, n), np.uint32) if probe_stats else None
        ln = np.zeros(max(1, n), np.uint64) if probe_stats else None
        _check(_lib.k2g_batch_fetch_counts(self._h, *[out[k].ctypes.data for k in
                                                      ("call", "hit_groups", "total_kmers", "n_runs")],
                                           lk.ctypes.data if lk is not None else None,
                                           ln.ctypes.data if ln is not None else None, n))
        res = {k: v[:n] for k, v in out.items()}
        if probe_stats:
            res["lookups"], res["probe_lines"] = lk[:n], ln[:n]
        return res

    def split_counts(self) -> tuple[int, int]:
        """(long fragments of the last call split over waves, their segments)."""
        a, b = C.c_uint64(), C.c_uint64()
        _check(_lib.k2g_batch_split_counts(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def redo_counts(self) -> tuple[int, int]:
        """(fragments redone into exact hitlist space, fragments redone with the dense table)."""
        a, b = C.c_uint64(), C.c_uint64()
        _check(_lib.k2g_batch_redo_counts(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value


def _arr(ptr, n, dtype):
    if n == 0 or not ptr:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


def _copy_results(res: Results, hll: bool, kmer_counts: bool) -> ClassifyOutput:
    n = res.n
    run_off = _arr(res.run_off, n + 1, np.uint64)
    m = int(run_off[-1]) if n else 0
    runs = _arr(res.runs, 2 * m, np.uint32).reshape(-1, 2)
    hoff = hrec = None
    if hll and res.hll_off:
        hoff = _arr(res.hll_off, n + 1, np.uint64)
        hrec = _arr(res.hll, 2 * int(hoff[-1]), np.uint32).reshape(-1, 2)
    kc = None
    if kmer_counts and res.kmer_counts:
        kc = _arr(res.kmer_counts, res.node_count, np.uint64)
    lk = _arr(res.lookups, n, np.uint32) if res.lookups else None
    ln = _arr(res.probe_lines, n, np.uint64) if res.probe_lines else None
    sets = None
    if res.sketch_off:
        g = int(res.n_sketches)
        soff = _arr(res.sketch_off, g + 1, np.uint64)
        sets = {"block": _arr(res.sketch_block, g, np.uint32), "taxon": _arr(res.sketch_taxon, g, np.uint32),
                "dense": _arr(res.sketch_dense, g, np.uint32), "n": _arr(res.sketch_n, g, np.uint64),
                "off": soff, "vals": _arr(res.sketch_vals, int(soff[-1]), np.uint32)}
    return ClassifyOutput(
        call=_arr(res.call, n, np.uint32),
        hit_groups=_arr(res.hit_groups, n, np.uint32),
        total_kmers=_arr(res.total_kmers, n, np.uint32),
        run_off=run_off,
        runs=runs,
        hll_off=hoff,
        hll=hrec,
        kmer_counts=kc,
        lookups=lk,
        probe_lines=ln,
        redone=int(res.redone),
        kernel_ms=float(res.kernel_ms),
        sets=sets,
    )


class ReportParams(C.Structure):
    _fields_ = [("report_zeros", C.c_int32), ("kmer_data", C.c_int32), ("paired", C.c_int32)]


def write_report(db: "Db", off: np.ndarray, out: ClassifyOutput, path: str, *, paired=False, paired_flags=None,
                 zeros=False, kmer_data=False):
    """kreport of one classify call's results (k2g_report_write): the
    fragments' offsets and paired flags cut them into the reference's input
    blocks; kmer_data needs results with HLL records (hll=1) or the
    device's per-(block, taxon) sets (hll=2)."""
    lib = load_library()
    lib.k2g_report_write.argtypes = [C.c_void_p, C.POINTER(Fragments), C.POINTER(Results),
                                     C.POINTER(ReportParams), C.c_char_p]
    off = np.ascontiguousarray(off, dtype=np.uint64)
    n = (off.shape[0] - 1) // 2
    flags = None if paired_flags is None else np.ascontiguousarray(paired_flags, dtype=np.uint8)
    fr = Fragments(None, off.ctypes.data, flags.ctypes.data if flags is not None else None, n)
    keep = []

    def ptr(a, ct):
        if a is None:
            return None
        a = np.ascontiguousarray(a)
        keep.append(a)
        return a.ctypes.data_as(C.POINTER(ct))

    r = Results()
    r.n = n
    r.call = ptr(out.call, C.c_uint32)
    r.hit_groups = ptr(out.hit_groups, C.c_uint32)
    r.total_kmers = ptr(out.total_kmers, C.c_uint32)
    r.run_off = ptr(out.run_off, C.c_uint64)
    r.runs = ptr(out.runs, C.c_uint32)
    r.hll_off = ptr(out.hll_off, C.c_uint64)
    r.hll = ptr(out.hll, C.c_uint32)
    r.kmer_counts = ptr(out.kmer_counts, C.c_uint64)
    if out.sets is not None:
        r.n_sketches = len(out.sets["block"])
        r.sketch_block = ptr(out.sets["block"], C.c_uint32)
        r.sketch_taxon = ptr(out.sets["taxon"], C.c_uint32)
        r.sketch_dense = ptr(out.sets["dense"], C.c_uint32)
        r.sketch_n = ptr(out.sets["n"], C.c_uint64)
        r.sketch_off = ptr(out.sets["off"], C.c_uint64)
        r.sketch_vals = ptr(out.sets["vals"] if len(out.sets["vals"]) else np.zeros(1, np.uint32), C.c_uint32)
    r.node_count = int(db.info.node_count)
    _check(lib.k2g_report_write(db._h, C.byref(fr), C.byref(r),
                                C.byref(ReportParams(int(zeros), int(kmer_data), int(paired))), path.encode()))


def format_hitlist(runs: np.ndarray, external_ids: np.ndarray) -> str:
    """Kraken hitlist text of one fragment's runs (src/classify.cc:950-991)."""
    if runs.shape[0] == 0:
        return "0:0"
    out = []
    last = runs.shape[0] - 1
    for i, (code, cnt) in enumerate(runs.tolist()):
        if code == CODE_MATE:
            out.append("|:|" if i == last else "|:| ")
        elif code == CODE_FRAME:
            out.append("-:-" if i == last else "-:- ")
        elif code == CODE_AMBIG:
            out.append(f"A:{cnt} ")
        else:
            out.append(f"{int(external_ids[code])}:{cnt}" + ("" if i == last else " "))
    return "".join(out)


class BuildParams(C.Structure):
    _fields_ = [
        ("k", C.c_uint64),
        ("l", C.c_uint64),
        ("spaced_seed_mask", C.c_uint64),
        ("toggle_mask", C.c_uint64),
        ("min_clear_hash_value", C.c_uint64),
        ("block_size", C.c_uint64),
        ("subblock_size", C.c_uint64),
        ("dna_db", C.c_int32),
        ("revcom_version", C.c_int32),
    ]


def expand_spaced_seed_mask(template: str, bits_per_char: int) -> int:
    """Template bit string -> mask with bits_per_char bits per template bit
    (src/utilities.cc:14-27)."""
    t = int(template, 2)
    out = 0
    ones = (1 << bits_per_char) - 1
    for i in range(64 // bits_per_char - 1, -1, -1):
        out <<= bits_per_char
        if (t >> i) & 1:
            out |= ones
    return out & 0xFFFFFFFFFFFFFFFF


class BuildCheck(C.Structure):
    _fields_ = [(name, C.c_uint64) for name in
                ("occurrences", "keys", "cells", "missing", "not_covering", "shared_cells", "shared_keys",
                 "differ", "unexplained", "orphan_cells")] + [("partitions", C.c_uint32), ("pad", C.c_uint32)]


class Builder:
    """GPU table construction (atomicCAS + LCA) into a resident cell array."""

    def __init__(self, taxo_k2d: str, capacity: int, *, k=35, l=31, spaced_seed_mask=0,
                 toggle_mask=0xE37E28C4271B5A2D, min_clear_hash_value=0, block_size=10 * 1024 * 1024,
                 subblock_size=1024, dna=True, value_bits=0):
        lib = load_library()
        lib.k2g_build_create.argtypes = [C.c_char_p, C.c_uint64, C.c_uint64, C.POINTER(BuildParams),
                                         C.POINTER(C.c_void_p)]
        lib.k2g_build_insert.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64]
        lib.k2g_build_insert_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64]
        lib.k2g_build_get_info.argtypes = [C.c_void_p] + [C.POINTER(C.c_uint64)] * 4
        lib.k2g_build_download.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64]
        lib.k2g_build_write.argtypes = [C.c_void_p, C.c_char_p]
        lib.k2g_build_destroy.argtypes = [C.c_void_p]
        p = BuildParams(k, l, spaced_seed_mask, toggle_mask, min_clear_hash_value, block_size,
                        subblock_size, int(dna), 1)
        h = C.c_void_p()
        _check(lib.k2g_build_create(taxo_k2d.encode(), capacity, value_bits, C.byref(p), C.byref(h)))
        self._h = h
        self.capacity = capacity

    def insert(self, seq: bytes, off: np.ndarray, taxid: np.ndarray):
        off = np.ascontiguousarray(off, dtype=np.uint64)
        taxid = np.ascontiguousarray(taxid, dtype=np.uint32)
        buf = np.frombuffer(seq, np.uint8) if len(seq) else np.zeros(1, np.uint8)
        _check(_lib.k2g_build_insert(self._h, buf.ctypes.data, off.ctypes.data, taxid.ctypes.data,
                                     taxid.shape[0]))

    def verify_device(self, d_seq, off: np.ndarray, taxid: np.ndarray) -> dict:
        """k2g_build_verify over device-resident sequences (those inserted)."""
        off = np.ascontiguousarray(off, dtype=np.uint64)
        taxid = np.ascontiguousarray(taxid, dtype=np.uint32)
        _lib.k2g_build_verify.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64,
                                          C.POINTER(BuildCheck)]
        chk = BuildCheck()
        _check(_lib.k2g_build_verify(self._h, d_seq, off.ctypes.data, taxid.ctypes.data, taxid.shape[0],
                                     C.byref(chk)))
        return {name: int(getattr(chk, name)) for name, _ in BuildCheck._fields_ if name != "pad"}

    def info(self):
        v = [C.c_uint64() for _ in range(4)]
        _check(_lib.k2g_build_get_info(self._h, *[C.byref(x) for x in v]))
        return dict(zip(("capacity", "size", "key_bits", "value_bits"), [x.value for x in v]))

    def cells(self) -> np.ndarray:
        out = np.zeros(self.capacity, np.uint32)
        _check(_lib.k2g_build_download(self._h, out.ctypes.data, self.capacity))
        return out

    def write(self, path: str):
        _check(_lib.k2g_build_write(self._h, path.encode()))

    def close(self):
        if self._h:
            _lib.k2g_build_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """The request server: requests admitted FIFO against a device-memory
    budget, run concurrently on `slots` streams, the table resident."""

    def __init__(self, db: Db, hbm_budget: int, slots: int = 4):
        self.db = db
        h = C.c_void_p()
        _check(_lib.k2g_server_create(db._h, C.byref(ServerParams(int(hbm_budget), int(slots))), C.byref(h)))
        self._h = h
        self._keep = {}

    def submit(self, seq: bytes, off: np.ndarray, *, paired=False, paired_flags=None, confidence=0.0,
               min_hit_groups=0, quick=False, mode=2, kmer_counts=False, hll=False, table_slots=0) -> int:
        off = np.ascontiguousarray(off, dtype=np.uint64)
        n = (off.shape[0] - 1) // 2
        p = ClassifyParams(confidence, min_hit_groups, int(quick), int(paired), mode,
                           int(kmer_counts), int(hll), table_slots, 0, 100)
        if isinstance(seq, np.ndarray):
            sbuf = np.ascontiguousarray(seq, dtype=np.uint8) if len(seq) else np.zeros(1, np.uint8)
        else:
            sbuf = np.frombuffer(seq, dtype=np.uint8) if len(seq) else np.zeros(1, np.uint8)
        flags = None if paired_flags is None else np.ascontiguousarray(paired_flags, dtype=np.uint8)
        fr = Fragments(sbuf.ctypes.data, off.ctypes.data, flags.ctypes.data if flags is not None else None, n)
        t = C.c_uint64()
        _check(_lib.k2g_server_submit(self._h, C.byref(p), C.byref(fr), C.byref(t)))
        self._keep[t.value] = (sbuf, off, flags, hll, kmer_counts)
        return t.value

    def wait(self, ticket: int, checksum_only: bool = False):
        """(ClassifyOutput, {"queued_ms", "gpu_ms", "total_ms"}); releases the request.
        With checksum_only the first element is the results' 64-bit checksum."""
        res = Results()
        times = (C.c_double * 3)()
        rc = _lib.k2g_server_wait(self._h, ticket, C.byref(res), times)
        try:
            _check(rc)
            _, _, _, hll, kc = self._keep[ticket]
            if checksum_only:
                h = C.c_uint64()
                _check(_lib.k2g_results_checksum(C.byref(res), C.byref(h)))
                out = h.value
            else:
                out = _copy_results(res, hll, kc)
        finally:
            _lib.k2g_server_release(self._h, ticket)
            self._keep.pop(ticket, None)
        return out, {"queued_ms": times[0], "gpu_ms": times[1], "total_ms": times[2]}

    def stats(self) -> dict:
        peak, conc, budget = C.c_uint64(), C.c_int32(), C.c_uint64()
        _check(_lib.k2g_server_stats(self._h, C.byref(peak), C.byref(conc), C.byref(budget)))
        return {"peak_bytes": peak.value, "max_concurrent": conc.value, "budget": budget.value}

    def close(self):
        if self._h:
            _lib.k2g_server_destroy(self._h)
            self._h = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PipelineParams(C.Structure):
    _fields_ = [("max_fragments", C.c_uint64), ("max_bases", C.c_uint64), ("host_io", C.c_int32),
                ("unit_elements", C.c_int32), ("bucket_shift", C.c_int32), ("probe_depth", C.c_int32),
                ("waves_per_cu", C.c_int32)]

    def tuning(self) -> dict:
        return {k: int(getattr(self, k)) for k in ("unit_elements", "bucket_shift", "probe_depth", "waves_per_cu")}


class TuneReport(C.Structure):
    _fields_ = [("sample_fragments", C.c_uint64), ("sample_bases", C.c_uint64), ("evaluated", C.c_uint32),
                ("mismatched", C.c_uint32), ("default_ms", C.c_double), ("best_ms", C.c_double)]


def _classify_params(kw) -> ClassifyParams:
    return ClassifyParams(kw.get("confidence", 0.0), kw.get("min_hit_groups", 0), int(kw.get("quick", False)),
                          int(kw.get("paired", False)), 0, int(kw.get("kmer_counts", False)),
                          int(kw.get("hll", False)), kw.get("table_slots", 0), 0, kw.get("run_cap_pct", 0), 0)


class Pipeline:
    """A batch pipeline recorded as one HIP graph (k2g_pipeline_*): batches up
    to max_fragments / max_bases replay it.  host_io=True takes host
    fragments and returns host results (run); host_io=False classifies
    device-resident fragments and leaves the results on the device
    (run_device, digests).  Tuning keys: unit_elements, bucket_shift,
    probe_depth, waves_per_cu."""

    def __init__(self, db: Db, max_fragments: int, max_bases: int, *, host_io=True, tuning=None, **kw):
        self.db = db
        self._kw = kw
        t = tuning or {}
        pp = PipelineParams(int(max_fragments), int(max_bases), int(host_io), int(t.get("unit_elements", 0)),
                            int(t.get("bucket_shift", 0)), int(t.get("probe_depth", 0)),
                            int(t.get("waves_per_cu", 0)))
        h = C.c_void_p()
        _check(load_library().k2g_pipeline_create(db._h, C.byref(_classify_params(kw)), C.byref(pp), C.byref(h)))
        self._h = h

    def _fragments(self, seq, off, paired_flags):
        off = np.ascontiguousarray(off, dtype=np.uint64)
        n = (off.shape[0] - 1) // 2
        if isinstance(seq, np.ndarray):
            sbuf = np.ascontiguousarray(seq, dtype=np.uint8) if len(seq) else np.zeros(1, np.uint8)
        else:
            sbuf = np.frombuffer(seq, dtype=np.uint8) if len(seq) else np.zeros(1, np.uint8)
        flags = None if paired_flags is None else np.ascontiguousarray(paired_flags, dtype=np.uint8)
        fr = Fragments(sbuf.ctypes.data, off.ctypes.data, flags.ctypes.data if flags is not None else None, n)
        return fr, (sbuf, off, flags)

    def run(self, seq, off, *, paired_flags=None, checksum_only=False):
        """ClassifyOutput of one host batch (or its 64-bit checksum)."""
        fr, keep = self._fragments(seq, off, paired_flags)
        res = Results()
        _check(_lib.k2g_pipeline_run(self._h, C.byref(fr), C.byref(res)))
        if checksum_only:
            h = C.c_uint64()
            _check(_lib.k2g_results_checksum(C.byref(res), C.byref(h)))
            return h.value
        return _copy_results(res, bool(self._kw.get("hll")), bool(self._kw.get("kmer_counts")))

    def staging(self, max_fragments: int, max_bases: int):
        """(bases, offsets, pair flags) numpy views of the host pipeline's
        pinned staging: a batch written there is run without a copy."""
        sq, so, sp = C.c_void_p(), C.c_void_p(), C.c_void_p()
        _check(_lib.k2g_pipeline_staging(self._h, C.byref(sq), C.byref(so), C.byref(sp)))
        seq = np.ctypeslib.as_array((C.c_uint8 * max(1, max_bases)).from_address(sq.value))
        off = np.ctypeslib.as_array((C.c_uint64 * (2 * max_fragments + 1)).from_address(so.value))
        pf = np.ctypeslib.as_array((C.c_uint8 * max(1, max_fragments)).from_address(sp.value))
        return seq, off, pf

    def run_device(self, d_seq, d_off, d_paired, n: int) -> float:
        """Classifies n device-resident fragments (device addresses); returns
        the graph's device time in ms."""
        ms = C.c_double()
        _check(_lib.k2g_pipeline_run_device(self._h, d_seq, d_off, d_paired, int(n), C.byref(ms)))
        return ms.value

    def digests(self, n: int) -> np.ndarray:
        """Per-fragment digests of the last batch (as Batch.digests)."""
        out = np.zeros(max(1, n), dtype=np.uint64)
        _check(_lib.k2g_pipeline_digests(self._h, out.ctypes.data, int(n)))
        return out[:n]

    def params(self) -> tuple[dict, int]:
        """(tuning parameters in effect, classify grid in workgroups)."""
        pp, grid = PipelineParams(), C.c_int32()
        _check(_lib.k2g_pipeline_get_params(self._h, C.byref(pp), C.byref(grid)))
        return pp.tuning(), grid.value

    def redo_counts(self) -> tuple[int, int]:
        """(fragments of the last batch redone into the overflow pool, redone with the dense table)."""
        a, b = C.c_uint64(), C.c_uint64()
        _check(_lib.k2g_pipeline_redo_counts(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def split_counts(self) -> tuple[int, int]:
        """(fragments of the last batch split over waves, their segments)."""
        f, g = C.c_uint64(), C.c_uint64()
        _check(_lib.k2g_pipeline_split_counts(self._h, C.byref(f), C.byref(g)))
        return f.value, g.value

    def close(self):
        if self._h:
            _lib.k2g_pipeline_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def tune_pipeline(db: Db, d_seq, d_off, d_paired, n: int, stride: int = 1, start=None, **kw) -> tuple[dict, dict]:
    """Autotunes a pipeline on every stride-th fragment of a device-resident
    set: (best tuning parameters, report)."""
    t = start or {}
    best = PipelineParams(0, 0, 0, int(t.get("unit_elements", 0)), int(t.get("bucket_shift", 0)),
                          int(t.get("probe_depth", 0)), int(t.get("waves_per_cu", 0)))
    rep = TuneReport()
    _check(load_library().k2g_pipeline_tune(db._h, C.byref(_classify_params(kw)), d_seq, d_off, d_paired, int(n),
                                            int(stride), C.byref(best), C.byref(rep)))
    return best.tuning(), {k: getattr(rep, k) for k, _ in TuneReport._fields_}


def pinned_array(nbytes: int) -> np.ndarray:
    """A uint8 array in pinned host memory (freed with the array)."""
    lib = load_library()
    p = C.c_void_p()
    _check(lib.k2g_host_alloc(max(1, int(nbytes)), C.byref(p)))
    buf = (C.c_uint8 * max(1, int(nbytes))).from_address(p.value)
    arr = np.frombuffer(buf, dtype=np.uint8)[:int(nbytes)]
    import weakref
    weakref.finalize(buf, lib.k2g_host_free, C.c_void_p(p.value))
    arr = arr.view()
    arr.flags.writeable = True
    return arr


@dataclass
class MergeOutput:
    call: np.ndarray       # int32 per read (resolve_tree's value)
    token_off: np.ndarray  # uint64, n + 1: read i prints tokens[token_off[i]:token_off[i+1]]
    tokens: np.ndarray     # (m, 2) int32: printed (taxid, count) pairs
    joined: np.ndarray     # uint8 per pair: 1 = no space after the pair
    kernel_ms: float
    text: list | None = None  # with text=True: each read's hit list as printed (bytes)


class Merger:
    """The merged taxonomy resident on the GPU and the per-read merge of two
    hit-list sets (k2g_merge_open / k2g_merge / k2g_merge_close)."""

    def __init__(self, taxo_k2d: str):
        self._lib = load_library()
        self._h = C.c_void_p()
        _check(self._lib.k2g_merge_open(taxo_k2d.encode(), C.byref(self._h)))

    def merge(self, items_a, off_a, items_b, off_b, confidence: float = 0.0, text: bool = False) -> MergeOutput:
        ia = np.ascontiguousarray(items_a, np.int32).reshape(-1)
        ib = np.ascontiguousarray(items_b, np.int32).reshape(-1)
        oa = np.ascontiguousarray(off_a, np.uint64)
        ob = np.ascontiguousarray(off_b, np.uint64)
        a = HitLists(ia.ctypes.data, oa.ctypes.data, len(oa) - 1)
        b = HitLists(ib.ctypes.data, ob.ctypes.data, len(ob) - 1)
        r = MergeResults()
        _check(self._lib.k2g_merge(self._h, C.byref(MergeParams(confidence, int(text))), C.byref(a), C.byref(b),
                                   C.byref(r)))
        n = int(r.n)

        def arr(ptr, count, dtype):
            return np.ctypeslib.as_array(ptr, (count,)).astype(dtype) if count else np.zeros(0, dtype)

        if text:
            off = arr(r.text_off, n + 1, np.uint64)
            raw = C.string_at(r.text, int(off[-1])) if int(off[-1]) else b""
            lines = [raw[int(off[i]):int(off[i + 1])] for i in range(n)]
            return MergeOutput(arr(r.call, n, np.int32), None, None, None, float(r.kernel_ms), lines)
        token_off = arr(r.token_off, n + 1, np.uint64)
        m = int(token_off[-1])
        return MergeOutput(arr(r.call, n, np.int32), token_off, arr(r.tokens, 2 * m, np.int32).reshape(-1, 2),
                           arr(r.joined, m, np.uint8), float(r.kernel_ms))

    def merge_text(self, lines_a, lines_b, confidence: float = 0.0, item_counts=None,
                   device_counts=(False, False)) -> MergeOutput:
        """Merges two lists of kraken lines (bytes, newline included) through
        k2g_merge_text; item_counts overrides side a's ':' counts (for tests);
        device_counts[s] passes no item offsets for side s, so the library
        counts the items on the GPU (as bin/k2g_merge does)."""
        def pack(lines):
            text = b"".join(lines)
            field, ln, off = [], [], [0]
            pos = 0
            for l in lines:
                parts = [i for i, c in enumerate(l) if c == 9]  # tab positions
                start = parts[3] + 1 if len(parts) >= 4 else 0
                field.append(pos + start)
                ln.append(len(l) - start)
                off.append(off[-1] + l[start:].count(b":"))
                pos += len(l)
            return (np.frombuffer(text, np.uint8) if text else np.zeros(1, np.uint8), len(text),
                    np.array(field, np.uint64), np.array(ln, np.uint32), np.array(off, np.uint64))
        ta, tb = pack(lines_a), pack(lines_b)
        if item_counts is not None:
            ta = ta[:4] + (np.concatenate([[0], np.cumsum(item_counts)]).astype(np.uint64),)
        a = HitListText(ta[0].ctypes.data, ta[1], ta[2].ctypes.data, ta[3].ctypes.data,
                        None if device_counts[0] else ta[4].ctypes.data, len(lines_a))
        b = HitListText(tb[0].ctypes.data, tb[1], tb[2].ctypes.data, tb[3].ctypes.data,
                        None if device_counts[1] else tb[4].ctypes.data, len(lines_b))
        r = MergeResults()
        _check(self._lib.k2g_merge_text(self._h, C.byref(MergeParams(confidence, 1)), C.byref(a), C.byref(b),
                                        C.byref(r)))
        n = int(r.n)
        off = np.ctypeslib.as_array(r.text_off, (n + 1,)).astype(np.uint64) if n else np.zeros(1, np.uint64)
        raw = C.string_at(r.text, int(off[-1])) if int(off[-1]) else b""
        call = np.ctypeslib.as_array(r.call, (n,)).astype(np.int32) if n else np.zeros(0, np.int32)
        return MergeOutput(call, None, None, None, float(r.kernel_ms),
                           [raw[int(off[i]):int(off[i + 1])] for i in range(n)])

    def close(self):
        if self._h:
            self._lib.k2g_merge_close(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
