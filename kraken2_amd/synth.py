"""Seeded synthetic workloads at the real shape, generated in HBM.

A workload is a synthetic taxonomy (written as taxo.k2d by the package's own
converter), a set of genomes generated on the GPU (libk2g_synth.so), a table
built from them on the GPU (k2g_estimate_capacity + k2g_build_*), and a read
set generated on the GPU.  Everything is a function of the seed.
"""
from __future__ import annotations

import ctypes as C
import os
import time

import numpy as np

from . import _lib
from .taxonomy import build_k2_taxonomy, read_k2_taxonomy

SYNTH_PATH = os.path.join(_lib.PKG_DIR, "lib", "libk2g_synth.so")
DNA_TEMPLATE = "1" * 17 + "01" * 7


class Hip:
    """Minimal HIP runtime access for device buffers (plumbing only)."""

    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so")
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.lib.hipFree.argtypes = [C.c_void_p]

    def malloc(self, n: int) -> C.c_void_p:
        p = C.c_void_p()
        rc = self.lib.hipMalloc(C.byref(p), max(1, int(n)))
        if rc != 0:
            raise MemoryError(f"hipMalloc({n}) failed ({rc})")
        return p

    def upload(self, arr: np.ndarray) -> C.c_void_p:
        p = self.malloc(arr.nbytes)
        if arr.nbytes and self.lib.hipMemcpy(p, arr.ctypes.data, arr.nbytes, 1) != 0:
            raise RuntimeError("hipMemcpy H2D failed")
        return p

    def download(self, p, nbytes: int, offset: int = 0) -> bytes:
        buf = np.empty(int(nbytes), np.uint8)
        if nbytes and self.lib.hipMemcpy(buf.ctypes.data, C.c_void_p(p.value + offset), int(nbytes), 2) != 0:
            raise RuntimeError("hipMemcpy D2H failed")
        return buf.tobytes()

    def free(self, p):
        if p is not None and p.value:
            self.lib.hipFree(p)


def _synth():
    lib = C.CDLL(SYNTH_PATH)
    u64, u32, dbl, vp = C.c_uint64, C.c_uint32, C.c_double, C.c_void_p
    lib.k2s_genome_lengths.argtypes = [u64, u32, dbl, dbl, dbl, dbl, vp]
    lib.k2s_genomes.argtypes = [u64, u32, u32, dbl, dbl, vp, vp]
    lib.k2s_read_layout.argtypes = [u64, u64, dbl, dbl, u32, u32, dbl, dbl, u32, vp, vp]
    lib.k2s_reads.argtypes = [u64, u64, vp, vp, vp, vp, u32, dbl, dbl, dbl, dbl, dbl, dbl, dbl, vp]
    return lib


def make_taxonomy(fanout=(5, 5, 5, 5, 4, 5), ranks=("phylum", "class", "order", "family", "genus", "species")):
    """Synthetic NCBI-style tree: parent/rank/name dicts keyed by external id."""
    parent, rank, name = {1: 0}, {1: "no rank"}, {1: "root"}
    level = [1]
    nxt = 2
    levels = []
    for f, r in zip(fanout, ranks):
        new = []
        for p in level:
            for _ in range(f):
                parent[nxt], rank[nxt], name[nxt] = p, r, f"{r} {nxt}"
                new.append(nxt)
                nxt += 1
        levels.append(new)
        level = new
    return parent, rank, name, levels


class Workload:
    """Genomes + table + reads for one configuration, resident on the GPU."""

    def __init__(self, out_dir: str, *, seed: int, n_genomes=4000, per_group=5, genome_median=2.5e6,
                 genome_sigma=0.3, genome_min=5e5, genome_max=6e6, divergent_frac=0.08, snp_rate=0.01,
                 k=35, l=31, load_factor=0.7, block=16384, subblock=2048, verify=False, log=print):
        self.hip = Hip()
        self.synth = _synth()
        self.dir = out_dir
        self.seed = seed
        self.log = log
        os.makedirs(out_dir, exist_ok=True)
        t0 = time.time()
        # taxonomy: genomes are the species of every third..nth genus
        parent, rank, name, levels = make_taxonomy()
        genera = levels[-2]
        n_groups = (n_genomes + per_group - 1) // per_group
        step = max(1, len(genera) // n_groups)
        chosen = genera[::step][:n_groups]
        kids = {}
        for sp in levels[-1]:
            kids.setdefault(parent[sp], []).append(sp)
        genome_ext = []
        for g in chosen:
            genome_ext.extend(kids[g][:per_group])
        genome_ext = genome_ext[:n_genomes]
        with open(os.path.join(out_dir, "taxo.k2d"), "wb") as f:
            f.write(build_k2_taxonomy(parent, rank, name, genome_ext))
        par, ext = read_k2_taxonomy(os.path.join(out_dir, "taxo.k2d"))
        e2i = {e: i for i, e in enumerate(ext)}
        self.node_count = len(ext)
        self.genome_taxid = np.array([e2i[e] for e in genome_ext], np.uint32)
        # genomes in HBM
        lens = np.zeros(n_genomes, np.uint64)
        self.synth.k2s_genome_lengths(seed, n_genomes, genome_median, genome_sigma, genome_min, genome_max,
                                      lens.ctypes.data)
        self.genome_off = np.zeros(n_genomes + 1, np.uint64)
        self.genome_off[1:] = np.cumsum(lens)
        self.genome_bases = int(self.genome_off[-1])
        self.d_genomes = self.hip.malloc(self.genome_bases)
        rc = self.synth.k2s_genomes(seed, n_genomes, per_group, divergent_frac, snp_rate,
                                    self.genome_off.ctypes.data, self.d_genomes)
        if rc:
            raise RuntimeError(f"genome generation failed ({rc})")
        log(f"# genomes: {n_genomes} x median {genome_median:.0f} bp = {self.genome_bases / 1e9:.2f} Gbp "
            f"in {time.time() - t0:.1f}s")
        # capacity estimate and table build on the GPU
        lib = _lib.load_library()
        mask = _lib.expand_spaced_seed_mask(DNA_TEMPLATE, 2)
        bp = _lib.BuildParams(k, l, mask, 0xE37E28C4271B5A2D, 0, block, subblock, 1, 1)
        lib.k2g_estimate_capacity_device.argtypes = [C.POINTER(_lib.BuildParams), C.c_void_p, C.c_void_p,
                                                     C.c_uint64, C.c_uint32, C.POINTER(C.c_uint64)]
        est = C.c_uint64()
        t1 = time.time()
        _lib._check(lib.k2g_estimate_capacity_device(C.byref(bp), self.d_genomes, self.genome_off.ctypes.data,
                                                     n_genomes, 4, C.byref(est)))
        self.estimate = est.value
        self.capacity = int((est.value + 8192) / load_factor)
        log(f"# capacity estimate {est.value} -> {self.capacity} cells in {time.time() - t1:.1f}s")
        t1 = time.time()
        bld = _lib.Builder(os.path.join(out_dir, "taxo.k2d"), self.capacity, k=k, l=l, spaced_seed_mask=mask,
                           block_size=block, subblock_size=subblock)
        lib.k2g_build_insert_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64]
        _lib._check(lib.k2g_build_insert_device(bld._h, self.d_genomes, self.genome_off.ctypes.data,
                                                self.genome_taxid.ctypes.data, n_genomes))
        info = bld.info()
        self.build_s = time.time() - t1
        log(f"# GPU build: {info['size']} cells filled of {info['capacity']} "
            f"({info['size'] / info['capacity']:.3f}) in {self.build_s:.1f}s")
        self.build_check = None
        if verify:  # every inserted minimizer looked up, compacted-key collisions counted
            t1 = time.time()
            self.build_check = bld.verify_device(self.d_genomes, self.genome_off, self.genome_taxid)
            self.build_check["seconds"] = round(time.time() - t1, 2)
            log(f"# build verification: {self.build_check}")
        t1 = time.time()
        bld.write(os.path.join(out_dir, "hash.k2d"))
        bld.close()
        opts = np.zeros(8, np.uint64)
        opts[0], opts[1], opts[2], opts[3] = k, l, mask, 0xE37E28C4271B5A2D
        raw = bytearray(opts.tobytes())
        raw[32] = 1                                   # dna_db
        raw[48:52] = (1).to_bytes(4, "little")        # revcom_version
        with open(os.path.join(out_dir, "opts.k2d"), "wb") as f:
            f.write(bytes(raw))
        log(f"# table written in {time.time() - t1:.1f}s")
        self.reads = None

    def free_genomes(self):
        self.hip.free(self.d_genomes)
        self.d_genomes = None

    def make_reads(self, n, *, f_single, f_paired, single_len=150, paired_len=150, long_median=8000,
                   long_sigma=0.8, long_cap=50000, sub=0.01, long_sub=0.025, long_indel=0.025,
                   insert_mean=350, insert_sd=50, off_frac=0.0, off_gc=0.41, seed=None, group_pairs=False):
        """group_pairs puts every unpaired read (short and long, still interleaved)
        before every pair: the two input streams a CPU classify run takes."""
        seed = self.seed if seed is None else seed
        off = np.zeros(2 * n + 1, np.uint64)
        kind = np.zeros(n, np.uint8)
        self.synth.k2s_read_layout(seed, n, f_single, f_paired, single_len, paired_len, long_median,
                                   long_sigma, long_cap, off.ctypes.data, kind.ctypes.data)
        if group_pairs:
            perm = np.argsort(kind == 1, kind="stable")
            l1 = (off[1:2 * n:2] - off[0:2 * n:2])[perm]
            l2 = (off[2::2] - off[1:2 * n:2])[perm]
            kind = np.ascontiguousarray(kind[perm])
            lens = np.empty(2 * n, np.uint64)
            lens[0::2], lens[1::2] = l1, l2
            off = np.zeros(2 * n + 1, np.uint64)
            off[1:] = np.cumsum(lens)
        bases = int(off[-1])
        d_seq = self.hip.malloc(bases)
        rc = self.synth.k2s_reads(seed, n, kind.ctypes.data, off.ctypes.data, self.d_genomes,
                                  self.genome_off.ctypes.data, len(self.genome_taxid), sub, long_sub,
                                  long_indel, insert_mean, insert_sd, off_frac, off_gc, d_seq)
        if rc:
            raise RuntimeError(f"read generation failed ({rc})")
        return ReadSet(self.hip, d_seq, off, kind)


class ReadSet:
    def __init__(self, hip: Hip, d_seq, off: np.ndarray, kind: np.ndarray):
        self.hip = hip
        self.d_seq = d_seq
        self.off = off
        self.kind = kind
        self.n = kind.shape[0]
        self.paired = (kind == 1).astype(np.uint8)
        self.d_off = hip.upload(off)
        self.d_paired = hip.upload(self.paired)

    @property
    def bases(self) -> int:
        return int(self.off[-1])

    def host_fragments(self, idx: np.ndarray):
        """(seq bytes, 2m+1 offsets, paired flags) of the selected fragments."""
        parts, off = [], np.zeros(2 * len(idx) + 1, np.uint64)
        pos = 0
        for j, i in enumerate(idx.tolist()):
            a, b, c = (int(x) for x in self.off[2 * i:2 * i + 3])
            s = self.hip.download(self.d_seq, c - a, a)
            parts.append(s)
            off[2 * j] = pos
            off[2 * j + 1] = pos + (b - a)
            pos += c - a
        off[-1] = pos
        return b"".join(parts), off, self.paired[idx]

    def free(self):
        for p in (self.d_seq, self.d_off, self.d_paired):
            self.hip.free(p)
        self.d_seq = self.d_off = self.d_paired = None


PROTEIN_TEMPLATE = "1" * 12


def _synth_protein(lib):
    u64, u32, dbl, vp = C.c_uint64, C.c_uint32, C.c_double, C.c_void_p
    lib.k2s_protein_lengths.argtypes = [u64, u64, dbl, dbl, dbl, dbl, vp]
    lib.k2s_proteomes.argtypes = [u64, u32, u32, u32, dbl, vp, vp]
    lib.k2s_protein_reads.argtypes = [u64, u64, u32, u32, u32, vp, vp, dbl, vp]
    return lib


class ProteinWorkload:
    """Translated search at the real shape: synthetic proteomes (taxa in
    sibling groups of `per_group`; each group's ancestral proteins have
    lognormal lengths and background amino-acid frequencies; each taxon keeps
    a residue with probability `identity`, so siblings share about
    identity^2 of their residues), a protein table (k=15, l=12) built from
    them on the GPU, and back-translated nucleotide reads generated on the
    GPU.  Everything is a function of the seed."""

    def __init__(self, out_dir: str, *, seed: int, n_taxa=2000, per_group=5, n_proteins=3000,
                 length_median=300.0, length_sigma=0.5, length_min=30.0, length_max=5000.0,
                 identity=0.8 ** 0.5, k=15, l=12, load_factor=0.7, block=16384, subblock=2048,
                 verify=False, log=print):
        self.hip = Hip()
        self.synth = _synth_protein(_synth())
        self.dir = out_dir
        self.seed = seed
        os.makedirs(out_dir, exist_ok=True)
        t0 = time.time()
        parent, rank, name, levels = make_taxonomy()
        genera = levels[-2]
        n_groups = (n_taxa + per_group - 1) // per_group
        step = max(1, len(genera) // n_groups)
        kids = {}
        for sp in levels[-1]:
            kids.setdefault(parent[sp], []).append(sp)
        taxa_ext = []
        for g in genera[::step][:n_groups]:
            taxa_ext.extend(kids[g][:per_group])
        taxa_ext = taxa_ext[:n_taxa]
        with open(os.path.join(out_dir, "taxo.k2d"), "wb") as f:
            f.write(build_k2_taxonomy(parent, rank, name, taxa_ext))
        _, ext = read_k2_taxonomy(os.path.join(out_dir, "taxo.k2d"))
        e2i = {e: i for i, e in enumerate(ext)}
        self.n_taxa, self.n_proteins = n_taxa, n_proteins
        # ancestral lengths per (group, protein); a taxon's proteins take its group's
        glens = np.zeros(n_groups * n_proteins, np.uint32)
        self.synth.k2s_protein_lengths(seed, glens.shape[0], length_median, length_sigma, length_min, length_max,
                                       glens.ctypes.data)
        lens = glens.reshape(n_groups, n_proteins)[np.arange(n_taxa) // per_group].reshape(-1).astype(np.uint64)
        self.seq_off = np.zeros(n_taxa * n_proteins + 1, np.uint64)
        self.seq_off[1:] = np.cumsum(lens + 1)  # each protein followed by '*'
        self.residues = int(self.seq_off[-1])
        self.seq_taxid = np.repeat(np.array([e2i[e] for e in taxa_ext], np.uint32), n_proteins)
        self.d_prot = self.hip.malloc(self.residues)
        rc = self.synth.k2s_proteomes(seed, n_taxa, per_group, n_proteins, identity, self.seq_off.ctypes.data,
                                      self.d_prot)
        if rc:
            raise RuntimeError(f"proteome generation failed ({rc})")
        log(f"# proteomes: {n_taxa} taxa x {n_proteins} proteins = {self.residues / 1e9:.2f} G residues "
            f"in {time.time() - t0:.1f}s")
        lib = _lib.load_library()
        mask = _lib.expand_spaced_seed_mask(PROTEIN_TEMPLATE, 4)
        toggle = 0xE37E28C4271B5A2D
        bp = _lib.BuildParams(k, l, mask, toggle, 0, block, subblock, 0, 1)
        lib.k2g_estimate_capacity_device.argtypes = [C.POINTER(_lib.BuildParams), C.c_void_p, C.c_void_p,
                                                     C.c_uint64, C.c_uint32, C.POINTER(C.c_uint64)]
        est = C.c_uint64()
        t1 = time.time()
        n_seq = n_taxa * n_proteins
        _lib._check(lib.k2g_estimate_capacity_device(C.byref(bp), self.d_prot, self.seq_off.ctypes.data, n_seq, 4,
                                                     C.byref(est)))
        self.estimate = est.value
        self.capacity = int((est.value + 8192) / load_factor)
        log(f"# capacity estimate {est.value} -> {self.capacity} cells in {time.time() - t1:.1f}s")
        t1 = time.time()
        bld = _lib.Builder(os.path.join(out_dir, "taxo.k2d"), self.capacity, k=k, l=l, spaced_seed_mask=mask,
                           toggle_mask=toggle, block_size=block, subblock_size=subblock, dna=False)
        lib.k2g_build_insert_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64]
        _lib._check(lib.k2g_build_insert_device(bld._h, self.d_prot, self.seq_off.ctypes.data,
                                                self.seq_taxid.ctypes.data, n_seq))
        info = bld.info()
        self.build_s = time.time() - t1
        log(f"# GPU protein build: {info['size']} cells filled of {info['capacity']} "
            f"({info['size'] / info['capacity']:.3f}) in {self.build_s:.1f}s")
        self.build_check = None
        if verify:
            t1 = time.time()
            self.build_check = bld.verify_device(self.d_prot, self.seq_off, self.seq_taxid)
            self.build_check["seconds"] = round(time.time() - t1, 2)
            log(f"# build verification: {self.build_check}")
        t1 = time.time()
        bld.write(os.path.join(out_dir, "hash.k2d"))
        bld.close()
        opts = np.zeros(8, np.uint64)
        opts[0], opts[1], opts[2], opts[3] = k, l, mask, toggle
        raw = bytearray(opts.tobytes())
        raw[32] = 0                                   # protein database
        raw[48:52] = (1).to_bytes(4, "little")        # revcom_version
        with open(os.path.join(out_dir, "opts.k2d"), "wb") as f:
            f.write(bytes(raw))
        log(f"# table written in {time.time() - t1:.1f}s")

    def free_proteomes(self):
        self.hip.free(self.d_prot)
        self.d_prot = None

    def make_reads(self, n, *, read_len=150, sub=0.01, seed=None):
        """n single-end nucleotide reads back-translated from random windows
        of the proteomes (uniformly random synonymous codons, `sub`
        substitutions, either strand)."""
        seed = self.seed if seed is None else seed
        off = np.zeros(2 * n + 1, np.uint64)
        ends = np.arange(1, n + 1, dtype=np.uint64) * np.uint64(read_len)
        off[1:2 * n:2] = ends
        off[2:2 * n + 1:2] = ends
        d_seq = self.hip.malloc(n * read_len)
        rc = self.synth.k2s_protein_reads(seed, n, read_len, self.n_taxa, self.n_proteins, self.d_prot,
                                          self.seq_off.ctypes.data, sub, d_seq)
        if rc:
            raise RuntimeError(f"read generation failed ({rc})")
        return ReadSet(self.hip, d_seq, off, np.zeros(n, np.uint8))
