"""Test helpers: synthetic DB construction and the CPU oracle's ctypes binding.

The oracle (oracle/k2cpu) is the checker only; product code never imports it.
"""
import ctypes as C
import os
import random
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
ORACLE_BIN = os.path.join(ROOT, "oracle", "bin")
K2SYNTH = os.path.join(ROOT, "kraken2_amd", "bin", "k2synth")
DNA_SEED_TEMPLATE = "1" * 17 + "01" * 7  # kraken2-build default for l=31, 7 spaces


def run(cmd, **kw):
    r = subprocess.run(cmd, capture_output=True, text=True, **kw)
    if r.returncode != 0:
        raise RuntimeError(f"{cmd} failed ({r.returncode}): {r.stderr}")
    return r


def make_synthetic_db(d, seed=1, levels="2,2,3", ranks=None, genome_len=20000, sub=0.05,
                      k=35, l=31, template=DNA_SEED_TEMPLATE, load_factor=0.7,
                      block=16384, subblock=2048, tools=None):
    """Writes taxonomy + genomes with k2synth and builds hash/taxo/opts with the
    oracle, or with tools = (estimate_capacity, build_db) executables of the
    same command lines (the GPU build front-ends)."""
    est_exe, build_exe = tools or (os.path.join(ORACLE_BIN, "k2cpu_estimate_capacity"),
                                   os.path.join(ORACLE_BIN, "k2cpu_build_db"))
    nlev = len(levels.split(","))
    if ranks is None:
        ranks = ",".join((["phylum", "class", "order", "family", "genus", "species"])[-nlev:])
    run([K2SYNTH, "db-inputs", "--out", d, "--seed", str(seed), "--levels", levels,
         "--ranks", ranks, "--genome-len", str(genome_len), "--sub", str(sub)])
    lib = os.path.join(d, "library.fna")
    with open(lib, "rb") as f:
        est = int(run([est_exe, "-k", str(k), "-l", str(l),
                       "-S", template], stdin=f).stdout.strip())
    cap = int((est + 8192) / load_factor)
    with open(lib, "rb") as f:
        run([build_exe, "-k", str(k), "-l", str(l), "-S", template,
             "-H", os.path.join(d, "hash.k2d"), "-t", os.path.join(d, "taxo.k2d"),
             "-o", os.path.join(d, "opts.k2d"), "-n", os.path.join(d, "taxonomy"),
             "-m", os.path.join(d, "seqid2taxid.map"), "-c", str(cap), "-B", str(block),
             "-b", str(subblock), "-r", "0"], stdin=f)
    return d


def write_custom_table(src, dst, capacity, key_bits, empty_cells, seed):
    """Copies DB directory src to dst with hash.k2d replaced by a random table
    of `capacity` cells: key_bits-bit compact keys, `empty_cells` empty cells
    at random positions and random valid taxa elsewhere.  Long probe chains,
    probes that wrap past the last cell and tables with no empty cell (every
    probe of an absent key visits all cells) are then common."""
    import shutil
    os.makedirs(dst, exist_ok=True)
    for f in ("taxo.k2d", "opts.k2d", "library.fna"):
        shutil.copy(os.path.join(src, f), os.path.join(dst, f))
    tx = open(os.path.join(src, "taxo.k2d"), "rb").read()
    nodes = int(np.frombuffer(tx[8:16], np.uint64)[0])
    vb = 32 - key_bits
    assert nodes < (1 << vb)
    rng = np.random.default_rng(seed)
    vals = rng.integers(1, nodes, capacity, dtype=np.uint64)
    keys = rng.integers(0, 1 << key_bits, capacity, dtype=np.uint64)
    cells = ((keys << np.uint64(vb)) | vals).astype(np.uint32)
    cells[rng.choice(capacity, empty_cells, replace=False)] = 0
    hdr = np.array([capacity, capacity - empty_cells, key_bits, vb], np.uint64)
    with open(os.path.join(dst, "hash.k2d"), "wb") as f:
        f.write(hdr.tobytes())
        f.write(cells.tobytes())
    return dst


def read_fasta(path):
    names, seqs = [], []
    with open(path) as f:
        for line in f:
            line = line.rstrip("\n")
            if line.startswith(">"):
                names.append(line[1:].split()[0])
                seqs.append([])
            elif seqs:
                seqs[-1].append(line)
    return names, ["".join(s) for s in seqs]


def revcomp(s):
    t = {"A": "T", "C": "G", "G": "C", "T": "A"}
    return "".join(t.get(c, "N") for c in reversed(s))


def chimeric_reads(genomes, n, seed, seg=40, length=2000):
    """Reads stitched from short segments of random genomes: many hitlist runs."""
    rng = random.Random(seed)
    out = []
    for _ in range(n):
        parts = []
        while sum(map(len, parts)) < length:
            g = genomes[rng.randrange(len(genomes))]
            p = rng.randrange(len(g) - seg)
            parts.append(g[p:p + seg])
        out.append("".join(parts)[:length])
    return out


def edge_reads(genomes, n, seed, min_len=0, max_len=400):
    """Reads sampled from genomes with the parity edge cases mixed in:
    N / IUPAC / lowercase bases, lengths around k, reverse strands, random reads."""
    rng = random.Random(seed)
    out = []
    for i in range(n):
        kind = rng.random()
        L = rng.choice([0, 1, 2, 3, 30, 34, 35, 36, 40, 64, 65, 100, 150, 151, 250, 400,
                        rng.randint(min_len, max_len)])
        if kind < 0.1:
            s = "".join(rng.choice("ACGT") for _ in range(L))
        else:
            g = genomes[rng.randrange(len(genomes))]
            L = min(L, len(g))
            p = rng.randrange(len(g) - L + 1)
            s = g[p:p + L]
            if rng.random() < 0.5:
                s = revcomp(s)
            s = list(s)
            for j in range(len(s)):
                r = rng.random()
                if r < 0.01:
                    s[j] = rng.choice("ACGT")
                elif r < 0.015:
                    s[j] = rng.choice("NNNRYKMSWnx-")
                elif r < 0.03:
                    s[j] = s[j].lower()
            s = "".join(s)
        out.append(s)
    return out


class Oracle:
    """ctypes binding of oracle/bin/libk2cpu.so (the CPU restatement)."""

    def __init__(self):
        lib = C.CDLL(os.path.join(ORACLE_BIN, "libk2cpu.so"))
        vp = C.c_void_p
        lib.k2cpu_db_open.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p]
        lib.k2cpu_db_open.restype = vp
        lib.k2cpu_db_close.argtypes = [vp]
        lib.k2cpu_get_batch.argtypes = [vp, vp, vp, C.c_long]
        lib.k2cpu_probe_lines.argtypes = [vp, vp, vp, C.c_long]
        lib.k2cpu_db_capacity.argtypes = [vp]
        lib.k2cpu_db_capacity.restype = C.c_uint64
        lib.k2cpu_classify.argtypes = [vp, C.c_int, C.c_int, C.c_double, C.c_int, vp, vp, C.c_long,
                                       vp, vp, vp, vp, vp, vp, C.c_long]
        lib.k2cpu_classify.restype = C.c_long
        lib.k2cpu_scan.argtypes = [C.c_long, C.c_long, C.c_uint64, C.c_uint64, C.c_int, C.c_long, C.c_long,
                                   C.c_char_p, C.c_long, vp, vp, C.c_long]
        lib.k2cpu_scan.restype = C.c_long
        lib.k2cpu_scan_rc.argtypes = [C.c_long, C.c_long, C.c_uint64, C.c_uint64, C.c_int, C.c_int, C.c_long,
                                      C.c_long, C.c_char_p, C.c_long, vp, vp, C.c_long]
        lib.k2cpu_scan_rc.restype = C.c_long
        lib.k2cpu_translate.argtypes = [C.c_char_p, C.c_long, vp, C.c_long]
        lib.k2cpu_translate.restype = C.c_long
        lib.k2cpu_hll_script.argtypes = [C.c_char_p, vp, C.c_long]
        lib.k2cpu_hll_script.restype = C.c_long
        lib.k2cpu_run_files.argtypes = [vp, C.c_int, C.c_int, C.c_double, C.c_int, C.c_char_p, C.c_int, C.c_int,
                                        C.c_char_p, C.c_char_p, C.c_char_p]
        lib.k2cpu_run_files.restype = C.c_int
        self.lib = lib

    def run_files(self, h, file1, file2=None, *, kraken_out, report=None, kmer_data=False, zeros=False,
                  confidence=0.0, min_hit_groups=0, threads=8):
        """The oracle's classify driver over FASTA/FASTQ files with the handle's
        index (blocks, counters, kraken output and kreport)."""
        enc = lambda x: x.encode() if x else None  # noqa: E731
        rc = self.lib.k2cpu_run_files(h, threads, int(file2 is not None), confidence, min_hit_groups, enc(report),
                                      int(kmer_data), int(zeros), enc(kraken_out), enc(file1), enc(file2))
        if rc:
            raise RuntimeError(f"oracle classify driver failed ({rc})")

    def open(self, d):
        h = self.lib.k2cpu_db_open(os.path.join(d, "hash.k2d").encode(), os.path.join(d, "taxo.k2d").encode(),
                                   os.path.join(d, "opts.k2d").encode())
        if not h:
            raise RuntimeError(f"oracle could not open DB in {d}")
        return h

    def get_batch(self, h, keys):
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        out = np.zeros(keys.shape[0], dtype=np.uint32)
        self.lib.k2cpu_get_batch(h, keys.ctypes.data, out.ctypes.data, keys.shape[0])
        return out

    def classify(self, h, seq, off, paired=False, quick=False, confidence=0.0, min_hit_groups=0,
                 lines=False):
        """Per fragment: call, groups, total, lookups, lines and the hitlist text."""
        off = np.ascontiguousarray(off, dtype=np.uint64)
        n = (off.shape[0] - 1) // 2
        res = {k: np.zeros(n, np.uint32) for k in ("call", "groups", "total", "lookups")}
        res["lines"] = np.zeros(n, np.uint64)
        cap = 6 * int(off[-1]) + 64 * n + 64
        text = C.create_string_buffer(cap)
        sbuf = np.frombuffer(seq if len(seq) else b"\0", dtype=np.uint8)
        used = self.lib.k2cpu_classify(h, int(paired), int(quick), confidence, min_hit_groups,
                                       sbuf.ctypes.data, off.ctypes.data, n,
                                       res["call"].ctypes.data, res["groups"].ctypes.data,
                                       res["total"].ctypes.data, res["lookups"].ctypes.data,
                                       res["lines"].ctypes.data if lines else None, text, cap)
        if used < 0:
            raise RuntimeError("oracle text buffer too small")
        res["text"] = [t.decode() for t in text.raw[:used].split(b"\0")[:n]]
        return res


def taxonomy_external_ids(taxo_k2d):
    """External ids of a taxo.k2d's nodes (node 0 first)."""
    raw = open(taxo_k2d, "rb").read()
    n = int(np.frombuffer(raw, np.uint64, 1, 8)[0])
    nodes = np.frombuffer(raw, np.uint64, 7 * n, 32).reshape(n, 7)
    return [int(x) for x in nodes[:, 5]]


def random_kraken_pairs(ext_ids, n, seed):
    """Two kraken outputs of the same n reads as two databases would give
    them: per read the same k-mer positions, ambiguous runs, mate ("|:|")
    and frame ("-:-") borders in both, each position's taxon drawn per
    output from a few taxa of the taxonomy (0 for no hit, sometimes an
    unknown id), so LCA ties and confidence climbs are frequent.  Also "0:0"
    lines of reads with no k-mer.  Returns the two texts."""
    rng = random.Random(seed)
    known = [e for e in ext_ids[1:] if e]
    out_a, out_b = [], []

    def rle(pos):
        toks, prev, cnt = [], None, 0
        for t in pos:
            if t == prev:
                cnt += 1
                continue
            if prev is not None:
                toks.append(prev + ":" + str(cnt))
            prev, cnt = t, 1
        if prev is not None:
            toks.append(prev + ":" + str(cnt))
        return toks

    for i in range(n):
        pool = rng.sample(known, min(len(known), rng.randrange(1, 5))) + ["0"] * rng.randrange(0, 3)
        if rng.random() < 0.05:
            pool.append(str(10 ** 9 + rng.randrange(100)))  # not in the taxonomy
        r = rng.random()
        if r < 0.03:
            lists = (["0:0"], ["0:0"])
            lens = "0"
        else:
            n_seg = 2 if r < 0.35 else (6 if r < 0.45 else 1)
            border = "|:|" if n_seg == 2 else "-:-"
            la, lb, seg_lens = [], [], []
            for s in range(n_seg):
                L = rng.randrange(1, 80)
                seg_lens.append(str(L + 30))
                pa, pb = [], []
                j = 0
                while j < L:
                    if rng.random() < 0.05:
                        run_len = min(L - j, rng.randrange(1, 30))
                        pa += ["A"] * run_len
                        pb += ["A"] * run_len
                        j += run_len
                        continue
                    pa.append(str(rng.choice(pool)))
                    pb.append(str(rng.choice(pool)))
                    j += 1
                if s:
                    la.append(border)
                    lb.append(border)
                la += rle(pa)
                lb += rle(pb)
            lists = (la, lb)
            lens = "|".join(seg_lens[:2]) if n_seg == 2 else seg_lens[0]
        for out, lst in zip((out_a, out_b), lists):
            st = "C" if rng.random() < 0.7 else "U"
            call = str(rng.choice(known)) if st == "C" else "0"
            out.append(f"{st}\tread{i}\t{call}\t{lens}\t{' '.join(lst)}\n")
    return "".join(out_a), "".join(out_b)


# Kraken lines whose hit-list totals differ between two outputs of the
# same reads (second list shorter or longer than the first), with count-0,
# unknown and marker items: edge cases of the merge walk.
RAGGED_MERGE_LINES = (
    ["C\tr0\t1\t50\t211044:5 0:3\n", "C\tr1\t1\t50\t211044:5\n", "U\tr2\t0\t50\t0:0\n",
     "C\tr3\t1\t50\t999999:4 11320:2 A:3\n", "C\tr4\t1\t50\t11320:7 |:| 211044:2\n"],
    ["C\tr0\t1\t50\t211044:5\n", "C\tr1\t1\t50\t211044:2 11320:9\n", "U\tr2\t0\t50\t11320:3\n",
     "U\tr3\t0\t50\t11320:1\n", "C\tr4\t1\t50\t11320:7 |:| 211044:2\n"],
)


def large_scratch_dir(prefix, need_gb=40):
    """A temporary directory for multi-GB databases: /dev/shm when it has
    need_gb free, else the default temporary directory."""
    import shutil
    import tempfile
    try:
        if shutil.disk_usage("/dev/shm").free >= (need_gb << 30):
            return tempfile.mkdtemp(prefix=prefix, dir="/dev/shm")
    except OSError:
        pass
    return tempfile.mkdtemp(prefix=prefix)


def fragment_lists(seq, off, idx):
    """(mates 1, mates 2) of fragments idx of a packed (seq, 2n+1 offsets) set."""
    m1, m2 = [], []
    for i in idx:
        a, b, c = (int(x) for x in off[2 * i:2 * i + 3])
        m1.append(seq[a:b])
        m2.append(seq[b:c])
    return m1, m2


def compare_with_oracle(k2, db, out, orc, h, seq, off, paired_flags, *, confidence=0.0, min_hit_groups=0,
                        lines=False):
    """Compares library results `out` of fragments (seq, off, paired_flags)
    with the oracle, classifying single and paired fragments as two oracle
    runs (one reference run cannot mix them).  Checks call, hit groups,
    total k-mers and the hitlist text of every fragment, and with lines=True
    the per-fragment lookups and distinct probe lines (out must carry probe
    statistics).  Returns the number of fragments compared."""
    n = len(paired_flags)
    checked = 0
    for pflag in (0, 1):
        sel = [i for i in range(n) if int(paired_flags[i]) == pflag]
        if not sel:
            continue
        m1, m2 = fragment_lists(seq, off, sel)
        s2, o2 = k2.pack_fragments(m1, m2 if pflag else None)
        ref = orc.classify(h, s2, o2, paired=bool(pflag), confidence=confidence,
                           min_hit_groups=min_hit_groups, lines=lines)
        for j, i in enumerate(sel):
            runs = out.runs[out.run_off[i]:out.run_off[i + 1]]
            got = (int(out.call[i]), int(out.hit_groups[i]), int(out.total_kmers[i]),
                   k2.format_hitlist(runs, db.external_ids))
            exp = (int(ref["call"][j]), int(ref["groups"][j]), int(ref["total"][j]), ref["text"][j])
            assert got == exp, f"fragment {i} (paired {pflag}): {got[:3]} != {exp[:3]}"
            if lines:
                assert (int(out.lookups[i]), int(out.probe_lines[i])) == \
                       (int(ref["lookups"][j]), int(ref["lines"][j])), f"probe statistics of fragment {i}"
            checked += 1
    return checked
