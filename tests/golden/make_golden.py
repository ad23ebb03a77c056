"""Writes the golden fixtures under tests/golden/ from the COMPILED REFERENCE.

Run in the development container only (it needs /root/reference and the
reference binaries that oracle/Makefile builds into oracle/_ref/):

    make -C oracle ref && python tests/golden/make_golden.py

Inputs are this repository's own seeded generators plus the reference's
sample data (tests/golden/toy_inputs, copied from the reference's data/
directory); every expected output is produced by the reference binaries or
by harness programs linking the reference sources (oracle/ref_harness).
"""
import hashlib
import json
import os
import random
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref")
TOY_IN = os.path.join(HERE, "toy_inputs")
DNA_TEMPLATE = "1" * 17 + "01" * 7


def run(cmd, stdin=None, **kw):
    r = subprocess.run(cmd, stdin=stdin, capture_output=True, **kw)
    if r.returncode != 0:
        sys.stderr.write(r.stderr.decode(errors="replace"))
        raise SystemExit(f"failed: {cmd}")
    return r


def fasta_records(path):
    recs = []
    with open(path) as f:
        for line in f:
            line = line.rstrip("\n")
            if line.startswith(">"):
                recs.append([line[1:], []])
            elif recs:
                recs[-1][1].append(line)
    return [(h, "".join(s)) for h, s in recs]


def build_db(out, library, nodes_dir, idmap, k, l, template, protein=False, load=0.7,
             block=16384, subblock=2048):
    os.makedirs(out, exist_ok=True)
    flags = ["-X"] if protein else []
    with open(library, "rb") as f:
        est = int(run([os.path.join(REF, "estimate_capacity"), "-k", str(k), "-l", str(l), "-S", template,
                       *flags], stdin=f).stdout)
    cap = int((est + 8192) / load)
    with open(library, "rb") as f:
        run([os.path.join(REF, "build_db"), "-k", str(k), "-l", str(l), "-S", template, *flags,
             "-H", f"{out}/hash.k2d", "-t", f"{out}/taxo.k2d", "-o", f"{out}/opts.k2d", "-n", nodes_dir,
             "-m", idmap, "-c", str(cap), "-p", "1", "-B", str(block), "-b", str(subblock), "-r", "0"],
            stdin=f)
    return est, cap


def mutate(rng, s, sub=0.01, amb=0.004, lower=0.01):
    s = list(s)
    for j in range(len(s)):
        r = rng.random()
        if r < sub:
            s[j] = rng.choice("ACGT")
        elif r < sub + amb:
            s[j] = rng.choice("NNNRYKMSWnx.")
        elif r < sub + amb + lower:
            s[j] = s[j].lower()
    return "".join(s)


def revcomp(s):
    t = {"A": "T", "C": "G", "G": "C", "T": "A", "a": "t", "c": "g", "g": "c", "t": "a"}
    return "".join(t.get(c, "N") for c in reversed(s))


def interleave(r1, r2, out):
    with open(r1) as a, open(r2) as b, open(out, "w") as o:
        la, lb = a.readlines(), b.readlines()
        for i in range(0, len(la), 4):
            o.writelines(la[i:i + 4] + lb[i:i + 4])


def toy():
    d = os.path.join(HERE, "toy")
    os.makedirs(d, exist_ok=True)
    lib = os.path.join(d, "library.fna.tmp")
    idmap = os.path.join(HERE, "toy_inputs", "seqid2taxid.map")
    genomes = []
    with open(lib, "w") as lf, open(idmap, "w") as mf:
        for name in sorted(os.listdir(TOY_IN)):
            if not name.endswith(".fa"):
                continue
            for h, s in fasta_records(os.path.join(TOY_IN, name)):
                sid = h.split()[0]
                taxid = sid.split("|")[1]
                lf.write(f">{h}\n{s}\n")
                mf.write(f"{sid}\t{taxid}\n")
                genomes.append(s)
    est, cap = build_db(d, lib, TOY_IN, idmap, 35, 31, DNA_TEMPLATE)
    os.remove(lib)
    rng = random.Random(42)
    # paired reads (FASTQ with varied qualities) and single edge-case reads
    with open(os.path.join(d, "reads_1.fq"), "w") as f1, open(os.path.join(d, "reads_2.fq"), "w") as f2:
        for i in range(2000):
            g = genomes[rng.randrange(len(genomes))]
            frag = max(150, int(rng.gauss(300, 30)))
            frag = min(frag, len(g))
            p = rng.randrange(len(g) - frag + 1)
            fr = g[p:p + frag]
            if rng.random() < 0.5:
                fr = revcomp(fr)
            L = min(100 + rng.randrange(60), len(fr))
            a = mutate(rng, fr[:L])
            b = mutate(rng, revcomp(fr[-L:]))
            qa = "".join(chr(33 + rng.randrange(2, 41)) for _ in a)
            qb = "".join(chr(33 + rng.randrange(2, 41)) for _ in b)
            f1.write(f"@frag{i}/1 c{i}\n{a}\n+\n{qa}\n")
            f2.write(f"@frag{i}/2\n{b}\n+\n{qb}\n")
    with open(os.path.join(d, "edge.fa"), "w") as f:
        n = 0
        for L in list(range(0, 80)) + [100, 149, 150, 151, 300, 1000, 5000]:
            g = genomes[rng.randrange(len(genomes))]
            p = rng.randrange(max(1, len(g) - L))
            s = g[p:p + L]
            f.write(f">len{L}_{n}\n{s}\n")
            n += 1
        base = genomes[7][1000:1150]
        for pos in [0, 1, 2, 3, 4, 5, 30, 31, 33, 34, 35, 36, 60, 100, 114, 115, 116, 145, 149]:
            s = base[:pos] + "N" + base[pos + 1:]
            f.write(f">N_at_{pos}\n{s}\n")
        for gap in [1, 2, 3, 4, 5, 6, 29, 30, 31, 33, 34, 35, 36]:
            s = base[:50] + "N" + base[51:50 + gap] + "N" + base[51 + gap:]
            f.write(f">NN_gap_{gap}\n{s}\n")
        f.write(f">allN\n{'N' * 150}\n>lower\n{base.lower()}\n>mixed\n{base[:75].lower() + base[75:]}\n")
        f.write(f">iupac\n{base[:40] + 'RYKMSWBDHV' + base[50:]}\n")
        f.write(f">multi_line\n{base[:60]}\n{base[60:120]}\n{base[120:]}\n")
        f.write(f">crlf\r\n{base}\r\n")
        f.write(f">withcomment some comment here\n{base}\n")
        f.write(f">random\n{''.join(rng.choice('ACGT') for _ in range(150))}\n")
        for i in range(300):
            g = genomes[rng.randrange(len(genomes))]
            L = rng.randrange(20, 400)
            p = rng.randrange(max(1, len(g) - L))
            s = mutate(rng, g[p:p + L], sub=0.02, amb=0.01)
            if rng.random() < 0.5:
                s = revcomp(s)
            f.write(f">rand{i}\n{s}\n")
    # interleaved copy of the pairs for -S (tests rebuild it the same way)
    interleave(os.path.join(d, "reads_1.fq"), os.path.join(d, "reads_2.fq"),
               os.path.join(d, "interleaved.fq.tmp"))
    # reference classify runs
    cl = os.path.join(REF, "classify")
    db = ["-H", f"{d}/hash.k2d", "-t", f"{d}/taxo.k2d", "-o", f"{d}/opts.k2d"]
    runs = {
        "edge_g0": ["-T", "0", "-g", "0", "-R", "REPORT", "-K", f"{d}/edge.fa"],
        "edge_g2_names": ["-T", "0", "-g", "2", "-n", f"{d}/edge.fa"],
        "edge_conf05": ["-T", "0.5", "-g", "0", "-R", "REPORT", "-z", f"{d}/edge.fa"],
        "edge_quick": ["-q", "-g", "2", f"{d}/edge.fa"],
        "paired": ["-P", "-T", "0.1", "-g", "2", "-R", "REPORT", "-K", f"{d}/reads_1.fq", f"{d}/reads_2.fq"],
        "paired_q20": ["-P", "-Q", "20", "-g", "2", f"{d}/reads_1.fq", f"{d}/reads_2.fq"],
        "interleaved": ["-S", "-g", "2", "-R", "REPORT", f"{d}/interleaved.fq.tmp"],
        "single_cu": ["-g", "2", "-C", "CLASSIFIED", "-U", "UNCLASSIFIED", f"{d}/reads_1.fq"],
    }
    for name, args in runs.items():
        args = [a.replace("REPORT", f"{d}/{name}.kreport").replace("CLASSIFIED", f"{d}/{name}.C.fq")
                .replace("UNCLASSIFIED", f"{d}/{name}.U.fq") for a in args]
        out = run([cl, *db, "-p", "2", *args]).stdout
        with open(os.path.join(d, f"{name}.kraken"), "wb") as f:
            f.write(out)
    os.remove(os.path.join(d, "interleaved.fq.tmp"))
    digests = {}
    for fn in sorted(os.listdir(d)):
        if fn.endswith((".C.fq", ".U.fq")):
            digests[fn] = hashlib.sha256(open(os.path.join(d, fn), "rb").read()).hexdigest()
            os.remove(os.path.join(d, fn))
    with open(os.path.join(d, "sequence_outputs.json"), "w") as f:
        json.dump(digests, f, indent=1, sort_keys=True)
    with open(os.path.join(d, "README"), "w") as f:
        f.write(f"Toy DB from the reference's data/ (estimate {est}, capacity {cap}); built and\n"
                "classified by the compiled reference; see tests/golden/make_golden.py.\n")


def scanner():
    rng = random.Random(7)
    cases = []
    dna_mask = int(DNA_TEMPLATE, 2)
    # expand 2 bits per template char (src/utilities.cc:14-27)
    m2 = 0
    for i in range(31, -1, -1):
        m2 <<= 2
        if (dna_mask >> i) & 1:
            m2 |= 3
    configs = [(35, 31, m2, 1), (35, 31, 0, 1), (31, 31, 0, 1), (31, 15, 0, 1), (8, 4, 0, 1),
               (15, 12, 0, 0), (15, 15, 0, 0), (10, 5, 0, 0)]
    for k, l, mask, dna in configs:
        alpha = "ACGT" if dna else "ACDEFGHIKLMNPQRSTVWY*"
        for t in range(40):
            L = rng.choice([0, 1, l - 1, l, k - 1, k, k + 1, 50, 64, 65, 128, 200, rng.randrange(0, 300)])
            s = [rng.choice(alpha) for _ in range(max(0, L))]
            for j in range(len(s)):
                r = rng.random()
                if r < 0.03:
                    s[j] = rng.choice("NXBZJ.-" if not dna else "NRYnx.")
                elif r < 0.06:
                    s[j] = s[j].lower()
            s = "".join(s) or "-"
            starts = [(0, -1)]
            if t % 3 == 0 and len(s) > 2:
                a = rng.randrange(0, len(s))
                starts.append((a, -1))
                starts.append((a, rng.randrange(a, len(s) + 1)))
            for a, b in starts:
                cases.append(f"{k} {l} {mask:x} e37e28c4271b5a2d {dna} {a} {b} {s}")
    inp = "\n".join(cases) + "\n"
    out = run([os.path.join(REF, "scan_dump")], input=inp.encode()).stdout.decode()
    with open(os.path.join(HERE, "scan_cases.txt"), "w") as f:
        f.write(inp)
    with open(os.path.join(HERE, "scan_expected.txt"), "w") as f:
        f.write(out)


def scanner_revcom0():
    """DNA scans under revcom_version 0 (databases built before the reverse
    complement fix, src/mmscanner.cc:193-215)."""
    rng = random.Random(17)
    cases = []
    for k, l in ((35, 31), (31, 31), (31, 15), (8, 4), (25, 20)):
        for t in range(30):
            L = rng.choice([0, l - 1, l, k, k + 1, 64, 65, 130, rng.randrange(0, 300)])
            s = "".join(rng.choice("ACGTACGTACGTNacgt") for _ in range(max(0, L))) or "-"
            cases.append(f"{k} {l} 0 e37e28c4271b5a2d 1 0 -1 {s}")
    inp = "\n".join(cases) + "\n"
    out = run([os.path.join(REF, "scan_dump"), "0"], input=inp.encode()).stdout.decode()
    with open(os.path.join(HERE, "scan_rc0_cases.txt"), "w") as f:
        f.write(inp)
    with open(os.path.join(HERE, "scan_rc0_expected.txt"), "w") as f:
        f.write(out)


def hll():
    rng = random.Random(13)
    lines = []
    # sparse only, around the 1024 threshold, dense, merges in every order
    for n in [0, 1, 2, 10, 1000, 1023, 1024, 1025, 1026, 3000, 6000]:
        lines.append("R")
        vals = [rng.getrandbits(62) for _ in range(n)]
        if n > 5:
            vals += vals[: n // 7]  # duplicates
        lines.append("I " + " ".join(map(str, vals)))
        lines.append("P")
    for na, nb in [(500, 600), (1100, 10), (10, 1100), (2000, 3000), (0, 50), (50, 0), (1024, 1024)]:
        lines.append("R")
        lines.append("I " + " ".join(str(rng.getrandbits(62)) for _ in range(na)))
        lines.append("J " + " ".join(str(rng.getrandbits(62)) for _ in range(nb)))
        lines.append("M")
        lines.append("P")
    # the exact switch rule: the last insert a duplicate vs a new value
    base = [rng.getrandbits(62) for _ in range(1024)]
    lines += ["R", "I " + " ".join(map(str, base + [base[3]])), "P"]
    lines += ["R", "I " + " ".join(map(str, base + [rng.getrandbits(62)])), "P"]
    small = [str(i) for i in range(0, 5000, 3)]
    lines += ["R", "I " + " ".join(small), "P"]
    inp = "\n".join(lines) + "\n"
    out = run([os.path.join(REF, "hll_dump")], input=inp.encode()).stdout.decode()
    with open(os.path.join(HERE, "hll_cases.txt"), "w") as f:
        f.write(inp)
    with open(os.path.join(HERE, "hll_expected.txt"), "w") as f:
        f.write(out)


def translate():
    rng = random.Random(17)
    cases = []
    for L in list(range(0, 12)) + [30, 31, 32, 150, 151, 152]:
        for t in range(3):
            s = [rng.choice("ACGT") for _ in range(L)]
            for j in range(len(s)):
                r = rng.random()
                if r < 0.05:
                    s[j] = rng.choice("NRx")
                elif r < 0.1:
                    s[j] = s[j].lower()
            cases.append("".join(s) or "-")
    inp = "\n".join(cases) + "\n"
    out = run([os.path.join(REF, "translate_dump")], input=inp.encode()).stdout.decode()
    with open(os.path.join(HERE, "translate_cases.txt"), "w") as f:
        f.write(inp)
    with open(os.path.join(HERE, "translate_expected.txt"), "w") as f:
        f.write(out)


CODONS = {}
AMINO = "ACDEFGHIKLMNPQRSTVWY"


def protein_library(taxa, rng):
    """Genus-level ancestral proteins (8 per genus of three taxa) and species
    copies at 80% identity: (FASTA text, seqid2taxid.map text, proteins).
    The golden protein database was built by the compiled reference from
    exactly this library: rng = random.Random(23), taxa as k2synth
    db-inputs --seed 21 --levels 2,2,3 lists them."""
    anc = [["".join(rng.choice(AMINO) for _ in range(rng.randrange(150, 400))) for _ in range(8)]
           for _ in range(4)]
    fa, mp, prots = [], [], []
    for ti, tax in enumerate(taxa):
        for pi, p in enumerate(anc[ti // 3]):
            q = "".join(c if rng.random() < 0.8 else rng.choice(AMINO) for c in p)
            name = f"p{ti}_{pi}"
            fa.append(f">{name}\n{q}\n")
            mp.append(f"{name}\t{tax}\n")
            prots.append(q)
    return "".join(fa), "".join(mp), prots


def protein_db():
    """Small translated-search DB (k=15, l=12) and back-translated reads."""
    d = os.path.join(HERE, "prot")
    tmp = os.path.join(d, "inputs")
    os.makedirs(tmp, exist_ok=True)
    run([os.path.join(ROOT, "kraken2_amd", "bin", "k2synth"), "db-inputs", "--out", tmp, "--seed", "21",
         "--levels", "2,2,3", "--ranks", "phylum,genus,species", "--genome-len", "10", "--sub", "0"])
    taxa = [line.split("\t")[1].strip() for line in open(os.path.join(tmp, "seqid2taxid.map"))]
    rng = random.Random(23)
    code = "KKNNRRSSTTTTIMIIEEDDGGGGAAAAVVVVQQHHRRRRPPPPLLLL**YY*WCCSSSSLLFF"
    bases = "AGCT"
    back = {}
    for i, c in enumerate(code):
        back.setdefault(c, []).append(bases[i >> 4] + bases[(i >> 2) & 3] + bases[i & 3])
    # genus-level ancestral proteins, species = 80% identity copies
    lib = os.path.join(tmp, "library.faa")
    idmap = os.path.join(tmp, "seqid2taxid.map")
    fa, mp, prots = protein_library(taxa, rng)
    with open(lib, "w") as lf, open(idmap, "w") as mf:
        lf.write(fa)
        mf.write(mp)
    est, cap = build_db(d, lib, os.path.join(tmp, "taxonomy"), idmap, 15, 12, "1" * 12, protein=True)
    shutil.copytree(os.path.join(tmp, "taxonomy"), os.path.join(d, "taxonomy"), dirs_exist_ok=True)
    shutil.rmtree(tmp)
    with open(os.path.join(d, "reads_1.fa"), "w") as f1, open(os.path.join(d, "reads_2.fa"), "w") as f2:
        for i in range(800):
            p = prots[rng.randrange(len(prots))]
            n_aa = rng.choice([0, 1, 5, 20, 50, 60])
            s = p[rng.randrange(max(1, len(p) - n_aa)):][:n_aa]
            dna = "".join(rng.choice(back[c]) for c in s)
            dna = mutate(rng, dna, sub=0.01, amb=0.005)
            if rng.random() < 0.5:
                dna = revcomp(dna)
            if rng.random() < 0.1:
                dna = "".join(rng.choice("ACGT") for _ in range(len(dna)))
            f1.write(f">t{i}/1\n{dna}\n")
            s2 = p[:60]
            f2.write(f">t{i}/2\n{revcomp(''.join(rng.choice(back[c]) for c in s2))}\n")
    cl = os.path.join(REF, "classify")
    db = ["-H", f"{d}/hash.k2d", "-t", f"{d}/taxo.k2d", "-o", f"{d}/opts.k2d"]
    for name, args in {"single": ["-g", "0", "-R", f"{d}/single.kreport", "-K", f"{d}/reads_1.fa"],
                       "paired": ["-P", "-T", "0.2", "-g", "2", f"{d}/reads_1.fa", f"{d}/reads_2.fa"]}.items():
        with open(os.path.join(d, f"{name}.kraken"), "wb") as f:
            f.write(run([cl, *db, "-p", "2", *args]).stdout)


# merge-classifications cases: (taxonomy dir, input 1, input 2, options)
MERGE_CASES = {
    "paired": ("toy", "merge/a_paired", "merge/b_paired", []),
    "paired_conf_names": ("toy", "merge/a_paired", "merge/b_paired", ["-i", "0.3", "-n"]),
    "edge": ("toy", "merge/a_edge", "merge/b_edge", []),
    "edge_conf03": ("toy", "merge/a_edge", "merge/b_edge", ["-i", "0.3"]),
    "edge_conf1_swapped": ("toy", "merge/b_edge", "merge/a_edge", ["-i", "1.0"]),
    "edge_names_zeros": ("toy", "merge/a_edge", "merge/b_edge", ["-i", "0.05", "-n", "-z"]),
    "edge_self": ("toy", "merge/a_edge", "merge/a_edge", ["-i", "0.7"]),
    "quick": ("toy", "merge/a_quick", "merge/b_quick", ["-i", "0.5"]),
    "prot_self": ("prot", "prot/single", "prot/single", ["-i", "0.2", "-n"]),
    "prot_paired_self": ("prot", "prot/paired", "prot/paired", []),
}


def merge():
    """Two databases over the two halves of the toy genomes (same k, l and
    seed), the toy reads classified against each by the reference classify,
    and the reference merge of those outputs under the full toy taxonomy (the
    merged taxonomy k2 would make from both seqid2taxid maps).  Expected
    outputs are recorded as digests."""
    d = os.path.join(HERE, "merge")
    os.makedirs(d, exist_ok=True)
    files = sorted(f for f in os.listdir(TOY_IN) if f.endswith(".fa"))
    toy_d = os.path.join(HERE, "toy")
    for part, sel in (("a", files[0::2]), ("b", files[1::2])):
        db = os.path.join(d, f"db_{part}.tmp")
        os.makedirs(db, exist_ok=True)
        lib, idmap = os.path.join(db, "library.fna"), os.path.join(db, "seqid2taxid.map")
        with open(lib, "w") as lf, open(idmap, "w") as mf:
            for name in sel:
                for h, s in fasta_records(os.path.join(TOY_IN, name)):
                    sid = h.split()[0]
                    lf.write(f">{h}\n{s}\n")
                    mf.write(f"{sid}\t{sid.split('|')[1]}\n")
        build_db(db, lib, TOY_IN, idmap, 35, 31, DNA_TEMPLATE)
        cl = [os.path.join(REF, "classify"), "-H", f"{db}/hash.k2d", "-t", f"{db}/taxo.k2d", "-o",
              f"{db}/opts.k2d", "-p", "2"]
        for name, args in {"paired": ["-P", "-g", "2", f"{toy_d}/reads_1.fq", f"{toy_d}/reads_2.fq"],
                           "edge": ["-g", "0", f"{toy_d}/edge.fa"],
                           "quick": ["-q", "-g", "2", f"{toy_d}/edge.fa"]}.items():
            with open(os.path.join(d, f"{part}_{name}.kraken"), "wb") as f:
                f.write(run([*cl, *args]).stdout)
        shutil.rmtree(db)
    expected = {}
    for case, (tax, in1, in2, opts) in MERGE_CASES.items():
        out, hdr, rep = (os.path.join(d, f"{case}.{x}.tmp") for x in ("out", "hdr", "rep"))
        run([os.path.join(REF, "merge"), "-T", "2", "-t", os.path.join(HERE, tax, "taxo.k2d"), "-o", out,
             "-c", hdr, "-r", rep, *opts, os.path.join(HERE, in1 + ".kraken"), os.path.join(HERE, in2 + ".kraken")])
        expected[case] = {}
        for key, path in (("output", out), ("headers", hdr), ("report", rep)):
            data = open(path, "rb").read()
            expected[case][key] = {"sha256": hashlib.sha256(data).hexdigest(), "lines": data.count(b"\n")}
            os.remove(path)
    with open(os.path.join(d, "expected.json"), "w") as f:
        json.dump(expected, f, indent=1, sort_keys=True)
    with open(os.path.join(d, "README"), "w") as f:
        f.write("Inputs: the toy reads classified by the reference against two databases built from the two\n"
                "halves of the toy genomes; expected.json: digests of the reference merge's output, classified\n"
                "headers and report for each case of MERGE_CASES; see tests/golden/make_golden.py.\n")


if __name__ == "__main__":
    if not os.path.exists(os.path.join(REF, "classify")):
        raise SystemExit("build the reference first: make -C oracle ref")
    if sys.argv[1:] == ["merge"]:  # only the merge fixtures
        merge()
        raise SystemExit(0)
    toy()
    scanner()
    scanner_revcom0()
    hll()
    translate()
    protein_db()
    merge()
    print("golden fixtures written")
