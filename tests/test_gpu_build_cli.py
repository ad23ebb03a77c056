"""k2g_build_db / k2g_estimate_capacity: the reference's build_db and
estimate_capacity command lines with the table built on the GPU, checked
against the golden toy database the compiled reference built from the
reference's own data/ (tests/golden/toy_inputs, tests/golden/toy):

* the estimate equals the reference's (66816);
* taxo.k2d is byte-identical and opts.k2d field-identical;
* hash.k2d has the same header, every minimizer the build windows emit Gets
  the same value from both tables, and classifying the golden reads with the
  GPU-built table reproduces the golden kraken output byte for byte."""
import os
import struct
import subprocess

import numpy as np
import pytest

from helpers import DNA_SEED_TEMPLATE, Oracle

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
TOY_IN = os.path.join(HERE, "golden", "toy_inputs")
TOY = os.path.join(HERE, "golden", "toy")


def toy_library():
    recs = []
    for name in sorted(os.listdir(TOY_IN)):
        if not name.endswith(".fa"):
            continue
        head, seq = None, []
        for line in open(os.path.join(TOY_IN, name)):
            line = line.rstrip("\n")
            if line.startswith(">"):
                if head is not None:
                    recs.append((head, "".join(seq)))
                head, seq = line[1:], []
            else:
                seq.append(line)
        if head is not None:
            recs.append((head, "".join(seq)))
    return "".join(f">{h}\n{s}\n" for h, s in recs).encode(), recs


def _bin(name):
    import kraken2_amd as k2
    return os.path.join(k2.BIN_DIR, name)


def test_estimate_capacity_cli():
    lib, _ = toy_library()
    r = subprocess.run([_bin("k2g_estimate_capacity"), "-k", "35", "-l", "31", "-S", DNA_SEED_TEMPLATE],
                       input=lib, capture_output=True)
    assert r.returncode == 0, r.stderr.decode()
    assert int(r.stdout) == 66816


def test_build_db_cli_matches_reference_build(tmp_path):
    lib, recs = toy_library()
    d = str(tmp_path)
    cmd = [_bin("k2g_build_db"), "-k", "35", "-l", "31", "-S", DNA_SEED_TEMPLATE, "-H", f"{d}/hash.k2d",
           "-t", f"{d}/taxo.k2d", "-o", f"{d}/opts.k2d", "-n", TOY_IN, "-m", os.path.join(TOY_IN, "seqid2taxid.map"),
           "-c", "107154", "-p", "1", "-B", "16384", "-b", "2048", "-r", "0"]
    r = subprocess.run(cmd, input=lib, capture_output=True)
    assert r.returncode == 0, r.stderr.decode()
    assert open(f"{d}/taxo.k2d", "rb").read() == open(os.path.join(TOY, "taxo.k2d"), "rb").read()
    mine, ref = open(f"{d}/opts.k2d", "rb").read(), open(os.path.join(TOY, "opts.k2d"), "rb").read()
    assert mine[:33] == ref[:33] and mine[40:60] == ref[40:60]  # fields; padding bytes are not data
    h_mine = struct.unpack_from("<4Q", open(f"{d}/hash.k2d", "rb").read(32))
    h_ref = struct.unpack_from("<4Q", open(os.path.join(TOY, "hash.k2d"), "rb").read(32))
    assert h_mine == h_ref  # capacity, cells filled, key bits, value bits

    # every minimizer of the library Gets the same value from both tables
    orc = Oracle()
    k, l = 35, 31
    import kraken2_amd as k2
    mask = k2.expand_spaced_seed_mask(DNA_SEED_TEMPLATE, 2)
    keys = []
    for _, s in recs:
        sb = s.encode()
        cap = len(sb) + 8
        amb = np.zeros(cap, np.uint8)
        mins = np.zeros(cap, np.uint64)
        n = orc.lib.k2cpu_scan(k, l, mask, 0xE37E28C4271B5A2D, 1, 0, -1, sb, len(sb), amb.ctypes.data,
                               mins.ctypes.data, cap)
        keys.append(mins[:n][amb[:n] == 0])
    keys = np.unique(np.concatenate(keys))
    h_ref_db = orc.open(TOY)
    h_mine_db = orc.open(d)
    assert np.array_equal(orc.get_batch(h_mine_db, keys), orc.get_batch(h_ref_db, keys))

    # the GPU-built table classifies the golden reads exactly as the reference's
    out = os.path.join(d, "paired.kraken")
    r = subprocess.run([_bin("k2g_classify"), "-H", f"{d}/hash.k2d", "-t", f"{d}/taxo.k2d", "-o", f"{d}/opts.k2d",
                        "-P", "-T", "0.1", "-g", "2", "-O", out, os.path.join(TOY, "reads_1.fq"),
                        os.path.join(TOY, "reads_2.fq")],
                       capture_output=True)
    assert r.returncode == 0, r.stderr.decode()
    assert open(out).read() == open(os.path.join(TOY, "paired.kraken")).read()


def test_protein_build_db_cli_matches_reference_build(tmp_path):
    """k2g_estimate_capacity -X / k2g_build_db -X on the library the
    compiled reference built the golden protein database from (k=15, l=12):
    the same estimate and header, byte-identical taxo.k2d, field-identical
    opts.k2d; every minimizer the build inserts is found in both tables with
    each cell the LCA of the minimizers ending there and equal values where
    neither table merges compacted keys; and the GPU-built table classifies
    the golden reads (translated search, -R -K report) byte for byte as the
    reference did."""
    import importlib.util
    import random
    import kraken2_amd as k2
    from helpers import K2SYNTH, taxonomy_external_ids
    from helpers import buildcheck as bc
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(HERE, "golden", "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    d = str(tmp_path)
    inp = os.path.join(d, "inputs")
    r = subprocess.run([K2SYNTH, "db-inputs", "--out", inp, "--seed", "21", "--levels", "2,2,3", "--ranks",
                        "phylum,genus,species", "--genome-len", "10", "--sub", "0"], capture_output=True)
    assert r.returncode == 0, r.stderr.decode()
    taxa = [line.split("\t")[1].strip() for line in open(os.path.join(inp, "seqid2taxid.map"))]
    fa, mp, prots = mg.protein_library(taxa, random.Random(23))
    open(os.path.join(d, "seqid2taxid.map"), "w").write(mp)
    golden = os.path.join(HERE, "golden", "prot")
    r = subprocess.run([_bin("k2g_estimate_capacity"), "-k", "15", "-l", "12", "-S", "1" * 12, "-X"],
                       input=fa.encode(), capture_output=True)
    assert r.returncode == 0, r.stderr.decode()
    cap = int((int(r.stdout) + 8192) / 0.7)
    h_ref = struct.unpack_from("<4Q", open(os.path.join(golden, "hash.k2d"), "rb").read(32))
    assert cap == h_ref[0]
    cmd = [_bin("k2g_build_db"), "-k", "15", "-l", "12", "-S", "1" * 12, "-X", "-H", f"{d}/hash.k2d",
           "-t", f"{d}/taxo.k2d", "-o", f"{d}/opts.k2d", "-n", os.path.join(golden, "taxonomy"), "-m",
           os.path.join(d, "seqid2taxid.map"), "-c", str(cap), "-p", "1", "-B", "16384", "-b", "2048", "-r", "0"]
    r = subprocess.run(cmd, input=fa.encode(), capture_output=True)
    assert r.returncode == 0, r.stderr.decode()
    assert open(f"{d}/taxo.k2d", "rb").read() == open(os.path.join(golden, "taxo.k2d"), "rb").read()
    mine, ref = open(f"{d}/opts.k2d", "rb").read(), open(os.path.join(golden, "opts.k2d"), "rb").read()
    assert mine[:33] == ref[:33] and mine[40:60] == ref[40:60]
    assert struct.unpack_from("<4Q", open(f"{d}/hash.k2d", "rb").read(32)) == h_ref

    raw = open(f"{d}/taxo.k2d", "rb").read()
    nn = int(np.frombuffer(raw, np.uint64, 1, 8)[0])
    parent = np.frombuffer(raw, np.uint64, 7 * nn, 32).reshape(nn, 7)[:, 0].astype(np.int64)
    e2i = {e: i for i, e in enumerate(taxonomy_external_ids(f"{d}/taxo.k2d"))}
    keys, tax = bc.inserted_pairs(Oracle(), prots, [e2i[int(t)] for t in taxa for _ in range(8)], 15, 12,
                                  (1 << 48) - 1, 0xE37E28C4271B5A2D, 16384, 2048, dna=False)
    uk, own = bc.key_lcas(keys, tax, parent)
    res = {}
    for name, dd in (("gpu", d), ("reference", golden)):
        cells, kb, vb, size = bc.read_table(os.path.join(dd, "hash.k2d"))
        res[name] = bc.check_table(cells, kb, vb, size, parent, uk, own)
        for f in ("missing", "not_covering", "unexplained", "orphan_cells"):
            assert res[name][f] == 0, (name, f, res[name][f])
    alone = ~res["gpu"]["shared_key_mask"] & ~res["reference"]["shared_key_mask"]
    assert alone.sum() > 0.99 * len(uk)
    assert np.array_equal(res["gpu"]["values"][alone], res["reference"]["values"][alone])

    out, rep = os.path.join(d, "single.kraken"), os.path.join(d, "single.kreport")
    r = subprocess.run([_bin("k2g_classify"), "-H", f"{d}/hash.k2d", "-t", f"{d}/taxo.k2d", "-o", f"{d}/opts.k2d",
                        "-g", "0", "-R", rep, "-K", "-O", out, os.path.join(golden, "reads_1.fa")],
                       capture_output=True)
    assert r.returncode == 0, r.stderr.decode()
    assert open(out).read() == open(os.path.join(golden, "single.kraken")).read()
    assert open(rep).read() == open(os.path.join(golden, "single.kreport")).read()
