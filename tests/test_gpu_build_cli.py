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
