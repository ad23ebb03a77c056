"""GPU table construction (atomicCAS + LCA merge) against the oracle's
deterministic build on the same inputs and window layout.

Cell placement may differ (it is order dependent); the key -> value map may
only differ where two minimizers share a compacted key, which a table of this
size does not produce: every minimizer the build windows yield must Get the
same value from both tables, and both tables must hold the same number of
cells."""
import os
import random

import numpy as np
import pytest

from helpers import DNA_SEED_TEMPLATE, ORACLE_BIN, Oracle, make_synthetic_db, read_fasta, run

pytestmark = pytest.mark.gpu


def _windows_keys(orc, seq, k, l, mask, B, b):
    """Every non-ambiguous minimizer the deterministic build's windows emit."""
    keys = []
    sb = seq.encode()
    L = len(sb)
    for j in range(0, L, B):
        bfin = min(j + B + k - 1, L)
        for i in range(j, bfin, b):
            fin = min(i + b + k - 1, bfin)
            cap = fin - i + 8
            amb = np.zeros(cap, np.uint8)
            mins = np.zeros(cap, np.uint64)
            n = orc.lib.k2cpu_scan(k, l, mask, 0xE37E28C4271B5A2D, 1, i, fin, sb, L, amb.ctypes.data,
                                   mins.ctypes.data, cap)
            keys.append(mins[:n][amb[:n] == 0])
    return np.concatenate(keys) if keys else np.zeros(0, np.uint64)


@pytest.mark.parametrize("B,b", [(16384, 2048), (5000, 700), (1000, 1)])
def test_gpu_build_matches_oracle_build(tmp_path, B, b):
    import kraken2_amd as k2
    d = str(tmp_path)
    make_synthetic_db(d, seed=17, levels="2,2,3", genome_len=6000 if b == 1 else 20000, block=B, subblock=b)
    cpu_cells = np.fromfile(os.path.join(d, "hash.k2d"), dtype=np.uint32)[8:]
    hdr = np.fromfile(os.path.join(d, "hash.k2d"), dtype=np.uint64, count=4)
    capacity, size, kb, vb = (int(x) for x in hdr)
    names, genomes = read_fasta(os.path.join(d, "library.fna"))
    idmap = dict(line.split("\t") for line in open(os.path.join(d, "seqid2taxid.map")).read().splitlines())
    db_cpu = k2.Db.from_dir(d)  # for the external -> internal id map
    ext = db_cpu.external_ids
    e2i = {int(e): i for i, e in enumerate(ext)}
    taxid = np.array([e2i[int(idmap[n])] for n in names], np.uint32)
    seq, off = b"", np.zeros(len(genomes) + 1, np.uint64)
    seq = "".join(genomes).encode()
    off[1:] = np.cumsum([len(g) for g in genomes])
    mask = k2.expand_spaced_seed_mask(DNA_SEED_TEMPLATE, 2)
    bld = k2.Builder(os.path.join(d, "taxo.k2d"), capacity, k=35, l=31, spaced_seed_mask=mask,
                     block_size=B, subblock_size=b)
    bld.insert(seq, off, taxid)
    info = bld.info()
    assert (info["key_bits"], info["value_bits"]) == (kb, vb)
    assert info["size"] == size
    gpu_path = os.path.join(d, "gpu_hash.k2d")
    bld.write(gpu_path)
    bld.close()
    gpu_cells = np.fromfile(gpu_path, dtype=np.uint32)[8:]
    assert np.count_nonzero(gpu_cells) == np.count_nonzero(cpu_cells)
    # Get of every inserted minimizer, through the GPU probe on the GPU-built
    # table and the oracle probe on the CPU-built table
    orc = Oracle()
    keys = np.concatenate([_windows_keys(orc, g, 35, 31, mask, B, b) for g in genomes])
    keys = np.unique(keys)
    h = orc.open(d)
    want = orc.get_batch(h, keys)
    assert (want != 0).all()
    gdir = tmp_path / "gpu"
    gdir.mkdir()
    os.replace(gpu_path, gdir / "hash.k2d")
    for fn in ("taxo.k2d", "opts.k2d"):
        os.link(os.path.join(d, fn), gdir / fn)
    db_gpu = k2.Db.from_dir(str(gdir))
    got = db_gpu.get_batch(keys)
    assert np.array_equal(got, want)
    rng = np.random.default_rng(3)
    miss = rng.integers(0, 2**62, 100000, dtype=np.uint64)
    assert np.array_equal(db_gpu.get_batch(miss), orc.get_batch(h, miss))
