"""The host-side table check (tests/helpers/buildcheck.py), which checks
k2g_build_verify on the GPU, run on tables the oracle's deterministic build
writes (byte-identical to the reference build_db on the golden inputs):
every inserted minimizer is found, each cell holds the LCA of the
minimizers ending there, no filled cell is left over.  A planted change to
one cell, or one minimizer left out of the enumeration, fails it."""
import os

import numpy as np

from helpers import DNA_SEED_TEMPLATE, Oracle, read_fasta, taxonomy_external_ids
from helpers import buildcheck as bc


def _inputs(d):
    import kraken2_amd as k2
    opts = np.fromfile(os.path.join(d, "opts.k2d"), np.uint64, 4)
    k, l, mask, toggle = (int(x) for x in opts)
    assert mask == k2.expand_spaced_seed_mask(DNA_SEED_TEMPLATE, 2)
    raw = open(os.path.join(d, "taxo.k2d"), "rb").read()
    nn = int(np.frombuffer(raw, np.uint64, 1, 8)[0])
    parent = np.frombuffer(raw, np.uint64, 7 * nn, 32).reshape(nn, 7)[:, 0].astype(np.int64)
    e2i = {e: i for i, e in enumerate(taxonomy_external_ids(os.path.join(d, "taxo.k2d")))}
    names, genomes = read_fasta(os.path.join(d, "library.fna"))
    idmap = dict(x.split("\t") for x in open(os.path.join(d, "seqid2taxid.map")).read().split("\n") if x)
    taxa = [e2i[int(idmap[nm])] for nm in names]
    keys, tax = bc.inserted_pairs(Oracle(), genomes, taxa, k, l, mask, toggle, 16384, 2048)
    uk, own = bc.key_lcas(keys, tax, parent)
    return parent, uk, own


def test_oracle_built_table_passes(synth_db):
    parent, uk, own = _inputs(synth_db)
    cells, kb, vb, size = bc.read_table(os.path.join(synth_db, "hash.k2d"))
    r = bc.check_table(cells, kb, vb, size, parent, uk, own)
    assert r["keys"] > 10000
    for f in ("missing", "not_covering", "unexplained", "orphan_cells"):
        assert r[f] == 0, (f, r[f])
    assert r["differ"] <= r["shared_keys"]
    assert (r["values"][~r["shared_key_mask"]].astype(np.int64) == own[~r["shared_key_mask"]]).all()


def test_planted_mistakes_fail(synth_db):
    parent, uk, own = _inputs(synth_db)
    cells, kb, vb, size = bc.read_table(os.path.join(synth_db, "hash.k2d"))
    cells = cells.copy()
    vmask = (1 << vb) - 1
    # a cell holding the wrong taxon (its parent instead)
    val, where = bc.probe(cells, kb, vb, uk[:1000])
    j = int(np.flatnonzero((val != 0) & (parent[val] != 0))[0])
    c = int(cells[where[j]])
    cells[where[j]] = np.uint32((c & ~vmask) | int(parent[c & vmask]))
    r = bc.check_table(cells, kb, vb, size, parent, uk, own)
    assert r["unexplained"] == 1
    # a minimizer missing from the enumeration leaves its cell an orphan
    cells, kb, vb, size = bc.read_table(os.path.join(synth_db, "hash.k2d"))
    r = bc.check_table(cells, kb, vb, size, parent, uk[1:], own[1:])
    assert r["orphan_cells"] + r["unexplained"] >= 1


def test_reference_built_protein_table_passes(tmp_path):
    """The golden protein table, written by the compiled reference's
    build_db -X from the library make_golden.py generates (k=15, l=12,
    '*' appended to every protein), passes the check."""
    import importlib.util
    import random
    import subprocess
    from helpers import K2SYNTH
    here = os.path.dirname(os.path.abspath(__file__))
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(here, "golden", "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    inp = str(tmp_path / "inputs")
    subprocess.run([K2SYNTH, "db-inputs", "--out", inp, "--seed", "21", "--levels", "2,2,3", "--ranks",
                    "phylum,genus,species", "--genome-len", "10", "--sub", "0"], check=True, capture_output=True)
    taxa = [line.split("\t")[1].strip() for line in open(os.path.join(inp, "seqid2taxid.map"))]
    _, _, prots = mg.protein_library(taxa, random.Random(23))
    golden = os.path.join(here, "golden", "prot")
    raw = open(os.path.join(golden, "taxo.k2d"), "rb").read()
    nn = int(np.frombuffer(raw, np.uint64, 1, 8)[0])
    parent = np.frombuffer(raw, np.uint64, 7 * nn, 32).reshape(nn, 7)[:, 0].astype(np.int64)
    e2i = {e: i for i, e in enumerate(taxonomy_external_ids(os.path.join(golden, "taxo.k2d")))}
    keys, tax = bc.inserted_pairs(Oracle(), prots, [e2i[int(t)] for t in taxa for _ in range(8)], 15, 12,
                                  (1 << 48) - 1, 0xE37E28C4271B5A2D, 16384, 2048, dna=False)
    uk, own = bc.key_lcas(keys, tax, parent)
    cells, kb, vb, size = bc.read_table(os.path.join(golden, "hash.k2d"))
    r = bc.check_table(cells, kb, vb, size, parent, uk, own)
    for f in ("missing", "not_covering", "unexplained", "orphan_cells"):
        assert r[f] == 0, (f, r[f])
    assert r["keys"] == size - r["shared_keys"] + r["shared_cells"]
