"""The package's taxo.k2d writer reproduces the reference build's taxonomy
file byte for byte (golden files written by the compiled reference)."""
import os

from kraken2_amd.taxonomy import write_k2_taxonomy

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_toy_taxonomy_matches_reference(tmp_path):
    inp = os.path.join(G, "toy_inputs")
    mapped = [int(l.split("\t")[1]) for l in open(os.path.join(inp, "seqid2taxid.map")).read().splitlines()]
    out = tmp_path / "taxo.k2d"
    write_k2_taxonomy(os.path.join(inp, "nodes.dmp"), os.path.join(inp, "names.dmp"), mapped, str(out))
    assert out.read_bytes() == open(os.path.join(G, "toy", "taxo.k2d"), "rb").read()


def test_protein_taxonomy_matches_reference(tmp_path):
    tx = os.path.join(G, "prot", "taxonomy")
    # the protein DB maps every species leaf of its generated tree
    species = [int(l.split("\t|\t")[0]) for l in open(os.path.join(tx, "nodes.dmp")) if "\tspecies\t" in l]
    out = tmp_path / "taxo.k2d"
    write_k2_taxonomy(os.path.join(tx, "nodes.dmp"), os.path.join(tx, "names.dmp"), species, str(out))
    assert out.read_bytes() == open(os.path.join(G, "prot", "taxo.k2d"), "rb").read()
