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


def test_build_cli_writes_reference_taxonomy_before_any_device_call(tmp_path):
    """k2g_build_db converts nodes.dmp/names.dmp to taxo.k2d on the host
    before it touches the GPU: on a machine without one the taxonomy file is
    already written (byte-identical to the reference build's) when the table
    step fails."""
    import subprocess
    import kraken2_amd as k2
    here = os.path.dirname(os.path.abspath(__file__))
    toy_in = os.path.join(here, "golden", "toy_inputs")
    d = str(tmp_path)
    exe = os.path.join(k2.BIN_DIR, "k2g_build_db")
    r = subprocess.run([exe, "-k", "35", "-l", "31", "-H", f"{d}/hash.k2d", "-t", f"{d}/taxo.k2d", "-o",
                        f"{d}/opts.k2d", "-n", toy_in, "-m", os.path.join(toy_in, "seqid2taxid.map"), "-c", "107154"],
                       input=b"", capture_output=True)
    assert os.path.exists(f"{d}/taxo.k2d"), r.stderr.decode()
    assert open(f"{d}/taxo.k2d", "rb").read() == open(os.path.join(here, "golden", "toy", "taxo.k2d"), "rb").read()
