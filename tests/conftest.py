import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built libk2g.so")


@pytest.fixture(scope="session")
def synth_db(tmp_path_factory):
    """A small synthetic nucleotide DB built by the oracle's deterministic build."""
    from helpers import make_synthetic_db
    d = tmp_path_factory.mktemp("synth_db")
    make_synthetic_db(str(d), seed=11, levels="2,2,3", genome_len=20000)
    return str(d)


@pytest.fixture(scope="session")
def synth_db_rc0(synth_db, tmp_path_factory):
    """synth_db with opts.k2d marked revcom_version 0 (the reverse complement
    of databases built before the fix), for scanner parity on that path."""
    import shutil
    d = tmp_path_factory.mktemp("synth_db_rc0")
    for f in ("hash.k2d", "taxo.k2d", "opts.k2d", "library.fna"):
        shutil.copy(os.path.join(synth_db, f), os.path.join(str(d), f))
    raw = bytearray(open(os.path.join(str(d), "opts.k2d"), "rb").read())
    raw[48:52] = (0).to_bytes(4, "little")
    open(os.path.join(str(d), "opts.k2d"), "wb").write(bytes(raw))
    return str(d)


@pytest.fixture(scope="session")
def edge_tables(synth_db, tmp_path_factory):
    """synth_db's scanner and taxonomy with random hash tables at the probe's
    edges: 8-bit compact keys (frequent key matches along long chains), one
    table with a single empty cell and a capacity not a multiple of 4 (probes
    wrap past the last cell), one with no empty cell, and one below 64 cells."""
    from helpers import write_custom_table
    base = tmp_path_factory.mktemp("edge_tables")
    specs = {"one_empty_102": (102, 1), "full_128": (128, 0), "small_37": (37, 15), "sparse_1001": (1001, 400)}
    return {name: write_custom_table(synth_db, os.path.join(str(base), name), cap, 8, empty, seed=i + 1)
            for i, (name, (cap, empty)) in enumerate(specs.items())}
