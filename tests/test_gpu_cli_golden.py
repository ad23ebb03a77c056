"""The k2g_classify front-end (GPU path) reproduces the compiled reference's
kraken output, kreport (with -K HLL estimates) and -C/-U sequence files byte
for byte on the golden runs."""
import os

import pytest

from helpers import run
from test_oracle_golden import PROT, TOY_RUNS, run_golden_case

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "kraken2_amd", "bin", "k2g_classify")


@pytest.mark.parametrize("mode", ["0", "1"])
@pytest.mark.parametrize("name", sorted(TOY_RUNS))
def test_cli_matches_reference(name, mode, tmp_path, monkeypatch):
    monkeypatch.setenv("K2G_MODE", mode)
    run_golden_case(CLI, name, str(tmp_path))


def test_cli_small_batches_keep_block_semantics(tmp_path, monkeypatch):
    """Blocks split over many GPU batches give the same output and -K report."""
    monkeypatch.setenv("K2G_BATCH_BASES", "20000")
    run_golden_case(CLI, "paired", str(tmp_path))
    run_golden_case(CLI, "edge_g0", str(tmp_path))


@pytest.mark.parametrize("name", ["single", "paired"])
def test_cli_translated_matches_reference(name, tmp_path):
    args = {"single": ["-g", "0", "-R", str(tmp_path / "r.kreport"), "-K", f"{PROT}/reads_1.fa"],
            "paired": ["-P", "-T", "0.2", "-g", "2", f"{PROT}/reads_1.fa", f"{PROT}/reads_2.fa"]}[name]
    out = run([CLI, "-H", f"{PROT}/hash.k2d", "-t", f"{PROT}/taxo.k2d", "-o", f"{PROT}/opts.k2d", *args]).stdout
    assert out == open(os.path.join(PROT, f"{name}.kraken")).read()
    if name == "single":
        assert open(tmp_path / "r.kreport").read() == open(os.path.join(PROT, "single.kreport")).read()
