"""Shared test fixtures.  GPU tests are marked ``@pytest.mark.gpu`` and run only on a machine with a GPU."""
import json
import lzma
import os
import subprocess
import sys
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parent.parent
GOLDEN = REPO / "tests" / "golden"
sys.path.insert(0, str(REPO))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


def _make_oracle():
    lib = REPO / "oracle" / "_build" / "liboracle.so"
    cli = REPO / "oracle" / "_build" / "tn93_oracle"
    src = REPO / "oracle" / "tn93_oracle.cpp"
    if not (lib.exists() and cli.exists()) or min(lib.stat().st_mtime, cli.stat().st_mtime) < src.stat().st_mtime:
        subprocess.run(["make", "-s", "-C", str(REPO / "oracle"), "oracle"], check=True)
    return cli, lib


@pytest.fixture(scope="session")
def oracle_bins():
    return _make_oracle()


@pytest.fixture(scope="session")
def fuzz_cases():
    d = json.loads(lzma.decompress((GOLDEN / "fuzz_cases.json.xz").read_bytes()))
    return d["cases"]


@pytest.fixture(scope="session")
def hiv8766(tmp_path_factory):
    """The reference's real 8,766-sequence HIV-1 pol input (dev + holdout concatenated), as one FASTA file."""
    p = tmp_path_factory.mktemp("real") / "hiv8766.fa"
    data = b"".join(lzma.decompress((GOLDEN / "real" / f).read_bytes())
                    for f in ["real_hiv_pol_dev.fa.xz", "real_hiv_pol_holdout.fa.xz"])
    p.write_bytes(data)
    return p


@pytest.fixture(scope="session")
def sars2(tmp_path_factory):
    """The reference's real SARS-CoV-2 sets (500 and 459 genomes x 29,903 columns) as FASTA files, and their pins:
    (dict name -> path, pins)."""
    pins = json.loads((GOLDEN / "real" / "sars2_pins.json").read_text())
    d = tmp_path_factory.mktemp("sars2")
    paths = {}
    for key, f in pins["files"].items():
        paths[key] = d / f"{key}.fa"
        paths[key].write_bytes(lzma.decompress((GOLDEN / "real" / f).read_bytes()))
    return paths, pins["pins"]


def sars2_argv(flags, paths):
    """A pin's flags with {second} replaced by the second file's path."""
    return [str(paths[f.strip("{}")]) if f.startswith("{") else f for f in flags]
