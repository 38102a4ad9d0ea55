import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the libmdk C ABI)")


@pytest.fixture(scope="session")
def golden_dir():
    return GOLDEN


@pytest.fixture(scope="session")
def synth_pair(tmp_path_factory):
    """Small synthetic receptor/ligand pair (PDB files) from the project generator."""
    from megadock_amd import synth
    d = tmp_path_factory.mktemp("synth")
    rec = str(d / "rec.pdb")
    lig = str(d / "lig.pdb")
    synth.write_pdb(rec, 120, seed=11)
    synth.write_pdb(lig, 40, seed=12)
    return rec, lig
