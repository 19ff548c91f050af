import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (ROCm device)")
    config.addinivalue_line("markers", "slow: long-running")


@pytest.fixture(scope="session")
def ubq_pdb():
    with open(os.path.join(GOLDEN, "1ubq.pdb")) as f:
        return f.read()


@pytest.fixture(scope="session")
def ljlk_term():
    from tmol_amd.score.ljlk_energy_term import LJLKEnergyTerm

    return LJLKEnergyTerm()


def load_term_baseline(name, term="ljlk"):
    """tmol's YAML term baselines (tmol/tests/data/term_baselines/
    LJLKEnergyTerm, ElecEnergyTerm) -> numpy array (term, pose[, res, res])."""
    import numpy
    import yaml

    with open(os.path.join(GOLDEN, term + "_term_baselines", name + ".yaml")) as f:
        data = yaml.safe_load(f)

    def unpack(d):
        if isinstance(d, dict):
            return [unpack(v) for v in d.values()]
        return d

    return numpy.array(unpack(data), dtype=numpy.float64)
