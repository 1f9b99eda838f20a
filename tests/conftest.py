import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


@pytest.fixture(scope="session")
def c1_system():
    """Config-1 input: seeded (seed 1) 20-residue extended chain."""
    from origami_amd import system as S
    return S.seeded_chain(1, 20)


@pytest.fixture(scope="session")
def c2_system():
    """Config-2 input: seeded (seed 2) 80-residue extended chain."""
    from origami_amd import system as S
    return S.seeded_chain(2, 80)


@pytest.fixture(scope="session")
def c3_system():
    """Config-3 replica: seeded (seed 3) 10-residue extended chain."""
    from origami_amd import system as S
    return S.seeded_chain(3, 10)
