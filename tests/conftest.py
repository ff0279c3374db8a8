import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("oracle", "imexlb-1.0_amd"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(scope="session")
def r_vec():
    """The 100 perturbation values InitUP draws (initialization.f90:264).

    tests/golden/r_vec_flang.txt holds the values the Fortran runtime's
    random_number produces as the first call of a process; fall back to a
    seeded draw if the fixture is absent."""
    path = os.path.join(ROOT, "tests", "golden", "r_vec_flang.txt")
    if os.path.exists(path):
        v = np.loadtxt(path)
        assert v.shape == (100,)
        return v
    return np.random.default_rng(1234).random(100)


@pytest.fixture(scope="session")
def root():
    return ROOT
