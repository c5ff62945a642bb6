import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP engine")


@pytest.fixture(scope="session")
def multipliers():
    """Engine default multiplier table (build-time cache)."""
    import ymc3d_amd as y

    return y.default_multipliers()


@pytest.fixture(scope="session")
def small_table():
    import ymc3d_amd as y

    return y.generate_multipliers(4097)


@pytest.fixture(scope="session")
def engine(multipliers):
    import ymc3d_amd as y

    e = y.Engine(0, multipliers)
    yield e
    e.close()
