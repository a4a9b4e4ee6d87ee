import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built product library")
    config.addinivalue_line("markers", "slow: takes more than a few seconds")


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def fci6():
    from psoci_amd import inputs
    return inputs.workload("fci6")


@pytest.fixture(scope="session")
def fci10():
    from psoci_amd import inputs
    return inputs.workload("fci10")


@pytest.fixture(scope="session")
def oracle_fci6(fci6):
    import oracle
    cfg, ints, _ = fci6
    sp = oracle.OracleSpace(cfg, ints)
    sp.build_h()
    return sp


@pytest.fixture(scope="session")
def oracle_fci10(fci10):
    import oracle
    cfg, ints, _ = fci10
    sp = oracle.OracleSpace(cfg, ints)
    sp.build_h()
    return sp


@pytest.fixture
def rng():
    return np.random.default_rng(1234)
