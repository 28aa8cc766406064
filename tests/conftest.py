"""Shared pytest configuration.

Markers:
  gpu  -- needs an AMD GPU and the built liblbm_hip.so (run with -m gpu)

CPU tests (-m "not gpu") check the oracle against the reference's fixtures,
golden vectors and closed forms, the host logic, and that the C ABI library
loads and exports what include/lbm_hip.h declares.
"""
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and liblbm_hip.so")
    config.addinivalue_line("markers", "slow: long runs")


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle as O
    O.set_threads(1)
    return O


@pytest.fixture(scope="session")
def lbm():
    import lbm_solver_3d_amd as M
    M.load_library()
    return M
