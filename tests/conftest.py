"""Shared test configuration: the ``gpu`` marker, import paths and helpers.

``-m "not gpu"`` tests run on any host (oracle, host logic, C-ABI symbols);
``-m gpu`` tests need a ROCm device and the built HIP library, and compare the
HIP extraction against the exact oracle.
"""

import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
PACKAGE = ROOT / "exact-depth-peeling_amd"
GOLDEN = ROOT / "tests" / "golden"
for p in (str(ROOT), str(PACKAGE)):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU and the built HIP library")


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle

    oracle.lib()
    return oracle
