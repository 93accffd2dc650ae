"""Test configuration.  `gpu` tests need an MI355X and the built HIP library;
everything else runs on the CPU (oracle, fixtures, host logic, C ABI load)."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: runs HIP kernels on an MI355X")


@pytest.fixture(scope="session")
def solver():
    import ispan_amd as I
    s = I.Solver(0)
    yield s
    s.close()
