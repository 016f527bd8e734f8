"""Shared test setup.

`-m "not gpu"` runs the oracle checks, host logic and ABI-surface tests on the
CPU; `-m gpu` runs the parity tests that call libpps through its C ABI on an
MI355X.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "parallelparticleswarms.jl_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (calls libpps kernels)")


@pytest.fixture(scope="session")
def oracle():
    import pps_oracle

    pps_oracle.lib()
    return pps_oracle


@pytest.fixture(scope="session")
def pps():
    import ppswarm

    ppswarm.lib()
    return ppswarm
