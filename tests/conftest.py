import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libotmap_amd.so")


@pytest.fixture(scope="session")
def oracle():
    from oracle import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def solver_factory():
    from otmap_amd import GridBasedTransportSolver
    cache = {}

    def make(n):
        if n not in cache:
            s = GridBasedTransportSolver(0)
            s.set_verbose_level(0)
            s.init(n)
            cache[n] = s
        return cache[n]
    return make


def pytest_collection_modifyitems(config, items):
    # unmarked tests run on the CPU; gpu tests carry the marker explicitly
    pass
