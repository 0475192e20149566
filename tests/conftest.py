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


@pytest.fixture(scope="session", autouse=True)
def _torch_runtime_first(request):
    """Some GPU tests hand torch device buffers to the library.  PyTorch-ROCm
    bundles its own HIP runtime, which only starts if it starts before the
    library's; import torch first in sessions that run GPU tests."""
    if any(item.get_closest_marker("gpu") for item in request.session.items):
        import torch  # noqa: F401  (otmap_amd._lib starts its runtime before loading the library)
    yield
