"""CPU checks of the C ABI: the library loads and exports every declared symbol.

No compute entry point is called (no GPU here); parity runs in the gpu tests.
"""
import ctypes
import os

import pytest

from otmap_amd import _lib


def test_library_built():
    assert os.path.exists(_lib.LIB_PATH), "run __graft_entry__.build() first"


def test_exports_every_declared_symbol():
    L = ctypes.CDLL(_lib.LIB_PATH)
    declared = _lib.declared_symbols()
    assert len(declared) >= 10
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, missing


def test_python_binding_covers_declared_symbols():
    assert set(_lib.declared_symbols()) <= set(_lib._SIGNATURES)


def test_options_defaults_match_reference():
    # SolverOptions defaults, otlib/otsolver_2dgrid.h:37-43
    o = _lib.OtmOptions()
    _lib.lib().otm_options_init(o)
    assert o.beta == _lib.OTM_BETA_CONJUGATE_JACOBIAN
    assert o.max_iter == 1000 and o.threshold == 1e-7
    assert o.max_ratio == pytest.approx(1.7976931348623157e308)
    assert o.precision == _lib.OTM_PRECISION_FP64
    # beyond the reference: BF16x3 options and the FP64 transform (fast DCTs where supported)
    assert o.refine_steps == 1 and o.mixed_min_n == 1024 and o.transform == _lib.OTM_TRANSFORM_FFT


def test_errors_without_device_are_reported():
    # null handles are rejected before any HIP call
    assert _lib.lib().otm_solver_init(None, 16) != 0
    assert b"null" in _lib.lib().otm_last_error()


def test_cpp_dropin_headers_build_and_link():
    # the C++ caller of include/otmap/*.h links against the library; without
    # arguments it prints its usage before touching a device
    import subprocess

    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tests", "cpp", "_build",
                       "dropin_example")
    assert os.path.exists(exe), "run __graft_entry__.build() first"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=30)
    assert r.returncode == 2 and "usage" in r.stderr


def test_off_round_trip(tmp_path):
    import numpy as np

    from otmap_amd import GridMesh

    n = 5
    m = GridMesh(n, np.random.default_rng(0).uniform(0, 1, 2 * (n + 1) ** 2))
    path = str(tmp_path / "m.off")
    m.write(path)
    r = GridMesh.read(path)
    assert r.n == n and np.abs(r.xy - m.xy).max() <= 5e-11


EIGEN_DIR = "/root/reference/extern"


@pytest.mark.skipif(not os.path.exists(os.path.join(EIGEN_DIR, "Eigen", "Core")),
                    reason="Eigen headers (the reference's vendored copy) not mounted")
@pytest.mark.parametrize("src", ["signatures.cpp", "refapi_example.cpp"])
def test_reference_signatures_compile(src):
    """The drop-in headers compile callers written against the reference's
    own declarations (otlib/otsolver_2dgrid.h:52-58, otlib/transport_map.h:
    29-70, otlib/utils/rasterizer.h:16-22, Surface_mesh points()/position()),
    with the reference's include layout and Eigen; signatures.cpp binds every
    entry point to a pointer of the reference's exact type."""
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-Wno-deprecated-copy",
                        "-Wno-class-memaccess", "-Wno-int-in-bool-context", "-I", os.path.join(root, "include", "otmap"),
                        "-isystem", EIGEN_DIR, os.path.join(root, "tests", "cpp", src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]


def test_dropin_headers_compile_without_eigen():
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-DOTMAP_NO_EIGEN",
                        "-I", os.path.join(root, "include"), os.path.join(root, "tests", "cpp", "dropin_example.cpp")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
