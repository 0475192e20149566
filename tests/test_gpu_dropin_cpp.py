"""GPU parity of the C++ drop-in headers (include/otmap/*.h).

tests/cpp/dropin_example is a C++ caller written against the reference's API
(GridBasedTransportSolver::init/solve, apply_inverse_map, rasterize_image,
Surface_mesh OFF output), built by __graft_entry__.build().  Its outputs are
checked against the oracle: positions within 1e-9 (FP64) / 1e-6 (mixed) and
iterations within 1; inverse images and raster bit-exact against the oracle
run on the binary's own forward mesh; the OFF file parses back to the mesh.
"""
import os
import subprocess

import numpy as np
import pytest

from otmap_amd import densities as D

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "_build", "dropin_example")


def read_off(path):
    with open(path) as f:
        assert f.readline().strip() == "OFF"
        nv, nf, _ = map(int, f.readline().split())
        xy = np.array([list(map(float, f.readline().split()))[:2] for _ in range(nv)])
        quads = np.array([list(map(int, f.readline().split())) for _ in range(nf)])
    return xy, quads


@pytest.mark.parametrize("n,mixed", [(64, False), (128, True)])
def test_cpp_dropin_vs_oracle(n, mixed, oracle, tmp_path):
    assert os.path.exists(EXE), "tests/cpp/_build/dropin_example missing: run __graft_entry__.build()"
    rho = D.gaussian_mixture(n, 7)
    pts = np.random.default_rng(3).uniform(0, 1, (5000, 2))
    dens_path, pts_path, out = tmp_path / "rho.f64", tmp_path / "pts.f64", str(tmp_path / "out")
    rho.astype(np.float64).tofile(dens_path)
    pts.astype(np.float64).tofile(pts_path)
    cmd = [EXE, str(n), str(dens_path), str(pts_path), out] + (["mixed"] if mixed else [])
    print("running", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr

    V = (n + 1) ** 2
    fwd = np.fromfile(out + ".fwd.f64")
    dens = np.fromfile(out + ".dens.f64")
    inv = np.fromfile(out + ".inv.f64").reshape(-1, 2)
    img = np.fromfile(out + ".img.f64").reshape(n, n)
    iters = int(open(out + ".stats.txt").read().split()[0])

    o = oracle.solve(n, rho)
    assert abs(iters - o["iters"]) <= 1, (iters, o["iters"])
    assert np.abs(fwd - o["fwd"]).max() < (1e-6 if mixed else 1e-9)
    assert np.allclose(dens, o["density"], rtol=1e-14, atol=0)

    oinv, _, _ = oracle.inverse(n, fwd, dens, pts)
    assert np.array_equal(inv, oinv, equal_nan=True)
    assert np.array_equal(img, oracle.rasterize(n, fwd, n, n, dens))

    off_xy, quads = read_off(out + ".fwd.off")
    assert off_xy.shape == (V, 2) and quads.shape == (n * n, 5)
    assert np.abs(off_xy[:, 0] - fwd[:V]).max() <= 5e-11 and np.abs(off_xy[:, 1] - fwd[V:]).max() <= 5e-11
    i, j = divmod(np.arange(n * n), n)
    assert np.array_equal(quads[:, 1], j + i * (n + 1)) and np.array_equal(quads[:, 3], j + 1 + (i + 1) * (n + 1))
