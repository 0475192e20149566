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


REFAPI = os.path.join(ROOT, "tests", "cpp", "_build", "refapi_example")


@pytest.mark.skipif(not os.path.exists(REFAPI), reason="refapi_example is built only where Eigen headers exist")
def test_cpp_reference_signature_caller_vs_oracle(oracle, tmp_path):
    """tests/cpp/refapi_example: a caller written against the reference's own
    signatures (Eigen types, surface_mesh::Surface_mesh, points(),
    Vertex handles), apps/barycenters.cpp's sequence for two densities plus
    transport_cost.  Forward meshes within 1e-9 of the oracle's solves;
    inverse meshes, the halfway mesh and its raster bit-exact against the
    oracle on the binary's own meshes (the image scale to the mean's sum
    order); costs per face bit-exact, total 1e-13."""
    import subprocess

    n = 64
    imgs = [D.gaussian_mixture(n, 41), D.gaussian_mixture(n, 42)]
    paths = [tmp_path / f"d{k}.f64" for k in range(2)]
    for p, d in zip(paths, imgs):
        d.astype(np.float64).tofile(p)  # face order j + i n == MatrixXd (row j, column i) column-major
    out = str(tmp_path / "out")
    print("running refapi_example", flush=True)
    r = subprocess.run([REFAPI, str(n), str(paths[0]), str(paths[1]), out], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    V = (n + 1) ** 2
    ident = np.concatenate([np.repeat(np.arange(n + 1.0), n + 1), np.tile(np.arange(n + 1.0), n + 1)]) / n
    grid_pts = np.stack([ident[:V], ident[V:]], 1)
    invs = []
    for k in range(2):
        fwd = np.fromfile(f"{out}.fwd{k}.f64")
        o = oracle.solve(n, imgs[k])
        assert np.abs(fwd - o["fwd"]).max() < 1e-9
        oi, _, _ = oracle.inverse(n, fwd, o["density"], grid_pts)
        inv = np.fromfile(f"{out}.inv{k}.f64")
        assert np.array_equal(inv, np.concatenate([oi[:, 0], oi[:, 1]]), equal_nan=True)
        invs.append(inv)
    mid = np.fromfile(f"{out}.mid.f64")
    assert np.array_equal(mid, (1. - 0.5) * invs[0] + 0.5 * invs[1])
    raster = oracle.rasterize(n, mid, n, n)
    expected = raster * ((0.5 * imgs[0].mean() + 0.5 * imgs[1].mean()) / raster.mean())
    img = np.fromfile(f"{out}.img.f64").reshape(n, n)
    assert np.allclose(img, expected, rtol=1e-13, atol=0)
    cost = np.fromfile(f"{out}.cost.f64")
    fwd0 = np.fromfile(f"{out}.fwd0.f64")
    oc, ocf = oracle.transport_cost(n, ident, fwd0, oracle.solve(n, imgs[0])["density"], per_face=True)
    assert abs(cost[0] - oc) <= 1e-13 * abs(oc)
    assert np.allclose(cost[1:], ocf, rtol=1e-13, atol=0)
