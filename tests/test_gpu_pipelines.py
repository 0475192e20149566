"""GPU parity of the map-level functions and the reference apps' pipelines.

  transport_cost, compute_transport_cost   per-face values bit-exact against
      the oracle (same operation order); totals 1e-13 relative (sum order),
      and against the reference's golden costs
  inverse mesh, composite map               bit-exact against the oracle's
      BVH inverse of the same points
  barycentres                               interpolated meshes bit-exact (same
      lerp / bilerp order), rasters bit-exact, the mean scale 1e-13 relative,
      so float32 images within one float32 rounding of the oracle's
  stippling                                 kept points identical, in order
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from otmap_amd import (BatchSolver, GridMesh, TransportMap, grid_transport_cost, pipelines,
                       transport_cost)
from otmap_amd import densities as D

pytestmark = pytest.mark.gpu


def solved(n, seed, oracle):
    o = oracle.solve(n, D.gaussian_mixture(n, seed))
    return TransportMap(GridMesh.regular(n), GridMesh(n, o["fwd"]), o["density"]), o


@pytest.mark.parametrize("n", [16, 32])
def test_transport_cost_vs_reference_golden(n):
    g = np.load(os.path.join(GOLDEN, "raster.npz"))
    ident = GridMesh.regular(n)
    c, cf = transport_cost(ident, GridMesh(n, g[f"n{n}_fwd"]), g[f"n{n}_density"], per_face=True)
    assert abs(c - float(g[f"n{n}_cost"])) <= 1e-13 * abs(c)
    assert np.allclose(cf, g[f"n{n}_cost_face"], rtol=1e-12, atol=1e-300)


@pytest.mark.parametrize("tag", ["rand5", "rand16", "rand33", "map16", "map64"])
def test_grid_transport_cost_vs_reference_golden(tag):
    """compute_transport_cost (otlib/otsolver_2dgrid.cpp:608-640) on device
    against the compiled reference's per-face costs: bit-exact (same
    operation order), total to the sum order."""
    g = np.load(os.path.join(GOLDEN, "costs.npz"))
    n = int(g[f"{tag}_rho"].size ** 0.5)
    t, cost = grid_transport_cost(n, g[f"{tag}_disp"], g[f"{tag}_rho"])
    assert np.array_equal(cost, g[f"{tag}_cost"])
    assert abs(t - float(g[f"{tag}_total"])) <= 1e-13 * abs(t)


def test_transport_costs_vs_oracle(oracle):
    n = 64
    tm, o = solved(n, 9, oracle)
    ident = GridMesh.regular(n)
    c, cf = transport_cost(ident, tm.fwd_mesh(), tm.density(), per_face=True)
    oc, ocf = oracle.transport_cost(n, ident.xy, tm.fwd_mesh().xy, tm.density(), per_face=True)
    assert abs(c - oc) <= 1e-13 * abs(oc)
    assert np.allclose(cf, ocf, rtol=1e-13, atol=0)
    disp = tm.fwd_mesh().xy - ident.xy
    t, cost = grid_transport_cost(n, disp, tm.density())
    ot, ocost = oracle.solver_transport_cost(n, disp, tm.density())
    assert np.array_equal(cost, ocost)
    assert abs(t - ot) <= 1e-13 * abs(ot)


def test_inverse_mesh_and_composite_map(oracle):
    n = 64
    tu, ou = solved(n, 9, oracle)
    tv, ov = solved(n, 10, oracle)
    V = (n + 1) ** 2
    ident = GridMesh.regular(n)
    grid_pts = np.stack([ident.xy[:V], ident.xy[V:]], 1)
    inv = tu.inverse_mesh()
    oinv, _, _ = oracle.inverse(n, ou["fwd"], ou["density"], grid_pts)
    assert np.array_equal(inv.xy, np.concatenate([oinv[:, 0], oinv[:, 1]]), equal_nan=True)
    uv = pipelines.composite_map(tu, tv)
    fwd_pts = np.stack([ou["fwd"][:V], ou["fwd"][V:]], 1)
    ouv, _, _ = oracle.inverse(n, ov["fwd"], ov["density"], fwd_pts)
    assert np.array_equal(uv.xy, np.concatenate([ouv[:, 0], ouv[:, 1]]), equal_nan=True)


@pytest.mark.parametrize("k", [2, 4])
def test_barycentres_vs_oracle(k, oracle):
    n, res, steps = 64, 64, 2
    dens = D.barycentre_inputs(n, count=k)
    b = BatchSolver([n] * k)
    maps = b.solve(dens)
    means = [float(d.mean()) for d in dens]
    print("barycentres", k, flush=True)
    imgs = pipelines.barycentres(maps, steps, res, means)
    V = (n + 1) ** 2
    ident = GridMesh.regular(n)
    grid_pts = np.stack([ident.xy[:V], ident.xy[V:]], 1)
    invs = []
    for tm in maps:
        oi, _, _ = oracle.inverse(n, tm.fwd_mesh().xy, tm.density(), grid_pts)
        invs.append(np.concatenate([oi[:, 0], oi[:, 1]]))
    for key, img in imgs.items():
        if k == 4:
            a, bb = key[0] / steps, key[1] / steps
            mesh = pipelines.bilerp(a, bb, *invs)
            expected = pipelines.bilerp(a, bb, *means)
        else:
            a = key / steps
            mesh = pipelines.lerp(a, *invs)
            expected = pipelines.lerp(a, *means)
        raster = oracle.rasterize(n, mesh, res, res)
        ref = (raster * (expected / raster.mean())).astype(np.float32)
        got = img.cpu().numpy()
        assert got.dtype == np.float32 and got.shape == (res, res)
        assert np.allclose(got, ref, rtol=3e-7, atol=0), key


def test_stippling_vs_oracle(oracle):
    n = 64
    rho = D.gaussian_mixture(n, 12)
    rho[(np.arange(n * n) % 7) == 0] = 0.0  # zero-density cells to prune
    o = oracle.solve(n, rho)
    tm = TransportMap(GridMesh.regular(n), GridMesh(n, o["fwd"]), o["density"])
    pts = np.random.default_rng(5).uniform(0, 1, (20000, 2))
    kept = pipelines.stippling(tm, rho, pts).cpu().numpy()
    adv, _, _ = oracle.inverse(n, o["fwd"], o["density"], pts)
    ix = np.clip(np.floor(adv[:, 0] * n).astype(int), 0, n - 1)
    iy = np.clip(np.floor(adv[:, 1] * n).astype(int), 0, n - 1)
    keep = np.isfinite(adv).all(1) & (rho[iy + ix * n] != 0)
    assert 0 < keep.sum() < len(pts)
    assert np.array_equal(kept, adv[keep])


@pytest.mark.parametrize("n", [100, 333, 1000])
def test_blue_noise_tile_vs_reference_golden(n):
    """generate_blue_noise_tile (apps/common/image_utils.cpp:148-173) on the
    device from the reference's own tile: same count, sampled points
    bit-exact in the reference's order, same coordinate sums."""
    g = np.load(os.path.join(GOLDEN, "blue_noise.npz"))
    pts = pipelines.blue_noise_tile(n, g["tile"]).cpu().numpy()
    assert len(pts) == int(g[f"n{n}_count"])
    assert np.array_equal(pts[g[f"n{n}_idx"]], g[f"n{n}_pts"])
    assert np.allclose(pts.sum(0), g[f"n{n}_sum"], rtol=1e-13, atol=0)


def test_stippling_from_blue_noise_tile_vs_oracle(oracle):
    """apps/stippling.cpp:177-197 with the reference's tile: the tile
    repeated at resolution 200, advected through the inverse map, pruned on
    zero density; kept points identical, in order, to the oracle's."""
    g = np.load(os.path.join(GOLDEN, "blue_noise.npz"))
    n = 64
    rho = D.gaussian_mixture(n, 13)
    rho[(np.arange(n * n) % 11) == 0] = 0.0
    o = oracle.solve(n, rho)
    tm = TransportMap(GridMesh.regular(n), GridMesh(n, o["fwd"]), o["density"])
    kept = pipelines.stippling(tm, rho, tile=g["tile"], res=200).cpu().numpy()
    pts = pipelines.blue_noise_tile(200, g["tile"]).cpu().numpy()
    adv, _, _ = oracle.inverse(n, o["fwd"], o["density"], pts)
    ix = np.clip(np.floor(adv[:, 0] * n).astype(int), 0, n - 1)
    iy = np.clip(np.floor(adv[:, 1] * n).astype(int), 0, n - 1)
    keep = np.isfinite(adv).all(1) & (rho[iy + ix * n] != 0)
    assert 0 < keep.sum() < len(pts)
    assert np.array_equal(kept, adv[keep])
