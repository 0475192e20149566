"""GPU parity against the compiled reference on the two full-size workloads.

Golden vectors (tests/golden/solve_gaussian256.npz, solve_shapes1024.npz) hold
the reference's iteration count, residual/alpha/beta traces, 8k sampled
forward positions and, for 16384 of the 1M seeded query points, its inverse
images, first-hit faces and hit counts.

Bars (the drop-in criteria): same iterations (+-1), forward positions within
1e-8 (FP64; 2.7e-9 measured at 1024^2) and 1e-6 (mixed), inverse-lookup cell
ids bit-exact for every sampled point farther than 1e-3 cell widths from a
cell edge of the reference's answer, and all 1M query points located.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from otmap_amd import Precision, SolverOptions, Transform
from otmap_amd import densities as D

pytestmark = pytest.mark.gpu

CASES = {
    "cfg0": (256, "solve_gaussian256.npz", lambda: D.gaussian_mixture(256)),
    "cfg1": (1024, "solve_shapes1024.npz", lambda: D.smoothed_shapes(1024)),
}


@pytest.mark.parametrize("precision", [Precision.FP64, Precision.MIXED], ids=["fp64", "mixed"])
@pytest.mark.parametrize("cfg", sorted(CASES))
def test_reference_workload(cfg, precision, solver_factory):
    """FP64 here runs the GEMM transform (the fast-DCT transform, the
    default, has the same test in test_gpu_fft.py)."""
    n, fname, make = CASES[cfg]
    g = np.load(os.path.join(GOLDEN, fname))
    V = (n + 1) ** 2
    print(f"{cfg} n={n} {precision.name}", flush=True)
    s = solver_factory(n)
    tm = s.solve(make(), SolverOptions(precision=precision, transform=Transform.GEMM))
    st = s.last_stats
    assert abs(st.iterations - int(g["iters"])) <= 1, (st.iterations, int(g["iters"]))
    assert st.final_residual <= 1e-7

    xy = tm.fwd_mesh().xy
    idx = g["vidx"]
    err = max(np.abs(xy[idx] - g["fx"]).max(), np.abs(xy[V + idx] - g["fy"]).max())
    assert err < (1e-8 if precision == Precision.FP64 else 1e-6), err

    # the step length is the minimiser of a quartic that is flat near its
    # minimum, so alpha moves by ~sqrt of the perturbation (2.2e-6 measured at
    # 1024^2 in FP64 with positions 2.7e-9 from the reference)
    res, alpha, beta = s.trace()
    k = min(len(alpha), len(g["alpha"]))
    tol = 2e-5 if precision == Precision.FP64 else 1e-3
    assert np.abs(alpha[:k] - g["alpha"][:k]).max() < tol

    # inverse cell ids at the sampled query points, away from cell edges
    pts = D.uniform_points(1 << 20, seed=2000 + n)
    q = pts[g["qidx"]]
    inv, face, nh = tm.inverse(q, fast=True)
    ref_inv = g["inv"]
    frac = np.concatenate([ref_inv[:, :1] * n % 1.0, ref_inv[:, 1:] * n % 1.0], 1)
    away = (np.minimum(frac, 1.0 - frac).min(1) > 1e-3) & (g["nhits"] == 1)
    assert away.mean() > 0.99
    assert np.array_equal(face[away], g["first_face"][away])
    # inverse images move by the vertex error over the local cell width, which
    # is as small as sqrt(1e-3)/n where the density sits at its floor
    assert np.abs(inv[away] - ref_inv[away]).max() < (1e-6 if precision == Precision.FP64 else 5e-5)

    # every one of the 1M query points is located (the forward mesh covers the square)
    inv_all, face_all, _ = tm.inverse(pts)
    assert np.all(face_all >= 0) and np.all(face_all < n * n)
    assert np.all(np.isfinite(inv_all))


# ---------------------------------------------------------------------------
# the three other workloads at their real sizes, against the compiled reference
# (tests/golden/batch64.npz, barycentre4x512.npz, stipple2048.npz)
# ---------------------------------------------------------------------------
def _sampled_err(xy, g, prefix, n):
    V = (n + 1) ** 2
    idx = g[f"{prefix}vidx"]
    return max(np.abs(xy[idx] - g[f"{prefix}fx"]).max(), np.abs(xy[V + idx] - g[f"{prefix}fy"]).max())


@pytest.mark.parametrize("precision", [Precision.FP64, Precision.MIXED], ids=["fp64", "mixed"])
def test_batch64_reference_workload(precision):
    """The 64 uneven densities (n in {128, ..., 1024}) in one BatchSolver,
    every grid against the reference's solve of its density: iterations
    within +-1 and 1024 sampled forward positions within 1e-6 in both modes
    (the default mixed mode: BF16x3 transform on the 1024^2 grids, FP64 on
    the smaller ones)."""
    from otmap_amd import BatchSolver

    g = np.load(os.path.join(GOLDEN, "batch64.npz"))
    sizes = [int(n) for n in g["n"]]
    dens = D.batch_densities(64)
    print(f"batch64 {precision.name}", flush=True)
    b = BatchSolver(sizes)
    maps = b.solve(dens, SolverOptions(precision=precision))
    worst = 0.0
    for p, n in enumerate(sizes):
        it, ref_it = b.last_stats[p].iterations, int(g[f"d{p}_iters"])
        assert abs(it - ref_it) <= 1, (p, n, it, ref_it)
        assert b.last_stats[p].final_residual <= 1e-7
        err = _sampled_err(maps[p].fwd_mesh().xy, g, f"d{p}_", n)
        worst = max(worst, err)
        assert err < 1e-6, (p, n, err)
    print(f"batch64 {precision.name}: largest sampled vertex error {worst:.3g}")


@pytest.mark.parametrize("precision", [Precision.FP64, Precision.MIXED], ids=["fp64", "mixed"])
def test_barycentre_reference_workload(precision):
    """apps/barycenters.cpp:167-290 on the four 512^2 shape densities: the
    four solves (iterations +-1, positions 1e-6), their inverse meshes
    (8192 sampled vertices), and the 25 bilerped rasters: FP32 images equal
    to the reference's raster scaled to the interpolated mean within one
    FP32 rounding at >= 99.9% of 4096 sampled pixels (a pixel centre within
    ~1e-8 of a face edge may change face) and image sums within 1e-6."""
    from otmap_amd import BatchSolver, pipelines

    g = np.load(os.path.join(GOLDEN, "barycentre4x512.npz"))
    n, steps = int(g["n"]), int(g["steps"])
    V = (n + 1) ** 2
    dens = D.barycentre_inputs(n, count=4)
    b = BatchSolver([n] * 4)
    maps = b.solve(dens, SolverOptions(precision=precision))
    for k in range(4):
        assert abs(b.last_stats[k].iterations - int(g[f"s{k}_iters"])) <= 1
        assert _sampled_err(maps[k].fwd_mesh().xy, g, f"s{k}_", n) < 1e-6
    invs = pipelines.inverse_meshes_device(maps)
    for k, inv in enumerate(invs):
        xy = inv.cpu().numpy()
        sel = np.arange(0, V, max(1, V // 8192))
        assert np.abs(xy[sel] - g[f"inv{k}_x"]).max() < 1e-5
        assert np.abs(xy[V + sel] - g[f"inv{k}_y"]).max() < 1e-5
    means = [float(m) for m in g["means"]]
    imgs = pipelines.barycentres(maps, steps, n, means)
    pix = g["pix"]
    for i in range(steps + 1):
        for j in range(steps + 1):
            img = imgs[(i, j)].cpu().numpy().reshape(-1)
            ref_sum = float(g[f"img{i}{j}_sum"])
            expected = pipelines.bilerp(i / steps, j / steps, *means)
            ref = (g[f"img{i}{j}_pix"] * (expected / (ref_sum / (n * n)))).astype(np.float32)
            got = img[pix]
            close = np.abs(got - ref) <= 3e-7 * np.abs(ref) + 1e-30
            assert close.mean() >= 0.999, ((i, j), close.mean())
            # the normalised image's mean is the interpolated mean by construction
            assert abs(float(img.astype(np.float64).mean()) - expected) <= 1e-6 * expected


@pytest.mark.parametrize("precision", [Precision.FP64, Precision.MIXED], ids=["fp64", "mixed"])
def test_stippling_reference_workload(precision, solver_factory):
    """apps/stippling.cpp:164-197 at 2048^2: the high-contrast log-normal
    density with max_ratio = 1e4 (iterations +-1, 16384 sampled positions
    within 1e-6), then 1M uniform points through the inverse map: cell ids
    bit-exact at the 16384 sampled points away from cell edges."""
    g = np.load(os.path.join(GOLDEN, "stipple2048.npz"))
    n = 2048
    print(f"stipple2048 {precision.name}", flush=True)
    s = solver_factory(n)
    tm = s.solve(D.hdr_lognormal(n), SolverOptions(precision=precision, max_ratio=float(g["max_ratio"])))
    st = s.last_stats
    assert abs(st.iterations - int(g["iters"])) <= 1, (st.iterations, int(g["iters"]))
    assert st.final_residual <= 1e-7
    err = _sampled_err(tm.fwd_mesh().xy, g, "", n)
    print(f"stipple2048 {precision.name}: sampled vertex error {err:.3g}")
    assert err < 1e-6, err
    pts = D.uniform_points(1 << 20, seed=2000 + n)
    q = pts[g["qidx"]]
    inv, face, nh = tm.inverse(q)
    ref_inv = g["inv"]
    frac = np.concatenate([ref_inv[:, :1] * n % 1.0, ref_inv[:, 1:] * n % 1.0], 1)
    away = (np.minimum(frac, 1.0 - frac).min(1) > 1e-3) & (g["nhits"] == 1)
    assert away.mean() > 0.98
    assert np.array_equal(face[away], g["first_face"][away])
    assert np.abs(inv[away] - ref_inv[away]).max() < 1e-4
