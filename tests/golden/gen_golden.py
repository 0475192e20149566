"""Writes the golden fixtures of tests/golden/ from the compiled reference.

The reference (ggael/otmap) is C++; oracle/Makefile compiles it from its own
sources under /root/reference into oracle/_ref/libotref.so, and this script
calls it through oracle/oracle.py's Reference wrapper on seeded inputs.  Only
the resulting vectors (inputs + outputs) are committed.  Run where the
reference is mounted:

    make -C oracle ref && python tests/golden/gen_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from oracle import Reference  # noqa: E402
from otmap_amd import densities as D  # noqa: E402

R = Reference()


def save(name, **arrays):
    path = os.path.join(HERE, name)
    np.savez_compressed(path, **arrays)
    print(f"{name}: {os.path.getsize(path) / 1024:.1f} KiB")


def stencils():
    out = {}
    for n in (3, 5, 16, 33):
        rng = np.random.default_rng(100 + n)
        psi = rng.standard_normal(n * n) * 1e-2
        rho_in = rng.uniform(0.2, 3.0, n * n)
        rho = R.adjust_density(n, rho_in)
        G = R.vertex_gradients(n, psi)
        e, r = R.residual(n, rho, psi)
        rhs = rng.standard_normal(n * n)
        dh = R.laplacian_solve(n, rhs)
        g0 = R.vertex_gradients(n, rng.standard_normal(n * n) * 1e-3)
        direction = rng.standard_normal(n * n) * 1e-3
        a, b = R.params_1d(n, g0, direction)
        s1 = R.solve_1d(n, g0, psi, direction, r, e)
        rkm1 = r + rng.standard_normal(n * n) * 1e-4
        dprev = rng.standard_normal(n * n) * 1e-3
        beta = R.cj_beta(n, rho, psi, rkm1, r, dh * 1e-3, dprev, 0.7)
        for k, v in dict(psi=psi, rho_in=rho_in, rho=rho, G=G, e=e, r=r, rhs=rhs, dhat=dh, g0=g0, dir=direction,
                         a=a, b=b, s1_alpha=s1["alpha"], s1_e=s1["e"], s1_xk1=s1["xk1"], s1_rk1=s1["rk1"],
                         s1_g0=s1["g0"], rkm1=rkm1, dprev=dprev, beta=beta).items():
            out[f"n{n}_{k}"] = np.asarray(v)
    for n in (8, 13, 32):
        out[f"nd{n}"] = R.nd_ordering(n)
    # ratio cap of adjust_density
    rng = np.random.default_rng(7)
    rho_in = rng.uniform(0.001, 1.0, 16 * 16)
    out["ratio_in"] = rho_in
    for mr in (2.0, 50.0, 1e300):
        out[f"ratio_{mr:g}"] = R.adjust_density(16, rho_in, mr)
    save("stencils.npz", **out)


def small_solves():
    out = {}
    cases = [(16, 7, True), (32, 8, True), (64, 9, True), (32, 4, False)]
    for n, seed, cj in cases:
        rho = D.gaussian_mixture(n, seed)
        r = R.solve(n, rho, beta_cj=cj)
        tag = f"n{n}_s{seed}_{'cj' if cj else 'zero'}"
        out[tag + "_density"] = rho
        out[tag + "_iters"] = np.int64(r["iters"])
        out[tag + "_res0"] = np.float64(r["res0"])
        out[tag + "_res"] = r["res"]
        out[tag + "_alpha"] = r["alpha"]
        out[tag + "_beta"] = r["beta"]
        out[tag + "_fwd"] = r["fwd"]
        out[tag + "_adj"] = r["density"]
    # max_ratio and max_iter variants
    n = 32
    rho = D.gaussian_mixture(n, 5, floor_rel=1e-6)
    r = R.solve(n, rho, max_ratio=50.0)
    out["ratio50_density"], out["ratio50_fwd"], out["ratio50_adj"] = rho, r["fwd"], r["density"]
    out["ratio50_iters"] = np.int64(r["iters"])
    for mi in (0, 1, 3):
        r = R.solve(n, D.gaussian_mixture(n, 3), max_iter=mi)
        out[f"maxit{mi}_fwd"] = r["fwd"]
        out[f"maxit{mi}_iters"] = np.int64(r["iters"])
    out["maxit_density"] = D.gaussian_mixture(n, 3)
    save("solve_small.npz", **out)


def sample_vertices(V, count):
    step = max(1, V // count)
    return np.arange(0, V, step, dtype=np.int64)


def big_solve(n, name, rho, nq):
    r = R.solve(n, rho)
    V = (n + 1) ** 2
    idx = sample_vertices(V, 8192)
    out = dict(iters=np.int64(r["iters"]), res0=np.float64(r["res0"]), res=r["res"], alpha=r["alpha"],
               beta=r["beta"], vidx=idx, fx=r["fwd"][idx], fy=r["fwd"][V + idx],
               fwd_sum=np.float64(r["fwd"].sum()), fwd_sq=np.float64((r["fwd"] ** 2).sum()),
               t_init=np.float64(r["t_init"]), t_solve=np.float64(r["t_solve"]))
    if nq:
        pts = D.uniform_points(1 << 20, seed=2000 + n)
        qi = np.arange(0, 1 << 20, (1 << 20) // nq, dtype=np.int64)
        inv, ff, nh = R.inverse(n, r["fwd"], r["density"], pts[qi])
        out.update(qidx=qi, inv=inv, first_face=ff, nhits=nh)
    save(name, **out)
    return r


def inverse_small():
    out = {}
    for n, seed in ((16, 7), (32, 8)):
        rho = D.gaussian_mixture(n, seed)
        r = R.solve(n, rho)
        rng = np.random.default_rng(300 + n)
        V = (n + 1) ** 2
        fwd = r["fwd"]
        verts = np.stack([fwd[:V], fwd[V:]], 1)
        pts = np.concatenate([
            rng.uniform(0, 1, (4000, 2)),                       # interior
            verts[rng.choice(V, 200, replace=False)],           # on forward vertices (multi-hit)
            rng.uniform(-0.2, 1.2, (300, 2)),                   # outside [0,1]^2 (clamped)
            np.stack([rng.uniform(0, 1, 100), np.zeros(100)], 1),  # on the border
        ])
        inv, ff, nh = R.inverse(n, fwd, r["density"], pts)
        invf, fff, nhf = R.inverse(n, fwd, r["density"], pts, fast=True)
        tag = f"n{n}"
        out.update({tag + "_fwd": fwd, tag + "_density": r["density"], tag + "_pts": pts, tag + "_inv": inv,
                    tag + "_first": ff, tag + "_nhits": nh, tag + "_invfast": invf})
    # identity map: inv(p) = p
    n = 16
    ident = np.concatenate([np.repeat(np.arange(n + 1.0), n + 1), np.tile(np.arange(n + 1.0), n + 1)]) / n
    pts = np.random.default_rng(5).uniform(0, 1, (1000, 2))
    inv, ff, nh = R.inverse(n, ident, np.ones(n * n), pts)
    out.update(ident_pts=pts, ident_inv=inv, ident_first=ff)
    save("inverse_small.npz", **out)


def raster():
    out = {}
    for n, seed in ((16, 7), (32, 8)):
        rho = D.gaussian_mixture(n, seed)
        r = R.solve(n, rho)
        fwd = r["fwd"]
        out[f"n{n}_fwd"] = fwd
        out[f"n{n}_density"] = r["density"]
        for rows, cols in ((n, n), (2 * n, 2 * n), (24, 40)):
            out[f"n{n}_img_{rows}x{cols}"] = R.rasterize(n, fwd, rows, cols, r["density"])
            out[f"n{n}_ones_{rows}x{cols}"] = R.rasterize(n, fwd, rows, cols)
        # inverse mesh (grid vertices through the inverse map), as the barycentre app rasterises
        V = (n + 1) ** 2
        ident = np.concatenate([np.repeat(np.arange(n + 1.0), n + 1), np.tile(np.arange(n + 1.0), n + 1)]) / n
        verts = np.stack([ident[:V], ident[V:]], 1)
        inv, _, _ = R.inverse(n, fwd, r["density"], verts)
        inv_xy = np.concatenate([inv[:, 0], inv[:, 1]])
        out[f"n{n}_invmesh"] = inv_xy
        out[f"n{n}_invimg"] = R.rasterize(n, inv_xy, n, n)
        src = ident
        c, cf = R.transport_cost(n, src, fwd, r["density"], per_face=True)
        out[f"n{n}_cost"] = np.float64(c)
        out[f"n{n}_cost_face"] = cf
    save("raster.npz", **out)


if __name__ == "__main__":
    which = sys.argv[1:] or ["stencils", "small", "inverse", "raster", "cfg0", "cfg1"]
    if "stencils" in which:
        stencils()
    if "small" in which:
        small_solves()
    if "inverse" in which:
        inverse_small()
    if "raster" in which:
        raster()
    if "cfg0" in which:
        big_solve(256, "solve_gaussian256.npz", D.gaussian_mixture(256), 16384)
    if "cfg1" in which:
        big_solve(1024, "solve_shapes1024.npz", D.smoothed_shapes(1024), 16384)
