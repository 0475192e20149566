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


def costs():
    """GridBasedTransportSolver::compute_transport_cost (reference
    otlib/otsolver_2dgrid.cpp:608-640) on random displacements and on the
    displacements of solved maps, with the adjusted densities."""
    out = {}
    for n in (5, 16, 33):
        rng = np.random.default_rng(400 + n)
        G = rng.standard_normal(2 * (n + 1) ** 2) * 1e-2
        rho = R.adjust_density(n, rng.uniform(0.1, 3.0, n * n))
        tot, cf = R.solver_transport_cost(n, G, rho)
        out.update({f"rand{n}_disp": G, f"rand{n}_rho": rho, f"rand{n}_total": np.float64(tot),
                    f"rand{n}_cost": cf})
    for n, seed in ((16, 7), (64, 9)):
        r = R.solve(n, D.gaussian_mixture(n, seed))
        ident = np.concatenate([np.repeat(np.arange(n + 1.0), n + 1), np.tile(np.arange(n + 1.0), n + 1)]) / n
        G = r["fwd"] - ident
        tot, cf = R.solver_transport_cost(n, G, r["density"])
        out.update({f"map{n}_disp": G, f"map{n}_rho": r["density"], f"map{n}_total": np.float64(tot),
                    f"map{n}_cost": cf})
    save("costs.npz", **out)


def _dyadic(x, maxbits=48):
    return x.denominator & (x.denominator - 1) == 0 and abs(x.numerator).bit_length() <= maxbits


def line_search_roots():
    """Step choices of the reference's solve_1D_problem (otlib/otsolver_2dgrid.cpp:
    541-606: stationary points of the quartic e(t) from the EigenSolver of the
    3 x 3 companion matrix, real (imag() == 0) roots in [-10, 10] of least e,
    else 0.5) on constructed inputs whose derivative cubic has a double root
    t0 (exact, or split by a relative perturbation of 2^-40 .. 2^-48 into two
    close real roots or a complex pair of tiny imaginary part), the third root
    inside or outside [-10, 10].  n = 4 grids with dyadic g0, direction and
    residual, chosen so that every dot product of z is exact in FP64 (checked
    with rationals): the recorded z is the one the reference used."""
    from fractions import Fraction as Fr
    import itertools

    n, ek = 4, 2.0
    F, V = n * n, (n + 1) ** 2
    rng = np.random.default_rng(77)
    cases = []
    for trial in range(4000):
        g0 = rng.integers(-4, 5, 2 * V) / 8.0 * float(2 ** rng.integers(0, 9)) * (-1) ** trial
        d = rng.integers(-2, 3, F) / 8.0
        a, b = R.params_1d(n, g0, d)
        fa, fb = [Fr(x) for x in a], [Fr(x) for x in b]
        z3, z4 = 2 * sum(x * y for x, y in zip(fa, fb)), sum(x * x for x in fa)
        if z4 == 0:
            continue
        pair = next(((k, m) for k, m in itertools.combinations(range(F), 2) if fa[k] * fb[m] != fa[m] * fb[k]), None)
        if pair is None:
            continue
        k, m = pair
        det = fa[k] * fb[m] - fa[m] * fb[k]
        bb = sum(x * x for x in fb)
        for t0n in rng.permutation(np.arange(-40, 41))[:12]:
            t0 = Fr(int(t0n), 8)
            t1 = -3 * z3 / (4 * z4) - 2 * t0
            z2, z1 = 2 * z4 * (t0 ** 2 + 2 * t0 * t1), -4 * z4 * t0 ** 2 * t1
            ta, tb = (z2 - bb) / 2, z1 / 2
            rk, rm = (ta * fb[m] - tb * fa[m]) / det, (tb * fa[k] - ta * fb[k]) / det
            for bits, jit in ((None, 0), (40, 1), (44, -1), (48, 1), (48, -1)):
                if bits is None:
                    if not (_dyadic(rk) and _dyadic(rm)):
                        continue
                    vk, vm = rk, rm
                else:
                    def rnd(x):
                        if x == 0:
                            return Fr(0)
                        q = Fr(2) ** (int(np.floor(np.log2(abs(float(x))))) - bits)
                        return (round(x / q) + jit) * q
                    vk, vm = rnd(rk), rnd(rm)
                r = np.zeros(F)
                r[k], r[m] = float(vk), float(vm)
                zx = [Fr(ek) / F, 2 * sum(x * Fr(y) for x, y in zip(fb, r)),
                      bb + 2 * sum(x * Fr(y) for x, y in zip(fa, r)), z3, z4]
                zf = np.array([float(v) for v in zx])
                if [Fr(v) for v in zf] != zx:
                    continue
                res = R.solve_1d(n, g0, np.zeros(F), d, r, ek)
                w = [zx[1], 2 * zx[2], 3 * zx[3], 4 * zx[4]]
                disc = (18 * w[3] * w[2] * w[1] * w[0] - 4 * w[2] ** 3 * w[0] + w[2] ** 2 * w[1] ** 2
                        - 4 * w[3] * w[1] ** 3 - 27 * w[3] ** 2 * w[0] ** 2)
                cases.append((zf, res["alpha"], res["e"], int(np.sign(disc)), float(t0), float(t1),
                              np.concatenate([g0, d, r])))
        if len(cases) >= 160:
            break
    out = dict(z=np.array([c[0] for c in cases]), alpha=np.array([c[1] for c in cases]),
               e=np.array([c[2] for c in cases]), disc_sign=np.array([c[3] for c in cases]),
               t0=np.array([c[4] for c in cases]), t1=np.array([c[5] for c in cases]),
               inputs=np.array([c[6] for c in cases]), n=np.int64(n), ek=np.float64(ek))
    save("line_search_roots.npz", **out)


def _sampled_solve(r, n, count):
    V = (n + 1) ** 2
    idx = sample_vertices(V, count)
    return dict(iters=np.int64(r["iters"]), res0=np.float64(r["res0"]), res=r["res"], alpha=r["alpha"],
                beta=r["beta"], vidx=idx, fx=r["fwd"][idx], fy=r["fwd"][V + idx],
                fwd_sum=np.float64(r["fwd"].sum()), fwd_sq=np.float64((r["fwd"] ** 2).sum()),
                adj_sum=np.float64(r["density"].sum()))


def _batch_job(args):
    p, n = args
    rho = D.batch_densities(64)[p]
    r = Reference().solve(n, rho)
    return p, _sampled_solve(r, n, 1024), r["t_solve"]


def cfg2():
    """The 64-density uneven batch (otmap_amd.densities.batch_densities): one
    reference solve per density (otlib/otsolver_2dgrid.cpp:106-245)."""
    from multiprocessing import Pool

    specs = D.batch_specs(64)
    out = {"n": np.array([n for n, _, _ in specs], dtype=np.int64)}
    order = sorted(range(64), key=lambda p: -specs[p][0])  # largest first
    with Pool(6) as pool:
        for p, d, t in pool.imap_unordered(_batch_job, [(p, specs[p][0]) for p in order]):
            print(f"  batch density {p} (n = {specs[p][0]}): {int(d['iters'])} iterations, {t:.1f} s", flush=True)
            out.update({f"d{p}_{k}": v for k, v in d.items()})
    save("batch64.npz", **out)


def cfg3(count=4, n=512, steps=4):
    """The bilinear barycentre pipeline of apps/barycenters.cpp:167-290 on the
    four 512^2 smoothed-shape densities: forward solves, inverse meshes (the
    origin grid through apply_inverse_map), bilerps of the inverse meshes at
    (i/steps, j/steps) and their rasters (2-argument rasterize_image, before
    the mean normalisation)."""
    dens = D.barycentre_inputs(n, count=count)
    V = (n + 1) ** 2
    ident = np.concatenate([np.repeat(np.arange(n + 1.0), n + 1), np.tile(np.arange(n + 1.0), n + 1)]) / n
    grid_pts = np.stack([ident[:V], ident[V:]], 1)
    out = dict(n=np.int64(n), steps=np.int64(steps), means=np.array([d.mean() for d in dens]))
    invs = []
    for k, rho in enumerate(dens):
        r = R.solve(n, rho)
        out.update({f"s{k}_{key}": v for key, v in _sampled_solve(r, n, 8192).items()})
        inv, _, _ = R.inverse(n, r["fwd"], r["density"], grid_pts)
        inv_xy = np.concatenate([inv[:, 0], inv[:, 1]])
        invs.append(inv_xy)
        idx = sample_vertices(V, 8192)
        out.update({f"inv{k}_x": inv_xy[idx], f"inv{k}_y": inv_xy[V + idx], f"inv{k}_sum": np.float64(inv_xy.sum())})
        print(f"  barycentre density {k}: {r['iters']} iterations", flush=True)
    rng = np.random.default_rng(3003)
    pix = rng.choice(n * n, 4096, replace=False)
    out["pix"] = pix
    for i in range(steps + 1):
        for j in range(steps + 1):
            a, b = i / steps, j / steps
            m = (1. - a) * (1. - b) * invs[0] + a * (1. - b) * invs[1] + a * b * invs[2] + (1. - a) * b * invs[3]
            img = R.rasterize(n, m, n, n)
            out[f"img{i}{j}_pix"] = img.reshape(-1)[pix]
            out[f"img{i}{j}_sum"] = np.float64(img.sum())
    save("barycentre4x512.npz", **out)


def cfg4(n=2048, nq=16384):
    """The 2048^2 high-contrast solve (max_ratio 1e4) and the inverse of a
    sample of its 1M advected points (apps/stippling.cpp:164-197)."""
    rho = D.hdr_lognormal(n)
    r = R.solve(n, rho, max_ratio=1e4)
    print(f"  2048^2: {r['iters']} iterations, init {r['t_init']:.1f} s, solve {r['t_solve']:.1f} s", flush=True)
    out = _sampled_solve(r, n, 16384)
    out.update(t_init=np.float64(r["t_init"]), t_solve=np.float64(r["t_solve"]), max_ratio=np.float64(1e4))
    pts = D.uniform_points(1 << 20, seed=2000 + n)
    qi = np.arange(0, 1 << 20, (1 << 20) // nq, dtype=np.int64)
    inv, ff, nh = R.inverse(n, r["fwd"], r["density"], pts[qi])
    out.update(qidx=qi, inv=inv, first_face=ff, nhits=nh)
    save("stipple2048.npz", **out)


def blue_noise():
    """The reference apps' blue-noise tile (data/uniform_pattern_sig2012.dat,
    read by load_point_cloud_dat, apps/common/image_utils.cpp:209-231) and
    generate_blue_noise_tile (:148-173) at three resolutions."""
    path = "/root/reference/data/uniform_pattern_sig2012.dat"
    tile = R.load_point_cloud_dat(path)
    with open(path) as f:
        text = f.read()
    import tempfile
    with tempfile.NamedTemporaryFile("w", suffix=".dat", delete=False) as f:
        f.write(text + "\n")
    tile_nl = R.load_point_cloud_dat(f.name)
    os.unlink(f.name)
    out = dict(tile=tile, tile_trailing_newline=tile_nl)
    rng = np.random.default_rng(5005)
    for n in (100, 333, 1000):
        p = R.blue_noise_tile(n, tile)
        out[f"n{n}_count"] = np.int64(len(p))
        out[f"n{n}_sum"] = p.sum(0)
        idx = np.sort(rng.choice(len(p), min(len(p), 4096), replace=False))
        out[f"n{n}_idx"] = idx
        out[f"n{n}_pts"] = p[idx]
    save("blue_noise.npz", **out)


def edge_cases():
    """The reference's input edge cases (otlib/otsolver_2dgrid.cpp):
      zero64   zero-density cells with the default max_ratio (DBL_MAX): the
               ratio is infinite and the shift beta' = max / DBL_MAX (:58-63)
      nan16    a NaN density: the initial residual is NaN, `residual >
               threshold` is false, 0 iterations and the identity map (:160)
      zdir16   constant density with threshold -1: r = 0, d_hat = 0 and the
               early break |d_hat| <= 2 DBL_MIN in the first iteration
               (:182-183)"""
    out = {}
    n = 64
    rho = D.gaussian_mixture(n, 21)
    rho[(np.arange(n * n) % 7) == 0] = 0.0
    cases = [("zero64", n, rho, {}),
             ("nan16", 16, np.where(np.arange(256) == 5, np.nan, D.gaussian_mixture(16, 3)), {}),
             ("zdir16", 16, np.ones(256), {"threshold": -1.0})]
    for tag, n, rho, kw in cases:
        r = R.solve(n, rho, **kw)
        out.update({f"{tag}_density": rho, f"{tag}_iters": np.int64(r["iters"]), f"{tag}_res0": np.float64(r["res0"]),
                    f"{tag}_fwd": r["fwd"], f"{tag}_adj": r["density"], f"{tag}_res": r["res"],
                    f"{tag}_alpha": r["alpha"], f"{tag}_threshold": np.float64(kw.get("threshold", 1e-7))})
    save("edge_cases.npz", **out)


if __name__ == "__main__":
    which = sys.argv[1:] or ["stencils", "small", "inverse", "raster", "costs", "edge", "roots", "blue", "cfg0", "cfg1",
                             "cfg2", "cfg3", "cfg4"]
    if "stencils" in which:
        stencils()
    if "small" in which:
        small_solves()
    if "inverse" in which:
        inverse_small()
    if "raster" in which:
        raster()
    if "costs" in which:
        costs()
    if "edge" in which:
        edge_cases()
    if "roots" in which:
        line_search_roots()
    if "blue" in which:
        blue_noise()
    if "cfg0" in which:
        big_solve(256, "solve_gaussian256.npz", D.gaussian_mixture(256), 16384)
    if "cfg1" in which:
        big_solve(1024, "solve_shapes1024.npz", D.smoothed_shapes(1024), 16384)
    if "cfg2" in which:
        cfg2()
    if "cfg3" in which:
        cfg3()
    if "cfg4" in which:
        cfg4()
