"""Pins the FP64 oracle (CPU) before it is trusted as the GPU checker.

Three independent sources:
  1. golden vectors recorded from the compiled reference (tests/golden/*.npz,
     written by tests/golden/gen_golden.py);
  2. closed forms / mathematical invariants (DCT-II eigen-decomposition of the
     pinned Laplacian, separable densities, exactness of the quartic, etc.);
  3. brute force on tiny inputs (line-search scan, all-faces point location).
"""
import os

import numpy as np
import pytest
import scipy.fft
import scipy.sparse as sp

from conftest import GOLDEN
from otmap_amd import densities as D


def load(name):
    return np.load(os.path.join(GOLDEN, name))


# ---------------------------------------------------------------- golden: stencils
@pytest.mark.parametrize("n", [3, 5, 16, 33])
def test_stencils_vs_reference(n, oracle):
    g = load("stencils.npz")
    k = lambda s: g[f"n{n}_{s}"]  # noqa: E731
    assert np.array_equal(oracle.vertex_gradients(n, k("psi")), k("G"))
    rho = oracle.adjust_density(n, k("rho_in"))
    assert np.allclose(rho, k("rho"), rtol=2e-15, atol=0)  # sum order of the integral differs
    e, r = oracle.residual(n, k("rho"), k("psi"))
    assert np.array_equal(r, k("r"))
    assert abs(e - k("e")) <= 1e-13 * abs(k("e"))
    dh = oracle.laplacian_solve(n, k("rhs"))
    assert np.abs(dh - k("dhat")).max() <= 1e-12 * np.abs(k("dhat")).max()
    a, b = oracle.params_1d(n, k("g0"), k("dir"))
    assert np.array_equal(a, k("a")) and np.array_equal(b, k("b"))
    s1 = oracle.solve_1d(n, k("g0"), k("psi"), k("dir"), k("r"), float(k("e")))
    assert abs(s1["alpha"] - k("s1_alpha")) <= 1e-12 * max(1.0, abs(k("s1_alpha")))
    assert abs(s1["e"] - k("s1_e")) <= 1e-12 * abs(k("s1_e"))
    assert np.allclose(s1["xk1"], k("s1_xk1"), rtol=1e-12, atol=1e-18)
    assert np.allclose(s1["g0"], k("s1_g0"), rtol=1e-12, atol=1e-18)
    assert np.allclose(s1["rk1"], k("s1_rk1"), rtol=1e-9, atol=1e-15 * np.abs(k("r")).max())
    beta = oracle.cj_beta(n, k("rho"), k("psi"), k("rkm1"), k("r"), k("dhat") * 1e-3, k("dprev"), 0.7)
    assert abs(beta - k("beta")) <= 1e-10 * max(1.0, abs(k("beta")))


@pytest.mark.parametrize("n", [8, 13, 32])
def test_nested_dissection_vs_reference(n, oracle):
    assert np.array_equal(oracle.nd_ordering(n), load("stencils.npz")[f"nd{n}"])


@pytest.mark.parametrize("mr", [2.0, 50.0, 1e300])
def test_adjust_density_ratio_vs_reference(mr, oracle):
    g = load("stencils.npz")
    got = oracle.adjust_density(16, g["ratio_in"], mr)
    assert np.allclose(got, g[f"ratio_{mr:g}"], rtol=2e-15, atol=0)


def test_adjust_density_closed_form(oracle):
    n = 24
    rho = np.random.default_rng(1).uniform(0.01, 5.0, n * n)
    out = oracle.adjust_density(n, rho, 10.0)
    assert abs(out.sum() / (n * n) - 1.0) < 1e-14          # integral 1
    assert abs(out.max() / out.min() - 10.0) < 1e-12        # ratio capped at max_ratio
    out2 = oracle.adjust_density(n, rho)                    # no cap: pure normalisation
    assert np.allclose(out2, rho / (rho.sum() / (n * n)), rtol=1e-15)


# ---------------------------------------------------------------- golden: full solves
@pytest.mark.parametrize("tag", ["n16_s7_cj", "n32_s8_cj", "n64_s9_cj", "n32_s4_zero"])
def test_small_solves_vs_reference(tag, oracle):
    g = load("solve_small.npz")
    n = int(tag.split("_")[0][1:])
    o = oracle.solve(n, g[tag + "_density"], beta_cj=not tag.endswith("zero"))
    assert o["iters"] == int(g[tag + "_iters"])
    assert abs(o["res0"] - float(g[tag + "_res0"])) <= 1e-12 * float(g[tag + "_res0"])
    assert np.abs(o["fwd"] - g[tag + "_fwd"]).max() < 1e-11
    assert np.allclose(o["density"], g[tag + "_adj"], rtol=2e-15, atol=0)
    assert np.abs(o["alpha"] - g[tag + "_alpha"]).max() < 1e-9
    assert np.abs(o["beta"] - g[tag + "_beta"]).max() < 1e-9
    assert np.allclose(o["res"], g[tag + "_res"], rtol=1e-7)


def test_ratio_and_max_iter_vs_reference(oracle):
    g = load("solve_small.npz")
    o = oracle.solve(32, g["ratio50_density"], max_ratio=50.0)
    assert o["iters"] == int(g["ratio50_iters"])
    assert np.abs(o["fwd"] - g["ratio50_fwd"]).max() < 1e-11
    for mi in (0, 1, 3):
        o = oracle.solve(32, g["maxit_density"], max_iter=mi)
        assert o["iters"] == int(g[f"maxit{mi}_iters"]) == mi
        assert np.abs(o["fwd"] - g[f"maxit{mi}_fwd"]).max() < 1e-13


def test_gaussian256_vs_reference(oracle):
    g = load("solve_gaussian256.npz")
    n = 256
    o = oracle.solve(n, D.gaussian_mixture(n))
    V = (n + 1) ** 2
    assert o["iters"] == int(g["iters"])
    idx = g["vidx"]
    err = max(np.abs(o["fwd"][idx] - g["fx"]).max(), np.abs(o["fwd"][V + idx] - g["fy"]).max())
    assert err < 1e-10, err
    assert np.abs(o["alpha"] - g["alpha"]).max() < 1e-8


# ---------------------------------------------------------------- closed forms
def pinned_laplacian(n):
    """Sparse M of otlib/otsolver_2dgrid.cpp:284-321 (unpinned M0 and pinned M)."""
    rows, cols, vals = [], [], []
    for i in range(n):
        for j in range(n):
            fid = j + i * n
            r1, c1 = (i if i == 0 else i - 1), (j if j == 0 else j - 1)
            r2, c2 = (i if i == n - 1 else i + 1), (j if j == n - 1 else j + 1)
            for (a, b) in ((r1, c1), (r1, c2), (r2, c1), (r2, c2)):
                rows.append(fid); cols.append(b + a * n); vals.append(-0.5)
            rows.append(fid); cols.append(fid); vals.append(2.0)
    M0 = sp.csr_matrix((vals, (rows, cols)), shape=(n * n, n * n))
    M = M0.tolil()
    M[0, 0] = M[0, 0] + abs(M[0, 0]) * 1e4
    return M0, M.tocsr()


@pytest.mark.parametrize("n", [4, 9, 16, 33])
def test_laplacian_is_dct_diagonal_and_oracle_solves_it(n, oracle):
    M0, M = pinned_laplacian(n)
    # DCT-II diagonalises M0 with lambda_kl = 2(1 - cos(pi k/n) cos(pi l/n))
    C = scipy.fft.dct(np.eye(n), type=2, norm="ortho", axis=0)
    C2 = np.kron(C, C)
    Mh = C2 @ M0.toarray() @ C2.T
    k = np.arange(n)
    lam = 2 * (1 - np.outer(np.cos(np.pi * k / n), np.cos(np.pi * k / n))).reshape(-1)
    assert np.abs(Mh - np.diag(lam)).max() < 1e-12
    r = np.random.default_rng(n).standard_normal(n * n) + 0.3
    d = oracle.laplacian_solve(n, r)
    # pinned solve + mean removal == pseudo-inverse of r - (sum r) e0
    rp = r.copy(); rp[0] -= r.sum()
    Rh = scipy.fft.dctn(rp.reshape(n, n), type=2, norm="ortho")
    lam2 = lam.reshape(n, n).copy(); lam2[0, 0] = np.inf
    x = scipy.fft.idctn(Rh / lam2, type=2, norm="ortho").reshape(-1)
    assert np.abs(d - x).max() <= 1e-12 * np.abs(x).max()
    assert abs(d.mean()) < 1e-15 * np.abs(d).max()
    assert np.abs(M0 @ d - rp).max() <= 1e-13 * np.abs(d).max()


def _positions(o, n):
    V = (n + 1) ** 2
    return o["fwd"][:V].reshape(n + 1, n + 1), o["fwd"][V:].reshape(n + 1, n + 1)


@pytest.mark.parametrize("n", [8, 16])
def test_separable_density_closed_form(n, oracle):
    """rho = f(x) g(y): the exact discrete map is the tensor grid of 1-D cumulative sums."""
    c = (np.arange(n) + 0.5) / n
    f = 1.0 + 0.8 * np.sin(2.0 * c) + c ** 2
    gy = 0.5 + np.exp(-((c - 0.3) ** 2) / 0.05)
    rho = np.outer(f, gy).reshape(-1)                     # face (i, j) -> f[i] g[j]
    o = oracle.solve(n, rho, threshold=1e-30, max_iter=200)
    X, Y = _positions(o, n)
    Xc = np.concatenate([[0.0], np.cumsum(f) / f.sum()])
    Yc = np.concatenate([[0.0], np.cumsum(gy) / gy.sum()])
    # the loop ends when e(alpha) reaches rounding level (~1e-17); the map is then
    # within ~2e-9 of the closed form (at the default 1e-7 threshold: ~3e-5)
    assert np.abs(X - Xc[:, None]).max() < 5e-9
    assert np.abs(Y - Yc[None, :]).max() < 5e-9


def test_constant_density_identity(oracle):
    n = 20
    o = oracle.solve(n, np.full(n * n, 2.5))
    assert o["iters"] == 0
    X, Y = _positions(o, n)
    assert np.abs(X - np.arange(n + 1)[:, None] * (1.0 / n)).max() == 0
    assert np.abs(Y - np.arange(n + 1)[None, :] * (1.0 / n)).max() == 0


def test_quartic_is_exact(oracle):
    """r(psi + t d) = r(psi) + b t + a t^2 for every t (the quad area is quadratic in psi)."""
    n = 12
    rng = np.random.default_rng(3)
    rho = oracle.adjust_density(n, rng.uniform(0.5, 2.0, n * n))
    psi = rng.standard_normal(n * n) * 1e-3
    d = rng.standard_normal(n * n) * 1e-3
    g0 = oracle.vertex_gradients(n, psi)
    a, b = oracle.params_1d(n, g0, d)
    _, r0 = oracle.residual(n, rho, psi)
    for t in (-0.7, 0.3, 2.0):
        _, rt = oracle.residual(n, rho, psi + t * d)
        assert np.abs(rt - (r0 + b * t + a * t * t)).max() < 1e-15


def test_line_search_picks_best_stationary_point(oracle):
    """Brute force: stationary points of e(t) in [-10, 10] from sign changes of e'
    on a dense grid (refined by bisection); the least e among them, else 0.5."""
    rng = np.random.default_rng(11)
    ts = np.linspace(-10, 10, 200001)
    for _ in range(200):
        z = rng.standard_normal(5)
        z[4] = abs(z[4]) + 0.05
        z[0] = abs(z[0])
        w = np.array([z[1], 2 * z[2], 3 * z[3], 4 * z[4]])
        de = np.polyval(w[::-1], ts)
        cands = []
        for k in np.nonzero(np.sign(de[:-1]) != np.sign(de[1:]))[0]:
            lo, hi = ts[k], ts[k + 1]
            for _ in range(80):
                mid = 0.5 * (lo + hi)
                if np.sign(np.polyval(w[::-1], mid)) == np.sign(np.polyval(w[::-1], lo)):
                    lo = mid
                else:
                    hi = mid
            cands.append(0.5 * (lo + hi))
        alpha, emin = oracle.pick_step(z)
        if not cands:
            assert alpha == 0.5
        else:
            best = min(cands, key=lambda t: np.polyval(z[::-1], t))
            assert abs(alpha - best) < 1e-7
        assert abs(emin - np.polyval(z[::-1], alpha)) < 1e-12 * max(1, abs(emin))
    # no real root inside [-10, 10] (or degenerate a = 0) -> alpha = 0.5
    assert oracle.pick_step(np.array([1.0, 0.0, 0.0, 0.0, 0.0]))[0] == 0.5
    assert oracle.pick_step(np.array([1.0, -1e6, 1.0, 0.0, 0.001]))[0] == 0.5


def test_cubic_roots_vs_numpy(oracle):
    rng = np.random.default_rng(5)
    for _ in range(500):
        c = rng.standard_normal(3) * rng.choice([0.1, 1, 10])
        got = np.sort(oracle.cubic_roots(c[2], c[1], c[0]))
        ref = np.roots([1.0, c[2], c[1], c[0]])
        real = np.sort(ref[np.abs(ref.imag) < 1e-9].real)
        if len(real) == len(got):
            assert np.allclose(got, real, rtol=1e-9, atol=1e-9)


# ---------------------------------------------------------------- inverse map / BVH
@pytest.mark.parametrize("n", [16, 32])
def test_inverse_vs_reference(n, oracle):
    g = load("inverse_small.npz")
    t = f"n{n}"
    fwd, dens, pts = g[t + "_fwd"], g[t + "_density"], g[t + "_pts"]
    inv, face, nh = oracle.inverse(n, fwd, dens, pts)
    ok = g[t + "_nhits"] > 0
    assert np.array_equal(nh, g[t + "_nhits"])
    assert np.array_equal(inv[ok], g[t + "_inv"][ok])
    invf, _, _ = oracle.inverse(n, fwd, dens, pts, fast=True)
    assert np.array_equal(invf[ok], g[t + "_invfast"][ok])
    # first hit in traversal order of the reference BVH
    with oracle.bvh(n, fwd) as b:
        _, first, _ = b.inverse(dens, pts, fast=True)
    assert np.array_equal(first[ok], g[t + "_first"][ok])


def test_inverse_identity_map(oracle):
    g = load("inverse_small.npz")
    n = 16
    ident = np.concatenate([np.repeat(np.arange(n + 1.0), n + 1), np.tile(np.arange(n + 1.0), n + 1)]) / n
    inv, face, _ = oracle.inverse(n, ident, np.ones(n * n), g["ident_pts"])
    assert np.array_equal(inv, g["ident_inv"])
    assert np.abs(inv - g["ident_pts"]).max() < 1e-15


def test_inverse_brute_force_location(oracle):
    n = 8
    rho = D.gaussian_mixture(n, 3)
    o = oracle.solve(n, rho)
    pts = np.random.default_rng(2).uniform(0, 1, (300, 2))
    _, face, nh = oracle.inverse(n, o["fwd"], o["density"], pts, fast=True)
    for k, q in enumerate(pts):
        allf = oracle.locate_all(n, o["fwd"], q)
        assert nh[k] == 1 and len(allf) >= 1
        assert face[k] in allf


def test_inverse_round_trip_interior_vertices(oracle):
    """inv_fast(fwd(v)) = origin(v) for interior vertices.  (inv() at such
    multi-hit points may switch to another face while keeping the first hit's
    weights -- a reference quirk, otlib/transport_map.cpp:81 -- checked against
    the golden vectors instead.)"""
    n = 16
    o = oracle.solve(n, D.gaussian_mixture(n, 7))
    V = (n + 1) ** 2
    i, j = np.meshgrid(np.arange(1, n), np.arange(1, n), indexing="ij")
    v = (j + i * (n + 1)).reshape(-1)
    pts = np.stack([o["fwd"][v], o["fwd"][V + v]], 1)
    inv, _, nh = oracle.inverse(n, o["fwd"], o["density"], pts, fast=True)
    exp = np.stack([i.reshape(-1) / n, j.reshape(-1) / n], 1)
    assert np.abs(inv - exp).max() < 1e-12


# ---------------------------------------------------------------- raster / costs
@pytest.mark.parametrize("n", [16, 32])
def test_raster_vs_reference(n, oracle):
    g = load("raster.npz")
    fwd, dens = g[f"n{n}_fwd"], g[f"n{n}_density"]
    for rows, cols in ((n, n), (2 * n, 2 * n), (24, 40)):
        assert np.array_equal(oracle.rasterize(n, fwd, rows, cols, dens), g[f"n{n}_img_{rows}x{cols}"])
        assert np.array_equal(oracle.rasterize(n, fwd, rows, cols), g[f"n{n}_ones_{rows}x{cols}"])
    assert np.array_equal(oracle.rasterize(n, g[f"n{n}_invmesh"], n, n), g[f"n{n}_invimg"])


def test_raster_identity_closed_form(oracle):
    n = 16
    ident = np.concatenate([np.repeat(np.arange(n + 1.0), n + 1), np.tile(np.arange(n + 1.0), n + 1)]) / n
    img = oracle.rasterize(n, ident, n, n)
    expect = float(np.float32(1.0 / (n * n))) / (1.0 / (n * n))
    assert np.allclose(img, expect, rtol=1e-12)


def test_raster_carried_weights_decide_reference_pixels(oracle):
    """On the reference's own golden meshes the bilinear solve fails at some
    pixel centres and the carried weights decide at least one pixel that the
    reference's image covers (the images are matched bit for bit by
    test_raster_vs_reference, so the carried-weights emulation is exercised);
    the weights never come from more than a few faces back, far inside the
    device's 64-face look-back.  An identity mesh never fails."""
    g = load("raster.npz")
    decided = 0
    for n in (16, 32):
        for rows, cols in ((n, n), (2 * n, 2 * n), (24, 40)):
            st = oracle.raster_stats(n, g[f"n{n}_fwd"], rows, cols)
            assert st["failed"] == st["carried_same_face"] + st["carried_prev_face"] + st["uncovered_no_weights"]
            assert st["max_faces_back"] < 64
            decided += st["failed_but_covered"]
    assert decided >= 1
    n = 16
    ident = np.concatenate([np.repeat(np.arange(n + 1.0), n + 1), np.tile(np.arange(n + 1.0), n + 1)]) / n
    assert oracle.raster_stats(n, ident, 37, 29)["failed"] == 0


@pytest.mark.parametrize("n", [16, 32])
def test_transport_cost_vs_reference(n, oracle):
    g = load("raster.npz")
    ident = np.concatenate([np.repeat(np.arange(n + 1.0), n + 1), np.tile(np.arange(n + 1.0), n + 1)]) / n
    c, cf = oracle.transport_cost(n, ident, g[f"n{n}_fwd"], g[f"n{n}_density"], per_face=True)
    assert abs(c - float(g[f"n{n}_cost"])) <= 1e-13 * abs(c)
    assert np.allclose(cf, g[f"n{n}_cost_face"], rtol=1e-12, atol=1e-300)


# ------------------------------------------------ compute_transport_cost (solver)
def _vertex_ij(n):
    i = np.repeat(np.arange(n + 1.0), n + 1)  # vertex j + i*(n+1)
    j = np.tile(np.arange(n + 1.0), n + 1)
    return i, j


@pytest.mark.parametrize("tag", ["rand5", "rand16", "rand33", "map16", "map64"])
def test_solver_transport_cost_vs_reference(tag, oracle):
    g = load("costs.npz")
    n = int(g[f"{tag}_rho"].size ** 0.5)
    tot, cost = oracle.solver_transport_cost(n, g[f"{tag}_disp"], g[f"{tag}_rho"])
    assert np.array_equal(cost, g[f"{tag}_cost"])
    assert abs(tot - float(g[f"{tag}_total"])) <= 1e-13 * abs(tot)


def test_solver_transport_cost_closed_forms(oracle):
    """2 x 2 Gauss quadrature of |displacement|^2 over each cell, times rho * A
    (otlib/otsolver_2dgrid.cpp:608-640): exact for displacements linear in the
    vertex coordinates.  A uniform displacement c gives rho * A * |c|^2; the
    field (alpha * x, gamma * y) gives rho * A * (alpha^2 (i^2 + i + 1/3) +
    gamma^2 (j^2 + j + 1/3)) / n^2 on face (i, j) -- a midpoint rule (1/4 in
    place of 1/3) or swapped face indices fail it."""
    n = 12
    A = 1.0 / (n * n)
    rho = np.random.default_rng(3).uniform(0.5, 2.0, n * n)
    V = (n + 1) ** 2
    c = np.array([0.013, -0.021])
    G = np.concatenate([np.full(V, c[0]), np.full(V, c[1])])
    tot, cost = oracle.solver_transport_cost(n, G, rho)
    assert np.allclose(cost, rho * A * (c @ c), rtol=1e-14, atol=0)
    assert abs(tot - (rho * A).sum() * (c @ c)) <= 1e-13 * tot
    vi, vj = _vertex_ij(n)
    al, ga = 0.03, -0.07
    G = np.concatenate([al * vi / n, ga * vj / n])
    tot, cost = oracle.solver_transport_cost(n, G, rho)
    fi = np.repeat(np.arange(n, dtype=float), n)  # face j + i*n
    fj = np.tile(np.arange(n, dtype=float), n)
    expect = rho * A * (al**2 * (fi**2 + fi + 1 / 3) + ga**2 * (fj**2 + fj + 1 / 3)) / n**2
    assert np.allclose(cost, expect, rtol=1e-13, atol=0)


# ------------------------------------------------------------ input edge cases
@pytest.mark.parametrize("tag", ["zero64", "nan16", "zdir16"])
def test_edge_cases_vs_reference(tag, oracle):
    """Zero-density cells with the default max_ratio, a NaN density (0
    iterations: NaN > threshold is false) and the |d_hat| <= 2 DBL_MIN early
    break, against the compiled reference."""
    g = load("edge_cases.npz")
    k = lambda s: g[f"{tag}_{s}"]  # noqa: E731
    n = int(k("density").size ** 0.5)
    o = oracle.solve(n, k("density"), threshold=float(k("threshold")))
    assert o["iters"] == int(k("iters"))
    assert np.abs(o["fwd"] - k("fwd")).max() <= 1e-9
    assert np.allclose(o["density"], k("adj"), rtol=1e-14, atol=0, equal_nan=True)
    if tag != "zero64":
        V = (n + 1) ** 2
        ident = np.concatenate([np.repeat(np.arange(n + 1.0), n + 1), np.tile(np.arange(n + 1.0), n + 1)]) / n
        assert o["iters"] == 0 and np.array_equal(o["fwd"], ident) and len(ident) == 2 * V


def test_line_search_step_choice_vs_reference(oracle):
    """The oracle's step choice against the reference on near-double-root
    quartics (tests/golden/line_search_roots.npz): same bars as the device
    test -- classification equal except near-double pairs whose third root
    lies outside [-10, 10], simple roots to 1e-11, near-double picks to 1e-5."""
    g = load("line_search_roots.npz")
    al = np.array([oracle.pick_step(z)[0] for z in g["z"]])
    ra = g["alpha"]
    outside = np.abs(g["t1"]) > 10
    diverge = np.flatnonzero((al == 0.5) != (ra == 0.5))
    assert np.all(outside[diverge]) and len(diverge) <= 3
    both = (al != 0.5) & (ra != 0.5)
    assert np.all(np.abs(al[both & ~outside] - ra[both & ~outside]) <= 1e-11 * np.abs(ra[both & ~outside]))
    assert np.all(np.abs(al[both] - ra[both]) <= 1e-5 * np.abs(ra[both]))


def test_point_cloud_loader_vs_reference(tmp_path):
    """load_point_cloud_dat (apps/common/image_utils.cpp:209-231), the
    Python mirror: the reference's tile file as it reads it, and the same
    text with a trailing newline (the reference then repeats the last point)."""
    from otmap_amd import pipelines
    g = load("blue_noise.npz")
    text = "\n".join(f"{float(x)!r} {float(y)!r}" for x, y in g["tile"])
    p0, p1 = tmp_path / "t0.dat", tmp_path / "t1.dat"
    p0.write_text(text)
    p1.write_text(text + "\n")
    assert np.array_equal(pipelines.load_point_cloud_dat(p0), g["tile"])
    assert np.array_equal(pipelines.load_point_cloud_dat(p1), g["tile_trailing_newline"])
    assert len(g["tile_trailing_newline"]) == len(g["tile"]) + 1


def test_fp32_quad_areas_break_the_map_bar_on_a_sensitive_density(oracle):
    """Why the quad areas stay FP64 in both precision modes.  Areas are
    evaluated per iteration only in the conjugate-Jacobian residual r(psi -
    eps d_hat) (the line search updates r exactly from the quartic); with
    those areas in FP32 (oracle switch orc_set_area_precision(1): FP32
    arithmetic on the FP64 gradients) the 192^2 batch density 29, whose
    trajectory amplifies perturbations ~3x per iteration, ends 1.2e-5 from the
    reference's positions (FP64: 3e-10) -- past the 1e-6 drop-in bar, with
    the same 37 iterations.  (cfg0 256^2: 1.3e-8, cfg1 1024^2: 9.2e-8 measured
    the same way.)"""
    g = load("batch64.npz")
    p = 29
    n = int(g["n"][p])
    rho = D.batch_densities(64)[p]
    V = (n + 1) ** 2
    idx = g[f"d{p}_vidx"]
    errs = {}
    try:
        for mode in (0, 1):
            oracle.L.orc_set_area_precision(mode)
            o = oracle.solve(n, rho)
            assert o["iters"] == int(g[f"d{p}_iters"])
            errs[mode] = max(np.abs(o["fwd"][idx] - g[f"d{p}_fx"]).max(), np.abs(o["fwd"][V + idx] - g[f"d{p}_fy"]).max())
    finally:
        oracle.L.orc_set_area_precision(0)
    assert errs[0] < 1e-8
    assert errs[1] > 1e-6
