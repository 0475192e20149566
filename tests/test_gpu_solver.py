"""GPU parity of the solver path against the FP64 oracle (HIP through the C ABI).

Bars:
  * stencils (vertex gradients, residual r, quartic coefficients a, b):
    bit-exact -- the kernels keep the reference's operation order and the
    library is built without FMA contraction;
  * scalar reductions (|r|^2/A): 1e-13 relative (summation order differs);
  * Laplacian solve: DCT-on-MFMA vs the oracle's sparse LDL^T, 1e-11 relative
    to max|d_hat| (both are O(1e-15) accurate; the bound leaves room for the
    conditioning ~4n^2/pi^2 at n = 256);
  * full solves (FP64 path): same iteration count (+-1 allowed, 0 observed),
    forward vertex positions within 1e-9 in unit-square coordinates (the
    drop-in bar is 1e-6), traces of alpha within 1e-7.
"""
import numpy as np
import pytest

from otmap_amd import densities as D

pytestmark = pytest.mark.gpu

SIZES = [1, 2, 3, 5, 16, 33, 64, 100, 256]


def _rng(n):
    return np.random.default_rng(12345 + n)


@pytest.mark.parametrize("n", SIZES)
def test_vertex_gradients_bitexact(n, oracle, solver_factory):
    s = solver_factory(n)
    psi = _rng(n).standard_normal(n * n) * 1e-2
    G = s.vertex_gradients(psi)
    Go = oracle.vertex_gradients(n, psi)
    assert np.array_equal(G, Go), np.abs(G - Go).max()


@pytest.mark.parametrize("n", SIZES)
def test_residual_bitexact(n, oracle, solver_factory):
    s = solver_factory(n)
    rng = _rng(n)
    rho = oracle.adjust_density(n, rng.uniform(0.2, 3.0, n * n))
    psi = rng.standard_normal(n * n) * 1e-3
    e, r = s.residual(rho, psi)
    eo, ro = oracle.residual(n, rho, psi)
    assert np.array_equal(r, ro), np.abs(r - ro).max()
    assert abs(e - eo) <= 1e-13 * abs(eo) + 1e-300


def _fft_dct_solve(n, r):
    """Closed form: pseudo-inverse of the pinned operator applied to r - (sum r) e0
    through scipy's FFT-based orthonormal DCT-II (accurate to ~n*eps)."""
    import scipy.fft
    k = np.arange(n)
    lam = 2 * (1 - np.outer(np.cos(np.pi * k / n), np.cos(np.pi * k / n)))
    lam[0, 0] = np.inf
    rp = r.copy()
    rp[0] -= r.sum()
    return scipy.fft.idctn(scipy.fft.dctn(rp.reshape(n, n), type=2, norm="ortho") / lam, type=2,
                           norm="ortho").reshape(-1)


@pytest.mark.parametrize("n", [3, 5, 16, 33, 64, 100, 128, 256, 512])
def test_laplacian_solve_matches_ldlt(n, oracle, solver_factory):
    """vs the closed form (FFT DCT): 5e-12 relative (2.1e-13 measured at n=256,
    a smooth low-mode difference); vs the oracle's sparse LDL^T:
    within the LDL^T's own error, which grows like the condition number ~ n^2
    (measured 2.8e-12 at n=128, 2.8e-11 at 256, 1.3e-10 at 512)."""
    s = solver_factory(n)
    r = _rng(n).standard_normal(n * n)
    for rhs in (r, r - r.mean()):
        d = s.laplacian_solve(rhs)
        d = d - d.mean()
        x = _fft_dct_solve(n, rhs)
        assert np.abs(d - x).max() / np.abs(x).max() < 5e-12
        if n <= 256:
            do = oracle.laplacian_solve(n, rhs)
            err = np.abs(d - do).max() / np.abs(do).max()
            assert err < 1e-15 * n * n + 1e-13, err


@pytest.mark.parametrize("n", [2, 3, 16, 33, 64, 256])
def test_line_search_coefficients_bitexact(n, oracle, solver_factory):
    s = solver_factory(n)
    rng = _rng(n)
    g0 = oracle.vertex_gradients(n, rng.standard_normal(n * n) * 1e-3)
    d = rng.standard_normal(n * n) * 1e-3
    a, b = s.line_search_coefficients(g0, d)
    ao, bo = oracle.params_1d(n, g0, d)
    assert np.array_equal(a, ao) and np.array_equal(b, bo)


@pytest.mark.parametrize("transform", ["gemm", "fft"])
@pytest.mark.parametrize("n,seed", [(16, 7), (32, 8), (64, 9), (128, 10), (256, 1000)])
def test_full_solve_fp64_matches_oracle(n, seed, transform, oracle, solver_factory):
    from otmap_amd import SolverOptions, Transform
    rho = D.gaussian_mixture(n, seed)
    s = solver_factory(n)
    tm = s.solve(rho, SolverOptions(transform=Transform.GEMM if transform == "gemm" else Transform.FFT))
    o = oracle.solve(n, rho)
    st = s.last_stats
    assert abs(st.iterations - o["iters"]) <= 1, (st.iterations, o["iters"])
    assert st.final_residual <= 1e-7
    err = np.abs(tm.fwd_mesh().xy - o["fwd"]).max()
    assert err < 1e-9, err
    assert np.allclose(tm.density(), o["density"], rtol=1e-14, atol=0)
    res, alpha, beta = s.trace()
    k = min(len(alpha), len(o["alpha"]))
    assert np.abs(alpha[:k] - o["alpha"][:k]).max() < 1e-7
    assert np.abs(beta[:k] - o["beta"][:k]).max() < 1e-7


def test_constant_density_zero_iterations(solver_factory):
    n = 64
    s = solver_factory(n)
    tm = s.solve(np.full(n * n, 3.0))
    assert s.last_stats.iterations == 0
    ref = np.concatenate([np.repeat(np.arange(n + 1.0), n + 1), np.tile(np.arange(n + 1.0), n + 1)]) * (1.0 / n)
    assert np.abs(tm.fwd_mesh().xy - ref).max() <= 1e-15


# 4, 5, 8, 9: the cap at and just past whole replays of the 4-iteration graph
@pytest.mark.parametrize("max_iter", [0, 1, 3, 4, 5, 8, 9])
def test_max_iter_cap(max_iter, oracle, solver_factory):
    from otmap_amd import SolverOptions
    n = 32
    rho = D.gaussian_mixture(n, 3)
    s = solver_factory(n)
    tm = s.solve(rho, SolverOptions(max_iter=max_iter))
    o = oracle.solve(n, rho, max_iter=max_iter)
    assert s.last_stats.iterations == o["iters"] == max_iter
    assert np.abs(tm.fwd_mesh().xy - o["fwd"]).max() < 1e-12


def test_beta_zero_mode(oracle, solver_factory):
    from otmap_amd import BetaOpt, SolverOptions
    n = 32
    rho = D.gaussian_mixture(n, 4)
    s = solver_factory(n)
    tm = s.solve(rho, SolverOptions(beta=BetaOpt.Zero))
    o = oracle.solve(n, rho, beta_cj=False)
    assert abs(s.last_stats.iterations - o["iters"]) <= 1
    assert np.abs(tm.fwd_mesh().xy - o["fwd"]).max() < 1e-9


def test_max_ratio_adjustment(oracle, solver_factory):
    from otmap_amd import SolverOptions
    n = 64
    rho = D.gaussian_mixture(n, 5, floor_rel=1e-6)
    s = solver_factory(n)
    tm = s.solve(rho, SolverOptions(max_ratio=50.0))
    o = oracle.solve(n, rho, max_ratio=50.0)
    assert np.allclose(tm.density(), o["density"], rtol=1e-14, atol=0)
    assert abs(tm.density().max() / tm.density().min() - 50.0) < 1e-9
    assert np.abs(tm.fwd_mesh().xy - o["fwd"]).max() < 1e-9


@pytest.mark.parametrize("n", [128, 256])
def test_folded_dct_tile_shapes_bit_identical(n, solver_factory):
    """Every candidate tile shape of the folded FP64 stages gives the same bits
    (the autotuner may pick any of them)."""
    s = solver_factory(n)
    chosen, count = s.gemm_config()
    assert count >= 2 and all(0 <= c < count for c in chosen)
    r = _rng(n).standard_normal(n * n)
    ref = s.laplacian_solve(r)
    for c in range(count):
        for st in range(4):
            s.set_gemm_config(st, c)
        assert np.array_equal(s.laplacian_solve(r), ref), c
    for st in range(4):
        s.set_gemm_config(st, chosen[st])


def test_unfolded_sizes_report_no_tile_choice(solver_factory):
    s = solver_factory(100)
    assert s.gemm_config()[1] == 0


# ------------------------------------------------------------ input edge cases
@pytest.mark.parametrize("tag,precision", [("zero64", 0), ("zero64", 1), ("nan16", 0), ("zdir16", 0)])
def test_edge_cases_vs_reference_golden(tag, precision):
    """otlib/otsolver_2dgrid.cpp edge cases against the compiled reference:
    zero-density cells with the default max_ratio = DBL_MAX (:58-63; both
    precision modes), a NaN density (NaN residual: 0 iterations, identity
    map, :160) and constant density with threshold -1 (r = 0: the early break
    |d_hat| <= 2 DBL_MIN before the first step, :182-183)."""
    import os

    from conftest import GOLDEN
    from otmap_amd import GridBasedTransportSolver, Precision, SolverOptions
    from otmap_amd._lib import OTM_EXIT_ZERO_DIRECTION

    g = np.load(os.path.join(GOLDEN, "edge_cases.npz"))
    k = lambda s: g[f"{tag}_{s}"]  # noqa: E731
    n = int(k("density").size ** 0.5)
    s = GridBasedTransportSolver(0)
    s.init(n)
    tm = s.solve(k("density"), SolverOptions(threshold=float(k("threshold")), precision=Precision(precision)))
    it = s.last_stats.iterations
    assert abs(it - int(k("iters"))) <= (0 if tag != "zero64" else 1), (it, int(k("iters")))
    assert np.abs(tm.fwd_mesh().xy - k("fwd")).max() <= (1e-9 if precision == 0 else 1e-6)
    assert np.allclose(tm.density(), k("adj"), rtol=1e-14, atol=0, equal_nan=True)
    if tag == "zdir16":
        assert s.last_stats.exit_reason == OTM_EXIT_ZERO_DIRECTION
    if tag == "nan16":
        assert np.isnan(s.last_stats.initial_residual)


# ------------------------------------------- line-search root classification
def test_line_search_step_choice_vs_reference_golden():
    """The device's step choice (otm_line_search_steps: the pick_step every
    solve runs) against the reference's solve_1D_problem on constructed
    quartics whose derivative cubic has a double root, exact or split by a
    2^-40..2^-48 relative perturbation into two close real roots or a complex
    pair of tiny imaginary part (reference otlib/otsolver_2dgrid.cpp:563-595;
    EigenSolver of the companion matrix, roots with imag() == 0 exactly).

    Bars: a pick of a simple, well separated root (the third root inside
    [-10, 10]) agrees to 1e-11 relative (the reference's QR iteration loses a
    few digits on such companion matrices); a pick inside a near-double pair
    agrees to 1e-5 relative (such roots are determined only to ~sqrt(eps));
    the classification (a root in [-10, 10] vs the 0.5 fallback) agrees
    except where the pair is double to within 2^-40 and the third root lies
    outside [-10, 10] -- there the reference's answer depends on the
    rounding inside its Schur iteration (it reports one exactly double root
    pair as complex and one as two reals); at most 3 of the 204 cases."""
    import os

    from conftest import GOLDEN
    from otmap_amd import _lib

    g = np.load(os.path.join(GOLDEN, "line_search_roots.npz"))
    z = np.ascontiguousarray(g["z"])
    m = len(z)
    al, em = np.zeros(m), np.zeros(m)
    assert _lib.lib().otm_line_search_steps(0, _lib.ptr(z), m, _lib.ptr(al), _lib.ptr(em)) == 0
    ra = g["alpha"]
    fb_dev, fb_ref = al == 0.5, ra == 0.5
    outside = np.abs(g["t1"]) > 10
    diverge = np.flatnonzero(fb_dev != fb_ref)
    print("classification differs at", [(int(i), int(g["disc_sign"][i]), float(g["t0"][i]), float(ra[i]),
                                          float(al[i])) for i in diverge])
    assert np.all(outside[diverge]) and len(diverge) <= 3
    both = ~fb_dev & ~fb_ref
    simple = both & ~outside
    assert np.all(np.abs(al[simple] - ra[simple]) <= 1e-11 * np.abs(ra[simple]))
    assert np.all(np.abs(al[both] - ra[both]) <= 1e-5 * np.abs(ra[both]))
    # emin = e(alpha) with the reference's Horner form
    zz = z.T
    e_al = zz[0] + al * (zz[1] + al * (zz[2] + al * (zz[3] + al * zz[4])))
    assert np.array_equal(em, e_al)
