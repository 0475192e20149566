"""GPU parity of the fast-DCT transform route (SolverOptions.transform = FFT).

The FFT route computes the same exact Laplacian solve as the GEMM routes
(otlib/otsolver_2dgrid.cpp:174-183, 284-321) with O(n^2 log n) work.  Bars:
the solve's residual M d_hat - R' to 1e-14 of |M| |d_hat|; against the closed
form (scipy's FFT DCT) and the GEMM route with the GEMM route's own bar
(5e-12, scaled by n^2 beyond 512) and against the oracle's sparse LDL^T
within that factor's own error; full solves against
the oracle (iterations +-1, positions 1e-9, alpha/beta traces) and against
the compiled reference's goldens of the bench workloads (iterations +-1,
positions 1e-8 single / 1e-6 batched and 2048^2); batches bit-identical to
single solves.  Sizes outside n = 2^a 3^b (8..2048, n % 4 == 0) keep the
GEMM route.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from otmap_amd import BatchSolver, GridBasedTransportSolver, SolverOptions, Transform
from otmap_amd import densities as D
from otmap_amd._lib import OtmError
from test_gpu_solver import _fft_dct_solve, _rng

pytestmark = pytest.mark.gpu

FFT = SolverOptions(transform=Transform.FFT)
SIZES = [8, 12, 16, 24, 48, 64, 96, 128, 192, 256, 384, 512, 768, 1024, 1152, 1536, 2048]


def _apply_laplacian(n, d):
    """M d with the reference's operator (otsolver_2dgrid.cpp:284-311): 2 on
    the diagonal, -1/2 at the four diagonal neighbours, indices clamped."""
    x = d.reshape(n, n)
    i = np.arange(n)
    lo, hi = np.maximum(i - 1, 0), np.minimum(i + 1, n - 1)
    nb = x[lo][:, lo] + x[lo][:, hi] + x[hi][:, lo] + x[hi][:, hi]
    return (2.0 * x - 0.5 * nb).reshape(-1)


@pytest.mark.parametrize("n", SIZES)
def test_laplacian_solve_fft_matches_closed_form(n, oracle, solver_factory):
    """Residual pin M d_hat = R' (R' is mean-free, M maps onto the mean-free
    fields): relative to |M| |d_hat| = 4 |d_hat| within 1e-14, independent of
    any other transform.  Against scipy's FFT DCT and the GEMM route: the
    GEMM route's bar 5e-12 up to 512 (both differ from scipy by a smooth
    low-mode term, 2.1e-13 at 256), scaled by the condition number ~n^2
    beyond."""
    s = solver_factory(n)
    print(f"fft laplacian n={n}", flush=True)
    r = _rng(n).standard_normal(n * n)
    tol = 5e-12 * max(1.0, (n / 512.0) ** 2)
    for rhs in (r, r - r.mean()):
        d = s.laplacian_solve_fft(rhs)
        rp = rhs.copy()
        rp[0] -= rhs.sum()
        # element (0,0) holds r - sum(r), whose host sum rounds differently
        # from numpy's (~n^2 eps); M's columns sum to zero, so the other
        # elements determine it
        res = _apply_laplacian(n, d) - rp
        assert np.abs(res[1:]).max() <= 1e-14 * 4 * np.abs(d).max()
        d = d - d.mean()
        x = _fft_dct_solve(n, rhs)
        scale = np.abs(x).max()
        assert np.abs(d - x).max() / scale < tol
        g = s.laplacian_solve(rhs)
        assert np.abs(g - g.mean() - d).max() / scale < tol
        if n <= 256:
            do = oracle.laplacian_solve(n, rhs)
            err = np.abs(d - do).max() / np.abs(do).max()
            assert err < 1e-15 * n * n + 1e-13, err


def test_fft_route_unsupported_sizes_keep_the_gemms(solver_factory):
    n = 100  # 4 * 25: no fast transform
    s = solver_factory(n)
    with pytest.raises(OtmError):
        s.laplacian_solve_fft(np.ones(n * n))
    rho = D.gaussian_mixture(n, 3)
    a = s.solve(rho, FFT).fwd_mesh().xy.copy()
    b = s.solve(rho).fwd_mesh().xy
    assert np.array_equal(a, b)


@pytest.mark.parametrize("n,seed", [(16, 7), (24, 3), (48, 5), (96, 6), (128, 10), (192, 11), (256, 1000)])
def test_full_solve_fft_matches_oracle(n, seed, oracle, solver_factory):
    rho = D.gaussian_mixture(n, seed)
    s = solver_factory(n)
    tm = s.solve(rho, FFT)
    o = oracle.solve(n, rho)
    st = s.last_stats
    assert abs(st.iterations - o["iters"]) <= 1, (st.iterations, o["iters"])
    assert st.final_residual <= 1e-7
    err = np.abs(tm.fwd_mesh().xy - o["fwd"]).max()
    assert err < 1e-9, err
    res, alpha, beta = s.trace()
    k = min(len(alpha), len(o["alpha"]))
    assert np.abs(alpha[:k] - o["alpha"][:k]).max() < 1e-7
    assert np.abs(beta[:k] - o["beta"][:k]).max() < 1e-7


def _sampled_err(xy, g, prefix, n):
    V = (n + 1) ** 2
    idx = g[f"{prefix}vidx"]
    return max(np.abs(xy[idx] - g[f"{prefix}fx"]).max(), np.abs(xy[V + idx] - g[f"{prefix}fy"]).max())


@pytest.mark.parametrize("n,fname,make", [(256, "solve_gaussian256.npz", lambda: D.gaussian_mixture(256)),
                                          (1024, "solve_shapes1024.npz", lambda: D.smoothed_shapes(1024))],
                         ids=["cfg0", "cfg1"])
def test_fft_reference_workload(n, fname, make, solver_factory):
    g = np.load(os.path.join(GOLDEN, fname))
    print(f"fft reference workload n={n}", flush=True)
    s = solver_factory(n)
    tm = s.solve(make(), FFT)
    st = s.last_stats
    assert abs(st.iterations - int(g["iters"])) <= 1, (st.iterations, int(g["iters"]))
    assert st.final_residual <= 1e-7
    err = _sampled_err(tm.fwd_mesh().xy, g, "", n)
    print(f"fft n={n}: sampled vertex error {err:.3g}")
    assert err < 1e-8, err


def test_fft_batch64_reference_workload():
    g = np.load(os.path.join(GOLDEN, "batch64.npz"))
    sizes = [int(n) for n in g["n"]]
    print("fft batch64", flush=True)
    b = BatchSolver(sizes)
    maps = b.solve(D.batch_densities(64), FFT)
    worst = 0.0
    for p, n in enumerate(sizes):
        assert abs(b.last_stats[p].iterations - int(g[f"d{p}_iters"])) <= 1, (p, n)
        assert b.last_stats[p].final_residual <= 1e-7
        err = _sampled_err(maps[p].fwd_mesh().xy, g, f"d{p}_", n)
        worst = max(worst, err)
        assert err < 1e-6, (p, n, err)
    print(f"fft batch64: largest sampled vertex error {worst:.3g}")


def test_fft_stippling_reference_workload(solver_factory):
    g = np.load(os.path.join(GOLDEN, "stipple2048.npz"))
    n = 2048
    print("fft stipple2048", flush=True)
    s = solver_factory(n)
    tm = s.solve(D.hdr_lognormal(n), SolverOptions(transform=Transform.FFT, max_ratio=float(g["max_ratio"])))
    st = s.last_stats
    assert abs(st.iterations - int(g["iters"])) <= 1, (st.iterations, int(g["iters"]))
    err = _sampled_err(tm.fwd_mesh().xy, g, "", n)
    print(f"fft stipple2048: sampled vertex error {err:.3g}")
    assert err < 1e-6, err


def test_fft_batch_bit_identical_to_single_solves():
    sizes = [64, 320, 192, 256, 128, 384]  # 320 = 64 * 5: GEMM route inside the same table
    dens = [D.gaussian_mixture(n, 60 + k, K=2 + k) for k, n in enumerate(sizes)]
    b = BatchSolver(sizes)
    maps = b.solve(dens, FFT)
    for k, (n, rho) in enumerate(zip(sizes, dens)):
        s = GridBasedTransportSolver(0)
        s.set_verbose_level(0)
        s.init(n)
        tm = s.solve(rho, FFT)
        assert b.last_stats[k].iterations == s.last_stats.iterations, (k, n)
        assert np.array_equal(maps[k].fwd_mesh().xy, tm.fwd_mesh().xy), (k, n)


def test_fft_pass_timing_entry_points(solver_factory):
    """otm_solver_profile_kernel times the whole fast-DCT solve and each of its
    three passes; sizes without the fast transform and unknown ids are
    rejected."""
    from otmap_amd import _lib

    ms = _lib.C.c_double()
    s = solver_factory(256)
    s.laplacian_solve_fft(_rng(256).standard_normal(256 * 256))
    for k in (_lib.OTM_KERNEL_FFT_SOLVE, _lib.OTM_KERNEL_FFT_ROWS_FWD, _lib.OTM_KERNEL_FFT_COLS,
              _lib.OTM_KERNEL_FFT_ROWS_INV, _lib.OTM_KERNEL_DIR_COEF_PASS):
        assert _lib.lib().otm_solver_profile_kernel(s._h, k, 3, _lib.C.byref(ms)) == 0
        assert ms.value > 0
    assert _lib.lib().otm_solver_profile_kernel(s._h, _lib.OTM_KERNEL_DIR_COEF_PASS + 1, 3, _lib.C.byref(ms)) != 0
    u = solver_factory(100)
    for k in (_lib.OTM_KERNEL_FFT_SOLVE, _lib.OTM_KERNEL_FFT_COLS):
        assert _lib.lib().otm_solver_profile_kernel(u._h, k, 3, _lib.C.byref(ms)) != 0


def test_fft_batch_mixing_both_workgroup_forms():
    """One job table with grids on both fast-DCT workgroup forms (2 rows per
    workgroup up to n = 1024, 4 rows above; 1152 = 2^7 3^2 needs a radix-3
    stage) and one on the GEMM route (320 = 64 * 5): bit-identical to single
    solves.  The batch has grids above 1024^2, so its scalar kernels run in
    the wide forms (thread groups folding more than 2048 partials), while the
    single solves of 192, 320 and 1024 run the 256-thread forms: the two
    forms give those grids the same bits."""
    sizes = [1152, 192, 2048, 320, 1024]
    dens = [D.gaussian_mixture(n, 70 + k, K=3) for k, n in enumerate(sizes)]
    print("fft batch of both forms", flush=True)
    b = BatchSolver(sizes)
    maps = b.solve(dens, SolverOptions(max_iter=12))
    for k, (n, rho) in enumerate(zip(sizes, dens)):
        s = GridBasedTransportSolver(0)
        s.set_verbose_level(0)
        s.init(n)
        tm = s.solve(rho, SolverOptions(max_iter=12))
        assert b.last_stats[k].iterations == s.last_stats.iterations, (k, n)
        assert np.array_equal(maps[k].fwd_mesh().xy, tm.fwd_mesh().xy), (k, n)
