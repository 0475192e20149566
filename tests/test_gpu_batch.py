"""GPU parity of the batched solver (otm_batch_*).

Every grid of a batch runs the single-grid FP64 computation over a shared job
table, so the bar is bit-identity with single-grid solves of the same
densities: same iteration counts, same alpha/beta traces, identical forward
positions and adjusted densities.  The single-grid path is itself checked
against the oracle and the reference elsewhere; here the smallest sizes are
also checked against the oracle directly (1e-9, as for single solves).
"""
import numpy as np
import pytest

from otmap_amd import BatchSolver, GridBasedTransportSolver, Precision, SolverOptions
from otmap_amd import densities as D

pytestmark = pytest.mark.gpu


def single(n, rho, opt=None):
    s = GridBasedTransportSolver(0)
    s.set_verbose_level(0)
    s.init(n)
    tm = s.solve(rho, opt)
    return tm, s.last_stats, s.trace()


def test_batch_bit_identical_to_single_solves(oracle):
    sizes = [64, 128, 192, 256, 128]
    dens = [D.gaussian_mixture(n, 50 + k, K=2 + k) for k, n in enumerate(sizes)]
    print("batch sizes", sizes, flush=True)
    b = BatchSolver(sizes)
    maps = b.solve(dens)
    for k, (n, rho) in enumerate(zip(sizes, dens)):
        tm, st, (res, alpha, beta) = single(n, rho)
        assert b.last_stats[k].iterations == st.iterations, (k, n)
        assert np.array_equal(maps[k].fwd_mesh().xy, tm.fwd_mesh().xy), (k, n)
        assert np.array_equal(maps[k].density(), tm.density())
        bres, balpha, bbeta = b.trace(k)
        assert np.array_equal(balpha, alpha) and np.array_equal(bbeta, beta) and np.array_equal(bres, res)
    o = oracle.solve(64, dens[0])
    assert abs(b.last_stats[0].iterations - o["iters"]) <= 1
    assert np.abs(maps[0].fwd_mesh().xy - o["fwd"]).max() < 1e-9


def test_batch_uneven_convergence_and_repeat():
    # a constant density stops at iteration 0 while the others keep iterating;
    # a second solve on the same handle (replayed graph) gives the same bits
    sizes = [128, 64, 256]
    dens = [D.gaussian_mixture(128, 7), np.full(64 * 64, 2.0), D.gaussian_mixture(256, 8, K=9)]
    b = BatchSolver(sizes)
    m1 = b.solve(dens)
    it1 = [s.iterations for s in b.last_stats]
    assert it1[1] == 0 and it1[0] > 0 and it1[2] > 0
    ref = np.concatenate([np.repeat(np.arange(65.0), 65), np.tile(np.arange(65.0), 65)]) / 64
    assert np.abs(m1[1].fwd_mesh().xy - ref).max() <= 1e-15
    m2 = b.solve(dens)
    assert [s.iterations for s in b.last_stats] == it1
    for a, c in zip(m1, m2):
        assert np.array_equal(a.fwd_mesh().xy, c.fwd_mesh().xy)


@pytest.mark.parametrize("max_iter", [0, 3])
def test_batch_max_iter(max_iter):
    sizes = [64, 128]
    dens = [D.gaussian_mixture(n, 11) for n in sizes]
    b = BatchSolver(sizes)
    opt = SolverOptions(max_iter=max_iter)
    maps = b.solve(dens, opt)
    for k, n in enumerate(sizes):
        tm, st, _ = single(n, dens[k], opt)
        assert b.last_stats[k].iterations == st.iterations == max_iter
        assert np.array_equal(maps[k].fwd_mesh().xy, tm.fwd_mesh().xy)


def test_batch_device_buffers():
    import torch

    sizes = [128, 64]
    dens = [D.gaussian_mixture(n, 21) for n in sizes]
    dev = torch.device("cuda:0")
    d_in = [torch.from_numpy(d).to(dev) for d in dens]
    d_out = [torch.empty(2 * (n + 1) ** 2, dtype=torch.float64, device=dev) for n in sizes]
    d_adj = [torch.empty(n * n, dtype=torch.float64, device=dev) for n in sizes]
    torch.cuda.synchronize()
    b = BatchSolver(sizes)
    b.solve_device([t.data_ptr() for t in d_in], [t.data_ptr() for t in d_out], [t.data_ptr() for t in d_adj])
    for k, n in enumerate(sizes):
        tm, st, _ = single(n, dens[k])
        assert np.array_equal(d_out[k].cpu().numpy(), tm.fwd_mesh().xy)
        assert np.array_equal(d_adj[k].cpu().numpy(), tm.density())


def test_batch_rejects_bad_input():
    from otmap_amd._lib import OtmError

    with pytest.raises(OtmError):
        BatchSolver([100])
    b = BatchSolver([64])
    with pytest.raises(OtmError):
        b.solve([np.ones(64 * 64)], SolverOptions(precision=Precision.MIXED, refine_steps=9))
    cfg, count = b.gemm_config()
    assert count > 0 and all(0 <= c < count for c in cfg)


def test_batch_mixed_bit_identical_to_single_mixed(oracle):
    sizes = [192, 128, 256]
    dens = [D.gaussian_mixture(n, 70 + k) for k, n in enumerate(sizes)]
    opt = SolverOptions(precision=Precision.MIXED, mixed_min_n=0)
    print("mixed batch sizes", sizes, flush=True)
    b = BatchSolver(sizes)
    maps = b.solve(dens, opt)
    for k, n in enumerate(sizes):
        tm, st, _ = single(n, dens[k], opt)
        assert b.last_stats[k].iterations == st.iterations, (k, n)
        assert np.array_equal(maps[k].fwd_mesh().xy, tm.fwd_mesh().xy), (k, n)
        o = oracle.solve(n, dens[k])
        assert abs(st.iterations - o["iters"]) <= 1
        assert np.abs(tm.fwd_mesh().xy - o["fwd"]).max() < 1e-6


def test_batch_tiles_tuned_per_grid_size():
    """Each grid size of a batch runs the tile configurations tuned for it
    (valid for that size, shared by every job of that size), and the results
    stay bit-identical to single solves, which tune their own size alone."""
    sizes = [192, 512, 192, 512]
    dens = [D.gaussian_mixture(n, 90 + k) for k, n in enumerate(sizes)]
    b = BatchSolver(sizes)
    for prec in (Precision.FP64, Precision.MIXED):
        opt = SolverOptions(precision=prec, mixed_min_n=0)
        maps = b.solve(dens, opt)
        for k, n in enumerate(sizes):
            tm, st, _ = single(n, dens[k], opt)
            assert b.last_stats[k].iterations == st.iterations, (prec, k)
            assert np.array_equal(maps[k].fwd_mesh().xy, tm.fwd_mesh().xy), (prec, k)
    h = {n: n // 2 for n in sizes}
    g = [b.gemm_config(k)[0] for k in range(len(sizes))]
    m = [b.mixed_config(k)[0] for k in range(len(sizes))]
    print("fp64 tiles", g, "mixed tiles", m, flush=True)
    assert g[0] == g[2] and g[1] == g[3] and m[0] == m[2] and m[1] == m[3]
    widths = [32, 64, 128, 128]
    for k, n in enumerate(sizes):
        assert all(h[n] % widths[c] == 0 for c in m[k]), (n, m[k])
    assert b.gemm_config()[0] == g[1] and b.mixed_config()[0] == m[1]  # -1: the largest grid
