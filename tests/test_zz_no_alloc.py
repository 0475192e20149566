"""The zero-allocation entry point otm_prune_points_device_noalloc against
otm_prune_points_device (the stippling prune, apps/stippling.cpp:184-197):
identical kept points, in order, and the same count, on two calls in a row
(the second with fewer points, so stale counts of the first call would
show), both through the C ABI; and the size limit of its fixed device
scratch is rejected on the host before any launch."""
import ctypes as C

import numpy as np
import pytest

from otmap_amd import _lib
from otmap_amd import densities as D

pytestmark = pytest.mark.gpu


def _prune(fn, n, d_rho, d_pts, m, d_out):
    kept = C.c_int64(-1)
    rc = fn(0, n, d_rho.data_ptr(), d_pts.data_ptr(), m, d_out.data_ptr(), C.byref(kept))
    assert rc == 0, _lib.lib().otm_last_error()
    return kept.value


def test_noalloc_prune_matches_allocating_prune():
    import torch

    L = _lib.lib()
    n = 128
    rho = D.gaussian_mixture(n, 12)
    rho[(np.arange(n * n) % 5) == 0] = 0.0  # zero-density cells to prune
    d_rho = torch.from_numpy(rho).to("cuda:0")
    rng = np.random.default_rng(9)
    for m in (300_000, 70_001):  # second call smaller than the first
        pts = rng.uniform(-0.05, 1.05, (m, 2))
        pts[::97] = np.nan  # unlocated points are dropped too
        d_pts = torch.from_numpy(pts.reshape(-1)).to("cuda:0")
        out_a = torch.full((2 * m,), -7.0, dtype=torch.float64, device="cuda:0")
        out_b = torch.full((2 * m,), -7.0, dtype=torch.float64, device="cuda:0")
        print("prune", m, flush=True)
        ka = _prune(L.otm_prune_points_device, n, d_rho, d_pts, m, out_a)
        kb = _prune(L.otm_prune_points_device_noalloc, n, d_rho, d_pts, m, out_b)
        assert ka == kb and 0 < ka < m
        assert torch.equal(out_a[: 2 * ka], out_b[: 2 * kb])
        assert torch.all(out_b[2 * kb:] == -7.0)  # nothing written past the kept points


def test_noalloc_prune_rejects_too_many_points():
    L = _lib.lib()
    kept = C.c_int64(-1)
    # the limit is checked before any pointer is used or any kernel launched
    rc = L.otm_prune_points_device_noalloc(0, 16, 1, 1, (1 << 24) + 1, 1, C.byref(kept))
    assert rc == _lib.OTM_ERR_ARG
    assert b"16777216" in L.otm_last_error()
