"""Batched solves of many densities of uneven grid sizes (C ABI otm_batch_*).

The reference solves one density per GridBasedTransportSolver
(otlib/otsolver_2dgrid.cpp:106-245); its apps loop over densities
(apps/barycenters.cpp:167-197).  BatchSolver solves a whole list together: one
kernel launch per step of the iteration for all grids, HIP-graph-captured
iterations, autotuned DCT tiles.  Each grid's result is bit-identical to a
single FP64 solve of its density.

    b = BatchSolver([256, 512, 1024])
    maps = b.solve([rho256, rho512, rho1024])
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import OtmStats, check, iptr
from .solver import SolverOptions, SolveStats, _f64


class BatchSolver:
    def __init__(self, sizes, device: int = 0):
        self.sizes = [int(n) for n in sizes]
        n_arr = np.asarray(self.sizes, dtype=np.int32)
        h = C.c_void_p()
        check(_lib.lib().otm_batch_create(C.byref(h), int(device), len(self.sizes), iptr(n_arr)), "otm_batch_create")
        self._h = h
        self.last_stats: list[SolveStats] = []

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._h = None
            try:
                _lib.lib().otm_batch_destroy(h)
            except Exception:  # interpreter shutdown: module globals already cleared
                pass

    def __len__(self):
        return len(self.sizes)

    def _stats(self, st):
        self.last_stats = [SolveStats(s.iterations, s.exit_reason, s.initial_residual, s.final_residual, s.ms_solve)
                           for s in st]
        return self.last_stats

    def solve(self, densities, opt: SolverOptions | None = None):
        """Returns one TransportMap per density (host buffers)."""
        from .transport_map import GridMesh, TransportMap

        if len(densities) != len(self.sizes):
            raise ValueError(f"expected {len(self.sizes)} densities")
        opt = opt or SolverOptions()
        dens = [_f64(d, n * n) for d, n in zip(densities, self.sizes)]
        fwd = [np.empty(2 * (n + 1) ** 2) for n in self.sizes]
        adj = [np.empty(n * n) for n in self.sizes]
        k = len(self.sizes)
        PA = _lib._dp * k
        st = (OtmStats * k)()
        check(_lib.lib().otm_batch_solve(self._h, PA(*[_lib.ptr(d) for d in dens]), opt.to_c(),
                                         PA(*[_lib.ptr(f) for f in fwd]), PA(*[_lib.ptr(a) for a in adj]), st),
              "batch solve")
        self._stats(st)
        return [TransportMap(GridMesh.regular(n), GridMesh(n, f), a) for n, f, a in zip(self.sizes, fwd, adj)]

    def solve_device(self, density_ptrs, fwd_ptrs, adjusted_ptrs=None, opt: SolverOptions | None = None):
        """Solve on device buffers (lists of raw HIP pointers, e.g. tensor.data_ptr())."""
        opt = opt or SolverOptions()
        k = len(self.sizes)
        PV = C.c_void_p * k
        st = (OtmStats * k)()
        adj = PV(*adjusted_ptrs) if adjusted_ptrs is not None else None
        check(_lib.lib().otm_batch_solve_device(self._h, PV(*density_ptrs), opt.to_c(), PV(*fwd_ptrs), adj, st),
              "batch solve_device")
        return self._stats(st)

    def trace(self, job: int, cap: int | None = None):
        cap = cap or max(1, self.last_stats[job].iterations if self.last_stats else 1)
        r, a, b = np.zeros(cap), np.zeros(cap), np.zeros(cap)
        k = _lib.lib().otm_batch_trace(self._h, int(job), _lib.ptr(r), _lib.ptr(a), _lib.ptr(b), cap)
        if k < 0:
            check(k, "batch trace")
        return r[:k], a[:k], b[:k]

    def gemm_config(self, job: int = -1):
        """(folded FP64 tile shape per DCT stage of grid `job` (-1: the
        largest grid; shapes are tuned per grid size), number of candidates)."""
        cfg = np.zeros(4, dtype=np.int32)
        count = _lib.lib().otm_batch_gemm_config(self._h, int(job), iptr(cfg))
        if count < 0:
            check(count, "batch gemm_config")
        return [int(c) for c in cfg], count

    def mixed_config(self, job: int = -1):
        """(tile configuration per mixed DCT stage of grid `job` (-1: the
        largest grid), number of candidates); the count is 0 before the first
        mixed solve."""
        cfg = np.zeros(4, dtype=np.int32)
        count = _lib.lib().otm_batch_mixed_config(self._h, int(job), iptr(cfg))
        if count < 0:
            check(count, "batch mixed_config")
        return [int(c) for c in cfg], count
