"""Python mirror of the reference's solver API (otlib/otsolver_2dgrid.h:33-58).

    solver = GridBasedTransportSolver()
    solver.init(n)
    tmap = solver.solve(density, SolverOptions())

Same names, argument meaning and non-throwing behaviour where the reference
has them; errors of the HIP layer raise OtmError.  Arrays follow the
reference's layouts (face order j + i*n, vertex order j + i*(n+1)).
"""
from __future__ import annotations

import enum
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import OtmOptions, OtmStats, check, iptr, ptr

DBL_MAX = float(np.finfo(np.float64).max)


class BetaOpt(enum.IntEnum):
    """otlib/otsolver_2dgrid.h:33-35"""
    Zero = 0
    ConjugateJacobian = 1


class Precision(enum.IntEnum):
    FP64 = _lib.OTM_PRECISION_FP64
    MIXED = _lib.OTM_PRECISION_MIXED


class Transform(enum.IntEnum):
    """How the FP64 Laplacian solves run: the DCT as dense GEMMs on FP64 MFMA,
    or the same DCTs by fast transforms (n = 2^a 3^b, 8 <= n <= 2048)."""
    GEMM = _lib.OTM_TRANSFORM_GEMM
    FFT = _lib.OTM_TRANSFORM_FFT


@dataclass
class SolverOptions:
    """otlib/otsolver_2dgrid.h:37-43, plus the arithmetic of the solve."""
    beta: BetaOpt = BetaOpt.ConjugateJacobian
    max_iter: int = 1000
    threshold: float = 1e-7
    max_ratio: float = DBL_MAX
    precision: Precision = Precision.FP64
    refine_steps: int = 1  # mixed precision: FP64 refinement steps per Laplacian solve
    # mixed precision: grids with n >= mixed_min_n (n % 64 == 0) run the BF16x3
    # transform, smaller ones the FP64 transform (0: BF16x3 on every supported size)
    mixed_min_n: int = 1024
    transform: Transform = Transform.FFT  # grids on the FP64 transform: fast DCTs (where supported) or GEMMs

    def to_c(self) -> OtmOptions:
        o = OtmOptions()
        _lib.lib().otm_options_init(o)
        o.beta = int(self.beta)
        o.max_iter = int(self.max_iter)
        o.threshold = float(self.threshold)
        o.max_ratio = float(self.max_ratio)
        o.precision = int(self.precision)
        o.refine_steps = int(self.refine_steps)
        o.mixed_min_n = int(self.mixed_min_n)
        o.transform = int(self.transform)
        return o


@dataclass
class SolveStats:
    iterations: int
    exit_reason: int
    initial_residual: float
    final_residual: float
    ms_solve: float


def _f64(a, size=None):
    a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
    if size is not None and a.size != size:
        raise ValueError(f"expected {size} values, got {a.size}")
    return a


class GridBasedTransportSolver:
    """otlib/otsolver_2dgrid.h:45-58 on one HIP device."""

    def __init__(self, device: int = 0):
        L = _lib.lib()
        h = _lib.C.c_void_p()
        check(L.otm_solver_create(_lib.C.byref(h), int(device)), "otm_solver_create")
        self._h = h
        self._n = 0
        self.last_stats: SolveStats | None = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._h = None
            try:
                _lib.lib().otm_solver_destroy(h)
            except Exception:  # interpreter shutdown: module globals already cleared
                pass

    def set_verbose_level(self, v: int) -> None:
        check(_lib.lib().otm_solver_set_verbose_level(self._h, int(v)), "set_verbose_level")

    def init(self, n: int) -> None:
        check(_lib.lib().otm_solver_init(self._h, int(n)), "init")
        self._n = int(n)

    @property
    def grid_size(self) -> int:
        return self._n

    def solve(self, density, opt: SolverOptions | None = None):
        """Returns a TransportMap (origin mesh, forward mesh, adjusted density)."""
        from .transport_map import TransportMap, GridMesh

        n = self._n
        if n <= 0:
            raise _lib.OtmError("solver not initialised: call init(n) first")
        opt = opt or SolverOptions()
        dens = _f64(density, n * n)
        fwd = np.empty(2 * (n + 1) ** 2)
        adj = np.empty(n * n)
        st = OtmStats()
        check(_lib.lib().otm_solve(self._h, ptr(dens), opt.to_c(), ptr(fwd), ptr(adj), st), "solve")
        self.last_stats = SolveStats(st.iterations, st.exit_reason, st.initial_residual, st.final_residual,
                                     st.ms_solve)
        return TransportMap(GridMesh.regular(n), GridMesh(n, fwd), adj)

    def solve_device(self, density_ptr: int, fwd_ptr: int, adjusted_ptr: int | None = None,
                     opt: SolverOptions | None = None) -> SolveStats:
        """Solve on device buffers (raw HIP pointers, e.g. tensor.data_ptr())."""
        opt = opt or SolverOptions()
        st = OtmStats()
        check(_lib.lib().otm_solve_device(self._h, density_ptr, opt.to_c(), fwd_ptr, adjusted_ptr, st),
              "solve_device")
        self.last_stats = SolveStats(st.iterations, st.exit_reason, st.initial_residual, st.final_residual,
                                     st.ms_solve)
        return self.last_stats

    def trace(self, cap: int | None = None):
        """Per-iteration residual, alpha, beta of the last solve."""
        cap = cap or max(1, self.last_stats.iterations if self.last_stats else 1)
        r, a, b = np.zeros(cap), np.zeros(cap), np.zeros(cap)
        k = _lib.lib().otm_solver_trace(self._h, ptr(r), ptr(a), ptr(b), cap)
        if k < 0:
            check(k, "trace")
        return r[:k], a[:k], b[:k]

    # -- component entry points (the reference's protected stencils) ---------------
    def vertex_gradients(self, psi):
        n = self._n
        G = np.empty(2 * (n + 1) ** 2)
        check(_lib.lib().otm_vertex_gradients(self._h, ptr(_f64(psi, n * n)), ptr(G)), "vertex_gradients")
        return G

    def residual(self, adjusted_density, psi):
        n = self._n
        r = np.empty(n * n)
        e = _lib.C.c_double()
        check(_lib.lib().otm_residual(self._h, ptr(_f64(adjusted_density, n * n)), ptr(_f64(psi, n * n)), ptr(r),
                                      _lib.C.byref(e)), "residual")
        return e.value, r

    def gemm_config(self):
        """(tile shape per folded DCT stage, number of candidate shapes); the
        count is 0 when the grid size does not fold (n % 64 != 0)."""
        cfg = np.zeros(4, dtype=np.int32)
        count = _lib.lib().otm_solver_gemm_config(self._h, iptr(cfg))
        if count < 0:
            check(count, "gemm_config")
        return [int(c) for c in cfg], count

    def set_gemm_config(self, stage: int, config: int) -> None:
        check(_lib.lib().otm_solver_set_gemm_config(self._h, stage, config), "set_gemm_config")

    def mixed_config(self):
        """(tile width config per mixed DCT stage, number of candidates); the
        count is 0 before the first mixed solve."""
        cfg = np.zeros(4, dtype=np.int32)
        count = _lib.lib().otm_solver_mixed_config(self._h, iptr(cfg))
        if count < 0:
            check(count, "mixed_config")
        return [int(c) for c in cfg], count

    def set_mixed_config(self, stage: int, config: int) -> None:
        check(_lib.lib().otm_solver_set_mixed_config(self._h, stage, config), "set_mixed_config")

    def laplacian_solve(self, r):
        n = self._n
        d = np.empty(n * n)
        check(_lib.lib().otm_laplacian_solve(self._h, ptr(_f64(r, n * n)), ptr(d)), "laplacian_solve")
        return d

    def laplacian_solve_mixed(self, r, refine_steps: int = 1):
        n = self._n
        d = np.empty(n * n)
        check(_lib.lib().otm_laplacian_solve_mixed(self._h, ptr(_f64(r, n * n)), ptr(d), int(refine_steps)),
              "laplacian_solve_mixed")
        return d

    def laplacian_solve_fft(self, r):
        n = self._n
        d = np.empty(n * n)
        check(_lib.lib().otm_laplacian_solve_fft(self._h, ptr(_f64(r, n * n)), ptr(d)), "laplacian_solve_fft")
        return d

    def line_search_coefficients(self, g0, direction):
        n = self._n
        a, b = np.empty(n * n), np.empty(n * n)
        check(_lib.lib().otm_line_search_coefficients(self._h, ptr(_f64(g0, 2 * (n + 1) ** 2)),
                                                      ptr(_f64(direction, n * n)), ptr(a), ptr(b)),
              "line_search_coefficients")
        return a, b
