"""ctypes front-end of the two CPU checkers -- TEST INFRASTRUCTURE ONLY.

* ``Oracle``    -- this project's FP64 restatement of the reference method
                   (``oracle/otmap_oracle.cpp`` -> ``oracle/liboracle.so``).
* ``Reference`` -- the reference itself (ggael/otmap) compiled from its own
                   sources by ``oracle/Makefile`` into ``oracle/_ref/libotref.so``
                   (present only where it was built; it travels with the repo).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg import
this module.  Array conventions follow the reference: face fields are flat
``j + i*n``, vertex fields ``j + i*(n+1)``, vertex pairs are SoA (all x, then
all y), point lists are interleaved (x0, y0, x1, y1, ...).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libotref.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_vp = C.c_void_p
_d, _i, _l = C.c_double, C.c_int, C.c_long


def build_oracle() -> None:
    subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _nullable(a):
    return None if a is None else _f64(a).ctypes.data_as(C.POINTER(C.c_double))


class Oracle:
    """FP64 restatement of the reference solver, BVH inverse and rasteriser."""

    def __init__(self, path: str = ORACLE_SO):
        if not os.path.exists(path):
            build_oracle()
        L = C.CDLL(path)
        self.L = L
        sig = {
            "orc_adjust_density": (None, [_i, _dp, _d]),
            "orc_vertex_gradients": (None, [_i, _dp, _dp]),
            "orc_face_area": (None, [_i, _dp, _dp]),
            "orc_residual": (_d, [_i, _dp, _dp, _dp]),
            "orc_nd_ordering": (None, [_i, _ip]),
            "orc_laplacian_create": (_vp, [_i]),
            "orc_laplacian_ok": (_i, [_vp]),
            "orc_laplacian_nnz": (_l, [_vp]),
            "orc_laplacian_solve": (None, [_vp, _dp, _dp]),
            "orc_laplacian_destroy": (None, [_vp]),
            "orc_params_1d": (None, [_i, _dp, _dp, _dp, _dp]),
            "orc_cubic_roots": (_i, [_d, _d, _d, _dp]),
            "orc_pick_step": (_d, [_dp, C.POINTER(C.c_double)]),
            "orc_solve_1d": (_d, [_i, _dp, _dp, _dp, _dp, _d, _dp, _dp, C.POINTER(C.c_double), _dp]),
            "orc_cj_beta": (_d, [_i, _dp, _dp, _dp, _dp, _dp, _dp, _d]),
            "orc_solve": (_i, [_vp, _i, _dp, _i, _i, _d, _d, _dp, _dp, C.POINTER(C.c_double),
                               _dp, _dp, _dp, _i]),
            "orc_bvh_build": (_vp, [_i, _dp, _i, _i]),
            "orc_bvh_destroy": (None, [_vp]),
            "orc_bvh_num_nodes": (_i, [_vp]),
            "orc_bvh_export": (None, [_vp, _dp, _ip, _ip]),
            "orc_inverse": (None, [_vp, _dp, _dp, _l, _i, _dp, _ip, _ip]),
            "orc_locate_all": (_i, [_i, _dp, _d, _d, _ip, _i]),
            "orc_rasterize": (None, [_i, _dp, C.POINTER(C.c_double), _i, _i, _dp]),
            "orc_raster_stats": (None, [_i, _dp, _i, _i, C.POINTER(C.c_longlong)]),
            "orc_transport_cost": (_d, [_i, _dp, _dp, _dp, C.POINTER(C.c_double)]),
            "orc_solver_transport_cost": (_d, [_i, _dp, _dp, C.POINTER(C.c_double)]),
            "orc_set_area_precision": (None, [_i]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        self._laps = {}

    # -- stencils ---------------------------------------------------------
    def adjust_density(self, n, rho, max_ratio=np.finfo(np.float64).max):
        r = _f64(rho).copy()
        self.L.orc_adjust_density(n, r, max_ratio)
        return r

    def vertex_gradients(self, n, psi):
        G = np.zeros(2 * (n + 1) ** 2)
        self.L.orc_vertex_gradients(n, _f64(psi), G)
        return G

    def face_area(self, n, G):
        a = np.zeros(n * n)
        self.L.orc_face_area(n, _f64(G), a)
        return a

    def residual(self, n, rho_adj, psi):
        r = np.zeros(n * n)
        e = self.L.orc_residual(n, _f64(rho_adj), _f64(psi), r)
        return e, r

    def nd_ordering(self, n):
        p = np.zeros(n * n, dtype=np.int32)
        self.L.orc_nd_ordering(n, p)
        return p

    # -- Laplacian ----------------------------------------------------------
    def laplacian(self, n):
        if n not in self._laps:
            h = self.L.orc_laplacian_create(n)
            if not self.L.orc_laplacian_ok(h):
                raise RuntimeError("oracle LDL^T factorisation failed")
            self._laps[n] = h
        return self._laps[n]

    def laplacian_solve(self, n, r):
        d = np.zeros(n * n)
        self.L.orc_laplacian_solve(self.laplacian(n), _f64(r), d)
        return d

    def laplacian_nnz(self, n):
        return int(self.L.orc_laplacian_nnz(self.laplacian(n)))

    # -- line search ----------------------------------------------------------
    def params_1d(self, n, g0, direction):
        a = np.zeros(n * n)
        b = np.zeros(n * n)
        self.L.orc_params_1d(n, _f64(g0), _f64(direction), a, b)
        return a, b

    def cubic_roots(self, c2, c1, c0):
        out = np.zeros(3)
        k = self.L.orc_cubic_roots(c2, c1, c0, out)
        return out[:k]

    def pick_step(self, z):
        al = C.c_double()
        e = self.L.orc_pick_step(_f64(z), C.byref(al))
        return al.value, e

    def solve_1d(self, n, g0, xk, direction, rk, ek):
        g = _f64(g0).copy()
        F = n * n
        xk1, rk1, z = np.zeros(F), np.zeros(F), np.zeros(5)
        al = C.c_double()
        e = self.L.orc_solve_1d(n, g, _f64(xk), _f64(direction), _f64(rk), ek, xk1, rk1, C.byref(al), z)
        return dict(e=e, alpha=al.value, xk1=xk1, rk1=rk1, g0=g, z=z)

    def cj_beta(self, n, rho_adj, xk, rkm1, rk, dhat, dprev, alpha):
        return self.L.orc_cj_beta(n, _f64(rho_adj), _f64(xk), _f64(rkm1), _f64(rk), _f64(dhat),
                                  _f64(dprev), alpha)

    # -- full solve -----------------------------------------------------------
    def solve(self, n, density, beta_cj=True, max_iter=1000, threshold=1e-7,
              max_ratio=np.finfo(np.float64).max, trace_cap=None):
        cap = max_iter if trace_cap is None else trace_cap
        cap = max(cap, 1)
        V = (n + 1) ** 2
        fwd = np.zeros(2 * V)
        adj = np.zeros(n * n)
        res0 = C.c_double()
        rt, at, bt = np.zeros(cap), np.zeros(cap), np.zeros(cap)
        it = self.L.orc_solve(self.laplacian(n), n, _f64(density), int(beta_cj), max_iter, threshold,
                              max_ratio, fwd, adj, C.byref(res0), rt, at, bt, cap)
        if it < 0:
            raise RuntimeError("oracle solve failed")
        k = min(it, cap)
        return dict(iters=it, fwd=fwd, density=adj, res0=res0.value, res=rt[:k], alpha=at[:k], beta=bt[:k])

    # -- inverse map ------------------------------------------------------------
    def bvh(self, n, fwd, leaf_size=4, max_depth=24):
        return _OracleBVH(self, n, fwd, leaf_size, max_depth)

    def inverse(self, n, fwd, density, pts, fast=False):
        with self.bvh(n, fwd) as b:
            return b.inverse(density, pts, fast)

    def locate_all(self, n, fwd, q, cap=16):
        out = np.zeros(cap, dtype=np.int32)
        k = self.L.orc_locate_all(n, _f64(fwd), float(q[0]), float(q[1]), out, cap)
        return out[:k]

    # -- raster / costs -------------------------------------------------------------
    def rasterize(self, n, xy, rows, cols, density=None):
        img = np.zeros(rows * cols)
        self.L.orc_rasterize(n, _f64(xy), _nullable(density), rows, cols, img)
        return img.reshape(rows, cols)

    def raster_stats(self, n, xy, rows, cols):
        """How often rasterize's carried weights (D < 0 pixels) decide a pixel."""
        out = (C.c_longlong * 7)()
        self.L.orc_raster_stats(n, _f64(xy), rows, cols, out)
        keys = ("pixels", "failed", "carried_same_face", "carried_prev_face", "uncovered_no_weights",
                "failed_but_covered", "max_faces_back")
        return dict(zip(keys, (int(v) for v in out)))

    def transport_cost(self, n, src, dst, density, per_face=False):
        cost = np.zeros(n * n) if per_face else None
        ptr = cost.ctypes.data_as(C.POINTER(C.c_double)) if per_face else None
        r = self.L.orc_transport_cost(n, _f64(src), _f64(dst), _f64(density), ptr)
        return (r, cost) if per_face else r

    def solver_transport_cost(self, n, G, rho):
        cost = np.zeros(n * n)
        r = self.L.orc_solver_transport_cost(n, _f64(G), _f64(rho), cost.ctypes.data_as(C.POINTER(C.c_double)))
        return r, cost


class _OracleBVH:
    def __init__(self, orc, n, fwd, leaf_size, max_depth):
        self.o, self.n = orc, n
        self.fwd = _f64(fwd)
        self.h = orc.L.orc_bvh_build(n, self.fwd, leaf_size, max_depth)

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def close(self):
        if self.h:
            self.o.L.orc_bvh_destroy(self.h)
            self.h = None

    def export(self):
        k = self.o.L.orc_bvh_num_nodes(self.h)
        boxes = np.zeros(4 * k)
        meta = np.zeros(3 * k, dtype=np.int32)
        faces = np.zeros(self.n * self.n, dtype=np.int32)
        self.o.L.orc_bvh_export(self.h, boxes, meta, faces)
        return boxes.reshape(k, 4), meta.reshape(k, 3), faces

    def inverse(self, density, pts, fast=False):
        p = _f64(pts).reshape(-1)
        m = p.size // 2
        out = np.zeros(2 * m)
        face = np.zeros(m, dtype=np.int32)
        nh = np.zeros(m, dtype=np.int32)
        self.o.L.orc_inverse(self.h, _f64(density), p, m, int(fast), out, face, nh)
        return out.reshape(m, 2), face, nh


class Reference:
    """The compiled reference (oracle/_ref/libotref.so)."""

    def __init__(self, path: str = REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        L = C.CDLL(path)
        self.L = L
        P = C.POINTER(C.c_double)
        sig = {
            "ref_solve": (_i, [_i, _dp, _i, _i, _d, _d, _dp, _dp, C.POINTER(C.c_int), P, _dp, _dp, _dp,
                               _i, _i, P, P]),
            "ref_adjust_density": (None, [_i, _dp, _d]),
            "ref_vertex_gradients": (None, [_i, _dp, _dp]),
            "ref_residual": (_d, [_i, _dp, _dp, _dp]),
            "ref_laplacian_solve": (None, [_i, _dp, _dp]),
            "ref_1d_params": (None, [_i, _dp, _dp, _dp, _dp]),
            "ref_solve_1d": (_d, [_i, _dp, _dp, _dp, _dp, _d, _dp, _dp, _dp, P]),
            "ref_cj_beta": (_d, [_i, _dp, _dp, _dp, _dp, _dp, _dp, _d]),
            "ref_nestdiss": (None, [_i, _ip]),
            "ref_inverse": (_i, [_i, _dp, _dp, _dp, _l, _i, _dp, _ip, _ip, P, P]),
            "ref_apply_inverse_map": (_i, [_i, _dp, _dp, _dp, _l, P]),
            "ref_rasterize": (None, [_i, _dp, P, _i, _i, _dp]),
            "ref_transport_cost": (_d, [_i, _dp, _dp, _dp, P]),
            "ref_solver_transport_cost": (_d, [_i, _dp, _dp, P]),
            "ref_solve_timed": (_i, [_i, _dp, _i, P, _dp]),
            "ref_load_point_cloud_dat": (_l, [C.c_char_p, _dp, _l]),
            "ref_blue_noise_tile": (_l, [_i, _dp, _l, _dp, _l]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args

    def solve(self, n, density, beta_cj=True, max_iter=1000, threshold=1e-7,
              max_ratio=np.finfo(np.float64).max, traced=True):
        V = (n + 1) ** 2
        fwd = np.zeros(2 * V)
        adj = np.zeros(n * n)
        cap = max(max_iter, 1)
        rt, at, bt = np.zeros(cap), np.zeros(cap), np.zeros(cap)
        iters = C.c_int()
        res0, ti, ts = C.c_double(), C.c_double(), C.c_double()
        self.L.ref_solve(n, _f64(density), int(beta_cj), max_iter, threshold, max_ratio, fwd, adj,
                         C.byref(iters), C.byref(res0), rt, at, bt, cap, int(traced), C.byref(ti), C.byref(ts))
        k = max(iters.value, 0)
        return dict(iters=iters.value, fwd=fwd, density=adj, res0=res0.value, res=rt[:k], alpha=at[:k],
                    beta=bt[:k], t_init=ti.value, t_solve=ts.value)

    def load_point_cloud_dat(self, path, cap=1 << 20):
        out = np.zeros(2 * cap)
        k = self.L.ref_load_point_cloud_dat(str(path).encode(), out, cap)
        if k < 0:
            raise FileNotFoundError(path)
        return out[: 2 * min(k, cap)].reshape(-1, 2)

    def blue_noise_tile(self, n, tile):
        t = _f64(tile).reshape(-1)
        nb = t.size // 2
        nt = (n + int(np.sqrt(nb)) - 1) // int(np.sqrt(nb))
        cap = nt * nt * nb
        out = np.zeros(2 * cap)
        k = self.L.ref_blue_noise_tile(n, t, nb, out, cap)
        return out[: 2 * k].reshape(-1, 2)

    def solve_timed(self, n, density, reps):
        """init(n) once, then `reps` solves of `density`: (init s, [solve s])."""
        ti = C.c_double()
        ts = np.zeros(reps)
        self.L.ref_solve_timed(n, _f64(density), reps, C.byref(ti), ts)
        return ti.value, ts

    def adjust_density(self, n, rho, max_ratio=np.finfo(np.float64).max):
        r = _f64(rho).copy()
        self.L.ref_adjust_density(n, r, max_ratio)
        return r

    def vertex_gradients(self, n, psi):
        G = np.zeros(2 * (n + 1) ** 2)
        self.L.ref_vertex_gradients(n, _f64(psi), G)
        return G

    def residual(self, n, rho_adj, psi):
        r = np.zeros(n * n)
        e = self.L.ref_residual(n, _f64(rho_adj), _f64(psi), r)
        return e, r

    def laplacian_solve(self, n, r):
        d = np.zeros(n * n)
        self.L.ref_laplacian_solve(n, _f64(r), d)
        return d

    def params_1d(self, n, g0, direction):
        a, b = np.zeros(n * n), np.zeros(n * n)
        self.L.ref_1d_params(n, _f64(g0), _f64(direction), a, b)
        return a, b

    def solve_1d(self, n, g0, xk, direction, rk, ek):
        F, V = n * n, (n + 1) ** 2
        xk1, rk1, go = np.zeros(F), np.zeros(F), np.zeros(2 * V)
        al = C.c_double()
        e = self.L.ref_solve_1d(n, _f64(g0), _f64(xk), _f64(direction), _f64(rk), ek, xk1, rk1, go, C.byref(al))
        return dict(e=e, alpha=al.value, xk1=xk1, rk1=rk1, g0=go)

    def cj_beta(self, n, rho_adj, xk, rkm1, rk, dhat, dprev, alpha):
        return self.L.ref_cj_beta(n, _f64(rho_adj), _f64(xk), _f64(rkm1), _f64(rk), _f64(dhat), _f64(dprev),
                                  alpha)

    def nd_ordering(self, n):
        p = np.zeros(n * n, dtype=np.int32)
        self.L.ref_nestdiss(n, p)
        return p

    def inverse(self, n, fwd, density, pts, fast=False):
        p = _f64(pts).reshape(-1)
        m = p.size // 2
        out = np.zeros(2 * m)
        ff = np.zeros(m, dtype=np.int32)
        nh = np.zeros(m, dtype=np.int32)
        tb, tq = C.c_double(), C.c_double()
        self.L.ref_inverse(n, _f64(fwd), _f64(density), p, m, int(fast), out, ff, nh, C.byref(tb), C.byref(tq))
        return out.reshape(m, 2), ff, nh

    def apply_inverse_map(self, n, fwd, density, pts):
        p = _f64(pts).reshape(-1).copy()
        t = C.c_double()
        self.L.ref_apply_inverse_map(n, _f64(fwd), _f64(density), p, p.size // 2, C.byref(t))
        return p.reshape(-1, 2), t.value

    def rasterize(self, n, xy, rows, cols, density=None):
        img = np.zeros(rows * cols)
        self.L.ref_rasterize(n, _f64(xy), _nullable(density), rows, cols, img)
        return img.reshape(rows, cols)

    def raster_stats(self, n, xy, rows, cols):
        """How often rasterize's carried weights (D < 0 pixels) decide a pixel."""
        out = (C.c_longlong * 7)()
        self.L.orc_raster_stats(n, _f64(xy), rows, cols, out)
        keys = ("pixels", "failed", "carried_same_face", "carried_prev_face", "uncovered_no_weights",
                "failed_but_covered", "max_faces_back")
        return dict(zip(keys, (int(v) for v in out)))

    def transport_cost(self, n, src, dst, density, per_face=False):
        cost = np.zeros(n * n) if per_face else None
        ptr = cost.ctypes.data_as(C.POINTER(C.c_double)) if per_face else None
        r = self.L.ref_transport_cost(n, _f64(src), _f64(dst), _f64(density), ptr)
        return (r, cost) if per_face else r

    def solver_transport_cost(self, n, G, rho_adj):
        """GridBasedTransportSolver::compute_transport_cost on vertex
        displacements G (SoA) with the adjusted density rho_adj: (sum, per face)."""
        cost = np.zeros(n * n)
        r = self.L.ref_solver_transport_cost(n, _f64(G), _f64(rho_adj), cost.ctypes.data_as(C.POINTER(C.c_double)))
        return r, cost


def reference_available() -> bool:
    return os.path.exists(REF_SO)
