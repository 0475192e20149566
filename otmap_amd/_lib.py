"""ctypes binding of libotmap_amd.so (the C ABI declared in include/otmap_capi.h).

The library is loaded from this package directory only (built in-tree by
otmap_amd/build.py).  There is no fallback: a missing or unloadable library
raises immediately, so nothing silently runs on another code path.
"""
from __future__ import annotations

import ctypes as C
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_PATH = os.path.join(HERE, "libotmap_amd.so")
HEADERS = [os.path.join(ROOT, "include", "otmap_capi.h")]

OTM_OK = 0
OTM_ERR_ARG, OTM_ERR_HIP, OTM_ERR_STATE, OTM_ERR_ALLOC = -1, -2, -3, -4
OTM_BETA_ZERO, OTM_BETA_CONJUGATE_JACOBIAN = 0, 1
OTM_PRECISION_FP64, OTM_PRECISION_MIXED = 0, 1
OTM_TRANSFORM_GEMM, OTM_TRANSFORM_FFT = 0, 1
OTM_EXIT_NONE, OTM_EXIT_CONVERGED_OR_MAX_ITER, OTM_EXIT_ZERO_DIRECTION = 0, 1, 2
(OTM_KERNEL_DCT_GEMM, OTM_KERNEL_BETA, OTM_KERNEL_DIR_COEF, OTM_KERNEL_UPDATE, OTM_KERNEL_DCT_SOLVE,
 OTM_KERNEL_MIXED_GEMM, OTM_KERNEL_MIXED_SOLVE, OTM_KERNEL_FFT_SOLVE, OTM_KERNEL_FFT_ROWS_FWD, OTM_KERNEL_FFT_COLS,
 OTM_KERNEL_FFT_ROWS_INV, OTM_KERNEL_DIR_COEF_PASS) = range(12)


class OtmOptions(C.Structure):
    _fields_ = [("beta", C.c_int), ("max_iter", C.c_int), ("threshold", C.c_double),
                ("max_ratio", C.c_double), ("precision", C.c_int), ("refine_steps", C.c_int),
                ("mixed_min_n", C.c_int), ("transform", C.c_int)]


class OtmStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("exit_reason", C.c_int), ("initial_residual", C.c_double),
                ("final_residual", C.c_double), ("ms_solve", C.c_double)]


class OtmError(RuntimeError):
    pass


_vp = C.c_void_p
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_SIGNATURES = {
    "otm_options_init": (None, [C.POINTER(OtmOptions)]),
    "otm_last_error": (C.c_char_p, []),
    "otm_solver_create": (C.c_int, [C.POINTER(_vp), C.c_int]),
    "otm_solver_destroy": (C.c_int, [_vp]),
    "otm_solver_set_verbose_level": (C.c_int, [_vp, C.c_int]),
    "otm_solver_init": (C.c_int, [_vp, C.c_int]),
    "otm_solver_grid_size": (C.c_int, [_vp]),
    "otm_solve": (C.c_int, [_vp, _dp, C.POINTER(OtmOptions), _dp, _dp, C.POINTER(OtmStats)]),
    "otm_solve_device": (C.c_int, [_vp, _vp, C.POINTER(OtmOptions), _vp, _vp, C.POINTER(OtmStats)]),
    "otm_solver_trace": (C.c_int, [_vp, _dp, _dp, _dp, C.c_int]),
    "otm_solver_profile_kernel": (C.c_int, [_vp, C.c_int, C.c_int, _dp]),
    "otm_solver_gemm_config": (C.c_int, [_vp, _ip]),
    "otm_solver_set_gemm_config": (C.c_int, [_vp, C.c_int, C.c_int]),
    "otm_solver_mixed_config": (C.c_int, [_vp, _ip]),
    "otm_solver_set_mixed_config": (C.c_int, [_vp, C.c_int, C.c_int]),
    "otm_vertex_gradients": (C.c_int, [_vp, _dp, _dp]),
    "otm_residual": (C.c_int, [_vp, _dp, _dp, _dp, _dp]),
    "otm_laplacian_solve": (C.c_int, [_vp, _dp, _dp]),
    "otm_map_bvh_export": (C.c_int, [_vp, _dp, _ip, _ip, _ip, _ip]),
    "otm_line_search_steps": (C.c_int, [C.c_int, _dp, C.c_int, _dp, _dp]),
    "otm_laplacian_solve_mixed": (C.c_int, [_vp, _dp, _dp, C.c_int]),
    "otm_laplacian_solve_fft": (C.c_int, [_vp, _dp, _dp]),
    "otm_line_search_coefficients": (C.c_int, [_vp, _dp, _dp, _dp, _dp]),
    "otm_map_create": (C.c_int, [C.POINTER(_vp), C.c_int, C.c_int, _dp, _dp]),
    "otm_map_create_device": (C.c_int, [C.POINTER(_vp), C.c_int, C.c_int, _vp, _vp]),
    "otm_map_destroy": (C.c_int, [_vp]),
    "otm_map_init_inverse": (C.c_int, [_vp]),
    "otm_map_bvh_size": (C.c_int, [_vp]),
    "otm_map_inverse": (C.c_int, [_vp, _dp, C.c_int64, C.c_int, _dp, _ip, _ip]),
    "otm_map_inverse_device": (C.c_int, [_vp, _vp, C.c_int64, C.c_int, _vp, _vp, _vp]),
    "otm_rasterize": (C.c_int, [C.c_int, C.c_int, _dp, _dp, C.c_int, C.c_int, _dp]),
    "otm_rasterize_device": (C.c_int, [C.c_int, C.c_int, _vp, _vp, C.c_int, C.c_int, C.c_int, _vp, _vp]),
    "otm_map_inverse_vertices": (C.c_int, [_vp, C.c_int, _dp]),
    "otm_map_inverse_vertices_device": (C.c_int, [_vp, C.c_int, _vp]),
    "otm_transport_cost": (C.c_int, [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp]),
    "otm_transport_cost_device": (C.c_int, [C.c_int, C.c_int, _vp, _vp, _vp, _vp, _dp]),
    "otm_grid_transport_cost": (C.c_int, [C.c_int, C.c_int, _dp, _dp, _dp, _dp]),
    "otm_mesh_lerp_device": (C.c_int, [C.c_int, C.c_long, C.c_int, C.POINTER(_vp), C.c_double, C.c_double, _vp]),
    "otm_image_normalize_mean_device": (C.c_int, [C.c_int, C.c_long, _vp, C.c_double, C.c_int, _vp, _dp]),
    "otm_prune_points_device": (C.c_int, [C.c_int, C.c_int, _vp, _vp, C.c_int64, _vp, C.POINTER(C.c_int64)]),
    "otm_blue_noise_tile_device": (C.c_int, [C.c_int, C.c_int, _vp, C.c_int64, _vp, C.c_int64,
                                             C.POINTER(C.c_int64)]),
    "otm_prune_points_device_noalloc": (C.c_int, [C.c_int, C.c_int, _vp, _vp, C.c_int64, _vp,
                                                  C.POINTER(C.c_int64)]),
    "otm_batch_create": (C.c_int, [C.POINTER(_vp), C.c_int, C.c_int, _ip]),
    "otm_batch_destroy": (C.c_int, [_vp]),
    "otm_batch_size": (C.c_int, [_vp]),
    "otm_batch_solve": (C.c_int, [_vp, C.POINTER(_dp), C.POINTER(OtmOptions), C.POINTER(_dp), C.POINTER(_dp),
                                  C.POINTER(OtmStats)]),
    "otm_batch_solve_device": (C.c_int, [_vp, C.POINTER(_vp), C.POINTER(OtmOptions), C.POINTER(_vp),
                                         C.POINTER(_vp), C.POINTER(OtmStats)]),
    "otm_batch_trace": (C.c_int, [_vp, C.c_int, _dp, _dp, _dp, C.c_int]),
    "otm_batch_gemm_config": (C.c_int, [_vp, C.c_int, _ip]),
    "otm_batch_mixed_config": (C.c_int, [_vp, C.c_int, _ip]),
}

_lib = None


def declared_symbols() -> list[str]:
    """Function names declared by the public C headers."""
    names = []
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"\b(otm_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def lib():
    """The loaded library (raises OtmError if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OtmError(f"{LIB_PATH} is missing: build it with `python -m otmap_amd.build` "
                       "(or __graft_entry__.build()); there is no CPU fallback")
    _init_torch_runtime_first()
    L = C.CDLL(LIB_PATH)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _init_torch_runtime_first() -> None:
    """PyTorch-ROCm wheels bundle their own HIP/HSA runtime, so a process using
    both torch and this library holds two runtimes; torch's can only start if
    it starts first.  When torch is already imported, start it before this
    library touches the device."""
    import sys

    torch = sys.modules.get("torch")
    if torch is not None and getattr(torch, "cuda", None) is not None and torch.cuda.device_count() > 0:
        torch.cuda.init()


def check(rc: int, what: str) -> None:
    if rc != OTM_OK:
        msg = lib().otm_last_error().decode(errors="replace")
        raise OtmError(f"{what} failed (code {rc}): {msg}")


def ptr(a):
    """double* of a C-contiguous float64 numpy array (None passes NULL)."""
    if a is None:
        return None
    return a.ctypes.data_as(_dp)


def iptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(_ip)
