"""The reference apps' pipelines around the solver, on device buffers.

barycentres   apps/barycenters.cpp:167-290: inverse map of each density as a
              mesh, linear (2 densities) or bilinear (4 densities)
              interpolation of the inverse meshes, each interpolated mesh
              rasterised (unit density per face) and scaled to the
              interpolated mean of the input densities.
stippling     apps/stippling.cpp:140-197: points (given, or a blue-noise tile
              repeated at a resolution, apps/common/image_utils.cpp:148-173)
              advected through the inverse map, then the points landing on
              zero density pruned.
composite_map apps/otmap.cpp:119-132: map u -> v = inverse map of v applied to
              the forward mesh of u.

Device buffers are torch tensors on cuda:0 (PyTorch-ROCm); every step is a
HIP kernel of libotmap_amd.so.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .transport_map import GridMesh, TransportMap


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise _lib.OtmError("the pipelines need a HIP device")
    return torch


def lerp(u, a0, a1):
    """apps/barycenters.cpp:137-140"""
    return (1. - u) * a0 + u * a1


def bilerp(u, v, a0, a1, a2, a3):
    """apps/barycenters.cpp:127-134"""
    return (1. - u) * (1. - v) * a0 + (u) * (1. - v) * a1 + (u) * (v) * a2 + (1. - u) * (v) * a3


def inverse_meshes_device(maps, device: int = 0):
    """Inverse map of each TransportMap as a mesh, in device memory (SoA)."""
    torch = _torch()
    out = []
    for tm in maps:
        n = tm.fwd_mesh().n
        d = torch.empty(2 * (n + 1) ** 2, dtype=torch.float64, device=f"cuda:{device}")
        _lib.check(_lib.lib().otm_map_inverse_vertices_device(tm._handle(), 0, d.data_ptr()), "inverse_vertices")
        out.append(d)
    return out


def barycentres(maps, steps: int, res: int, means, f32: bool = True, device: int = 0):
    """Images of the interpolated inverse maps at alpha (, beta) = k / steps.

    maps:  2 or 4 TransportMaps of one grid size; means: the mean of each
    input density.  Returns {(i, j): image} (4 maps) or {i: image} (2 maps),
    res x res tensors (float32 when f32, else float64) on the device, row
    y / column x as the reference's img(y, x)."""
    torch = _torch()
    k = len(maps)
    if k not in (2, 4):
        raise ValueError("barycentres interpolate 2 or 4 maps")
    n = maps[0].fwd_mesh().n
    if any(tm.fwd_mesh().n != n for tm in maps):
        raise ValueError("maps of different grid sizes")
    L = _lib.lib()
    dev = f"cuda:{device}"
    inv = inverse_meshes_device(maps, device)
    count = 2 * (n + 1) ** 2
    mesh = torch.empty(count, dtype=torch.float64, device=dev)
    img = torch.empty(res * res, dtype=torch.float64, device=dev)
    winner = torch.empty(res * res, dtype=torch.int32, device=dev)
    ptrs = (C.c_void_p * k)(*[t.data_ptr() for t in inv])
    out = {}
    for i in range(steps + 1):
        alpha = i / steps
        for j in range(steps + 1 if k == 4 else 1):
            beta = j / steps
            _lib.check(L.otm_mesh_lerp_device(device, count, k, ptrs, alpha, beta, mesh.data_ptr()), "mesh_lerp")
            _lib.check(L.otm_rasterize_device(device, n, mesh.data_ptr(), None, res, res, 0, img.data_ptr(),
                                              winner.data_ptr()), "rasterize")
            expected = bilerp(alpha, beta, *means) if k == 4 else lerp(alpha, *means)
            dst = torch.empty(res * res, dtype=torch.float32 if f32 else torch.float64, device=dev)
            _lib.check(L.otm_image_normalize_mean_device(device, res * res, img.data_ptr(), float(expected),
                                                         1 if f32 else 0, dst.data_ptr(), None), "normalize")
            out[(i, j) if k == 4 else i] = dst.view(res, res)
    return out


def load_point_cloud_dat(path):
    """A point tile as apps/common/image_utils.cpp:209-231 reads it: pairs of
    numbers until the end of the file.  The reference's loop tests for the
    end before reading, so a file that ends in whitespace after its last
    number yields that last point twice (and an empty file one (0, 0)
    point); the same points result here.  Returns (m, 2) float64."""
    with open(path) as f:
        text = f.read()
    vals = [float(t) for t in text.split()]
    if len(vals) % 2:
        raise ValueError(f"{path}: odd number of coordinates")
    pts = np.array(vals, dtype=np.float64).reshape(-1, 2)
    if not text or text[-1].isspace():
        last = pts[-1:] if len(pts) else np.zeros((1, 2))
        pts = np.concatenate([pts, last])
    return pts


def blue_noise_tile(n: int, tile, device: int = 0):
    """generate_blue_noise_tile(n, pts, tile) (apps/common/image_utils.cpp:
    148-173) on the device: the tile repeated over ceil(n / N)^2 cells of size
    N / n (N = floor(sqrt(len(tile)))), points with both coordinates <= 1, in
    the reference's order.  Returns an (m, 2) float64 device tensor."""
    torch = _torch()
    dev = f"cuda:{device}"
    t = torch.as_tensor(np.ascontiguousarray(tile, dtype=np.float64).reshape(-1), device=dev)
    nb = t.numel() // 2
    N = int(np.sqrt(nb))
    nt = (n + N - 1) // N
    cap = nt * nt * nb
    out = torch.empty(2 * cap, dtype=torch.float64, device=dev)
    count = C.c_int64()
    _lib.check(_lib.lib().otm_blue_noise_tile_device(device, int(n), t.data_ptr(), nb, out.data_ptr(), cap,
                                                     C.byref(count)), "blue_noise_tile")
    return out[: 2 * count.value].view(-1, 2)


def stippling(tmap: TransportMap, density, points=None, device: int = 0, tile=None, res: int | None = None):
    """Advects points through the inverse map and drops those landing on
    zero density of `density` (n x n, face order).  The points are `points`
    ((m, 2) in [0,1]^2, host array or device tensor) or, as the reference app
    makes them, the blue-noise `tile` repeated at resolution `res`
    (apps/stippling.cpp:177-181).  Returns the kept points as an (k, 2) device
    tensor."""
    torch = _torch()
    n = tmap.fwd_mesh().n
    dev = f"cuda:{device}"
    if points is None:
        if tile is None or res is None:
            raise ValueError("stippling needs points, or a tile and a resolution")
        points = blue_noise_tile(res, tile, device)
    if isinstance(points, torch.Tensor):
        pts = points.to(device=dev, dtype=torch.float64).contiguous().reshape(-1)
    else:
        pts = torch.as_tensor(np.ascontiguousarray(points, dtype=np.float64).reshape(-1), device=dev)
    m = pts.numel() // 2
    adv = torch.empty_like(pts)
    L = _lib.lib()
    tmap.init_inverse()
    _lib.check(L.otm_map_inverse_device(tmap._handle(), pts.data_ptr(), m, 0, adv.data_ptr(), None, None), "inverse")
    dens = torch.as_tensor(np.ascontiguousarray(density, dtype=np.float64).reshape(-1), device=dev)
    kept_pts = torch.empty_like(pts)
    kept = C.c_int64()
    _lib.check(L.otm_prune_points_device(device, n, dens.data_ptr(), adv.data_ptr(), m, kept_pts.data_ptr(),
                                         C.byref(kept)), "prune")
    return kept_pts[: 2 * kept.value].view(-1, 2)


def composite_map(tm_u: TransportMap, tm_v: TransportMap) -> GridMesh:
    """Map u -> v (apps/otmap.cpp:119-124): the inverse map of v applied to the
    forward mesh of u."""
    fwd = tm_u.fwd_mesh()
    V = (fwd.n + 1) ** 2
    pts = np.stack([fwd.xy[:V], fwd.xy[V:]], 1)
    inv, _, _ = tm_v.inverse(pts)
    return GridMesh(fwd.n, np.concatenate([inv[:, 0], inv[:, 1]]))
