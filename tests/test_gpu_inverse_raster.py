"""GPU parity of TransportMap inverse (BVH) and rasterize_image.

Bars: bit-exact against the reference's golden vectors and the oracle --
positions, selected faces and hit counts of inv()/inv_fast(), every pixel of
the rasterised images.  Predicates run in FP64 in the reference's operation
order without FMA contraction, so equality is exact, including the
reference's multi-hit selection and its carried-over weights in the raster.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from otmap_amd import GridMesh, TransportMap, rasterize_image
from otmap_amd import densities as D

pytestmark = pytest.mark.gpu


def load(name):
    return np.load(os.path.join(GOLDEN, name))


@pytest.mark.parametrize("n", [16, 32])
def test_inverse_vs_reference_golden(n):
    g = load("inverse_small.npz")
    t = f"n{n}"
    tm = TransportMap(GridMesh.regular(n), GridMesh(n, g[t + "_fwd"]), g[t + "_density"])
    inv, face, nh = tm.inverse(g[t + "_pts"])
    ok = g[t + "_nhits"] > 0
    assert np.array_equal(nh, g[t + "_nhits"])
    assert np.array_equal(inv[ok], g[t + "_inv"][ok])
    assert np.all(np.isnan(inv[~ok]))
    invf, ff, _ = tm.inverse(g[t + "_pts"], fast=True)
    assert np.array_equal(invf[ok], g[t + "_invfast"][ok])
    assert np.array_equal(ff[ok], g[t + "_first"][ok])


def test_inverse_identity():
    g = load("inverse_small.npz")
    n = 16
    tm = TransportMap(GridMesh.regular(n), GridMesh.regular(n), np.ones(n * n))
    inv, face, _ = tm.inverse(g["ident_pts"])
    assert np.array_equal(inv, g["ident_inv"])


# the 60000-point case adds points outside the unit square (the query clamps them)
@pytest.mark.parametrize("n,seed,npts", [(64, 9, 20000), (256, 1000, 20000), (256, 1001, 60000)])
def test_inverse_vs_oracle(n, seed, npts, oracle):
    o = oracle.solve(n, D.gaussian_mixture(n, seed))
    rng = np.random.default_rng(n + npts)
    pts = np.concatenate([rng.uniform(0, 1, (npts, 2)), rng.uniform(-0.1, 1.1, (npts // 20, 2)),
                          np.stack([o["fwd"][: (n + 1) ** 2], o["fwd"][(n + 1) ** 2:]], 1)[::7]])
    tm = TransportMap(GridMesh.regular(n), GridMesh(n, o["fwd"]), o["density"])
    for fast in (False, True):
        inv, face, nh = tm.inverse(pts, fast=fast)
        oinv, oface, onh = oracle.inverse(n, o["fwd"], o["density"], pts, fast=fast)
        assert np.array_equal(face, oface)
        assert np.array_equal(nh, onh)
        assert np.array_equal(inv, oinv, equal_nan=True)


def _warped_grid(n):
    """Regular grid under a smooth displacement that fixes the square's
    boundary (quads stay convex)."""
    V = (n + 1) ** 2
    xy = GridMesh.regular(n).xy.copy()
    x, y = xy[:V].copy(), xy[V:].copy()
    xy[:V] = x + 0.01 * np.sin(2 * np.pi * x) * np.sin(2 * np.pi * y)
    xy[V:] = y + 0.01 * np.sin(4 * np.pi * x) * np.sin(2 * np.pi * y)
    return xy


def _bvh_export(tm):
    import ctypes as C

    from otmap_amd import _lib
    tm.init_inverse()
    h = tm._handle()
    k = _lib.lib().otm_map_bvh_size(h)
    n = tm.fwd_mesh().n
    boxes = np.zeros(4 * k)
    first, count, esc = (np.zeros(k, dtype=np.int32) for _ in range(3))
    faces = np.zeros(n * n, dtype=np.int32)
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))  # noqa: E731
    _lib.check(_lib.lib().otm_map_bvh_export(h, _lib.ptr(boxes), ip(first), ip(count), ip(esc), ip(faces)), "export")
    return boxes.reshape(k, 4), first, count, esc, faces


def _expected_escape(first, count):
    """Depth-first escape links of a node array in the reference's order."""
    esc = np.full(len(first), -1, dtype=np.int64)
    stack = [0]
    while stack:
        p = stack.pop()
        if count[p] >= 0:
            continue
        c1 = first[p]
        esc[c1], esc[c1 + 1] = c1 + 1, esc[p]
        stack += [c1 + 1, c1]
    return esc


@pytest.mark.parametrize("case", ["ident16", "ident33", "solved64", "warped100", "warped512", "solved1024"])
def test_device_bvh_build_node_array_vs_oracle(case, oracle):
    """BVH2D::build (otlib/utils/bvh2d.cpp:29-243) on the device: the node
    array (inflated boxes bit-exact, children, leaf face ranges, escape links)
    and the face order of the leaves equal the oracle's sequential
    restatement of the reference build (itself pinned by the reference's
    inverse goldens) -- on regular grids, where centroids fall exactly on
    split values, on warped grids and on solved maps."""
    if case.startswith("ident"):
        n = int(case[5:])
        xy = GridMesh.regular(n).xy
    elif case.startswith("warped"):
        n = int(case[6:])
        xy = _warped_grid(n)
    else:
        n = int(case[6:])
        rho = D.smoothed_shapes(n) if n == 1024 else D.gaussian_mixture(n, 9)
        xy = oracle.solve(n, rho)["fwd"]
    print(case, flush=True)
    tm = TransportMap(GridMesh.regular(n), GridMesh(n, xy), np.ones(n * n))
    boxes, first, count, esc, faces = _bvh_export(tm)
    with oracle.bvh(n, xy) as b:
        oboxes, ometa, ofaces = b.export()
    assert len(boxes) == len(oboxes)
    assert np.array_equal(boxes, oboxes)
    leaf = ometa[:, 2] != 0
    assert np.array_equal(count < 0, ~leaf)
    assert np.array_equal(first, ometa[:, 0])
    assert np.array_equal(count[leaf], ometa[leaf, 1])
    assert np.array_equal(faces, ofaces)
    assert np.array_equal(esc, _expected_escape(first, count))


def test_inverse_parallel_bvh_build_vs_oracle(oracle):
    """At n = 512 the queries against the device-built BVH give the hit
    order and multi-hit selections of the oracle's sequential build."""
    n = 512
    xy = _warped_grid(n)
    dens = np.random.default_rng(3).uniform(0.5, 2.0, n * n)
    V = (n + 1) ** 2
    pts = np.concatenate([np.random.default_rng(4).uniform(0, 1, (30000, 2)),
                          np.stack([xy[:V], xy[V:]], 1)[::37]])
    tm = TransportMap(GridMesh.regular(n), GridMesh(n, xy), dens)
    for fast in (False, True):
        inv, face, nh = tm.inverse(pts, fast=fast)
        oinv, oface, onh = oracle.inverse(n, xy, dens, pts, fast=fast)
        assert np.array_equal(face, oface)
        assert np.array_equal(nh, onh)
        assert np.array_equal(inv, oinv, equal_nan=True)


def test_inverse_cfg0_vs_reference_golden():
    g = load("solve_gaussian256.npz")
    # the map itself comes from the GPU solve (FP64 path); cell ids must match
    # the reference's away from cell edges, positions to the map tolerance
    from otmap_amd import GridBasedTransportSolver
    n = 256
    s = GridBasedTransportSolver(0)
    s.init(n)
    tm = s.solve(D.gaussian_mixture(n))
    pts = D.uniform_points(1 << 20, seed=2000 + n)[g["qidx"]]
    inv, face, nh = tm.inverse(pts)
    ref_cell = np.floor(g["inv"] * n).astype(np.int64)
    cell = np.floor(inv * n).astype(np.int64)
    frac = g["inv"] * n - np.floor(g["inv"] * n)
    away = np.all((frac > 1e-6) & (frac < 1 - 1e-6), axis=1)
    assert away.mean() > 0.99
    assert np.array_equal(cell[away], ref_cell[away])
    assert np.abs(inv - g["inv"]).max() < 1e-6


@pytest.mark.parametrize("n", [16, 32])
def test_raster_vs_reference_golden(n):
    g = load("raster.npz")
    mesh = GridMesh(n, g[f"n{n}_fwd"])
    for rows, cols in ((n, n), (2 * n, 2 * n), (24, 40)):
        img = rasterize_image(mesh, g[f"n{n}_density"], shape=(rows, cols))
        assert np.array_equal(img, g[f"n{n}_img_{rows}x{cols}"])
        img1 = rasterize_image(mesh, None, shape=(rows, cols))
        assert np.array_equal(img1, g[f"n{n}_ones_{rows}x{cols}"])
    assert np.array_equal(rasterize_image(GridMesh(n, g[f"n{n}_invmesh"])), g[f"n{n}_invimg"])


@pytest.mark.parametrize("n,R", [(64, 64), (64, 200), (256, 256), (256, 97)])
def test_raster_vs_oracle(n, R, oracle):
    o = oracle.solve(n, D.gaussian_mixture(n, n + 1))
    img = rasterize_image(GridMesh(n, o["fwd"]), o["density"], shape=(R, R + 3))
    assert np.array_equal(img, oracle.rasterize(n, o["fwd"], R, R + 3, o["density"]))


def test_raster_identity_closed_form():
    n = 64
    img = rasterize_image(GridMesh.regular(n))
    expect = float(np.float32(1.0 / (n * n))) / (1.0 / (n * n))
    assert np.allclose(img, expect, rtol=1e-12)


def test_raster_carried_weights_on_barycentre_meshes(oracle):
    """The rasteriser's carried weights (pixels whose bilinear solve fails,
    otlib/utils/rasterizer.cpp:169-171) on the meshes the barycentre workload
    rasterises: the bilerped inverse meshes of the four 512^2 shape densities
    (apps/barycenters.cpp:187-245) at five of its 25 weights, 512 x 512
    pixels.  Every pixel equals the oracle's raster of the same mesh, and the
    carried weights never come from more than 64 faces back (the device's
    look-back bound); the counts are printed."""
    import ctypes as C

    import torch

    from otmap_amd import BatchSolver, SolverOptions, _lib, pipelines

    n = 512
    maps = BatchSolver([n] * 4).solve(D.barycentre_inputs(n, count=4), SolverOptions())
    invs = pipelines.inverse_meshes_device(maps)
    ptrs = (C.c_void_p * 4)(*[t.data_ptr() for t in invs])
    count = 2 * (n + 1) ** 2
    mesh = torch.empty(count, dtype=torch.float64, device="cuda:0")
    L = _lib.lib()
    for a, b in [(0.0, 0.0), (1.0, 0.0), (0.0, 1.0), (1.0, 1.0), (0.5, 0.5)]:
        _lib.check(L.otm_mesh_lerp_device(0, count, 4, ptrs, a, b, mesh.data_ptr()), "mesh_lerp")
        xy = mesh.cpu().numpy()
        img = rasterize_image(GridMesh(n, xy), None, shape=(n, n))
        assert np.array_equal(img, oracle.rasterize(n, xy, n, n)), (a, b)
        st = oracle.raster_stats(n, xy, n, n)
        print(f"carried weights at ({a}, {b}): {st}")
        assert st["max_faces_back"] < 64, st
