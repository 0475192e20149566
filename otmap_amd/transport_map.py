"""Grid meshes and transport maps (otlib/transport_map.h:29-70).

GridMesh stands in for the subset of surface_mesh::Surface_mesh that the
reference's callers use on a regular (n+1) x (n+1) quad grid: vertex
positions, the vertices of a quad, sizes and the OFF writer.  Vertex (i, j) has
index j + i*(n+1); quad (i, j) has index j + i*n and vertices
(i,j), (i+1,j), (i+1,j+1), (i,j+1) (otlib/utils/mesh_utils.cpp:90-122).
"""
from __future__ import annotations

import numpy as np

from . import _lib


class GridMesh:
    def __init__(self, n: int, xy):
        """xy: 2*(n+1)^2 SoA positions (all x, then all y)."""
        self.n = int(n)
        V = (self.n + 1) ** 2
        xy = np.ascontiguousarray(xy, dtype=np.float64).reshape(-1)
        if xy.size != 2 * V:
            raise ValueError(f"expected {2 * V} coordinates, got {xy.size}")
        self.xy = xy

    @staticmethod
    def regular(n: int) -> "GridMesh":
        """generate_quad_mesh(n+1, n+1): vertex (i, j) at (i*(1/n), j*(1/n))."""
        dx = 1.0 / float(n)
        i = np.repeat(np.arange(n + 1, dtype=np.float64), n + 1)
        j = np.tile(np.arange(n + 1, dtype=np.float64), n + 1)
        return GridMesh(n, np.concatenate([i * dx, j * dx]))

    # -- Surface_mesh subset -------------------------------------------------
    def vertices_size(self) -> int:
        return (self.n + 1) ** 2

    def faces_size(self) -> int:
        return self.n * self.n

    n_vertices = vertices_size
    n_faces = faces_size

    def points(self) -> np.ndarray:
        """(V, 2) positions (a copy; assign through set_points)."""
        V = self.vertices_size()
        return np.stack([self.xy[:V], self.xy[V:]], axis=1)

    def set_points(self, pts) -> None:
        pts = np.asarray(pts, dtype=np.float64).reshape(-1, 2)
        V = self.vertices_size()
        self.xy[:V] = pts[:, 0]
        self.xy[V:] = pts[:, 1]

    def position(self, v: int):
        V = self.vertices_size()
        return np.array([self.xy[v], self.xy[V + v]])

    def vertices(self, f: int):
        n = self.n
        i, j = divmod(int(f), n)
        return [j + i * (n + 1), j + (i + 1) * (n + 1), (j + 1) + (i + 1) * (n + 1), (j + 1) + i * (n + 1)]

    def copy(self) -> "GridMesh":
        return GridMesh(self.n, self.xy.copy())

    @classmethod
    def read(cls, path: str) -> "GridMesh":
        """Reads an OFF file written by write() (or by the reference for a
        grid map, otlib/surface_mesh/IO_off.cpp): the faces must be the
        regular grid's quads in face order."""
        with open(path) as fh:
            tokens = fh.read().split()
        if not tokens or tokens[0] != "OFF":
            raise ValueError(f"{path}: not an OFF file")
        nv, nf = int(tokens[1]), int(tokens[2])
        n = int(round(np.sqrt(nf)))
        if n * n != nf or (n + 1) ** 2 != nv:
            raise ValueError(f"{path}: {nv} vertices / {nf} faces is not an n x n quad grid")
        pos = np.asarray(tokens[4:4 + 3 * nv], dtype=np.float64).reshape(nv, 3)
        faces = np.asarray(tokens[4 + 3 * nv:4 + 3 * nv + 5 * nf], dtype=np.int64).reshape(nf, 5)
        mesh = cls(n, np.concatenate([pos[:, 0], pos[:, 1]]))
        i, j = np.divmod(np.arange(nf), n)
        expect = np.stack([j + i * (n + 1), j + (i + 1) * (n + 1), j + 1 + (i + 1) * (n + 1), j + 1 + i * (n + 1)], 1)
        if not (np.all(faces[:, 0] == 4) and np.array_equal(faces[:, 1:], expect)):
            raise ValueError(f"{path}: faces are not the regular grid's quads")
        return mesh

    def write(self, path: str) -> None:
        """OFF file with z = 0, "%.10f" coordinates (otlib/surface_mesh/IO_off.cpp)."""
        n = self.n
        V, F = self.vertices_size(), self.faces_size()
        with open(path, "w") as fh:
            fh.write(f"OFF\n{V} {F} 0\n")
            x, y = self.xy[:V], self.xy[V:]
            for k in range(V):
                fh.write("%.10f %.10f %.10f\n" % (x[k], y[k], 0.0))
            for f in range(F):
                a, b, c, d = self.vertices(f)
                fh.write(f"4 {a} {b} {c} {d}\n")


class TransportMap:
    """origin mesh, forward mesh and adjusted density of one solve
    (otlib/transport_map.h:29-63).  The inverse runs on the GPU through a BVH
    built lazily by init_inverse(), as in the reference."""

    def __init__(self, origin: GridMesh, fwd: GridMesh, density, device: int = 0):
        self._origin = origin
        self._fwd = fwd
        self._density = np.ascontiguousarray(density, dtype=np.float64).reshape(-1)
        self._device = device
        self._h = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._h = None
            try:
                _lib.lib().otm_map_destroy(h)
            except Exception:  # interpreter shutdown: module globals already cleared
                pass

    def origin_mesh(self) -> GridMesh:
        return self._origin

    def fwd_mesh(self) -> GridMesh:
        return self._fwd

    def density(self) -> np.ndarray:
        return self._density

    fwd_mesh_ptr = fwd_mesh
    density_ptr = density

    def _handle(self):
        if self._h is None:
            h = _lib.C.c_void_p()
            _lib.check(_lib.lib().otm_map_create(_lib.C.byref(h), self._device, self._fwd.n, _lib.ptr(self._fwd.xy),
                                                 _lib.ptr(self._density)), "otm_map_create")
            self._h = h
        return self._h

    def init_inverse(self) -> None:
        """otlib/transport_map.cpp:36-43"""
        _lib.check(_lib.lib().otm_map_init_inverse(self._handle()), "init_inverse")

    def inverse(self, points, fast: bool = False):
        """Batched inv / inv_fast of (m, 2) points -> (positions (m, 2), face (m,), nhits (m,))."""
        p = np.ascontiguousarray(points, dtype=np.float64).reshape(-1)
        m = p.size // 2
        out = np.empty(2 * m)
        face = np.empty(m, dtype=np.int32)
        nh = np.empty(m, dtype=np.int32)
        _lib.check(_lib.lib().otm_map_inverse(self._handle(), _lib.ptr(p), m, 1 if fast else 0, _lib.ptr(out),
                                              _lib.iptr(face), _lib.iptr(nh)), "inverse")
        return out.reshape(m, 2), face, nh

    def inverse_mesh(self, fast: bool = False) -> GridMesh:
        """The inverse map as a mesh: inverse images of the origin grid
        vertices (apply_inverse_map on origin_mesh().points(),
        apps/barycenters.cpp:187-190)."""
        n = self._fwd.n
        xy = np.empty(2 * (n + 1) ** 2)
        _lib.check(_lib.lib().otm_map_inverse_vertices(self._handle(), 1 if fast else 0, _lib.ptr(xy)),
                   "inverse_vertices")
        return GridMesh(n, xy)

    def inv(self, p):
        """otlib/transport_map.h:44"""
        return self.inverse(np.asarray(p, dtype=np.float64).reshape(1, 2))[0][0]

    def inv_fast(self, p):
        """otlib/transport_map.h:45"""
        return self.inverse(np.asarray(p, dtype=np.float64).reshape(1, 2), fast=True)[0][0]


def apply_inverse_map(tmap: TransportMap, points, verbose_level: int = 2):
    """otlib/transport_map.cpp:105-128: returns the inverse-mapped (m, 2) points."""
    tmap.init_inverse()
    return tmap.inverse(points)[0]


def rasterize_image(mesh: GridMesh, density=None, shape=None, device: int = 0):
    """otlib/utils/rasterizer.cpp:188-308 (per-face density).  Returns a
    (rows, cols) float64 image, img[y, x]; shape defaults to (n, n)."""
    rows, cols = shape if shape is not None else (mesh.n, mesh.n)
    img = np.empty(rows * cols)
    dens = None if density is None else np.ascontiguousarray(density, dtype=np.float64).reshape(-1)
    if dens is not None and dens.size != mesh.n * mesh.n:
        raise ValueError("density must hold one value per face")
    _lib.check(_lib.lib().otm_rasterize(device, mesh.n, _lib.ptr(mesh.xy), _lib.ptr(dens), int(rows), int(cols),
                                        _lib.ptr(img)), "rasterize_image")
    return img.reshape(rows, cols)


def transport_cost(src_mesh: GridMesh, dst_mesh: GridMesh, density, per_face: bool = False, device: int = 0):
    """otlib/transport_map.cpp:131-215 on quad grid meshes: the total cost, and
    with per_face the per-face costs (normalised as in the reference)."""
    n = src_mesh.n
    if dst_mesh.n != n:
        raise ValueError("meshes of different sizes")
    dens = np.ascontiguousarray(density, dtype=np.float64).reshape(-1)
    cost = np.empty(n * n) if per_face else None
    total = _lib.C.c_double()
    _lib.check(_lib.lib().otm_transport_cost(device, n, _lib.ptr(src_mesh.xy), _lib.ptr(dst_mesh.xy), _lib.ptr(dens),
                                             _lib.ptr(cost), _lib.C.byref(total)), "transport_cost")
    return (total.value, cost) if per_face else total.value


def grid_transport_cost(n: int, displacement, density, device: int = 0):
    """compute_transport_cost (otlib/otsolver_2dgrid.cpp:608-640): per-face
    cost and total for vertex displacements (SoA, 2 (n+1)^2)."""
    disp = np.ascontiguousarray(displacement, dtype=np.float64).reshape(-1)
    dens = np.ascontiguousarray(density, dtype=np.float64).reshape(-1)
    cost = np.empty(n * n)
    total = _lib.C.c_double()
    _lib.check(_lib.lib().otm_grid_transport_cost(device, n, _lib.ptr(disp), _lib.ptr(dens), _lib.ptr(cost),
                                                  _lib.C.byref(total)), "grid_transport_cost")
    return total.value, cost
