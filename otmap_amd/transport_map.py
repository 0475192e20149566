"""Grid meshes and transport maps (otlib/transport_map.h:29-70).

GridMesh stands in for the subset of surface_mesh::Surface_mesh that the
reference's callers use on a regular (n+1) x (n+1) quad grid: vertex
positions, the vertices of a quad, sizes and the OFF writer.  Vertex (i, j) has
index j + i*(n+1); quad (i, j) has index j + i*n and vertices
(i,j), (i+1,j), (i+1,j+1), (i,j+1) (otlib/utils/mesh_utils.cpp:90-122).
"""
from __future__ import annotations

import numpy as np


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
    """origin mesh, forward mesh and adjusted density of one solve."""

    def __init__(self, origin: GridMesh, fwd: GridMesh, density):
        self._origin = origin
        self._fwd = fwd
        self._density = np.ascontiguousarray(density, dtype=np.float64).reshape(-1)

    def origin_mesh(self) -> GridMesh:
        return self._origin

    def fwd_mesh(self) -> GridMesh:
        return self._fwd

    def density(self) -> np.ndarray:
        return self._density

    fwd_mesh_ptr = fwd_mesh
    density_ptr = density
