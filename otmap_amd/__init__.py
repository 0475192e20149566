"""MI355X (gfx950) implementation of otmap's grid transport solver.

Drop-in mirror of the reference's API (ggael/otmap):
  GridBasedTransportSolver, SolverOptions, BetaOpt      otlib/otsolver_2dgrid.h
  TransportMap, apply_inverse_map, transport_cost        otlib/transport_map.h
  rasterize_image                                        otlib/utils/rasterizer.h
plus BatchSolver (many densities of uneven sizes solved together), backed by
hand-written HIP kernels in libotmap_amd.so (C ABI: include/otmap_capi.h).
"""
from .batch import BatchSolver
from .solver import BetaOpt, GridBasedTransportSolver, Precision, SolverOptions, SolveStats, Transform
from .transport_map import (GridMesh, TransportMap, apply_inverse_map, grid_transport_cost, rasterize_image,
                            transport_cost)

__all__ = ["BatchSolver", "BetaOpt", "GridBasedTransportSolver", "Precision", "SolverOptions", "SolveStats", "Transform",
           "GridMesh", "TransportMap", "apply_inverse_map", "grid_transport_cost", "rasterize_image",
           "transport_cost"]
