// Drop-in for the reference's otlib/surface_mesh/Surface_mesh.h as its
// callers of the grid transport path use it: the regular quad grid mesh
// otmap::GridMesh under the reference's name (positions, points(), Vertex /
// Face handles, face circulator, sizes, OFF writer).
#pragma once

#include "../transport_map.h"

namespace surface_mesh {
typedef otmap::GridMesh Surface_mesh;
typedef otmap::Point Point;
}  // namespace surface_mesh
