// Compile-time check that the drop-in headers declare the reference's entry
// points with the reference's exact signatures (ggael/otmap), Eigen types
// included: each line binds the drop-in function to a pointer of the type
// the reference's declaration has.  Compiled with -fsyntax-only by the CPU
// tests where Eigen is available; nothing runs.
#include <type_traits>
#include <vector>

#include <Eigen/Core>

#include "otsolver_2dgrid.h"
#include "surface_mesh/Surface_mesh.h"
#include "transport_map.h"
#include "utils/rasterizer.h"

using otmap::GridBasedTransportSolver;
using otmap::TransportMap;
using surface_mesh::Surface_mesh;

// otlib/otsolver_2dgrid.h:52-56
void (GridBasedTransportSolver::*sig_verbose)(int) = &GridBasedTransportSolver::set_verbose_level;
void (GridBasedTransportSolver::*sig_init)(int) = &GridBasedTransportSolver::init;
TransportMap (GridBasedTransportSolver::*sig_solve)(Eigen::Ref<const Eigen::VectorXd>, otmap::SolverOptions) =
    &GridBasedTransportSolver::solve;
// otlib/transport_map.h:32-35, 40-53
static_assert(std::is_constructible<TransportMap, std::shared_ptr<Surface_mesh>, std::shared_ptr<Surface_mesh>,
                                    std::shared_ptr<Eigen::VectorXd>>::value,
              "TransportMap(origin, fwd, density)");
static_assert(std::is_copy_constructible<TransportMap>::value, "TransportMap is copyable");
void (TransportMap::*sig_init_inverse)() const = &TransportMap::init_inverse;
Eigen::Vector2d (TransportMap::*sig_inv)(const Eigen::Vector2d&) const = &TransportMap::inv;
Eigen::Vector2d (TransportMap::*sig_inv_fast)(const Eigen::Vector2d&) const = &TransportMap::inv_fast;
std::shared_ptr<Surface_mesh> (TransportMap::*sig_fwd_ptr)() = &TransportMap::fwd_mesh_ptr;
std::shared_ptr<Eigen::VectorXd> (TransportMap::*sig_density_ptr)() = &TransportMap::density_ptr;
const Surface_mesh& (TransportMap::*sig_origin)() const = &TransportMap::origin_mesh;
const Surface_mesh& (TransportMap::*sig_fwd)() const = &TransportMap::fwd_mesh;
const Eigen::VectorXd& (TransportMap::*sig_density)() const = &TransportMap::density;
// otlib/transport_map.h:66-70
void (*sig_apply)(const TransportMap&, std::vector<Eigen::Vector2d>&, int) = &otmap::apply_inverse_map;
double (*sig_cost)(const Surface_mesh&, const Surface_mesh&, const Eigen::VectorXd&, Eigen::VectorXd*) =
    &otmap::transport_cost;
// otlib/utils/rasterizer.h:21-22
void (*sig_raster)(const Surface_mesh&, const Eigen::VectorXd&, Eigen::MatrixXd&, otmap::RasterImageOption) =
    &otmap::rasterize_image;
void (*sig_raster1)(const Surface_mesh&, Eigen::MatrixXd&, otmap::RasterImageOption) = &otmap::rasterize_image;
// the Surface_mesh members the callers use (apps/barycenters.cpp, apps/stippling.cpp)
std::vector<Eigen::Vector2d>& (Surface_mesh::*sig_points)() = &Surface_mesh::points;
static_assert(std::is_same<decltype(std::declval<Surface_mesh&>().position(Surface_mesh::Vertex(0))),
                           Eigen::Vector2d&>::value,
              "position(Vertex) -> Point&");
