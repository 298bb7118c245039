// replicated_mesh.h -- flat-array stand-in for libMesh's ReplicatedMesh on the
// assembly path: one element type per mesh, nodes in id order, element
// connectivity in libMesh local node order, nodeset and sideset boundary ids.
// MeshBase accessors used by the drop-in (n_nodes, n_elem, point, node ids of
// an element) keep their libMesh meaning.
#pragma once

#include <array>
#include <vector>

#include "libmesh/libmesh_common.h"

namespace libMesh {

struct Point {
  Real x = 0, y = 0, z = 0;
  Real operator()(unsigned i) const { return i == 0 ? x : (i == 1 ? y : z); }
};

class ReplicatedMesh {
 public:
  ReplicatedMesh() = default;
  explicit ReplicatedMesh(int /*comm*/) {}

  dof_id_type n_nodes() const { return (dof_id_type)(_xyz.size() / 3); }
  dof_id_type n_elem() const { return _npe ? (dof_id_type)(_conn.size() / _npe) : 0; }
  unsigned mesh_dimension() const { return 3; }
  ElemType elem_type() const { return _type; }
  unsigned nodes_per_elem() const { return _npe; }
  unsigned sides_per_elem() const { return _nsides; }

  Point point(dof_id_type n) const { return {_xyz[3 * n], _xyz[3 * n + 1], _xyz[3 * n + 2]}; }
  void set_point(dof_id_type n, const Point &p) {
    _xyz[3 * n] = p.x;
    _xyz[3 * n + 1] = p.y;
    _xyz[3 * n + 2] = p.z;
  }
  dof_id_type node_id(dof_id_type e, unsigned a) const { return (dof_id_type)_conn[(size_t)e * _npe + a]; }
  // boundary ids: nodeset bit mask per node; per-element side ids (-1 = none)
  unsigned node_boundary_mask(dof_id_type n) const { return _bmask[n]; }
  int side_boundary_id(dof_id_type e, unsigned s) const { return _sbid[(size_t)e * _nsides + s]; }

  // raw arrays (what the C ABI uploads)
  const std::vector<double> &coords() const { return _xyz; }
  std::vector<double> &coords() { return _xyz; }
  const std::vector<int32_t> &connectivity() const { return _conn; }
  // lattice spacing per axis of a generated cube (for MeshTools perturbations)
  std::array<int, 3> lattice_cells() const { return _cells; }

  // filled by MeshTools::Generation::build_cube
  ElemType _type = INVALID_ELEM;
  unsigned _npe = 0, _nsides = 0;
  std::array<int, 3> _cells = {0, 0, 0};
  std::vector<double> _xyz;
  std::vector<int32_t> _conn;
  std::vector<uint8_t> _bmask;
  std::vector<int8_t> _sbid;
};

typedef ReplicatedMesh Mesh;
typedef ReplicatedMesh MeshBase;

namespace MeshTools {
namespace Generation {
// MeshTools::Generation::build_cube (src/mesh/mesh_generation.C:317), 3D
// HEX8/HEX27/TET4/TET10, followed by prepare_for_use's first-touch node
// renumbering (src/mesh/replicated_mesh.C:700-830).
void build_cube(ReplicatedMesh &mesh, unsigned nx, unsigned ny, unsigned nz, Real xmin = 0., Real xmax = 1.,
                Real ymin = 0., Real ymax = 1., Real zmin = 0., Real zmax = 1., ElemType type = INVALID_ELEM);
}  // namespace Generation
namespace Modification {
// Seeded uniform perturbation of every node that carries no boundary id:
// each coordinate moves by U[-frac*h_c, frac*h_c], h_c = lattice node spacing
// along axis c; draws in node id order from mt19937_64(seed).
void perturb_interior(ReplicatedMesh &mesh, uint64_t seed, Real frac);
}  // namespace Modification
// kappa_e = exp(z_e), z_e ~ N(0,1) from Box-Muller over mt19937_64(seed), in element order.
std::vector<Real> lognormal_coefficients(dof_id_type n_elem, uint64_t seed);
// Batch of cube problems: for each problem, from mt19937_64(seed), edge
// N = round(nmin (nmax/nmin)^u) (log-uniform in [nmin, nmax]) and element type
// uniform over HEX8, HEX27, TET4, TET10 (two draws per problem, in order).
void batch_recipe(uint64_t seed, int count, int nmin, int nmax, std::vector<int> &edge, std::vector<ElemType> &type);
}  // namespace MeshTools

}  // namespace libMesh
