// replicated_mesh.h -- flat-array stand-in for libMesh's ReplicatedMesh on the
// assembly path: one element type per mesh, nodes in id order, element
// connectivity in libMesh local node order, nodeset and sideset boundary ids.
// The MeshBase / Elem accessors the drop-in classes read (n_nodes, n_elem,
// point, elem_ref, Elem::type / n_nodes / node_id) keep their libMesh
// signatures and meaning.
#pragma once

#include <array>
#include <vector>

#include "libmesh/libmesh_common.h"
#include "libmesh/parallel_object.h"

namespace libMesh {

struct Point {
  Real x = 0, y = 0, z = 0;
  Real operator()(unsigned i) const { return i == 0 ? x : (i == 1 ? y : z); }
};

class ReplicatedMesh;

// Elem (include/geom/elem.h): the accessors the assembly path reads, as a
// view of one element of a flat-array mesh
class Elem {
 public:
  dof_id_type id() const { return _id; }
  inline ElemType type() const;
  inline unsigned int n_nodes() const;
  inline unsigned int n_sides() const;
  inline dof_id_type node_id(const unsigned int i) const;
  inline const Point &point(const unsigned int i) const;

 private:
  friend class ReplicatedMesh;
  const ReplicatedMesh *_mesh = nullptr;
  dof_id_type _id = 0;
};

class ReplicatedMesh : public ParallelObject {
 public:
  explicit ReplicatedMesh(const Parallel::Communicator &comm_in, unsigned char dim = 1)
      : ParallelObject(comm_in), _dim(dim) {}
  ReplicatedMesh(const ReplicatedMesh &other) = default;

  // MeshBase accessors (include/mesh/mesh_base.h:570 n_nodes, :641 n_elem,
  // :738 point, :788 elem_ref)
  dof_id_type n_nodes() const { return (dof_id_type)(_xyz.size() / 3); }
  dof_id_type n_elem() const { return _npe ? (dof_id_type)(_conn.size() / _npe) : 0; }
  unsigned int mesh_dimension() const { return 3; }
  const Point &point(const dof_id_type i) const { return reinterpret_cast<const Point *>(_xyz.data())[i]; }
  const Elem &elem_ref(const dof_id_type i) const {
    std::vector<Elem> &v = _elems.v;
    if (v.size() != n_elem()) {
      v.resize(n_elem());
      for (dof_id_type e = 0; e < n_elem(); ++e) {
        v[e]._mesh = this;
        v[e]._id = e;
      }
    }
    return v.at(i);
  }
  const Elem *elem_ptr(const dof_id_type i) const { return &elem_ref(i); }

  // one element type per mesh on this path
  ElemType elem_type() const { return _type; }
  unsigned nodes_per_elem() const { return _npe; }
  unsigned sides_per_elem() const { return _nsides; }
  void set_point(dof_id_type n, const Point &p) {
    _xyz[3 * n] = p.x;
    _xyz[3 * n + 1] = p.y;
    _xyz[3 * n + 2] = p.z;
  }
  dof_id_type node_id(dof_id_type e, unsigned a) const { return (dof_id_type)_conn[(size_t)e * _npe + a]; }
  // boundary ids: nodeset bit mask per node; per-element side ids (-1 = none)
  unsigned node_boundary_mask(dof_id_type n) const { return _bmask[n]; }
  int side_boundary_id(dof_id_type e, unsigned s) const { return _sbid[(size_t)e * _nsides + s]; }

  // flat arrays of the generated mesh (mesh generation, perturbation, files)
  const std::vector<double> &coords() const { return _xyz; }
  std::vector<double> &coords() { return _xyz; }
  const std::vector<int32_t> &connectivity() const { return _conn; }
  // lattice spacing per axis of a generated cube (for MeshTools perturbations)
  std::array<int, 3> lattice_cells() const { return _cells; }
  // empties the mesh (MeshBase::clear)
  void clear() {
    _type = INVALID_ELEM;
    _npe = _nsides = 0;
    _cells = {0, 0, 0};
    _xyz.clear();
    _conn.clear();
    _bmask.clear();
    _sbid.clear();
    _elems.v.clear();
  }

  // filled by MeshTools::Generation::build_cube
  unsigned char _dim = 1;
  ElemType _type = INVALID_ELEM;
  unsigned _npe = 0, _nsides = 0;
  std::array<int, 3> _cells = {0, 0, 0};
  std::vector<double> _xyz;
  std::vector<int32_t> _conn;
  std::vector<uint8_t> _bmask;
  std::vector<int8_t> _sbid;

 private:
  // element views, rebuilt on demand (a copied mesh builds its own)
  struct ElemCache {
    mutable std::vector<Elem> v;
    ElemCache() = default;
    ElemCache(const ElemCache &) {}
    ElemCache &operator=(const ElemCache &) {
      v.clear();
      return *this;
    }
  };
  ElemCache _elems;
};

static_assert(sizeof(Point) == 3 * sizeof(Real), "Point must alias three coordinates");

ElemType Elem::type() const { return _mesh->elem_type(); }
unsigned int Elem::n_nodes() const { return _mesh->nodes_per_elem(); }
unsigned int Elem::n_sides() const { return _mesh->sides_per_elem(); }
dof_id_type Elem::node_id(const unsigned int i) const { return _mesh->node_id(_id, i); }
const Point &Elem::point(const unsigned int i) const { return _mesh->point(_mesh->node_id(_id, i)); }

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
