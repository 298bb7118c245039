// dof_map.h -- DofMap subset on the assembly path (include/base/dof_map.h):
// variables of one Lagrange type, var-major first-touch numbering, element
// DoF indices, and homogeneous Dirichlet boundaries by boundary id.
#pragma once

#include <set>
#include <vector>

#include "libmesh/libmesh_common.h"
#include "libmesh/replicated_mesh.h"

namespace libMesh {

struct FEType {
  Order order = FIRST;
  FEFamily family = LAGRANGE;
  FEType(Order o = FIRST, FEFamily f = LAGRANGE) : order(o), family(f) {}
  // include/fe/fe_type.h:415-418
  Order default_quadrature_order() const { return static_cast<Order>(2 * static_cast<int>(order) + 1); }
};

// DirichletBoundary with a zero function (include/base/dirichlet_boundaries.h)
struct DirichletBoundary {
  std::set<boundary_id_type> b;
  std::vector<unsigned> variables;
  DirichletBoundary(const std::set<boundary_id_type> &ids, const std::vector<unsigned> &vars)
      : b(ids), variables(vars) {}
};

class DofMap {
 public:
  explicit DofMap(unsigned sys_number = 0) : _sys(sys_number) {}

  unsigned add_variable(const std::string &name, const FEType &type);
  unsigned n_variables() const { return (unsigned)_names.size(); }
  const FEType &variable_type(unsigned v) const { return _types.at(v); }
  void add_dirichlet_boundary(const DirichletBoundary &d) { _dirichlet.push_back(d); }

  // DofMap::distribute_dofs (dof_map.C:949) -> distribute_local_dofs_var_major
  // (dof_map.C:1417-1530), then create_dof_constraints for the Dirichlet
  // boundaries (dof_map_constraints.C:1736, 592-740).
  void distribute_dofs(const MeshBase &mesh);
  dof_id_type n_dofs() const { return _n_dofs; }
  unsigned dofs_per_elem() const { return _dpe; }
  // DofMap::dof_indices (dof_map.C:2201-2316): variable-major local order
  void dof_indices(dof_id_type e, std::vector<dof_id_type> &di) const;
  bool is_constrained_dof(dof_id_type d) const { return _cons[d] != 0; }
  dof_id_type n_constrained_dofs() const;

  const std::vector<int32_t> &elem_dofs() const { return _edofs; }
  const std::vector<uint8_t> &constrained() const { return _cons; }
  const MeshBase *mesh() const { return _mesh; }

 private:
  unsigned _sys;
  std::vector<std::string> _names;
  std::vector<FEType> _types;
  std::vector<DirichletBoundary> _dirichlet;
  const MeshBase *_mesh = nullptr;
  dof_id_type _n_dofs = 0;
  unsigned _dpe = 0;
  std::vector<int32_t> _edofs;
  std::vector<uint8_t> _cons;
};

}  // namespace libMesh
