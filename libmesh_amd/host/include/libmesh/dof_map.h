// dof_map.h -- DofMap subset on the assembly path (include/base/dof_map.h):
// variables of one Lagrange type, var-major first-touch numbering, element
// DoF indices, and homogeneous Dirichlet boundaries by boundary id.
#pragma once

#include <map>
#include <memory>
#include <set>
#include <tuple>
#include <vector>

#include "libmesh/libmesh_common.h"
#include "libmesh/replicated_mesh.h"
#include "libmesh/static_condensation_dof_map.h"

namespace libMesh {

struct FEType {
  Order order = FIRST;
  FEFamily family = LAGRANGE;
  FEType(Order o = FIRST, FEFamily f = LAGRANGE) : order(o), family(f) {}
  // include/fe/fe_type.h:415-418
  Order default_quadrature_order() const { return static_cast<Order>(2 * static_cast<int>(order) + 1); }
};

class System;
template <typename T>
class NumericVector;

// DofConstraintRow, DofConstraints, DofConstraintValueMap
// (include/base/dof_map.h:97-125): constraining DoF -> coefficient; the rows
// by constrained DoF; the nonzero right-hand sides by constrained DoF
typedef std::map<dof_id_type, Real> DofConstraintRow;
class DofConstraints : public std::map<dof_id_type, DofConstraintRow> {};
class DofConstraintValueMap : public std::map<dof_id_type, Number> {};

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
  unsigned int n_variables() const { return (unsigned)_names.size(); }
  const std::string &variable_name(const unsigned int i) const { return _names.at(i); }
  const FEType &variable_type(const unsigned int i) const { return _types.at(i); }
  void add_dirichlet_boundary(const DirichletBoundary &dirichlet_boundary) { _dirichlet.push_back(dirichlet_boundary); }

  // DofMap::distribute_dofs (dof_map.C:949) -> distribute_local_dofs_var_major
  // (dof_map.C:1417-1530), then create_dof_constraints for the Dirichlet
  // boundaries (dof_map_constraints.C:1736, 592-740): an empty row with a
  // zero right-hand side per boundary DoF.
  void distribute_dofs(MeshBase &mesh);
  dof_id_type n_dofs() const { return _n_dofs; }
  // DofMap::dof_indices (dof_map.C:2201-2316): all variables, variable-major
  // local order; or one variable's DoFs
  void dof_indices(const Elem *const elem, std::vector<dof_id_type> &di) const;
  void dof_indices(const Elem *const elem, std::vector<dof_id_type> &di, const unsigned int vn) const;
  bool is_constrained_dof(const dof_id_type dof) const { return _dof_constraints.count(dof) != 0; }
  dof_id_type n_constrained_dofs() const { return (dof_id_type)_dof_constraints.size(); }

  // DofMap::add_constraint_row (dof_map_constraints.C:2180-2208, declared
  // dof_map.h:1130, 1156): x_dof = constraint_rhs + sum of row coefficients
  // times their DoFs.  Rows are added after distribute_dofs, from a
  // System::Constraint object as libMesh applications do, and processed by
  // process_constraints.
  void add_constraint_row(const dof_id_type dof_number, const DofConstraintRow &constraint_row,
                          const Number constraint_rhs, const bool forbid_constraint_overwrite);
  void add_constraint_row(const dof_id_type dof_number, const DofConstraintRow &constraint_row,
                          const bool forbid_constraint_overwrite = true) {
    add_constraint_row(dof_number, constraint_row, 0., forbid_constraint_overwrite);
  }
  // dof_map.h:1164-1175, 1263
  DofConstraints::const_iterator constraint_rows_begin() const { return _dof_constraints.begin(); }
  DofConstraints::const_iterator constraint_rows_end() const { return _dof_constraints.end(); }
  const DofConstraints &get_dof_constraints() const { return _dof_constraints; }
  DofConstraintValueMap &get_primal_constraint_values() { return _primal_constraint_values; }
  // DofMap::set_error_on_constraint_loop (check_for_constraint_loops, :4509-4597)
  void set_error_on_constraint_loop(bool error_on_constraint_loop) { _error_on_loop = error_on_constraint_loop; }
  // DofMap::process_constraints (dof_map_constraints.C:4341-4497): rows are
  // substituted until they refer to unconstrained DoFs only (on the device);
  // with set_error_on_constraint_loop(true) a constraint loop is an error.
  void process_constraints(MeshBase &mesh);
  // DofMap::enforce_constraints_exactly (:3088-3170) on the system's solution
  // (or v), on the device, with the processed rows of the system matrix.
  void enforce_constraints_exactly(const System &system, NumericVector<Number> *v = nullptr,
                                   bool homogeneous = false) const;
  // DofMap::enforce_adjoint_constraints_exactly (dof_map.h:1557): the
  // homogeneous form of the same rows on an adjoint solution
  void enforce_adjoint_constraints_exactly(NumericVector<Number> &v, unsigned int q) const;

  // DofMap::create_static_condensation (dof_map.h:1792-1809): the map of the
  // DoFs a StaticCondensation eliminates, filled when the DoFs are distributed
  void create_static_condensation(MeshBase &mesh, System &system) {
    _sc = std::make_unique<StaticCondensationDofMap>(mesh, system, *this);
  }
  bool has_static_condensation() const { return _sc != nullptr; }
  StaticCondensationDofMap &get_static_condensation() {
    if (!_sc) libmesh_error_msg("DofMap: no static condensation");
    return *_sc;
  }
  const StaticCondensationDofMap &get_static_condensation() const {
    if (!_sc) libmesh_error_msg("DofMap: no static condensation");
    return *_sc;
  }

  // This library's mesh generation, files and C host API read the numbering
  // as flat arrays; the drop-in matrix reads it through the accessors above.
  unsigned dofs_per_elem() const { return _dpe; }
  const std::vector<int32_t> &elem_dofs() const { return _edofs; }
  std::vector<uint8_t> constrained() const;
  const MeshBase *mesh() const { return _mesh; }
  bool has_general_constraints() const;

 private:
  unsigned _sys;
  std::vector<std::string> _names;
  std::vector<FEType> _types;
  std::vector<DirichletBoundary> _dirichlet;
  const MeshBase *_mesh = nullptr;
  const System *_system = nullptr;
  dof_id_type _n_dofs = 0;
  unsigned _dpe = 0;
  std::vector<int32_t> _edofs;
  DofConstraints _dof_constraints;
  DofConstraintValueMap _primal_constraint_values;
  bool _error_on_loop = false;
  std::unique_ptr<StaticCondensationDofMap> _sc;

  friend class System;
};

}  // namespace libMesh
