// dof_map.h -- DofMap subset on the assembly path (include/base/dof_map.h):
// variables of one Lagrange type, var-major first-touch numbering, element
// DoF indices, and homogeneous Dirichlet boundaries by boundary id.
#pragma once

#include <map>
#include <set>
#include <tuple>
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

class System;
template <typename T>
class NumericVector;

// DofConstraintRow (include/base/dof_map.h): constraining DoF -> coefficient
typedef std::map<dof_id_type, Real> DofConstraintRow;

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
  const std::string &variable_name(unsigned v) const { return _names.at(v); }
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

  // DofMap::add_constraint_row (dof_map_constraints.C:2180-2208): x_dof =
  // constraint_rhs + sum of row coefficients times their DoFs.  Rows are
  // processed (process_constraints, :4341-4497) on the device when the system
  // matrix is initialised, so rows are added after distribute_dofs and before
  // the matrices are built: from a System::Constraint object, as libMesh
  // applications do.
  void add_constraint_row(dof_id_type dof_number, const DofConstraintRow &constraint_row, Number constraint_rhs,
                          bool forbid_constraint_overwrite);
  void add_constraint_row(dof_id_type dof_number, const DofConstraintRow &constraint_row,
                          bool forbid_constraint_overwrite = true) {
    add_constraint_row(dof_number, constraint_row, 0, forbid_constraint_overwrite);
  }
  // DofMap::set_error_on_constraint_loop (check_for_constraint_loops, :4509-4597)
  void set_error_on_constraint_loop(bool error_on_constraint_loop) { _error_on_loop = error_on_constraint_loop; }
  bool get_error_on_constraint_loop() const { return _error_on_loop; }
  // The asymmetric_constraint_rows argument that the device assembly passes to
  // heterogeneously_constrain_element_matrix_and_vector (:2518-2616); default
  // true as there.  Symmetric rows keep the system SPD for CG; then call
  // enforce_constraints_exactly after the solve.
  void set_asymmetric_constraint_rows(bool asymmetric) { _asymmetric = asymmetric; }
  bool asymmetric_constraint_rows() const { return _asymmetric; }
  // DofMap::enforce_constraints_exactly (:3088-3170) on the system's solution
  // (or v), on the device, with the processed rows of the system matrix.
  void enforce_constraints_exactly(const System &system, NumericVector<Number> *v = nullptr,
                                   bool homogeneous = false) const;
  // rows added with add_constraint_row, in order (DoF, row, value)
  const std::vector<std::tuple<dof_id_type, DofConstraintRow, Number>> &added_constraint_rows() const {
    return _rows;
  }

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
  std::vector<uint8_t> _dirichlet_cons;  // the Dirichlet set alone
  std::vector<std::tuple<dof_id_type, DofConstraintRow, Number>> _rows;
  bool _error_on_loop = false, _asymmetric = true;

 public:
  const std::vector<uint8_t> &dirichlet_constrained() const { return _dirichlet_cons; }
};

}  // namespace libMesh
