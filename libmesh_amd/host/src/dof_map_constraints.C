// dof_map_constraints.C -- DofMap::enforce_constraints_exactly on the device
// (dof_map_constraints.C:3088-3170 of libMesh): constrained entries of the
// solution (or another vector) are set from the processed constraint rows that
// the system matrix's device DoF map holds.
#include "libmesh/dof_map.h"
#include "libmesh/hip_csr_matrix.h"
#include "libmesh/system.h"

namespace libMesh {

void DofMap::enforce_constraints_exactly(const System &system, NumericVector<Number> *v, bool homogeneous) const {
  if (!n_constrained_dofs()) return;
  System &sys = const_cast<System &>(system);
  if (&sys.get_dof_map() != this) libmesh_error_msg("enforce_constraints_exactly: DofMap of another system");
  auto *A = dynamic_cast<HipCsrMatrix<Number> *>(&sys.get_matrix("System Matrix"));
  if (!A || !A->device_dofmap()) libmesh_error_msg("enforce_constraints_exactly: needs the system's HipCsrMatrix");
  auto *x = dynamic_cast<HipVector<Number> *>(v ? v : sys.solution.get());
  if (!x) libmesh_error_msg("enforce_constraints_exactly: needs a HipVector");
  hip_check(lmh_enforce_constraints(A->device_dofmap(), (double *)x->device_data(), homogeneous ? 1 : 0),
            "DofMap::enforce_constraints_exactly");
  x->mark_device_modified();
}

// homogeneous rows on an adjoint solution; with empty rows only (homogeneous
// Dirichlet) the adjoint solve already leaves those entries at zero
void DofMap::enforce_adjoint_constraints_exactly(NumericVector<Number> &v, unsigned int) const {
  if (!has_general_constraints()) return;
  if (!_system) libmesh_error_msg("enforce_adjoint_constraints_exactly: DofMap of no system");
  enforce_constraints_exactly(*_system, &v, true);
}

}  // namespace libMesh
