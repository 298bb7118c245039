// static_condensation.h -- StaticCondensation (include/numerics/static_condensation.h:64):
// the system matrix of a statically condensed system.  Assembly writes the
// full operator into it (it is a HipCsrMatrix, so the device assembly fills
// it as usual); close() forms, on the device, the reduced matrix
// S = A_uu - A_uc A_cc^-1 A_cu, which for DoFs that each belong to one element
// is the sum of the element Schur complements (static_condensation.C:185-227);
// apply() is forward elimination, a solve of the reduced system with the
// reduced system solver and back substitution (:383-473).
//
// libMesh hands apply() to PETSc as the preconditioner of the full solve
// (StaticCondensationPreconditioner); here HipLinearSolver recognises the
// matrix and runs the condensed solve itself at the tolerance it was given.
#pragma once

#include <memory>
#include <unordered_set>

#include "libmesh/hip_csr_matrix.h"
#include "libmesh/linear_solver.h"
#include "libmesh/static_condensation_dof_map.h"

namespace libMesh {

class System;

class StaticCondensation : public HipCsrMatrix<Number> {
 public:
  StaticCondensation(const MeshBase &mesh, System &system, const DofMap &full_dof_map,
                     StaticCondensationDofMap &reduced_dof_map);
  ~StaticCondensation() override;

  // SparseMatrix overrides: the full operator; init also prepares the
  // condensation (StaticCondensation::init(ParallelType) -> init())
  void init(numeric_index_type m, numeric_index_type n, numeric_index_type m_l, numeric_index_type n_l,
            numeric_index_type nnz = 30, numeric_index_type noz = 10, numeric_index_type blocksize = 1) override;
  void init(ParallelType type) override;
  void clear() override;
  // the full values, then the reduced matrix from them
  void close() override;

  // StaticCondensation::init: the reduced DoF map and the reduced solver
  void init();
  // StaticCondensation::setup (the reduced values are formed by close())
  void setup();
  // StaticCondensation::apply: full_sol from full_rhs, the reduced system
  // solved from zero to the reference's tolerance (1e-5 relative, 300 its)
  void apply(const NumericVector<Number> &full_rhs, NumericVector<Number> &full_sol);
  // The same solve at a given tolerance: (reduced solver iterations, its final
  // relative residual)
  std::pair<unsigned int, Real> condensed_solve(const NumericVector<Number> &full_rhs,
                                                NumericVector<Number> &full_sol, double rtol, unsigned int max_its);
  // the reduced matrix (rows: the reduced DoFs, in full order)
  const SparseMatrix<Number> &get_condensed_mat() const;
  // the device assembly addresses elements itself: nothing to record
  void set_current_elem(const Elem &) {}
  LinearSolver<Number> &reduced_system_solver();
  void dont_condense_vars(const std::unordered_set<unsigned int> &vars);
  // LinearSolver::get_converged_reason of the last condensed solve
  LinearConvergenceReason reduced_converged_reason() const;

 private:
  void condense() const;
  const MeshBase &_sc_mesh;
  System &_system;
  StaticCondensationDofMap &_reduced_dof_map;
  mutable lmh_condensation *_sc = nullptr;
  mutable bool _sc_stale = true;
  mutable std::unique_ptr<HipCsrMatrix<Number>> _reduced_view;
  std::unique_ptr<LinearSolver<Number>> _reduced_solver;
};

}  // namespace libMesh
