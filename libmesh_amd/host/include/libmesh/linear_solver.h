// linear_solver.h -- LinearSolver<T> (include/solvers/linear_solver.h:58-330):
// the four pure-virtual solve overloads and the two that dispatch on an
// optional preconditioning matrix, with the reference's signatures.  solve
// returns (iterations, final relative residual); non-convergence is reported
// through get_converged_reason, not by an exception.
#pragma once

#include <optional>
#include <utility>

#include "libmesh/parallel_object.h"
#include "libmesh/shell_matrix.h"
#include "libmesh/sparse_matrix.h"

namespace libMesh {

class System;

template <typename T>
class LinearSolver : public ParallelObject {
 public:
  LinearSolver(const libMesh::Parallel::Communicator &comm_in) : ParallelObject(comm_in) {}
  virtual ~LinearSolver() = default;

  // LinearSolver::build (linear_solver.C:60-110): the GPU solver for HIP_SOLVERS
  static std::unique_ptr<LinearSolver<T>> build(const libMesh::Parallel::Communicator &comm_in,
                                                const SolverPackage solver_package = HIP_SOLVERS);

  bool initialized() const { return _is_initialized; }
  virtual void clear() {}
  virtual void init(const char *name = nullptr) = 0;
  virtual void init_systems(const System &) {}
  SolverType solver_type() const { return _solver_type; }
  void set_solver_type(const SolverType st) { _solver_type = st; }
  PreconditionerType preconditioner_type() const { return _preconditioner_type; }
  void set_preconditioner_type(const PreconditionerType pct) { _preconditioner_type = pct; }
  // LinearSolver::restrict_solve_to (linear_solver.C:136-143): subset solves
  // are not implemented by this backend, as by the base class
  virtual void restrict_solve_to(const std::vector<unsigned int> *const dofs,
                                 const SubsetSolveMode subset_solve_mode = SUBSET_ZERO);

  virtual std::pair<unsigned int, Real> solve(SparseMatrix<T> &, NumericVector<T> &, NumericVector<T> &,
                                              const std::optional<double> tol = std::nullopt,
                                              const std::optional<unsigned int> m_its = std::nullopt) = 0;
  // LinearSolver::adjoint_solve (src/solvers/linear_solver.C:147-171): solve with the transpose
  virtual std::pair<unsigned int, Real> adjoint_solve(SparseMatrix<T> &, NumericVector<T> &, NumericVector<T> &,
                                                      const std::optional<double> tol = std::nullopt,
                                                      const std::optional<unsigned int> m_its = std::nullopt);
  virtual std::pair<unsigned int, Real> solve(SparseMatrix<T> &, SparseMatrix<T> &, NumericVector<T> &,
                                              NumericVector<T> &, const std::optional<double> tol = std::nullopt,
                                              const std::optional<unsigned int> m_its = std::nullopt) = 0;
  // linear_solver.h:335-346: the preconditioning matrix if given, else the
  // system matrix
  std::pair<unsigned int, Real> solve(SparseMatrix<T> &matrix, SparseMatrix<T> *precond_matrix, NumericVector<T> &,
                                      NumericVector<T> &, const std::optional<double> tol = std::nullopt,
                                      const std::optional<unsigned int> m_its = std::nullopt);
  virtual std::pair<unsigned int, Real> solve(const ShellMatrix<T> &shell_matrix, NumericVector<T> &,
                                              NumericVector<T> &, const std::optional<double> tol = std::nullopt,
                                              const std::optional<unsigned int> m_its = std::nullopt) = 0;
  virtual std::pair<unsigned int, Real> solve(const ShellMatrix<T> &shell_matrix,
                                              const SparseMatrix<T> &precond_matrix, NumericVector<T> &,
                                              NumericVector<T> &, const std::optional<double> tol = std::nullopt,
                                              const std::optional<unsigned int> m_its = std::nullopt) = 0;
  std::pair<unsigned int, Real> solve(const ShellMatrix<T> &matrix, const SparseMatrix<T> *precond_matrix,
                                      NumericVector<T> &, NumericVector<T> &,
                                      const std::optional<double> tol = std::nullopt,
                                      const std::optional<unsigned int> m_its = std::nullopt);
  virtual LinearConvergenceReason get_converged_reason() const = 0;

 protected:
  // defaults of EquationSystems parameters (src/systems/equation_systems.C:59-60)
  double default_tolerance() const { return 1e-12; }
  unsigned int default_max_iterations() const { return 5000; }
  SolverType _solver_type = CG;
  PreconditionerType _preconditioner_type = JACOBI_PRECOND;
  bool _is_initialized = false;
};

template <typename T>
inline std::pair<unsigned int, Real> LinearSolver<T>::solve(SparseMatrix<T> &mat, SparseMatrix<T> *pc_mat,
                                                            NumericVector<T> &sol, NumericVector<T> &rhs,
                                                            const std::optional<double> tol,
                                                            const std::optional<unsigned int> n_iter) {
  if (pc_mat) return this->solve(mat, *pc_mat, sol, rhs, tol, n_iter);
  return this->solve(mat, sol, rhs, tol, n_iter);
}

template <typename T>
inline std::pair<unsigned int, Real> LinearSolver<T>::solve(const ShellMatrix<T> &mat,
                                                            const SparseMatrix<T> *pc_mat, NumericVector<T> &sol,
                                                            NumericVector<T> &rhs, const std::optional<double> tol,
                                                            const std::optional<unsigned int> n_iter) {
  if (pc_mat) return this->solve(mat, *pc_mat, sol, rhs, tol, n_iter);
  return this->solve(mat, sol, rhs, tol, n_iter);
}

}  // namespace libMesh
