// linear_solver.h -- LinearSolver<T> interface (include/solvers/linear_solver.h):
// solve returns (iterations, final relative residual); non-convergence is
// reported through get_converged_reason, not by an exception.
#pragma once

#include <optional>
#include <utility>

#include "libmesh/sparse_matrix.h"

namespace libMesh {

class System;

template <typename T>
class LinearSolver {
 public:
  virtual ~LinearSolver() = default;
  virtual void init(const char *name = nullptr) = 0;
  virtual void clear() {}
  virtual void init_systems(const System &) {}
  virtual std::pair<unsigned int, Real> solve(SparseMatrix<T> &matrix, NumericVector<T> &solution,
                                              NumericVector<T> &rhs, std::optional<double> tol = std::nullopt,
                                              std::optional<unsigned int> m_its = std::nullopt) = 0;
  // LinearSolver::adjoint_solve (src/solvers/linear_solver.C:147-171): solve with the transpose
  virtual std::pair<unsigned int, Real> adjoint_solve(SparseMatrix<T> &matrix, NumericVector<T> &solution,
                                                      NumericVector<T> &rhs, std::optional<double> tol = std::nullopt,
                                                      std::optional<unsigned int> m_its = std::nullopt) = 0;
  virtual LinearConvergenceReason get_converged_reason() const = 0;

  SolverType solver_type() const { return _solver_type; }
  void set_solver_type(SolverType t) { _solver_type = t; }
  PreconditionerType preconditioner_type() const { return _preconditioner_type; }
  void set_preconditioner_type(PreconditionerType p) { _preconditioner_type = p; }
  bool initialized() const { return _is_initialized; }

 protected:
  // defaults of EquationSystems parameters (src/systems/equation_systems.C:59-60)
  double default_tolerance() const { return 1e-12; }
  unsigned int default_max_iterations() const { return 5000; }
  SolverType _solver_type = CG;
  PreconditionerType _preconditioner_type = JACOBI_PRECOND;
  bool _is_initialized = false;
};

}  // namespace libMesh
