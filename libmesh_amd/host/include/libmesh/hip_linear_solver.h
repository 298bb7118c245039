// hip_linear_solver.h -- HipLinearSolver<T>: CG or BiCGSTAB on the GPU with
// JACOBI_PRECOND (default) or IDENTITY_PRECOND (EigenSparseLinearSolver,
// src/solvers/eigen_sparse_linear_solver.C:82-205), CG optionally as FP32
// inner CG with FP64 reliable-update refinement to the same FP64 tolerance.
// The matrix must be a HipCsrMatrix; solution and right-hand side may be any
// NumericVector (copied to and from the device when they are not HipVectors).
#pragma once

#include "libmesh/hip_csr_matrix.h"
#include "libmesh/linear_solver.h"

namespace libMesh {

template <typename T>
class HipLinearSolver : public LinearSolver<T> {
 public:
  explicit HipLinearSolver(const libMesh::Parallel::Communicator &comm_in) : LinearSolver<T>(comm_in) {}
  void init(const char *name = nullptr) override;
  std::pair<unsigned int, Real> solve(SparseMatrix<T> &matrix, NumericVector<T> &solution, NumericVector<T> &rhs,
                                      const std::optional<double> tol = std::nullopt,
                                      const std::optional<unsigned int> m_its = std::nullopt) override;
  // CG: the assembled operator is symmetric (C^T K C with unit diagonals on
  // constrained rows), so the adjoint solve is the same solve; BiCGSTAB
  // solves with the transpose.
  std::pair<unsigned int, Real> adjoint_solve(SparseMatrix<T> &matrix, NumericVector<T> &solution,
                                              NumericVector<T> &rhs, const std::optional<double> tol = std::nullopt,
                                              const std::optional<unsigned int> m_its = std::nullopt) override;
  // A separate preconditioning matrix is refused, as by
  // EigenSparseLinearSolver (eigen_sparse_linear_solver.h:186-196); passing
  // the system matrix itself solves normally.
  std::pair<unsigned int, Real> solve(SparseMatrix<T> &matrix, SparseMatrix<T> &pc, NumericVector<T> &solution,
                                      NumericVector<T> &rhs, const std::optional<double> tol = std::nullopt,
                                      const std::optional<unsigned int> m_its = std::nullopt) override;
  // Shell matrices are not implemented (eigen_sparse_linear_solver.C:355-390)
  std::pair<unsigned int, Real> solve(const ShellMatrix<T> &shell_matrix, NumericVector<T> &solution,
                                      NumericVector<T> &rhs, const std::optional<double> tol = std::nullopt,
                                      const std::optional<unsigned int> m_its = std::nullopt) override;
  std::pair<unsigned int, Real> solve(const ShellMatrix<T> &shell_matrix, const SparseMatrix<T> &precond_matrix,
                                      NumericVector<T> &solution, NumericVector<T> &rhs,
                                      const std::optional<double> tol = std::nullopt,
                                      const std::optional<unsigned int> m_its = std::nullopt) override;
  using LinearSolver<T>::solve;
  LinearConvergenceReason get_converged_reason() const override { return _reason; }

  // mixed precision: FP32 inner CG, FP64 true residual refreshed when the FP32
  // residual drops by `inner` relative to the last refresh
  void set_mixed_precision(bool on, double inner = 0.1) {
    _mixed = on;
    _inner = inner;
  }
  unsigned int n_refreshes() const { return _nref; }

 private:
  std::pair<unsigned int, Real> run(SparseMatrix<T> &matrix, NumericVector<T> &solution, NumericVector<T> &rhs,
                                    std::optional<double> tol, std::optional<unsigned int> m_its, bool transpose);
  LinearConvergenceReason _reason = UNKNOWN_FLAG;
  bool _mixed = false;
  double _inner = 0.1;
  unsigned int _nref = 0;
};

}  // namespace libMesh
