// hip_linear_solver.C -- LinearSolver backend: CG (FP64 or mixed) and
// BiCGSTAB (FP64) on the GPU, with JACOBI_PRECOND or IDENTITY_PRECOND (the two
// preconditioners EigenSparseLinearSolver::solve runs without a factorisation,
// eigen_sparse_linear_solver.C:136-170, 177-205).
#include "libmesh/hip_linear_solver.h"

namespace libMesh {

template <typename T>
void HipLinearSolver<T>::init(const char *) {
  if (this->_solver_type != CG && this->_solver_type != BICGSTAB) libmesh_not_implemented();
  if (this->_preconditioner_type != JACOBI_PRECOND && this->_preconditioner_type != IDENTITY_PRECOND)
    libmesh_not_implemented();
  this->_is_initialized = true;
}

namespace {
LinearConvergenceReason to_reason(int r) {
  switch (r) {
    case LMH_CONVERGED_RTOL: return CONVERGED_RTOL;
    case LMH_DIVERGED_ITS: return DIVERGED_ITS;
    case LMH_DIVERGED_NANORINF: return DIVERGED_NANORINF;
    case LMH_DIVERGED_BREAKDOWN: return DIVERGED_BREAKDOWN;
  }
  return UNKNOWN_FLAG;
}
}  // namespace

template <typename T>
std::pair<unsigned int, Real> HipLinearSolver<T>::solve(SparseMatrix<T> &matrix_in, NumericVector<T> &solution_in,
                                                        NumericVector<T> &rhs_in, std::optional<double> tol,
                                                        std::optional<unsigned int> m_its) {
  this->init();
  auto *A = dynamic_cast<HipCsrMatrix<T> *>(&matrix_in);
  auto *x = dynamic_cast<HipVector<T> *>(&solution_in);
  auto *b = dynamic_cast<HipVector<T> *>(&rhs_in);
  if (!A || !x || !b) libmesh_error_msg("HipLinearSolver needs a HipCsrMatrix and HipVectors");
  // close everything, as EigenSparseLinearSolver::solve does (eigen_sparse_linear_solver.C:97-100)
  A->close();
  x->close();
  b->close();
  hip_check(lmh_matrix_set_preconditioner(A->device_matrix(), this->_preconditioner_type == IDENTITY_PRECOND
                                                                  ? LMH_PRECOND_IDENTITY
                                                                  : LMH_PRECOND_JACOBI),
            "HipLinearSolver (preconditioner)");
  const double rtol = tol.value_or(this->default_tolerance());
  const unsigned maxits = m_its.value_or(this->default_max_iterations());
  int64_t its = 0;
  double res = 0;
  int reason = 0;
  if (this->_solver_type == BICGSTAB) {
    // EigenSparseLinearSolver::solve, case BICGSTAB (eigen_sparse_linear_solver.C:175-205)
    if (_mixed) libmesh_not_implemented();
    hip_check(lmh_bicgstab_solve(A->device_matrix(), (const double *)b->device_data(), (double *)x->device_data(),
                                 rtol, maxits, &its, &res, &reason),
              "HipLinearSolver::solve (BiCGSTAB)");
  } else if (_mixed) {
    hip_check(lmh_matrix_make_fp32(A->device_matrix()), "HipLinearSolver (FP32 operator)");
    double stats[3];
    hip_check(lmh_mixed_solve(A->device_matrix(), (const double *)b->device_data(), (double *)x->device_data(), rtol,
                              _inner, 0, maxits, &its, &res, &reason, stats),
              "HipLinearSolver::solve (mixed)");
    _nref = (unsigned)stats[1];
  } else {
    hip_check(lmh_cg_solve(A->device_matrix(), (const double *)b->device_data(), (double *)x->device_data(), rtol,
                           maxits, &its, &res, &reason),
              "HipLinearSolver::solve");
  }
  x->mark_device_modified();
  _reason = to_reason(reason);
  return std::make_pair((unsigned int)its, (Real)res);
}

template <typename T>
std::pair<unsigned int, Real> HipLinearSolver<T>::adjoint_solve(SparseMatrix<T> &matrix, NumericVector<T> &solution,
                                                                NumericVector<T> &rhs, std::optional<double> tol,
                                                                std::optional<unsigned int> m_its) {
  // CG: the system matrices are symmetric, so A^T = A.  BiCGSTAB serves the
  // nonsymmetric ones (asymmetric constraint rows): A^T x = b on A's
  // structurally symmetric pattern with mirrored values, what
  // LinearSolver::adjoint_solve's transpose-solve-transpose does
  // (src/solvers/linear_solver.C:147-171).
  if (this->_solver_type != BICGSTAB) return solve(matrix, solution, rhs, tol, m_its);
  this->init();
  auto *A = dynamic_cast<HipCsrMatrix<T> *>(&matrix);
  auto *x = dynamic_cast<HipVector<T> *>(&solution);
  auto *b = dynamic_cast<HipVector<T> *>(&rhs);
  if (!A || !x || !b) libmesh_error_msg("HipLinearSolver needs a HipCsrMatrix and HipVectors");
  if (_mixed) libmesh_not_implemented();
  A->close();
  x->close();
  b->close();
  hip_check(lmh_matrix_set_preconditioner(A->device_matrix(), this->_preconditioner_type == IDENTITY_PRECOND
                                                                  ? LMH_PRECOND_IDENTITY
                                                                  : LMH_PRECOND_JACOBI),
            "HipLinearSolver (preconditioner)");
  int64_t its = 0;
  double res = 0;
  int reason = 0;
  hip_check(lmh_bicgstab_solve_transpose(A->device_matrix(), (const double *)b->device_data(),
                                         (double *)x->device_data(), tol.value_or(this->default_tolerance()),
                                         m_its.value_or(this->default_max_iterations()), &its, &res, &reason),
            "HipLinearSolver::adjoint_solve (BiCGSTAB)");
  x->mark_device_modified();
  _reason = to_reason(reason);
  return std::make_pair((unsigned int)its, (Real)res);
}

template class HipLinearSolver<double>;

}  // namespace libMesh
