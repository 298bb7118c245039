// hip_linear_solver.C -- LinearSolver backend: CG (FP64 or mixed) and
// BiCGSTAB (FP64) on the GPU, with JACOBI_PRECOND or IDENTITY_PRECOND (the two
// preconditioners EigenSparseLinearSolver::solve runs without a factorisation,
// eigen_sparse_linear_solver.C:136-170, 177-205); plus the LinearSolver
// members that the interface implements itself (src/solvers/linear_solver.C).
#include "libmesh/hip_linear_solver.h"

#include <type_traits>

#include "libmesh/static_condensation.h"

namespace libMesh {

template <typename T>
std::unique_ptr<LinearSolver<T>> LinearSolver<T>::build(const libMesh::Parallel::Communicator &comm_in,
                                                        const SolverPackage solver_package) {
  if (solver_package != HIP_SOLVERS) libmesh_not_implemented();
  return std::make_unique<HipLinearSolver<T>>(comm_in);
}

template <typename T>
void LinearSolver<T>::restrict_solve_to(const std::vector<unsigned int> *const dofs, const SubsetSolveMode) {
  if (dofs != nullptr) libmesh_not_implemented();
}

// linear_solver.C:147-171: transpose in place, solve, transpose back
template <typename T>
std::pair<unsigned int, Real> LinearSolver<T>::adjoint_solve(SparseMatrix<T> &mat, NumericVector<T> &sol,
                                                             NumericVector<T> &rhs, const std::optional<double> tol,
                                                             const std::optional<unsigned int> n_its) {
  mat.close();
  mat.get_transpose(mat);
  const auto r = this->solve(mat, sol, rhs, tol, n_its);
  mat.get_transpose(mat);
  return r;
}

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

// A HipVector view of a solve operand: the vector itself, or a device copy
// of another NumericVector that copy_back() writes back.
template <typename T>
struct DeviceOperand {
  NumericVector<T> &v;
  HipVector<T> *hv;
  std::unique_ptr<HipVector<T>> tmp;
  explicit DeviceOperand(NumericVector<T> &vec) : v(vec), hv(dynamic_cast<HipVector<T> *>(&vec)) {
    if (!hv) {
      tmp = std::make_unique<HipVector<T>>(vec.comm(), vec.size(), vec.size());
      std::vector<T> vals;
      vec.localize(vals);
      *tmp = vals;
      hv = tmp.get();
    }
    hv->close();
  }
  void copy_back() {
    if (!tmp) return;
    std::vector<T> vals;
    tmp->localize(vals);
    v = vals;
    v.close();
  }
};
}  // namespace

template <typename T>
std::pair<unsigned int, Real> HipLinearSolver<T>::run(SparseMatrix<T> &matrix_in, NumericVector<T> &solution_in,
                                                      NumericVector<T> &rhs_in, std::optional<double> tol,
                                                      std::optional<unsigned int> m_its, bool transpose) {
  this->init();
  auto *A = dynamic_cast<HipCsrMatrix<T> *>(&matrix_in);
  if (!A) libmesh_error_msg("HipLinearSolver needs a HipCsrMatrix");
  if constexpr (std::is_same<T, Number>::value) {
    // a statically condensed system: forward elimination, the reduced solve
    // with the condensation's own solver at this tolerance, back substitution
    // (the StaticCondensationPreconditioner apply of libMesh, used as the whole solve)
    if (auto *sc = dynamic_cast<StaticCondensation *>(&matrix_in)) {
      if (transpose && this->_solver_type == BICGSTAB) libmesh_not_implemented();
      const auto r = sc->condensed_solve(rhs_in, solution_in, tol.value_or(this->default_tolerance()),
                                         m_its.value_or(this->default_max_iterations()));
      _reason = sc->reduced_converged_reason();
      return r;
    }
  }
  // close everything, as EigenSparseLinearSolver::solve does (eigen_sparse_linear_solver.C:97-100)
  A->close();
  DeviceOperand<T> x(solution_in), b(rhs_in);
  hip_check(lmh_matrix_set_preconditioner(A->device_matrix(), this->_preconditioner_type == IDENTITY_PRECOND
                                                                  ? LMH_PRECOND_IDENTITY
                                                                  : LMH_PRECOND_JACOBI),
            "HipLinearSolver (preconditioner)");
  const double rtol = tol.value_or(this->default_tolerance());
  const unsigned maxits = m_its.value_or(this->default_max_iterations());
  const double *bd = (const double *)b.hv->device_data();
  double *xd = (double *)x.hv->device_data();
  int64_t its = 0;
  double res = 0;
  int reason = 0;
  if (this->_solver_type == BICGSTAB) {
    // EigenSparseLinearSolver::solve, case BICGSTAB (eigen_sparse_linear_solver.C:175-205)
    if (_mixed) libmesh_not_implemented();
    if (transpose)
      hip_check(lmh_bicgstab_solve_transpose(A->device_matrix(), bd, xd, rtol, maxits, &its, &res, &reason),
                "HipLinearSolver::adjoint_solve (BiCGSTAB)");
    else
      hip_check(lmh_bicgstab_solve(A->device_matrix(), bd, xd, rtol, maxits, &its, &res, &reason),
                "HipLinearSolver::solve (BiCGSTAB)");
  } else if (_mixed) {
    hip_check(lmh_matrix_make_fp32(A->device_matrix()), "HipLinearSolver (FP32 operator)");
    double stats[3];
    hip_check(lmh_mixed_solve(A->device_matrix(), bd, xd, rtol, _inner, 0, maxits, &its, &res, &reason, stats),
              "HipLinearSolver::solve (mixed)");
    _nref = (unsigned)stats[1];
  } else {
    hip_check(lmh_cg_solve(A->device_matrix(), bd, xd, rtol, maxits, &its, &res, &reason), "HipLinearSolver::solve");
  }
  x.hv->mark_device_modified();
  x.copy_back();
  _reason = to_reason(reason);
  return std::make_pair((unsigned int)its, (Real)res);
}

template <typename T>
std::pair<unsigned int, Real> HipLinearSolver<T>::solve(SparseMatrix<T> &matrix, NumericVector<T> &solution,
                                                        NumericVector<T> &rhs, const std::optional<double> tol,
                                                        const std::optional<unsigned int> m_its) {
  return run(matrix, solution, rhs, tol, m_its, false);
}

template <typename T>
std::pair<unsigned int, Real> HipLinearSolver<T>::adjoint_solve(SparseMatrix<T> &matrix, NumericVector<T> &solution,
                                                                NumericVector<T> &rhs,
                                                                const std::optional<double> tol,
                                                                const std::optional<unsigned int> m_its) {
  // CG: the system matrices are symmetric, so A^T = A.  BiCGSTAB serves the
  // nonsymmetric ones (asymmetric constraint rows): A^T x = b on A's
  // structurally symmetric pattern with mirrored values, what
  // LinearSolver::adjoint_solve's transpose-solve-transpose does
  // (src/solvers/linear_solver.C:147-171).
  return run(matrix, solution, rhs, tol, m_its, this->_solver_type == BICGSTAB);
}

template <typename T>
std::pair<unsigned int, Real> HipLinearSolver<T>::solve(SparseMatrix<T> &matrix, SparseMatrix<T> &pc,
                                                        NumericVector<T> &solution, NumericVector<T> &rhs,
                                                        const std::optional<double> tol,
                                                        const std::optional<unsigned int> m_its) {
  if (&pc != &matrix) libmesh_error_msg("ERROR: HipLinearSolver does not support a user-supplied preconditioner!");
  return run(matrix, solution, rhs, tol, m_its, false);
}

template <typename T>
std::pair<unsigned int, Real> HipLinearSolver<T>::solve(const ShellMatrix<T> &, NumericVector<T> &,
                                                        NumericVector<T> &, const std::optional<double>,
                                                        const std::optional<unsigned int>) {
  libmesh_not_implemented();
  return std::make_pair(0, 0.0);
}

template <typename T>
std::pair<unsigned int, Real> HipLinearSolver<T>::solve(const ShellMatrix<T> &, const SparseMatrix<T> &,
                                                        NumericVector<T> &, NumericVector<T> &,
                                                        const std::optional<double>,
                                                        const std::optional<unsigned int>) {
  libmesh_not_implemented();
  return std::make_pair(0, 0.0);
}

template class LinearSolver<double>;
template class HipLinearSolver<double>;

}  // namespace libMesh
