// eigen_ref.cpp -- drives the reference's own vendored Eigen 3.3.9 CG with
// DiagonalPreconditioner, exactly as EigenSparseLinearSolver::solve does for
// CG + JACOBI_PRECOND (src/solvers/eigen_sparse_linear_solver.C:106-160:
// ConjugateGradient<EigenSM, Lower|Upper, DiagonalPreconditioner<Number>>,
// setMaxIterations, setTolerance, solveWithGuess).  EigenSM is a RowMajor
// SparseMatrix<Number, RowMajor, eigen_idx_type> (include/numerics/eigen_core_support.h).
//
// TEST INFRASTRUCTURE ONLY: built by oracle/Makefile against the headers under
// the reference tree into oracle/_ref/, and used to pin the oracle's CG
// restatement and as the "reference" CPU baseline.
#include <omp.h>

#include <Eigen/IterativeLinearSolvers>
#include <Eigen/Sparse>
#include <chrono>
#include <cstdint>
#include <cstring>

typedef Eigen::SparseMatrix<double, Eigen::RowMajor, int> EigenSM;

extern "C" int ref_eigen_cg(int64_t n, const int32_t *rp, const int32_t *ci, const double *v, const double *b,
                            double *x, double tol, int64_t maxits, int64_t *iters, double *err) {
  Eigen::Map<const EigenSM> view((Eigen::Index)n, (Eigen::Index)n, (Eigen::Index)rp[n], rp, ci, v);
  EigenSM A(view);
  Eigen::Map<const Eigen::VectorXd> bb(b, n);
  Eigen::VectorXd xx = Eigen::Map<const Eigen::VectorXd>(x, n);
  Eigen::ConjugateGradient<EigenSM, Eigen::Lower | Eigen::Upper, Eigen::DiagonalPreconditioner<double>> solver(A);
  solver.setMaxIterations((Eigen::Index)maxits);
  solver.setTolerance(tol);
  xx = solver.solveWithGuess(bb, xx);
  std::memcpy(x, xx.data(), sizeof(double) * n);
  *iters = solver.iterations();
  *err = solver.error();
  return (int)solver.info();
}

extern "C" void ref_eigen_spmv(int64_t n, const int32_t *rp, const int32_t *ci, const double *v, const double *x,
                               double *y) {
  Eigen::Map<const EigenSM> A((Eigen::Index)n, (Eigen::Index)n, (Eigen::Index)rp[n], rp, ci, v);
  Eigen::Map<const Eigen::VectorXd> xx(x, n);
  Eigen::Map<Eigen::VectorXd> yy(y, n);
  yy.noalias() = A * xx;
}

// BICGSTAB + JACOBI_PRECOND of EigenSparseLinearSolver::solve
// (eigen_sparse_linear_solver.C:175-205): BiCGSTAB<EigenSM, DiagonalPreconditioner>.
extern "C" int ref_eigen_bicgstab(int64_t n, const int32_t *rp, const int32_t *ci, const double *v, const double *b,
                                  double *x, double tol, int64_t maxits, int64_t *iters, double *err) {
  Eigen::Map<const EigenSM> view((Eigen::Index)n, (Eigen::Index)n, (Eigen::Index)rp[n], rp, ci, v);
  EigenSM A(view);
  Eigen::Map<const Eigen::VectorXd> bb(b, n);
  Eigen::VectorXd xx = Eigen::Map<const Eigen::VectorXd>(x, n);
  Eigen::BiCGSTAB<EigenSM, Eigen::DiagonalPreconditioner<double>> solver(A);
  solver.setMaxIterations((Eigen::Index)maxits);
  solver.setTolerance(tol);
  xx = solver.solveWithGuess(bb, xx);
  std::memcpy(x, xx.data(), sizeof(double) * n);
  *iters = solver.iterations();
  *err = solver.error();
  return (int)solver.info();
}

// CG and BICGSTAB with IDENTITY_PRECOND (eigen_sparse_linear_solver.C:136-143,
// 177-183): ConjugateGradient / BiCGSTAB with IdentityPreconditioner.
extern "C" int ref_eigen_cg_identity(int64_t n, const int32_t *rp, const int32_t *ci, const double *v,
                                     const double *b, double *x, double tol, int64_t maxits, int64_t *iters,
                                     double *err) {
  Eigen::Map<const EigenSM> view((Eigen::Index)n, (Eigen::Index)n, (Eigen::Index)rp[n], rp, ci, v);
  EigenSM A(view);
  Eigen::Map<const Eigen::VectorXd> bb(b, n);
  Eigen::VectorXd xx = Eigen::Map<const Eigen::VectorXd>(x, n);
  Eigen::ConjugateGradient<EigenSM, Eigen::Lower | Eigen::Upper, Eigen::IdentityPreconditioner> solver(A);
  solver.setMaxIterations((Eigen::Index)maxits);
  solver.setTolerance(tol);
  xx = solver.solveWithGuess(bb, xx);
  std::memcpy(x, xx.data(), sizeof(double) * n);
  *iters = solver.iterations();
  *err = solver.error();
  return (int)solver.info();
}

extern "C" int ref_eigen_bicgstab_identity(int64_t n, const int32_t *rp, const int32_t *ci, const double *v,
                                           const double *b, double *x, double tol, int64_t maxits, int64_t *iters,
                                           double *err) {
  Eigen::Map<const EigenSM> view((Eigen::Index)n, (Eigen::Index)n, (Eigen::Index)rp[n], rp, ci, v);
  EigenSM A(view);
  Eigen::Map<const Eigen::VectorXd> bb(b, n);
  Eigen::VectorXd xx = Eigen::Map<const Eigen::VectorXd>(x, n);
  Eigen::BiCGSTAB<EigenSM, Eigen::IdentityPreconditioner> solver(A);
  solver.setMaxIterations((Eigen::Index)maxits);
  solver.setTolerance(tol);
  xx = solver.solveWithGuess(bb, xx);
  std::memcpy(x, xx.data(), sizeof(double) * n);
  *iters = solver.iterations();
  *err = solver.error();
  return (int)solver.info();
}

// The CG + JACOBI_PRECOND solve above with the matrix copy kept out of the
// clock: the solve alone (solveWithGuess) is timed with steady_clock on
// `threads` OpenMP threads (Eigen's row-major sparse * dense product is the
// threaded part, SparseDenseProduct.h:41-51; vector operations stay serial).
// Used as the CPU baseline's solve phase.

extern "C" int ref_eigen_cg_timed(int64_t n, const int32_t *rp, const int32_t *ci, const double *v, const double *b,
                                  double *x, double tol, int64_t maxits, int threads, int64_t *iters, double *err,
                                  double *seconds) {
  Eigen::Map<const EigenSM> view((Eigen::Index)n, (Eigen::Index)n, (Eigen::Index)rp[n], rp, ci, v);
  EigenSM A(view);
  Eigen::Map<const Eigen::VectorXd> bb(b, n);
  Eigen::VectorXd xx = Eigen::Map<const Eigen::VectorXd>(x, n);
  const int before = omp_get_max_threads();
  omp_set_num_threads(threads > 0 ? threads : before);
  Eigen::setNbThreads(threads > 0 ? threads : before);
  Eigen::ConjugateGradient<EigenSM, Eigen::Lower | Eigen::Upper, Eigen::DiagonalPreconditioner<double>> solver(A);
  solver.setMaxIterations((Eigen::Index)maxits);
  solver.setTolerance(tol);
  const auto t0 = std::chrono::steady_clock::now();
  xx = solver.solveWithGuess(bb, xx);
  const auto t1 = std::chrono::steady_clock::now();
  omp_set_num_threads(before);
  Eigen::setNbThreads(before);
  std::memcpy(x, xx.data(), sizeof(double) * n);
  *iters = solver.iterations();
  *err = solver.error();
  *seconds = std::chrono::duration<double>(t1 - t0).count();
  return (int)solver.info();
}
