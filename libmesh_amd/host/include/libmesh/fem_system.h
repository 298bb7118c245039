// fem_system.h -- FEMSystem with a steady Newton solve, the subset of
// include/systems/fem_system.h, diff_system.h, include/solvers/steady_solver.h,
// newton_solver.h and diff_solver.h that the assembly-plus-solve path uses.
// The element physics runs on the device: the attached physics object is the
// HipAssembler of -div(kappa grad u) = f (per-element kappa), whose residual
// R_e = K_e u_e - F_e and Jacobian K_e are what element_time_derivative of
// such a physics returns (the Laplace physics of adjoints_ex1/ex2 with a
// coefficient).
#pragma once

#include <memory>

#include "libmesh/device_physics.h"
#include "libmesh/hip_assembler.h"
#include "libmesh/system.h"

namespace libMesh {

// DiffSolver (include/solvers/diff_solver.h): parameters, defaults
// (src/solvers/diff_solver.C:30-50) and result flags
class NewtonSolver {
 public:
  enum SolveResult {
    INVALID_SOLVE_RESULT = 0,
    CONVERGED_NO_REASON = 1,
    CONVERGED_ABSOLUTE_RESIDUAL = 2,
    CONVERGED_RELATIVE_RESIDUAL = 4,
    CONVERGED_ABSOLUTE_STEP = 8,
    CONVERGED_RELATIVE_STEP = 16,
    DIVERGED_NO_REASON = 32,
    DIVERGED_MAX_NONLINEAR_ITERATIONS = 64,
    DIVERGED_BACKTRACKING_FAILURE = 128,
    DIVERGED_LINEAR_SOLVER_FAILURE = 256
  };

  explicit NewtonSolver(ImplicitSystem &s) : _system(s) {}

  unsigned int max_linear_iterations = 1000;
  unsigned int max_nonlinear_iterations = 100;
  bool quiet = true, verbose = false;
  bool continue_after_max_iterations = true, continue_after_backtrack_failure = false;
  Real absolute_residual_tolerance = 0, relative_residual_tolerance = 0;
  Real absolute_step_tolerance = 0, relative_step_tolerance = 0;
  double initial_linear_tolerance = 1e-12, minimum_linear_tolerance = TOLERANCE * TOLERANCE;
  // NewtonSolver (src/solvers/newton_solver.C:237-247)
  bool require_residual_reduction = true, require_finite_residual = true;
  bool brent_line_search = true, track_linear_convergence = false;
  Real minsteplength = 1e-5;
  double linear_tolerance_multiplier = 1e-3;

  // NewtonSolver::solve (newton_solver.C:280-580); returns the SolveResult flags
  unsigned int solve();
  unsigned int total_outer_iterations() const { return _outer_iterations; }
  unsigned int total_inner_iterations() const { return _inner_iterations; }
  unsigned int solve_result() const { return _solve_result; }

 private:
  bool test_convergence(Real current_residual, Real step_norm, bool linear_solve_finished);
  Real line_search(Real tol, Real last_residual, Real &current_residual, NumericVector<Number> &newton_iterate,
                   const NumericVector<Number> &linear_solution);
  ImplicitSystem &_system;
  Real max_solution_norm = 0, max_residual_norm = 0;
  unsigned int _outer_iterations = 0, _inner_iterations = 0, _solve_result = INVALID_SOLVE_RESULT;
};

// SteadySolver (include/solvers/steady_solver.h): a single nonlinear solve
class SteadySolver {
 public:
  explicit SteadySolver(ImplicitSystem &s) : _diff_solver(std::make_unique<NewtonSolver>(s)) {}
  NewtonSolver &diff_solver() { return *_diff_solver; }
  void solve() { _diff_solver->solve(); }

 private:
  std::unique_ptr<NewtonSolver> _diff_solver;
};

class FEMSystem : public ImplicitSystem {
 public:
  using ImplicitSystem::ImplicitSystem;
  ~FEMSystem() override;
  // DifferentiableSystem::attach_physics (diff_system.h:200): the built-in
  // device Poisson physics, or a user physics whose element_time_derivative
  // is device code (DevicePhysicsFor<P>, device_physics_hip.h)
  void attach_physics(HipAssembler *physics) {
    _physics = physics;
    _device_physics = nullptr;
  }
  void attach_physics(DevicePhysics *physics) {
    _device_physics = physics;
    _physics = nullptr;
  }
  // DifferentiableSystem::init_data builds the default time solver
  void init() override;
  // FEMSystem::assembly (fem_system.C:875-1010): residual at the current
  // solution into rhs, Jacobian into the system matrix.  A DevicePhysics runs
  // element_time_derivative on every element (quadrature: the variable's
  // default QGauss rule plus extra_quadrature_order, fem_context.C:155-175),
  // the element constraint of the Dirichlet rows, and the element-order sum
  // into the HipCsrMatrix and rhs.  The built-in physics is constrained as with
  // constrain_in_solver (element constraint; the Newton solver enforces the
  // constraints exactly).  With general constraint rows the residual is
  // C^T (K u - F) at the enforced solution and the Jacobian C^T K C in the
  // DofMap's row form (CG for symmetric rows, BICGSTAB for asymmetric).
  void assembly(bool get_residual, bool get_jacobian, bool apply_heterogeneous_constraints = false,
                bool apply_no_constraints = false) override;
  // DifferentiableSystem::solve -> time_solver->solve()
  void solve() override;
  std::unique_ptr<SteadySolver> time_solver;

 protected:
  HipAssembler *device_assembler() const override { return _physics; }

 private:
  void device_physics_assembly(bool get_residual, bool get_jacobian);
  HipAssembler *_physics = nullptr;
  DevicePhysics *_device_physics = nullptr;
  // element results and coefficients of the device physics (kept between assemblies)
  double *_Fe_d = nullptr, *_Ke_d = nullptr, *_coef_d = nullptr;
  size_t _Fe_n = 0, _Ke_n = 0, _coef_n = 0;
};

}  // namespace libMesh
