// fem_system.C -- FEMSystem assembly through the device physics, and the
// steady Newton solve (src/systems/fem_system.C:875-1010,
// src/solvers/newton_solver.C:280-631 of libMesh).
#include "libmesh/fem_system.h"

#include <algorithm>
#include <cmath>
#include <cstdio>

#include "libmesh/hip_csr_matrix.h"
#include "libmesh/hip_device.h"
#include "libmesh/quadrature_gauss.h"

namespace libMesh {

void FEMSystem::init() {
  ImplicitSystem::init();
  if (!time_solver) time_solver = std::make_unique<SteadySolver>(*this);
}

FEMSystem::~FEMSystem() {
  lmh_context *ctx = HipDevice::context();
  for (double *p : {_Fe_d, _Ke_d, _coef_d})
    if (p) lmh_free(ctx, p);
}

// (re)allocates a device buffer of at least n doubles
static double *device_buffer(double *&p, size_t &have, size_t n, const char *what) {
  if (p && have >= n) return p;
  lmh_context *ctx = HipDevice::context();
  if (p) lmh_free(ctx, p);
  p = nullptr;
  have = 0;
  void *q = nullptr;
  hip_check(lmh_malloc(ctx, sizeof(double) * (n ? n : 1), &q), what);
  p = (double *)q;
  have = n;
  return p;
}

// FEMSystem::assembly for a DevicePhysics: the element loop of
// fem_system.C:875-1010 (FEMContext::pre_fe_reinit / elem_fe_reinit,
// element_time_derivative, the element constraint, add_matrix / add_vector)
// with all elements on the device, then their sums in element order.
void FEMSystem::device_physics_assembly(bool get_residual, bool get_jacobian) {
  auto *A = dynamic_cast<HipCsrMatrix<Number> *>(&get_system_matrix());
  if (!A || !A->device_matrix() || !A->device_mesh() || !A->device_dofmap())
    libmesh_error_msg("FEMSystem::assembly (device physics): the system matrix must be a HipCsrMatrix "
                      "initialised from the DofMap");
  const MeshBase *mesh = get_dof_map().mesh();
  if (!mesh) libmesh_error_msg("FEMSystem::assembly (device physics): DoFs not distributed");
  // FEMContext::use_default_quadrature_rules (fem_context.C:155-175)
  const FEType fe_type = get_dof_map().variable_type(0);
  const QGauss qrule(3, static_cast<Order>(fe_type.default_quadrature_order() + extra_quadrature_order));
  lmh_fem_data d;
  hip_check(lmh_fem_element_data(A->device_mesh(), A->device_dofmap(), qrule.kernel_arg(mesh->elem_type()), &d),
            "FEMSystem::assembly (element data)");
  const size_t ne = (size_t)d.n_elem, nd = (size_t)d.n_dofs;
  double *Fe = device_buffer(_Fe_d, _Fe_n, ne * nd, "FEMSystem::assembly (element residuals)");
  double *Ke = get_jacobian ? device_buffer(_Ke_d, _Ke_n, ne * nd * nd, "FEMSystem::assembly (element Jacobians)")
                            : nullptr;
  const std::vector<Real> &coef = _device_physics->coefficient();
  const double *coef_d = nullptr;
  if (!coef.empty()) {
    if (coef.size() != ne) libmesh_error_msg("DevicePhysics::coefficient needs one value per element");
    coef_d = device_buffer(_coef_d, _coef_n, ne, "FEMSystem::assembly (coefficients)");
    hip_check(lmh_memcpy_h2d(HipDevice::context(), (void *)coef_d, coef.data(), sizeof(double) * ne),
              "FEMSystem::assembly (coefficients)");
  }
  solution->close();
  auto *u = dynamic_cast<HipVector<Number> *>(solution.get());
  std::unique_ptr<HipVector<Number>> u_copy;
  if (!u) {  // any other NumericVector: a device copy of its values
    u_copy = std::make_unique<HipVector<Number>>(solution->comm(), solution->size(), solution->size());
    std::vector<Number> vals;
    solution->localize(vals);
    *u_copy = vals;
    u_copy->close();
    u = u_copy.get();
  }
  hip_check(_device_physics->element_loop(d, u->device_data(), coef_d, get_jacobian, Fe, Ke),
            "FEMSystem::assembly (element_time_derivative)");
  auto *R = dynamic_cast<HipVector<Number> *>(rhs);
  if (get_residual && !R) libmesh_error_msg("FEMSystem::assembly (device physics): rhs must be a HipVector");
  double *R_d = get_residual ? (double *)R->device_data() : nullptr;
  hip_check(lmh_matrix_scatter_elements(A->device_matrix(), Ke, get_residual ? Fe : nullptr, R_d),
            "FEMSystem::assembly (element sums)");
  if (get_jacobian) A->mark_device_modified();
  if (get_residual) R->mark_device_modified();
}

void FEMSystem::assembly(bool get_residual, bool get_jacobian, bool apply_heterogeneous_constraints,
                         bool apply_no_constraints) {
  if (!get_residual && !get_jacobian) libmesh_error_msg("FEMSystem::assembly: nothing requested");
  if (_device_physics) {
    // Dirichlet rows are empty with value 0, so the heterogeneous and the
    // homogeneous element constraint coincide
    (void)apply_heterogeneous_constraints;
    if (apply_no_constraints) libmesh_not_implemented();
    device_physics_assembly(get_residual, get_jacobian);
    return;
  }
  if (!_physics) libmesh_error_msg("FEMSystem::assembly: no physics attached");
  // Dirichlet rows are empty with value 0, so the heterogeneous and the
  // homogeneous element constraint coincide; unconstrained assembly is not
  // offered by the device kernels
  (void)apply_heterogeneous_constraints;
  if (apply_no_constraints) libmesh_not_implemented();
  // general constraint rows: the Jacobian is C^T K C with the DofMap's row
  // form, the residual C^T (K u - F) at the solution, whose constraints the
  // Newton solver keeps enforced
  if (get_jacobian) _physics->assemble_jacobian();
  if (get_residual) {
    solution->close();
    _physics->assemble_residual(*solution, *rhs);
  }
}

void FEMSystem::solve() {
  if (!time_solver) libmesh_error_msg("FEMSystem::solve: call init() first");
  time_solver->solve();
}

bool NewtonSolver::test_convergence(Real current_residual, Real step_norm, bool linear_solve_finished) {
  bool has_converged = false;
  if (current_residual < absolute_residual_tolerance) {
    _solve_result |= CONVERGED_ABSOLUTE_RESIDUAL;
    has_converged = true;
  }
  if ((current_residual / max_residual_norm) < relative_residual_tolerance) {
    _solve_result |= CONVERGED_RELATIVE_RESIDUAL;
    has_converged = true;
  }
  if (!linear_solve_finished) return has_converged;
  if (step_norm < absolute_step_tolerance) {
    _solve_result |= CONVERGED_ABSOLUTE_STEP;
    has_converged = true;
  }
  if (max_solution_norm && (step_norm / max_solution_norm < relative_step_tolerance)) {
    _solve_result |= CONVERGED_RELATIVE_STEP;
    has_converged = true;
  }
  return has_converged;
}

// Backtracking of NewtonSolver::line_search (newton_solver.C:660-720): the step
// is shortened while the residual grows.  The Brent refinement that follows it
// in libMesh is not restated: the returned step is the first one that reduces
// the residual.
Real NewtonSolver::line_search(Real tol, Real last_residual, Real &current_residual,
                               NumericVector<Number> &newton_iterate, const NumericVector<Number> &linear_solution) {
  if ((current_residual < last_residual) ||
      (!require_residual_reduction && (!require_finite_residual || !std::isnan(current_residual))))
    return 1.;
  auto &sys = static_cast<FEMSystem &>(_system);
  Real bx = 1.;
  while (std::isnan(current_residual) || (current_residual > last_residual && require_residual_reduction)) {
    Real sub;
    if (brent_line_search && !std::isnan(current_residual)) {
      sub = std::min(0.5, last_residual / current_residual);
      sub = std::max(sub, tol * 2.);
    } else {
      sub = 0.5;
    }
    newton_iterate.add(bx * (1. - sub), linear_solution);
    newton_iterate.close();
    bx *= sub;
    if (verbose) std::printf("  Shrinking Newton step to %g\n", bx);
    sys.assembly(true, false);
    current_residual = _system.rhs->l2_norm();
    if (bx / 2. < minsteplength && (std::isnan(current_residual) || current_residual > last_residual)) {
      std::printf("Inexact Newton step FAILED at step %u\n", _outer_iterations);
      if (!continue_after_backtrack_failure) libmesh_error_msg("Newton line search failed (convergence failure)");
      _solve_result = DIVERGED_BACKTRACKING_FAILURE;
      return bx;
    }
  }
  return bx;
}

unsigned int NewtonSolver::solve() {
  auto &sys = dynamic_cast<FEMSystem &>(_system);
  _solve_result = INVALID_SOLVE_RESULT;
  NumericVector<Number> &newton_iterate = *sys.solution;
  std::unique_ptr<NumericVector<Number>> linear_solution_ptr = newton_iterate.zero_clone();
  NumericVector<Number> &linear_solution = *linear_solution_ptr;
  NumericVector<Number> &rhs = *sys.rhs;
  newton_iterate.close();
  rhs.close();
  sys.get_dof_map().enforce_constraints_exactly(sys);
  SparseMatrix<Number> &matrix = sys.get_system_matrix();
  if (!sys.linear_solver) libmesh_error_msg("NewtonSolver: the system has no linear solver");
  LinearSolver<Number> &lin = *sys.linear_solver;
  lin.init();
  lin.init_systems(sys);
  double current_linear_tolerance = initial_linear_tolerance;
  _inner_iterations = 0;
  max_solution_norm = max_residual_norm = 0;
  for (_outer_iterations = 0; _outer_iterations < max_nonlinear_iterations; ++_outer_iterations) {
    sys.assembly(true, true);
    Real current_residual = rhs.l2_norm();
    if (std::isnan(current_residual)) {
      std::printf("  Nonlinear solver DIVERGED at step %u with residual Not-a-Number\n", _outer_iterations);
      libmesh_error_msg("Newton residual is NaN (convergence failure)");
    }
    if (current_residual <= absolute_residual_tolerance) {
      _solve_result |= CONVERGED_ABSOLUTE_RESIDUAL;
      if (current_residual == 0) {
        if (relative_residual_tolerance > 0) _solve_result |= CONVERGED_RELATIVE_RESIDUAL;
        if (absolute_step_tolerance > 0) _solve_result |= CONVERGED_ABSOLUTE_STEP;
        if (relative_step_tolerance > 0) _solve_result |= CONVERGED_RELATIVE_STEP;
      }
      break;
    }
    const Real last_residual = current_residual;
    max_residual_norm = std::max(current_residual, max_residual_norm);
    Real norm_total = newton_iterate.l2_norm();
    max_solution_norm = std::max(max_solution_norm, norm_total);
    if (verbose) std::printf("Nonlinear Residual: %.6e\n", current_residual);
    current_linear_tolerance = std::min(current_linear_tolerance, current_residual * linear_tolerance_multiplier);
    if (current_linear_tolerance < minimum_linear_tolerance) current_linear_tolerance = minimum_linear_tolerance;
    current_linear_tolerance = std::max(current_linear_tolerance, absolute_residual_tolerance / current_residual / 10.0);
    linear_solution.zero();
    const std::pair<unsigned int, Real> rval =
        lin.solve(matrix, linear_solution, rhs, current_linear_tolerance, max_linear_iterations);
    if (track_linear_convergence && lin.get_converged_reason() < 0) {
      _solve_result |= DIVERGED_LINEAR_SOLVER_FAILURE;
      std::printf("Linear solver failed during Newton step, dropping out.\n");
      break;
    }
    sys.get_dof_map().enforce_constraints_exactly(sys, &linear_solution, /* homogeneous = */ true);
    const unsigned int linear_steps = rval.first;
    _inner_iterations += linear_steps;
    const bool linear_solve_finished = !(linear_steps == max_linear_iterations);
    Real norm_delta = linear_solution.l2_norm();
    newton_iterate.add(-1., linear_solution);
    newton_iterate.close();
    if (require_residual_reduction || require_finite_residual || _outer_iterations + 1 < max_nonlinear_iterations ||
        !continue_after_max_iterations) {
      sys.assembly(true, false);
      current_residual = rhs.l2_norm();
      if (verbose) std::printf("  Current Residual: %.6e\n", current_residual);
      if (test_convergence(current_residual, norm_delta, linear_solve_finished && current_residual <= last_residual)) {
        if (!quiet)
          std::printf("  Nonlinear solver converged, step %u, residual %.6e\n", _outer_iterations, current_residual);
        _outer_iterations++;
        break;
      }
    }
    Real steplength = line_search(std::sqrt(TOLERANCE), last_residual, current_residual, newton_iterate,
                                  linear_solution);
    norm_delta *= steplength;
    if (_solve_result == DIVERGED_BACKTRACKING_FAILURE) {
      _outer_iterations++;
      break;
    }
    if (_outer_iterations + 1 >= max_nonlinear_iterations) {
      std::printf("  Nonlinear solver reached maximum step %u, latest evaluated residual %.6e\n",
                  max_nonlinear_iterations, current_residual);
      if (continue_after_max_iterations) {
        _solve_result = DIVERGED_MAX_NONLINEAR_ITERATIONS;
        std::printf("  Continuing...\n");
      } else {
        libmesh_error_msg("Newton reached the maximum number of steps (convergence failure)");
      }
      continue;
    }
    norm_total = newton_iterate.l2_norm();
    max_solution_norm = std::max(max_solution_norm, norm_total);
    if (verbose) std::printf("  Nonlinear step: |du|/|u| = %.6e, |du| = %.6e\n", norm_delta / norm_total, norm_delta);
    if (test_convergence(current_residual, norm_delta / steplength, linear_solve_finished)) {
      if (!quiet)
        std::printf("  Nonlinear solver converged, step %u, residual %.6e\n", _outer_iterations, current_residual);
      _outer_iterations++;
      break;
    }
  }
  sys.get_dof_map().enforce_constraints_exactly(sys);
  return _solve_result;
}

}  // namespace libMesh
