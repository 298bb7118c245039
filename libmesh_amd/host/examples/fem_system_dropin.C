// fem_system_dropin.C -- a steady FEMSystem solved by NewtonSolver, with the
// Poisson physics -div(kappa grad u) = f evaluated on the GPU (residual and
// Jacobian through FEMSystem::assembly), the way FEMSystem applications
// (adjoints_ex1/ex2, fem_system_ex*) drive assembly and solve.
// Usage: fem_system_dropin n elem_type(10|12|8|9) out_prefix [perturb] [kappa] [sine] [rtol=R] [adjoint]
//                          [periodic] [bicgstab]
// Writes <prefix>_u.bin; prints Newton and linear iteration counts.  `adjoint`:
// adjoint_qoi_parameter_sensitivity of J = integral of u over [0.25,0.75]^3 for
// every per-element kappa through the FEMSystem's Jacobian, written to
// <prefix>_grad.bin.  `periodic`: Dirichlet on the y and z faces only and the
// rows u(x=1) = u(x=0) plus u = 0.3 at one x=0 face DoF from a
// System::Constraint object, symmetric rows with CG; with `bicgstab` the
// default asymmetric rows and BiCGSTAB.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <string>

#include "libmesh/fem_system.h"
#include "libmesh/hip_assembler.h"
#include "libmesh/hip_csr_matrix.h"
#include "libmesh/hip_linear_solver.h"
#include "libmesh/libmesh.h"
#include "box_qoi.h"
#include "periodic_x.h"

using namespace libMesh;

int main(int argc, char **argv) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: %s n elem_type out_prefix [perturb] [kappa] [sine] [rtol]\n", argv[0]);
    return 2;
  }
  const unsigned n = std::atoi(argv[1]);
  const ElemType type = (ElemType)std::atoi(argv[2]);
  const std::string prefix = argv[3];
  bool perturb = false, kappa = false, sine = false, adjoint = false, periodic = false, bicgstab = false;
  Real rtol = 1e-9;
  for (int a = 4; a < argc; ++a) {
    perturb |= !std::strcmp(argv[a], "perturb");
    kappa |= !std::strcmp(argv[a], "kappa");
    sine |= !std::strcmp(argv[a], "sine");
    adjoint |= !std::strcmp(argv[a], "adjoint");
    periodic |= !std::strcmp(argv[a], "periodic");
    bicgstab |= !std::strcmp(argv[a], "bicgstab");
    if (!std::strncmp(argv[a], "rtol=", 5)) rtol = std::atof(argv[a] + 5);
  }
  LibMeshInit init(argc, argv);
  try {
    Mesh mesh(init.comm());
    MeshTools::Generation::build_cube(mesh, n, n, n, 0., 1., 0., 1., 0., 1., type);
    if (perturb) MeshTools::Modification::perturb_interior(mesh, 0x5EED0001, 0.2);
    EquationSystems es(mesh);
    FEMSystem &sys = es.add_system<FEMSystem>("Laplace");
    const Order order = (type == HEX8 || type == TET4) ? FIRST : SECOND;
    sys.add_variable("u", order, LAGRANGE);
    PeriodicX periodic_rows(sys, mesh, n);
    if (periodic) {
      sys.get_dof_map().add_dirichlet_boundary(DirichletBoundary({0, 1, 3, 5}, {0}));
      sys.get_dof_map().set_error_on_constraint_loop(true);
      sys.attach_constraint_object(periodic_rows);
    } else {
      sys.get_dof_map().add_dirichlet_boundary(DirichletBoundary({0, 1, 2, 3, 4, 5}, {0}));
    }
    auto matrix = std::make_unique<HipCsrMatrix<Number>>(init.comm());
    matrix->set_asymmetric_constraint_rows(bicgstab);
    sys.add_matrix("System Matrix", std::move(matrix));
    sys.linear_solver = std::make_unique<HipLinearSolver<Number>>(init.comm());
    if (bicgstab) sys.linear_solver->set_solver_type(BICGSTAB);
    const FEType fe_type(order, LAGRANGE);
    const QGauss qrule(3, fe_type.default_quadrature_order());
    HipAssembler physics(sys, mesh, sys.get_dof_map(), fe_type, qrule, sine ? HipAssembler::SINE : HipAssembler::UNIT);
    sys.attach_physics(&physics);
    es.init();
    if (kappa) physics.set_coefficient(MeshTools::lognormal_coefficients(mesh.n_elem(), 0x5EED0002));
    NewtonSolver &newton = sys.time_solver->diff_solver();
    newton.relative_residual_tolerance = rtol;
    newton.max_linear_iterations = 20000;
    const auto t0 = std::chrono::steady_clock::now();
    sys.solve();
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    std::printf("dofs %u newton %u linear %u result %u solve_ms %.3f\n", sys.get_dof_map().n_dofs(),
                newton.total_outer_iterations(), newton.total_inner_iterations(), newton.solve_result(), ms);
    std::vector<double> u;
    sys.solution->localize(u);
    std::ofstream f(prefix + "_u.bin", std::ios::binary);
    f.write((const char *)u.data(), sizeof(double) * u.size());
    if (adjoint) {
      // ImplicitSystem::adjoint_qoi_parameter_sensitivity through the
      // FEMSystem's Jacobian, J = integral of u over [0.25,0.75]^3
      BoxQoIDerivative qoi_derivative(sys, physics);
      sys.init_qois(1);
      sys.attach_QOI_derivative_object(qoi_derivative);
      if (physics.coefficient().empty()) physics.set_coefficient(std::vector<Real>(mesh.n_elem(), 1.0));
      const auto r = sys.adjoint_solve();
      sys.set_adjoint_already_solved(true);
      std::printf("adjoint iterations %u residual %.6e\n", r.first, r.second);
      ParameterVector params;
      for (auto &k : physics.coefficient()) params.push_back(std::make_unique<ParameterPointer<Number>>(&k));
      SensitivityData sens;
      sys.adjoint_qoi_parameter_sensitivity(QoISet(), params, sens);
      std::vector<double> g(params.size());
      for (unsigned int j = 0; j < params.size(); ++j) g[j] = sens[0][j];
      std::ofstream fg(prefix + "_grad.bin", std::ios::binary);
      fg.write((const char *)g.data(), sizeof(double) * g.size());
    }
  } catch (const std::exception &e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
