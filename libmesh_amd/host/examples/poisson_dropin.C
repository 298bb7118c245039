// poisson_dropin.C -- introduction_ex4-style Poisson solve written against the
// libMesh API, with the GPU backends injected at libMesh's own extension points:
//   sys.add_matrix("System Matrix", make_unique<HipCsrMatrix>)   (System::add_matrix)
//   sys.linear_solver = make_unique<HipLinearSolver>             (ImplicitSystem::linear_solver)
//   sys.attach_assemble_object(HipAssembler(...))                (System::attach_assemble_object)
// Usage: poisson_dropin n elem_type(10|12|8|9) out_prefix [perturb] [kappa] [sine] [mixed] [adjoint] [distort]
//                      [periodic] [bicgstab] [identity]
// Writes <prefix>_u.bin (solution) and, with `adjoint`, <prefix>_grad.bin.
// `periodic`: Dirichlet on the y and z faces only, u(x=1) = u(x=0) through
// constraint rows from a System::Constraint object (plus one heterogeneous
// row u = 0.3 on an x=0 face DoF), symmetric constraint rows, and
// enforce_constraints_exactly after the solve.  `periodic bicgstab`: the
// default asymmetric constraint rows instead, solved by Jacobi-BiCGSTAB
// (LinearSolver::set_solver_type(BICGSTAB)).  `identity`: IDENTITY_PRECOND
// instead of JACOBI_PRECOND (LinearSolver::set_preconditioner_type).
// `condense`: System::create_static_condensation before init, as
// miscellaneous_ex16 does: the system matrix is a StaticCondensation and the
// solve runs on the reduced system (its size is printed).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <string>

#include "libmesh/hip_assembler.h"
#include "libmesh/hip_csr_matrix.h"
#include "libmesh/hip_linear_solver.h"
#include "libmesh/libmesh.h"
#include "libmesh/static_condensation.h"
#include "libmesh/system.h"
#include "box_qoi.h"
#include "periodic_x.h"

using namespace libMesh;

static void write_bin(const std::string &path, const std::vector<double> &v) {
  std::ofstream f(path, std::ios::binary);
  f.write((const char *)v.data(), sizeof(double) * v.size());
}

int main(int argc, char **argv) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: %s n elem_type out_prefix [perturb] [kappa] [sine] [mixed] [adjoint]\n", argv[0]);
    return 2;
  }
  const unsigned n = std::atoi(argv[1]);
  const ElemType type = (ElemType)std::atoi(argv[2]);
  const std::string prefix = argv[3];
  bool perturb = false, kappa = false, sine = false, mixed = false, adjoint = false, distort = false,
       periodic = false, bicgstab = false, identity = false, condense = false;
  for (int a = 4; a < argc; ++a) {
    perturb |= !std::strcmp(argv[a], "perturb");
    kappa |= !std::strcmp(argv[a], "kappa");
    sine |= !std::strcmp(argv[a], "sine");
    mixed |= !std::strcmp(argv[a], "mixed");
    adjoint |= !std::strcmp(argv[a], "adjoint");
    distort |= !std::strcmp(argv[a], "distort");
    periodic |= !std::strcmp(argv[a], "periodic");
    bicgstab |= !std::strcmp(argv[a], "bicgstab");
    identity |= !std::strcmp(argv[a], "identity");
    condense |= !std::strcmp(argv[a], "condense");
  }
  LibMeshInit init(argc, argv);
  try {
    Mesh mesh(init.comm());
    MeshTools::Generation::build_cube(mesh, n, n, n, 0., 1., 0., 1., 0., 1., type);
    if (perturb) MeshTools::Modification::perturb_interior(mesh, 0x5EED0001, 0.2);
    if (distort) {
      // push the centre node of element 0 through its +x face: det J < 0
      const dof_id_type c = mesh.node_id(0, 26);
      Point p = mesh.point(c);
      p.x += 1.5 / n;
      mesh.set_point(c, p);
    }

    EquationSystems es(mesh);
    LinearImplicitSystem &sys = es.add_system<LinearImplicitSystem>("Poisson");
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
    if (condense) {
      sys.create_static_condensation();  // the system matrix becomes a StaticCondensation
    } else {
      auto matrix = std::make_unique<HipCsrMatrix<Number>>(init.comm());
      matrix->set_asymmetric_constraint_rows(bicgstab);
      sys.add_matrix("System Matrix", std::move(matrix));
    }
    auto solver = std::make_unique<HipLinearSolver<Number>>(init.comm());
    solver->set_mixed_precision(mixed);
    if (bicgstab) solver->set_solver_type(BICGSTAB);
    if (identity) solver->set_preconditioner_type(IDENTITY_PRECOND);
    sys.linear_solver = std::move(solver);
    es.linear_solver_tolerance = 1e-10;
    es.linear_solver_maximum_iterations = 20000;

    const FEType fe_type(order, LAGRANGE);
    const QGauss qrule(3, fe_type.default_quadrature_order());
    HipAssembler assembler(sys, mesh, sys.get_dof_map(), fe_type, qrule, sine ? HipAssembler::SINE : HipAssembler::UNIT);
    sys.attach_assemble_object(assembler);
    es.init();
    if (kappa) assembler.set_coefficient(MeshTools::lognormal_coefficients(mesh.n_elem(), 0x5EED0002));
    if (condense && mixed)  // the reduced system has its own solver (the "<system>_condensed_" one)
      static_cast<HipLinearSolver<Number> &>(sys.get_static_condensation().reduced_system_solver())
          .set_mixed_precision(true);

    sys.solve();
    if (periodic && !bicgstab) sys.get_dof_map().enforce_constraints_exactly(sys);
    if (condense)
      std::printf("Static condensation reduced problem size to %u\n",
                  (unsigned)sys.get_static_condensation().get_condensed_mat().m());
    auto *lin = static_cast<HipLinearSolver<Number> *>(sys.linear_solver.get());
    std::printf("dofs %u iterations %u residual %.6e reason %d\n", sys.get_dof_map().n_dofs(),
                sys.n_linear_iterations(), sys.final_linear_residual(), (int)lin->get_converged_reason());
    std::vector<double> u;
    sys.solution->localize(u);
    write_bin(prefix + "_u.bin", u);

    if (adjoint) {
      // ImplicitSystem::adjoint_solve and adjoint_qoi_parameter_sensitivity
      // for every per-element kappa, J = integral of u over [0.25,0.75]^3
      BoxQoIDerivative qoi_derivative(sys, assembler);
      sys.init_qois(1);
      sys.attach_QOI_derivative_object(qoi_derivative);
      if (assembler.coefficient().empty()) assembler.set_coefficient(std::vector<Real>(mesh.n_elem(), 1.0));
      const auto r = sys.adjoint_solve();
      sys.set_adjoint_already_solved(true);
      std::printf("adjoint iterations %u residual %.6e\n", r.first, r.second);
      ParameterVector params;
      for (auto &k : assembler.coefficient()) params.push_back(std::make_unique<ParameterPointer<Number>>(&k));
      SensitivityData sens;
      sys.adjoint_qoi_parameter_sensitivity(QoISet(), params, sens);
      std::vector<double> g(params.size());
      for (unsigned int j = 0; j < params.size(); ++j) g[j] = sens[0][j];
      write_bin(prefix + "_grad.bin", g);
      std::vector<double> zh;
      sys.get_adjoint_solution(0).localize(zh);
      write_bin(prefix + "_z.bin", zh);
    }
  } catch (const DegenerateMap &e) {
    std::fprintf(stderr, "DegenerateMap: %s\n", e.what());
    return 3;
  } catch (const std::exception &e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
