// fem_physics_dropin.hip -- a FEMSystem whose physics is user code run on the
// GPU: the steady reaction-diffusion -div(kappa grad u) + sigma u^3 = f (f = 1)
// written as an element_time_derivative, the way FEMSystem applications write
// theirs (adjoints_ex1/L-shaped.C:58-108: F(i) += JxW (...), K(i,j) += JxW (...),
// u and grad u from interior_value / interior_gradient), here as a device
// member that one wavefront runs per element.  Compiled with hipcc.
//
// Usage: fem_physics_dropin n elem_type(10|12|8|9) out_prefix mode [perturb] [kappa]
//                           [sigma=S] [u=FILE] [rtol=R] [extra_q=K] [elasticity]
//   elasticity  : the three-variable LinearElasticity physics below (E = 1,
//                 nu = 0.3, body force (0, 0, -1), clamped x = 0 face)
//   mode eval   : FEMSystem::assembly(true, true) at the solution read from FILE
//                 (n_dofs doubles; default 0): writes <prefix>_R.bin (residual)
//                 and <prefix>_K.bin (CSR values of the Jacobian, pattern order)
//   mode newton : SteadySolver / NewtonSolver from u = 0: writes <prefix>_u.bin
// Prints the DoF count, Newton and linear iteration counts and the times.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <string>
#include <vector>

#include "libmesh/device_physics_hip.h"
#include "libmesh/fem_system.h"
#include "libmesh/hip_csr_matrix.h"
#include "libmesh/hip_linear_solver.h"
#include "libmesh/libmesh.h"

using namespace libMesh;

// The user physics.  Every lane runs this code; lane l owns the test
// functions i = l, l + 64, ...: their residual entries F(i) and Jacobian rows
// K(i, .).
struct ReactionDiffusion {
  double sigma = 100.0;  // reaction coefficient
  double f = 1.0;        // source

  __device__ void element_time_derivative(bool request_jacobian, lmh_fem::Context &c) const {
    const int n = c.n_dofs(0), nq = c.n_qpoints();
    const double kappa = c.coefficient();
    for (int i = c.lane(); i < n; i += c.width()) {
      double Fi = 0;
      for (int qp = 0; qp < nq; ++qp) {
        const double u = c.interior_value(0, qp);
        const double grad_u_dphi = c.interior_gradient(0, qp, 0) * c.dphi(i, qp, 0) +
                                   c.interior_gradient(0, qp, 1) * c.dphi(i, qp, 1) +
                                   c.interior_gradient(0, qp, 2) * c.dphi(i, qp, 2);
        Fi += c.JxW(qp) * (kappa * grad_u_dphi + (sigma * u * u * u - f) * c.phi(i, qp));
      }
      c.get_elem_residual(0, i) += Fi;
      if (!request_jacobian) continue;
      for (int j = 0; j < n; ++j) {
        double Kij = 0;
        for (int qp = 0; qp < nq; ++qp) {
          const double u = c.interior_value(0, qp);
          const double dd = c.dphi(i, qp, 0) * c.dphi(j, qp, 0) + c.dphi(i, qp, 1) * c.dphi(j, qp, 1) +
                            c.dphi(i, qp, 2) * c.dphi(j, qp, 2);
          Kij += c.JxW(qp) * (kappa * dd + 3.0 * sigma * u * u * c.phi(j, qp) * c.phi(i, qp));
        }
        c.get_elem_jacobian(0, 0, i, j) += Kij;
      }
    }
  }
};

// A second user physics with three variables: linear elasticity
// (systems_of_equations_ex6.C's operator), E, nu and a body force (0, 0, fz),
// written as the residual of the stress, R_(i,a) = int sigma_ij(u) dphi_a/dx_j
// - f_i phi_a, and its Jacobian, the elasticity tensor
// C_ijkl = lam1 d_ij d_kl + lam2 (d_ik d_jl + d_il d_jk) between test and
// trial gradients.  Lanes own the (variable, node) test functions.
struct LinearElasticity {
  double E = 1.0, nu = 0.3, fz = -1.0;

  __device__ void element_time_derivative(bool request_jacobian, lmh_fem::Context &c) const {
    const int n = c.n_dofs(0), nq = c.n_qpoints();
    const double lam1 = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)), lam2 = E / (2.0 * (1.0 + nu));
    for (int t = c.lane(); t < 3 * n; t += c.width()) {
      const int i = t / n, a = t - (t / n) * n;
      double Fi = 0;
      for (int qp = 0; qp < nq; ++qp) {
        const double divu = c.interior_gradient(0, qp, 0) + c.interior_gradient(1, qp, 1) +
                            c.interior_gradient(2, qp, 2);
        double s = 0;
        for (int j = 0; j < 3; ++j) {
          const double sij = (i == j ? lam1 * divu : 0.0) +
                             lam2 * (c.interior_gradient(i, qp, j) + c.interior_gradient(j, qp, i));
          s += sij * c.dphi(a, qp, j);
        }
        Fi += c.JxW(qp) * (s - (i == 2 ? fz : 0.0) * c.phi(a, qp));
      }
      c.get_elem_residual(i, a) += Fi;
      if (!request_jacobian) continue;
      for (int k = 0; k < 3; ++k)
        for (int b = 0; b < n; ++b) {
          double Kab = 0;
          for (int qp = 0; qp < nq; ++qp) {
            double dd = 0;
            for (int j = 0; j < 3; ++j) dd += c.dphi(a, qp, j) * c.dphi(b, qp, j);
            Kab += c.JxW(qp) * (lam1 * c.dphi(a, qp, i) * c.dphi(b, qp, k) +
                                lam2 * ((i == k ? dd : 0.0) + c.dphi(a, qp, k) * c.dphi(b, qp, i)));
          }
          c.get_elem_jacobian(i, k, a, b) += Kab;
        }
    }
  }
};

static bool read_doubles(const std::string &path, std::vector<double> &v) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  f.read((char *)v.data(), sizeof(double) * v.size());
  return (size_t)f.gcount() == sizeof(double) * v.size();
}

static void write_doubles(const std::string &path, const std::vector<double> &v) {
  std::ofstream f(path, std::ios::binary);
  f.write((const char *)v.data(), sizeof(double) * v.size());
}

int main(int argc, char **argv) {
  if (argc < 5) {
    std::fprintf(stderr, "usage: %s n elem_type out_prefix eval|newton [perturb] [kappa] [sigma=S] [u=FILE]\n",
                 argv[0]);
    return 2;
  }
  const unsigned n = std::atoi(argv[1]);
  const ElemType type = (ElemType)std::atoi(argv[2]);
  const std::string prefix = argv[3], mode = argv[4];
  bool perturb = false, kappa = false, elasticity = false;
  double sigma = 100.0, rtol = 1e-10;
  int extra_q = 0;
  std::string u_file;
  for (int a = 5; a < argc; ++a) {
    perturb |= !std::strcmp(argv[a], "perturb");
    kappa |= !std::strcmp(argv[a], "kappa");
    elasticity |= !std::strcmp(argv[a], "elasticity");
    if (!std::strncmp(argv[a], "sigma=", 6)) sigma = std::atof(argv[a] + 6);
    if (!std::strncmp(argv[a], "u=", 2)) u_file = argv[a] + 2;
    if (!std::strncmp(argv[a], "rtol=", 5)) rtol = std::atof(argv[a] + 5);
    if (!std::strncmp(argv[a], "extra_q=", 8)) extra_q = std::atoi(argv[a] + 8);
  }
  if (mode != "eval" && mode != "newton") {
    std::fprintf(stderr, "mode must be eval or newton\n");
    return 2;
  }
  LibMeshInit init(argc, argv);
  try {
    Mesh mesh(init.comm());
    MeshTools::Generation::build_cube(mesh, n, n, n, 0., 1., 0., 1., 0., 1., type);
    if (perturb) MeshTools::Modification::perturb_interior(mesh, 0x5EED0001, 0.2);
    EquationSystems es(mesh);
    FEMSystem &sys = es.add_system<FEMSystem>(elasticity ? "Elasticity" : "ReactionDiffusion");
    const Order order = (type == HEX8 || type == TET4) ? FIRST : SECOND;
    if (elasticity) {  // displacement (u, v, w), clamped on the x = 0 face (boundary id 4)
      sys.add_variable("u", order, LAGRANGE);
      sys.add_variable("v", order, LAGRANGE);
      sys.add_variable("w", order, LAGRANGE);
      sys.get_dof_map().add_dirichlet_boundary(DirichletBoundary({4}, {0, 1, 2}));
    } else {
      sys.add_variable("u", order, LAGRANGE);
      sys.get_dof_map().add_dirichlet_boundary(DirichletBoundary({0, 1, 2, 3, 4, 5}, {0}));
    }
    sys.extra_quadrature_order = extra_q;
    sys.add_matrix("System Matrix", std::make_unique<HipCsrMatrix<Number>>(init.comm()));
    sys.linear_solver = std::make_unique<HipLinearSolver<Number>>(init.comm());
    ReactionDiffusion rd;
    rd.sigma = sigma;
    DevicePhysicsFor<ReactionDiffusion> rd_physics(rd);
    DevicePhysicsFor<LinearElasticity> el_physics;
    DevicePhysics &physics = elasticity ? (DevicePhysics &)el_physics : (DevicePhysics &)rd_physics;
    sys.attach_physics(&physics);
    es.init();
    if (kappa) physics.coefficient() = MeshTools::lognormal_coefficients(mesh.n_elem(), 0x5EED0002);
    const unsigned int nd = sys.get_dof_map().n_dofs();
    if (mode == "eval") {
      std::vector<double> u(nd, 0.0);
      if (!u_file.empty() && !read_doubles(u_file, u)) libmesh_error_msg("cannot read " + u_file);
      *sys.solution = u;
      sys.solution->close();
      sys.assembly(true, true);  // warm-up: device buffers, FE tables
      const auto t0 = std::chrono::steady_clock::now();
      sys.assembly(true, true);
      std::vector<double> R;
      sys.rhs->localize(R);
      const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
      auto &A = static_cast<HipCsrMatrix<Number> &>(sys.get_system_matrix());
      std::vector<double> vals(lmh_matrix_nnz(A.device_matrix()));
      hip_check(lmh_matrix_get_values(A.device_matrix(), vals.data()), "get_values");
      write_doubles(prefix + "_R.bin", R);
      write_doubles(prefix + "_K.bin", vals);
      std::printf("dofs %u nnz %zu assembly_ms %.3f\n", nd, vals.size(), ms);
    } else {
      NewtonSolver &newton = sys.time_solver->diff_solver();
      newton.relative_residual_tolerance = rtol;
      newton.max_linear_iterations = 20000;
      const auto t0 = std::chrono::steady_clock::now();
      sys.solve();
      const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
      std::printf("dofs %u newton %u linear %u result %u solve_ms %.3f\n", nd, newton.total_outer_iterations(),
                  newton.total_inner_iterations(), newton.solve_result(), ms);
      std::vector<double> u;
      sys.solution->localize(u);
      write_doubles(prefix + "_u.bin", u);
    }
  } catch (const std::exception &e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
