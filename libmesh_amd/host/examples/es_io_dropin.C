// es_io_dropin.C -- EquationSystems::write / read as a libMesh application
// uses them for restart files: two systems on a perturbed HEX27 cube (a
// LinearImplicitSystem "Poisson" with u, a basic System "aux" with a and b),
// solutions set from the node coordinates, and an additional vector "u_old" =
// 2u + 1 of the Poisson system, written as <prefix>.xda and <prefix>.xdr with
// WRITE_ADDITIONAL_DATA, read back with READ_ADDITIONAL_DATA into fresh
// EquationSystems and compared.
// Usage: es_io_dropin n out_prefix
// Prints: nodes, DoFs, the largest difference after each round trip, and
// whether a header that does not describe the systems is rejected.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "libmesh/libmesh.h"
#include "libmesh/system.h"

using namespace libMesh;

namespace {

// u = 1 + x + 2y + 3z; a = xy - z; b = x^2 - yz + 1/8 (exact polynomials, so
// a reader of the file can recompute them from the coordinates)
Real f_u(const Point &p) { return 1. + p.x + 2. * p.y + 3. * p.z; }
Real f_a(const Point &p) { return p.x * p.y - p.z; }
Real f_b(const Point &p) { return p.x * p.x - p.y * p.z + 0.125; }

void declare(EquationSystems &es, const char *aux_var_b) {
  LinearImplicitSystem &p = es.add_system<LinearImplicitSystem>("Poisson");
  p.add_variable("u", SECOND, LAGRANGE);
  p.add_vector("u_old", false);
  System &aux = es.add_system<System>("aux");
  aux.add_variable("a", SECOND, LAGRANGE);
  aux.add_variable(aux_var_b, SECOND, LAGRANGE);
  es.init();
}

void fill(EquationSystems &es, const Mesh &mesh) {
  System &p = es.get_system<System>("Poisson"), &aux = es.get_system<System>("aux");
  const auto &dp = p.get_dof_map(), &da = aux.get_dof_map();
  std::vector<dof_id_type> du, dab;
  for (dof_id_type e = 0; e < mesh.n_elem(); ++e) {
    const Elem &elem = mesh.elem_ref(e);
    dp.dof_indices(&elem, du);
    da.dof_indices(&elem, dab);
    const unsigned npe = elem.n_nodes();
    for (unsigned k = 0; k < npe; ++k) {
      const Point &x = elem.point(k);
      p.solution->set(du[k], f_u(x));
      p.get_vector("u_old").set(du[k], 2. * f_u(x) + 1.);
      aux.solution->set(dab[k], f_a(x));
      aux.solution->set(dab[npe + k], f_b(x));
    }
  }
  p.solution->close();
  p.get_vector("u_old").close();
  aux.solution->close();
}

double max_diff(EquationSystems &a, EquationSystems &b) {
  double d = 0;
  for (const char *name : {"Poisson", "aux"}) {
    std::vector<Number> x, y;
    a.get_system<System>(name).solution->localize(x);
    b.get_system<System>(name).solution->localize(y);
    if (x.size() != y.size()) return INFINITY;
    for (size_t i = 0; i < x.size(); ++i) d = std::fmax(d, std::fabs(x[i] - y[i]));
  }
  std::vector<Number> x, y;
  a.get_system<System>("Poisson").get_vector("u_old").localize(x);
  b.get_system<System>("Poisson").get_vector("u_old").localize(y);
  if (x.size() != y.size()) return INFINITY;
  for (size_t i = 0; i < x.size(); ++i) d = std::fmax(d, std::fabs(x[i] - y[i]));
  return d;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: %s n out_prefix\n", argv[0]);
    return 2;
  }
  const unsigned n = std::atoi(argv[1]);
  const std::string prefix = argv[2];
  LibMeshInit init(argc, argv);
  try {
    Mesh mesh(init.comm());
    MeshTools::Generation::build_cube(mesh, n, n, n, 0., 1., 0., 1., 0., 1., HEX27);
    MeshTools::Modification::perturb_interior(mesh, 0x5EED0001, 0.2);
    EquationSystems es(mesh);
    declare(es, "b");
    fill(es, mesh);
    const unsigned wflags = EquationSystems::WRITE_DATA | EquationSystems::WRITE_ADDITIONAL_DATA;
    es.write(prefix + ".xda", wflags);
    es.write(prefix + ".xdr", wflags);
    EquationSystems from_xdr(mesh), from_xda(mesh);
    declare(from_xdr, "b");
    declare(from_xda, "b");
    const unsigned rflags =
        EquationSystems::READ_HEADER | EquationSystems::READ_DATA | EquationSystems::READ_ADDITIONAL_DATA;
    from_xdr.read(prefix + ".xdr", rflags);
    from_xda.read(prefix + ".xda", rflags);
    int rejected = 0;
    try {
      EquationSystems other(mesh);
      declare(other, "c");
      other.read(prefix + ".xda");
    } catch (const std::exception &e) {
      rejected = 1;
      std::printf("mismatch: %s\n", e.what());
    }
    std::printf("nodes %u poisson_dofs %u aux_dofs %u xdr_maxdiff %.3e xda_maxdiff %.3e mismatch_rejected %d\n",
                mesh.n_nodes(), es.get_system<System>("Poisson").get_dof_map().n_dofs(),
                es.get_system<System>("aux").get_dof_map().n_dofs(), max_diff(es, from_xdr), max_diff(es, from_xda),
                rejected);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
