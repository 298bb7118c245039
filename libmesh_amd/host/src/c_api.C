// c_api.C -- flat C entry points over the host mesh/DoF classes (lmesh_host.h).
#include <cstring>
#include <memory>
#include <string>

#include "../../../include/lmesh_host.h"
#include <fstream>

#include "libmesh/dof_map.h"
#include "libmesh/equation_systems_io.h"
#include "libmesh/matrix_io.h"
#include "libmesh/replicated_mesh.h"

using namespace libMesh;

// one serial communicator for the meshes of this API
static const Parallel::Communicator &serial_comm() {
  static const Parallel::Communicator comm;
  return comm;
}

struct lmesh_problem {
  ReplicatedMesh mesh{serial_comm()};
  DofMap dof_map;
};

static thread_local std::string g_lmesh_err;

extern "C" {

const char *lmesh_last_error(void) { return g_lmesh_err.c_str(); }

lmesh_problem *lmesh_build_cube(int nx, int ny, int nz, int elem_type, int n_vars, unsigned mask) {
  try {
    if (n_vars < 1 || n_vars > 3) libmesh_error_msg("n_vars must be 1..3");
    std::unique_ptr<lmesh_problem> p(new lmesh_problem);
    if (nx <= 0 || ny <= 0 || nz <= 0) libmesh_error_msg("mesh sizes must be positive");
    MeshTools::Generation::build_cube(p->mesh, nx, ny, nz, 0., 1., 0., 1., 0., 1., (ElemType)elem_type);
    const Order o = (elem_type == HEX8 || elem_type == TET4) ? FIRST : SECOND;
    const char *names[3] = {"u", "v", "w"};
    std::vector<unsigned> vars;
    for (int v = 0; v < n_vars; ++v) vars.push_back(p->dof_map.add_variable(names[v], FEType(o, LAGRANGE)));
    std::set<boundary_id_type> ids;
    for (int b = 0; b < 6; ++b)
      if (mask & (1u << b)) ids.insert((boundary_id_type)b);
    if (!ids.empty()) p->dof_map.add_dirichlet_boundary(DirichletBoundary(ids, vars));
    p->dof_map.distribute_dofs(p->mesh);
    return p.release();
  } catch (const std::exception &e) {
    g_lmesh_err = e.what();
    return nullptr;
  }
}

void lmesh_free(lmesh_problem *p) { delete p; }

void lmesh_perturb(lmesh_problem *p, uint64_t seed, double frac) {
  MeshTools::Modification::perturb_interior(p->mesh, seed, frac);
}

void lmesh_lognormal(int64_t n, uint64_t seed, double *out) {
  const std::vector<Real> k = MeshTools::lognormal_coefficients((dof_id_type)n, seed);
  std::memcpy(out, k.data(), sizeof(double) * k.size());
}

void lmesh_batch_recipe(uint64_t seed, int count, int nmin, int nmax, int *edge, int *elem_type) {
  std::vector<int> e;
  std::vector<ElemType> t;
  MeshTools::batch_recipe(seed, count, nmin, nmax, e, t);
  for (int p = 0; p < count; ++p) {
    edge[p] = e[p];
    elem_type[p] = (int)t[p];
  }
}

int lmesh_matrix_read(const char *filename, int64_t *m, int64_t *n, int64_t *nnz, int64_t *row, int64_t *col,
                      double *val) {
  try {
    const MatrixIO::Entries E = MatrixIO::read(filename);
    *m = (int64_t)E.m;
    *n = (int64_t)E.n;
    if (row && col && val) {
      if (*nnz < (int64_t)E.ijv.size()) libmesh_error_msg("lmesh_matrix_read: output arrays too small");
      for (std::size_t k = 0; k < E.ijv.size(); ++k) {
        row[k] = (int64_t)std::get<0>(E.ijv[k]);
        col[k] = (int64_t)std::get<1>(E.ijv[k]);
        val[k] = std::get<2>(E.ijv[k]);
      }
    }
    *nnz = (int64_t)E.ijv.size();
    return 0;
  } catch (const std::exception &e) {
    g_lmesh_err = e.what();
    return -1;
  }
}

int lmesh_matrix_write_matlab(const char *filename, int64_t m, int64_t n, const int32_t *row_ptr,
                              const int32_t *col_idx, const double *vals) {
  try {
    std::vector<int32_t> rp(row_ptr, row_ptr + m + 1), ci(col_idx, col_idx + row_ptr[m]);
    std::vector<double> v(vals, vals + row_ptr[m]);
    std::ofstream os(filename);
    if (!os.good()) libmesh_error_msg("cannot write " << filename);
    MatrixIO::write_matlab(os, (std::size_t)m, (std::size_t)n, rp, ci, v);
    return 0;
  } catch (const std::exception &e) {
    g_lmesh_err = e.what();
    return -1;
  }
}

int lmesh_es_file_convert(const char *in_name, const char *out_name) {
  try {
    XdaFile f;
    {
      Xdr in(in_name, xdr_mode_for(in_name, false));
      read_equation_systems_file(in, f, true);
      in.close();
    }
    Xdr out(out_name, xdr_mode_for(out_name, true));
    write_equation_systems_file(out, f, true);
    out.close();
    return 0;
  } catch (const std::exception &e) {
    g_lmesh_err = e.what();
    return -1;
  }
}

int64_t lmesh_n_nodes(const lmesh_problem *p) { return p->mesh.n_nodes(); }
int64_t lmesh_n_elem(const lmesh_problem *p) { return p->mesh.n_elem(); }
int64_t lmesh_n_dofs(const lmesh_problem *p) { return p->dof_map.n_dofs(); }
int lmesh_nodes_per_elem(const lmesh_problem *p) { return (int)p->mesh.nodes_per_elem(); }
int lmesh_dofs_per_elem(const lmesh_problem *p) { return (int)p->dof_map.dofs_per_elem(); }
int lmesh_elem_type(const lmesh_problem *p) { return (int)p->mesh.elem_type(); }
void lmesh_coords(const lmesh_problem *p, double *xyz) {
  std::memcpy(xyz, p->mesh.coords().data(), sizeof(double) * p->mesh.coords().size());
}
void lmesh_conn(const lmesh_problem *p, int32_t *conn) {
  std::memcpy(conn, p->mesh.connectivity().data(), sizeof(int32_t) * p->mesh.connectivity().size());
}
void lmesh_elem_dofs(const lmesh_problem *p, int32_t *dofs) {
  std::memcpy(dofs, p->dof_map.elem_dofs().data(), sizeof(int32_t) * p->dof_map.elem_dofs().size());
}
void lmesh_constrained(const lmesh_problem *p, uint8_t *c) {
  std::memcpy(c, p->dof_map.constrained().data(), p->dof_map.constrained().size());
}

}  // extern "C"
