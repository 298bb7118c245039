// sparse_matrix_contract.C -- the SparseMatrix contract checks of the
// reference's tests/numerics/sparse_matrix_test.h (setValues/testValues with
// (i+1)(j+1), testTransposeNorms l1(A) = linf(A^T), testClone, get_diagonal,
// add(a, X)) run on HipCsrMatrix, plus add_matrix on a DofMap-built pattern,
// and, given the directory of the reference's matrix fixtures as argv[1],
// testReadMatlab (the .m and .m.gz files give the same norms), the PETSc
// binary reads and a print_matlab/read round trip on rectangular matrices.
// Prints one line per check and exits non-zero on any failure.
#include <cmath>
#include <cstdio>
#include <numeric>
#include <string>

#include "libmesh/hip_csr_matrix.h"
#include "libmesh/libmesh.h"

using namespace libMesh;

static int failures = 0;
// the world communicator of LibMeshInit, for the helpers below
static const Parallel::Communicator *g_comm = nullptr;
static void check(bool ok, const char *what) {
  std::printf("%s %s\n", ok ? "PASS" : "FAIL", what);
  if (!ok) ++failures;
}

static bool same_entries(const SparseMatrix<Number> &A, const SparseMatrix<Number> &B, bool skip_zeros,
                         double rtol = 0.0) {
  if (A.m() != B.m() || A.n() != B.n()) return false;
  double amax = 0;
  for (numeric_index_type i = 0; i < A.m(); ++i) {
    std::vector<numeric_index_type> ia;
    std::vector<Number> va;
    A.get_row(i, ia, va);
    for (Number v : va) amax = std::max(amax, std::abs(v));
  }
  for (numeric_index_type i = 0; i < A.m(); ++i) {
    std::vector<numeric_index_type> ia, ib;
    std::vector<Number> va, vb;
    A.get_row(i, ia, va);
    B.get_row(i, ib, vb);
    std::vector<std::pair<numeric_index_type, Number>> a, b;
    for (size_t k = 0; k < ia.size(); ++k)
      if (!skip_zeros || va[k] != 0.0) a.emplace_back(ia[k], va[k]);
    for (size_t k = 0; k < ib.size(); ++k)
      if (!skip_zeros || vb[k] != 0.0) b.emplace_back(ib[k], vb[k]);
    if (a.size() != b.size()) return false;
    for (size_t k = 0; k < a.size(); ++k)  // normwise: relative to the largest entry
      if (a[k].first != b[k].first || std::abs(a[k].second - b[k].second) > rtol * amax) return false;
  }
  return true;
}

static void matrix_files(const std::string &dir) {
  for (const char *g : {"1", "2", "4", "8"}) {
    const std::string base = dir + "/geom_" + g + "_extraction_op.m";
    HipCsrMatrix<Number> A(*g_comm), B(*g_comm);
    A.read(base);
    char what[160];
    if (std::string(g) != "8") {
      B.read(base + ".gz");
      std::snprintf(what, sizeof(what), "testReadMatlab geom_%s: .m and .m.gz norms equal (%zu x %zu)", g,
                    (size_t)A.m(), (size_t)A.n());
      check(A.l1_norm() == B.l1_norm() && A.linfty_norm() == B.linfty_norm() && same_entries(A, B, false), what);
    }
    // print_matlab then read: the same nonzeros, exactly
    const std::string out = std::string("roundtrip_") + g + ".m";
    A.print_matlab(out);
    HipCsrMatrix<Number> C(*g_comm);
    C.read_matlab(out);
    std::snprintf(what, sizeof(what), "print_matlab/read_matlab round trip geom_%s", g);
    check(same_entries(A, C, true), what);
    // rectangular SpMV on the device: y = A 1 equals the row sums
    HipVector<Number> x(*g_comm, A.n()), y(*g_comm, A.m());
    for (numeric_index_type j = 0; j < A.n(); ++j) x.set(j, 1.0);
    x.close();
    A.vector_mult(y, x);
    bool ok = true;
    for (numeric_index_type i = 0; i < A.m(); ++i) {
      std::vector<numeric_index_type> idx;
      std::vector<Number> val;
      A.get_row(i, idx, val);
      double s = 0;
      for (Number v : val) s += v;
      ok &= std::abs(y(i) - s) <= 1e-14 * std::max(1.0, std::abs(s));
    }
    std::snprintf(what, sizeof(what), "vector_mult on the %zu x %zu matrix geom_%s", (size_t)A.m(), (size_t)A.n(), g);
    check(ok, what);
  }
  HipCsrMatrix<Number> M(*g_comm), P32(*g_comm), P64(*g_comm);
  M.read(dir + "/geom_1_extraction_op.m");
  P32.read(dir + "/geom_1_extraction_op.petsc32");
  P64.read(dir + "/geom_1_extraction_op.petsc64");
  // the reference's binary and text files differ in the last bit of some entries
  check(same_entries(M, P32, true, 2.3e-16) && same_entries(M, P64, true, 2.3e-16) && same_entries(P32, P64, true),
        "PETSc binary (32/64-bit) equals the Matlab file");
  bool threw = false;
  try {
    HipCsrMatrix<Number> H(*g_comm);
    H.read(dir + "/geom_1_extraction_op.h5");
  } catch (const LogicError &) {
    threw = true;
  }
  check(threw, "unsupported extension throws LogicError");
}

int main(int argc, char **argv) {
  LibMeshInit init(argc, argv);
  g_comm = &init.comm();
  try {
    // one rank: global_m = global_n = 1 * 5.75 - 1 = 4, locally dense block
    const numeric_index_type m = 4;
    HipCsrMatrix<Number> A(*g_comm);
    A.init(m, m, m, m, m, 5);
    std::vector<numeric_index_type> rows(m), cols(m);
    std::iota(rows.begin(), rows.end(), A.row_start());
    std::iota(cols.begin(), cols.end(), A.col_start());
    DenseMatrix<Number> local(m, m);
    for (unsigned i = 0; i < m; ++i)
      for (unsigned j = 0; j < m; ++j) local(i, j) = (i + 1.0) * (j + 1.0);
    A.zero();
    A.add_matrix(local, rows, cols);
    A.close();

    bool ok = true;
    for (numeric_index_type i = A.row_start(); i < A.row_stop(); ++i) {
      std::vector<numeric_index_type> idx;
      std::vector<Number> val;
      A.get_row(i, idx, val);
      for (size_t j = 0; j < idx.size(); ++j)
        ok &= idx[j] == A.col_start() + j && std::abs(val[j] - (i + 1.0) * (j + 1.0)) < 1e-12;
      ok &= idx.size() == m;
    }
    check(ok, "testValues: get_row returns (i+1)(j+1)");
    check(A(2, 3) == 12.0 && A.m() == m && A.n() == m, "operator() and sizes");

    HipCsrMatrix<Number> At(*g_comm);
    A.get_transpose(At);
    check(std::abs(A.l1_norm() - At.linfty_norm()) < 1e-12 && std::abs(A.linfty_norm() - At.l1_norm()) < 1e-12,
          "testTransposeNorms: l1(A) = linf(A^T)");
    check(std::abs(A.l1_norm() - 40.0) < 1e-12, "l1 norm closed form (max column sum 4*10)");

    auto C = A.clone();
    bool cl = true;
    for (numeric_index_type i = 0; i < m; ++i)
      for (numeric_index_type j = 0; j < m; ++j) cl &= (*C)(i, j) == A(i, j);
    check(cl, "testClone: clone equals original");
    auto Z = A.zero_clone();
    check(Z->l1_norm() == 0.0 && Z->n_nonzeros() == A.n_nonzeros(), "zero_clone keeps the pattern, zero values");

    A.add(2.0, *C);
    check(A(3, 3) == 48.0, "add(a, X): A + 2A");

    HipVector<Number> d(*g_comm, m);
    A.get_diagonal(d);
    check(d(0) == 3.0 && d(3) == 48.0, "get_diagonal");

    HipVector<Number> x(*g_comm, m), y(*g_comm, m);
    for (numeric_index_type i = 0; i < m; ++i) x.set(i, 1.0);
    x.close();
    A.vector_mult(y, x);
    check(std::abs(y(0) - 3.0 * 10.0) < 1e-12, "vector_mult on the device");

    // DofMap-built pattern: element matrices in call order
    Mesh mesh(*g_comm);
    MeshTools::Generation::build_cube(mesh, 3, 3, 3, 0., 1., 0., 1., 0., 1., HEX8);
    DofMap dm;
    dm.add_variable("u", FEType(FIRST, LAGRANGE));
    dm.distribute_dofs(mesh);
    HipCsrMatrix<Number> B(*g_comm);
    B.attach_dof_map(dm);
    check(B.need_full_sparsity_pattern(), "need_full_sparsity_pattern without a mesh");
    B.attach_mesh(mesh);
    check(!B.need_full_sparsity_pattern(), "the GPU builds the pattern once a mesh is attached");
    B.init();
    check(B.n_nonzeros() == 10 * 10 * 10, "DofMap pattern nnz = (3n+1)^3");
    B.zero();
    std::vector<dof_id_type> di;
    DenseMatrix<Number> Ke(8, 8);
    for (unsigned a = 0; a < 8; ++a)
      for (unsigned b = 0; b < 8; ++b) Ke(a, b) = (a == b) ? 1.0 : 0.0;
    for (dof_id_type e = 0; e < mesh.n_elem(); ++e) {
      dm.dof_indices(&mesh.elem_ref(e), di);
      std::vector<numeric_index_type> dd(di.begin(), di.end());
      B.add_matrix(Ke, dd);
    }
    B.close();
    // libMesh's graph path: SparsityPattern::Graph handed to
    // update_sparsity_pattern (as DofMap::update_sparsity_pattern does for
    // matrices that need the full pattern) gives the same matrix
    SparsityPattern::Graph graph;
    graph.resize(B.m());
    for (numeric_index_type i = 0; i < B.m(); ++i) {
      std::vector<numeric_index_type> idx;
      std::vector<Number> val;
      B.get_row(i, idx, val);
      graph[i].assign(idx.begin(), idx.end());
    }
    HipCsrMatrix<Number> G(*g_comm);
    G.attach_dof_map(dm);
    G.update_sparsity_pattern(graph);
    G.init();
    check(G.n_nonzeros() == B.n_nonzeros(), "update_sparsity_pattern: the graph's nnz");
    for (dof_id_type e = 0; e < mesh.n_elem(); ++e) {
      dm.dof_indices(&mesh.elem_ref(e), di);
      std::vector<numeric_index_type> dd(di.begin(), di.end());
      G.add_matrix(Ke, dd);
    }
    G.close();
    check(G.l1_norm_diff(B) == 0.0, "update_sparsity_pattern: element matrices sum as on the GPU pattern");
    // SparseMatrix::vector_mult / vector_mult_add through NumericVector::add_vector(v, A)
    HipVector<Number> ones(*g_comm, B.n()), Bx(*g_comm, B.m());
    ones = 1.0;
    B.vector_mult(Bx, ones);
    B.vector_mult_add(Bx, ones);
    check(Bx(0) == 2.0, "vector_mult then vector_mult_add: twice the row sum");
    // scale on the device, zero_rows on the host mirror
    G.scale(3.0);
    check(G(0, 0) == 3.0, "scale");
    std::vector<numeric_index_type> zr{0};
    G.zero_rows(zr, 5.0);
    check(G(0, 0) == 5.0 && G.linfty_norm() == 24.0, "zero_rows with a diagonal value");
    // PETSc binary written and read back
    G.print_petsc_binary("g_roundtrip.petsc32");
    HipCsrMatrix<Number> R(*g_comm);
    R.read_petsc_binary("g_roundtrip.petsc32");
    check(R.l1_norm_diff(G) == 0.0 && R.n_nonzeros() == G.n_nonzeros(), "print_petsc_binary/read_petsc_binary round trip");
    // diagonal = number of elements touching each node: corner 1, interior 8
    check(B(0, 0) == 1.0, "add_matrix batch: corner diagonal 1");
    double mx = 0;
    for (numeric_index_type i = 0; i < B.m(); ++i) mx = std::max(mx, (double)B(i, i));
    check(mx == 8.0, "add_matrix batch: interior diagonal 8");
    bool threw = false;
    try {
      B.set(0, B.m() - 1, 1.0);
    } catch (const LogicError &) {
      threw = true;
    }
    check(threw, "set outside the pattern throws LogicError");

    if (argc > 1) matrix_files(argv[1]);
  } catch (const std::exception &e) {
    std::printf("FAIL exception: %s\n", e.what());
    return 1;
  }
  std::printf("%d failures\n", failures);
  return failures ? 1 : 0;
}
