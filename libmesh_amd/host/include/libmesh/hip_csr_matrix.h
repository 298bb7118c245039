// hip_csr_matrix.h -- HipCsrMatrix<T>: a SparseMatrix backend whose values
// live in device memory.  Its CSR pattern comes from one of:
//  * the GPU, from the attached DofMap and the mesh given to attach_mesh (the
//    rule of SparsityPattern::Build, src/base/sparsity_pattern.C:96-365),
//    read through libMesh's MeshBase / Elem / DofMap accessors; this is the
//    pattern device assembly (HipAssembler) writes into;
//  * libMesh's own graph: without a mesh the matrix needs the full sparsity
//    pattern and takes the one DofMap hands to update_sparsity_pattern;
//  * the entries inserted before the first close() (init(m, n, ...) mode).
// Host-side set/add/add_matrix calls are staged and applied at close() in
// call order; element matrices go through the batched deterministic device
// scatter.
#pragma once

#include <map>
#include <mutex>

#include "libmesh/dof_map.h"
#include "libmesh/sparse_matrix.h"

namespace libMesh {

template <typename T>
class HipCsrMatrix : public SparseMatrix<T> {
 public:
  explicit HipCsrMatrix(const Parallel::Communicator &comm, int device = 0);
  ~HipCsrMatrix() override;
  HipCsrMatrix(const HipCsrMatrix &) = delete;
  HipCsrMatrix &operator=(const HipCsrMatrix &) = delete;

  SolverPackage solver_package() override { return HIP_SOLVERS; }
  // Without a mesh the pattern must come from DofMap's graph.
  bool need_full_sparsity_pattern() const override { return _mesh_src == nullptr; }
  // SparseMatrix::update_sparsity_pattern (sparse_matrix.h:175): the rows of
  // libMesh's graph, used by init() when no mesh is attached
  void update_sparsity_pattern(const SparsityPattern::Graph &sparsity_pattern) override;
  // The mesh whose elements the attached DofMap numbers: init() then builds
  // the pattern, the element scatter maps and the processed constraint rows
  // on the GPU.  Not part of libMesh's SparseMatrix: the drop-in System and
  // HipAssembler call it.
  void attach_mesh(const MeshBase &mesh) { _mesh_src = &mesh; }
  // The asymmetric_constraint_rows argument that libMesh assembly passes to
  // heterogeneously_constrain_element_matrix_and_vector
  // (dof_map_constraints.C:2518-2616), here fixed per matrix before init():
  // true (default) keeps the constraint equations in the constrained rows;
  // false keeps only their diagonals, so the system stays SPD for CG (call
  // DofMap::enforce_constraints_exactly after the solve).
  void set_asymmetric_constraint_rows(bool asymmetric) { _asymmetric = asymmetric; }
  bool asymmetric_constraint_rows() const { return _asymmetric; }

  void init(numeric_index_type m, numeric_index_type n, numeric_index_type m_l, numeric_index_type n_l,
            numeric_index_type nnz = 30, numeric_index_type noz = 10, numeric_index_type blocksize = 1) override;
  void init(ParallelType type = PARALLEL) override;
  void clear() override;
  void zero() override;
  std::unique_ptr<SparseMatrix<T>> zero_clone() const override;
  std::unique_ptr<SparseMatrix<T>> clone() const override;
  void close() override;
  numeric_index_type m() const override { return _rows; }
  numeric_index_type n() const override { return _cols; }
  numeric_index_type row_start() const override { return 0; }
  numeric_index_type row_stop() const override { return _rows; }
  numeric_index_type col_start() const override { return 0; }
  numeric_index_type col_stop() const override { return _cols; }
  void set(numeric_index_type i, numeric_index_type j, T value) override;
  void add(numeric_index_type i, numeric_index_type j, T value) override;
  void add_matrix(const DenseMatrix<T> &dm, const std::vector<numeric_index_type> &rows,
                  const std::vector<numeric_index_type> &cols) override;
  void add_matrix(const DenseMatrix<T> &dm, const std::vector<numeric_index_type> &dof_indices) override;
  void add(T a, const SparseMatrix<T> &X) override;
  T operator()(numeric_index_type i, numeric_index_type j) const override;
  Real l1_norm() const override;
  Real linfty_norm() const override;
  bool closed() const override { return _closed; }
  void print_personal(std::ostream &os = std::cout) const override;
  void get_diagonal(NumericVector<T> &dest) const override;
  void get_transpose(SparseMatrix<T> &dest) const override;
  void get_row(numeric_index_type i, std::vector<numeric_index_type> &indices, std::vector<T> &values) const override;
  void zero_rows(std::vector<numeric_index_type> &rows, T diag_value = 0.0) override;
  void scale(const T scale) override;
  std::size_t n_nonzeros() const override { return (std::size_t)_nnz; }
  // dest = A arg on the device (dest sized to the rows); dest = A^T arg on the host mirror
  void device_vector_mult(HipVector<T> &dest, const HipVector<T> &arg) const;
  void vector_mult_transpose(NumericVector<T> &dest, const NumericVector<T> &arg) const;
  // Matlab text (17 significant digits), Matlab text or PETSc binary in
  void print_matlab(const std::string &name = "") const override;
  void print_petsc_binary(const std::string &filename) const override;
  void read(const std::string &filename) override;
  void read_matlab(const std::string &filename) override;
  void read_petsc_binary(const std::string &filename) override;

  // Makes this matrix a view of a device matrix owned elsewhere (the reduced
  // matrix of a StaticCondensation): the pattern is copied to the host, the
  // values are read from the device on demand, and clear() leaves A alone.
  void view_device_matrix(lmh_matrix *A);

  // device handles for the GPU assembly and solver
  lmh_matrix *device_matrix() const { return _A; }
  lmh_mesh *device_mesh() const { return _mesh; }
  lmh_dofmap *device_dofmap() const { return _dm; }
  lmh_context *context() const { return _ctx; }
  const DofMap *dof_map() const { return this->_dof_map; }
  const MeshBase *attached_mesh() const { return _mesh_src; }
  // values were written on the device (host mirror stale)
  void mark_device_modified() {
    _vals_host_valid = false;
    _closed = true;
  }

 private:
  int64_t slot(numeric_index_type i, numeric_index_type j) const;  // -1 if absent
  void pull_values() const;
  void flush_elements();
  void flush_entries();
  void push_values();

  int _device;
  lmh_context *_ctx;
  const MeshBase *_mesh_src = nullptr;
  bool _asymmetric = true;
  std::vector<int32_t> _graph_rp, _graph_ci;  // libMesh's graph (update_sparsity_pattern)
  lmh_mesh *_mesh = nullptr;
  lmh_dofmap *_dm = nullptr;
  lmh_matrix *_A = nullptr;
  bool _owns_A = true;
  numeric_index_type _rows = 0, _cols = 0;
  int64_t _nnz = 0;
  std::vector<int32_t> _rp, _ci;        // host copy of the pattern
  mutable std::vector<T> _vals;         // host mirror of the values
  mutable bool _vals_host_valid = false;
  bool _vals_host_dirty = false;
  bool _closed = false;
  // staged element matrices (square, add_matrix(dm, dofs)) in call order
  int _stage_n = 0;
  std::vector<T> _stage_K;
  std::vector<int32_t> _stage_dofs;
  // entries inserted before the pattern exists (init(m, n, ...) mode)
  std::map<std::pair<numeric_index_type, numeric_index_type>, T> _open;
  std::mutex _mutex;  // host add/add_matrix may be called from several threads
};

}  // namespace libMesh
