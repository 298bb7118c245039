// hip_csr_matrix.h -- HipCsrMatrix<T>: a SparseMatrix backend whose CSR
// pattern is built on the GPU from the attached DofMap (the rule of
// SparsityPattern::Build, src/base/sparsity_pattern.C:96-365) and whose values
// live in device memory.  Host-side set/add/add_matrix calls are staged and
// applied at close() in call order; element matrices go through the batched
// deterministic device scatter.  Device assembly (HipAssembler) writes the
// values directly.
#pragma once

#include <map>
#include <mutex>

#include "libmesh/dof_map.h"
#include "libmesh/sparse_matrix.h"

namespace libMesh {

template <typename T>
class HipCsrMatrix : public SparseMatrix<T> {
 public:
  explicit HipCsrMatrix(int device = 0);
  ~HipCsrMatrix() override;
  HipCsrMatrix(const HipCsrMatrix &) = delete;
  HipCsrMatrix &operator=(const HipCsrMatrix &) = delete;

  SolverPackage solver_package() override { return HIP_SOLVERS; }
  bool need_full_sparsity_pattern() const override { return true; }

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
  void vector_mult(NumericVector<T> &dest, const NumericVector<T> &arg) const override;
  std::size_t n_nonzeros() const override { return (std::size_t)_nnz; }
  // Matlab text (17 significant digits), Matlab text or PETSc binary in
  void print_matlab(const std::string &name = "") const override;
  void read(const std::string &filename) override;
  void read_matlab(const std::string &filename) override;
  void read_petsc_binary(const std::string &filename) override;

  // device handles for the GPU assembly and solver
  lmh_matrix *device_matrix() const { return _A; }
  lmh_mesh *device_mesh() const { return _mesh; }
  lmh_dofmap *device_dofmap() const { return _dm; }
  lmh_context *context() const { return _ctx; }
  const DofMap *dof_map() const { return this->_dof_map; }
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
  lmh_mesh *_mesh = nullptr;
  lmh_dofmap *_dm = nullptr;
  lmh_matrix *_A = nullptr;
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
