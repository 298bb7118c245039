// sparse_matrix.h -- SparseMatrix<T> (include/numerics/sparse_matrix.h:69-690):
// the pure virtuals a backend implements, with the reference's signatures,
// the sparsity-pattern hooks DofMap calls, and the members the interface
// implements on top of the pure virtuals (src/numerics/sparse_matrix.C).
#pragma once

#include <iostream>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "libmesh/dense_matrix.h"
#include "libmesh/numeric_vector.h"
#include "libmesh/parallel_object.h"
#include "libmesh/sparsity_pattern.h"

namespace libMesh {

class DofMap;

// include/enums/enum_matrix_build_type.h
enum class MatrixBuildType { AUTOMATIC = 0, DIAGONAL };

template <typename T>
class SparseMatrix : public ParallelObject {
 public:
  explicit SparseMatrix(const Parallel::Communicator &comm) : ParallelObject(comm) {}
  virtual SparseMatrix<T> &operator=(const SparseMatrix<T> &) {
    libmesh_not_implemented();
    return *this;
  }
  virtual ~SparseMatrix() = default;

  // SparseMatrix::build (sparse_matrix.C:140-200): the GPU CSR matrix for
  // HIP_SOLVERS, the only package this library provides
  static std::unique_ptr<SparseMatrix<T>> build(const Parallel::Communicator &comm,
                                                const SolverPackage solver_package = HIP_SOLVERS,
                                                const MatrixBuildType matrix_build_type = MatrixBuildType::AUTOMATIC);

  virtual SolverPackage solver_package() = 0;
  virtual bool initialized() const { return _is_initialized; }
  void attach_dof_map(const DofMap &dof_map) { _dof_map = &dof_map; }
  // SparseMatrix::need_full_sparsity_pattern (sparse_matrix.h:162): backends
  // that return true receive the pattern through update_sparsity_pattern
  virtual bool need_full_sparsity_pattern() const { return false; }
  virtual void update_sparsity_pattern(const SparsityPattern::Graph &) {}

  virtual void init(const numeric_index_type m, const numeric_index_type n, const numeric_index_type m_l,
                    const numeric_index_type n_l, const numeric_index_type nnz = 30, const numeric_index_type noz = 10,
                    const numeric_index_type blocksize = 1) = 0;
  virtual void init(ParallelType type = PARALLEL) = 0;
  virtual void clear() = 0;
  virtual void zero() = 0;
  virtual std::unique_ptr<SparseMatrix<T>> zero_clone() const = 0;
  virtual std::unique_ptr<SparseMatrix<T>> clone() const = 0;
  // SparseMatrix::zero_rows (sparse_matrix.C:388-395 of the default): rows
  // emptied, diag_value on their diagonals
  virtual void zero_rows(std::vector<numeric_index_type> &rows, T diag_value = 0.0);
  virtual void close() = 0;
  virtual void flush() { close(); }
  virtual numeric_index_type m() const = 0;
  virtual numeric_index_type local_m() const { return row_stop() - row_start(); }
  virtual numeric_index_type local_n() const { return col_stop() - col_start(); }
  virtual numeric_index_type n() const = 0;
  virtual numeric_index_type row_start() const = 0;
  virtual numeric_index_type row_stop() const = 0;
  virtual numeric_index_type col_start() const = 0;
  virtual numeric_index_type col_stop() const = 0;
  virtual void set(const numeric_index_type i, const numeric_index_type j, const T value) = 0;
  virtual void add(const numeric_index_type i, const numeric_index_type j, const T value) = 0;
  virtual void add_matrix(const DenseMatrix<T> &dm, const std::vector<numeric_index_type> &rows,
                          const std::vector<numeric_index_type> &cols) = 0;
  virtual void add_matrix(const DenseMatrix<T> &dm, const std::vector<numeric_index_type> &dof_indices) = 0;
  // SparseMatrix::add_block_matrix (sparse_matrix.C:93-124): block indices
  // expanded by the block size dm.m() / brows.size()
  virtual void add_block_matrix(const DenseMatrix<T> &dm, const std::vector<numeric_index_type> &brows,
                                const std::vector<numeric_index_type> &bcols);
  virtual void add_block_matrix(const DenseMatrix<T> &dm, const std::vector<numeric_index_type> &dof_indices) {
    this->add_block_matrix(dm, dof_indices, dof_indices);
  }
  virtual void add(const T a, const SparseMatrix<T> &X) = 0;
  virtual void matrix_matrix_mult(SparseMatrix<T> &, SparseMatrix<T> &, bool) { libmesh_not_implemented(); }
  virtual void add_sparse_matrix(const SparseMatrix<T> &, const std::map<numeric_index_type, numeric_index_type> &,
                                 const std::map<numeric_index_type, numeric_index_type> &, const T) {
    libmesh_not_implemented();
  }
  virtual T operator()(const numeric_index_type i, const numeric_index_type j) const = 0;
  virtual Real l1_norm() const = 0;
  virtual Real linfty_norm() const = 0;
  // |this - other|_1 (sparse_matrix.C:640-660)
  Real l1_norm_diff(const SparseMatrix<T> &other_mat) const;
  virtual bool closed() const = 0;
  virtual std::size_t n_nonzeros() const;
  void print(std::ostream &os = std::cout, const bool sparse = false) const;
  virtual void print_personal(std::ostream &os = std::cout) const = 0;
  void print(const std::string &filename) const;
  // On-disk formats (sparse_matrix.C: print_matlab :398, read :686,
  // read_matlab :1050, read_petsc_binary :1332, print_petsc_binary :1260)
  virtual void print_matlab(const std::string & = "") const { libmesh_not_implemented(); }
  virtual void print_petsc_binary(const std::string &) const { libmesh_not_implemented(); }
  virtual void read(const std::string &) { libmesh_not_implemented(); }
  virtual void read_matlab(const std::string &) { libmesh_not_implemented(); }
  virtual void read_petsc_binary(const std::string &) { libmesh_not_implemented(); }
  virtual void create_submatrix(SparseMatrix<T> &submatrix, const std::vector<numeric_index_type> &rows,
                                const std::vector<numeric_index_type> &cols) const {
    this->_get_submatrix(submatrix, rows, cols, false);
  }
  virtual void reinit_submatrix(SparseMatrix<T> &submatrix, const std::vector<numeric_index_type> &rows,
                                const std::vector<numeric_index_type> &cols) const {
    this->_get_submatrix(submatrix, rows, cols, true);
  }
  // SparseMatrix::vector_mult / vector_mult_add (sparse_matrix.C:237-253):
  // dest = A arg, dest += A arg, through NumericVector::add_vector(arg, A)
  void vector_mult(NumericVector<T> &dest, const NumericVector<T> &arg) const;
  void vector_mult_add(NumericVector<T> &dest, const NumericVector<T> &arg) const;
  virtual void get_diagonal(NumericVector<T> &dest) const = 0;
  virtual void get_transpose(SparseMatrix<T> &dest) const = 0;
  virtual void get_row(numeric_index_type i, std::vector<numeric_index_type> &indices,
                       std::vector<T> &values) const = 0;
  // SparseMatrix::scale (sparse_matrix.C:700-716): every stored entry times `scale`
  virtual void scale(const T scale);

 protected:
  virtual void _get_submatrix(SparseMatrix<T> &, const std::vector<numeric_index_type> &,
                              const std::vector<numeric_index_type> &, const bool) const {
    libmesh_not_implemented();
  }
  const DofMap *_dof_map = nullptr;
  bool _is_initialized = false;
};

}  // namespace libMesh
