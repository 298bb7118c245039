// sparse_matrix.h -- SparseMatrix<T> interface (include/numerics/sparse_matrix.h):
// the pure virtuals a backend implements, the sparsity-pattern hooks DofMap
// calls, and the defaults the drop-in keeps.
#pragma once

#include <string>

#include <iostream>
#include <memory>
#include <vector>

#include "libmesh/dense_matrix.h"
#include "libmesh/numeric_vector.h"

namespace libMesh {

class DofMap;

template <typename T>
class SparseMatrix {
 public:
  virtual ~SparseMatrix() = default;

  virtual SolverPackage solver_package() = 0;
  virtual bool initialized() const { return _is_initialized; }
  void attach_dof_map(const DofMap &dof_map) { _dof_map = &dof_map; }
  // SparseMatrix::need_full_sparsity_pattern (sparse_matrix.h:162): backends
  // that store the exact pattern return true and receive it via
  // update_sparsity_pattern.
  virtual bool need_full_sparsity_pattern() const { return false; }

  virtual void init(numeric_index_type m, numeric_index_type n, numeric_index_type m_l, numeric_index_type n_l,
                    numeric_index_type nnz = 30, numeric_index_type noz = 10, numeric_index_type blocksize = 1) = 0;
  virtual void init(ParallelType type = PARALLEL) = 0;
  virtual void clear() = 0;
  virtual void zero() = 0;
  virtual std::unique_ptr<SparseMatrix<T>> zero_clone() const = 0;
  virtual std::unique_ptr<SparseMatrix<T>> clone() const = 0;
  virtual void close() = 0;
  virtual void flush() { close(); }
  virtual numeric_index_type m() const = 0;
  virtual numeric_index_type n() const = 0;
  virtual numeric_index_type row_start() const = 0;
  virtual numeric_index_type row_stop() const = 0;
  virtual numeric_index_type col_start() const = 0;
  virtual numeric_index_type col_stop() const = 0;
  virtual void set(numeric_index_type i, numeric_index_type j, T value) = 0;
  virtual void add(numeric_index_type i, numeric_index_type j, T value) = 0;
  virtual void add_matrix(const DenseMatrix<T> &dm, const std::vector<numeric_index_type> &rows,
                          const std::vector<numeric_index_type> &cols) = 0;
  virtual void add_matrix(const DenseMatrix<T> &dm, const std::vector<numeric_index_type> &dof_indices) = 0;
  virtual void add(T a, const SparseMatrix<T> &X) = 0;
  virtual T operator()(numeric_index_type i, numeric_index_type j) const = 0;
  virtual Real l1_norm() const = 0;
  virtual Real linfty_norm() const = 0;
  virtual bool closed() const = 0;
  virtual void print_personal(std::ostream &os = std::cout) const = 0;
  virtual void get_diagonal(NumericVector<T> &dest) const = 0;
  virtual void get_transpose(SparseMatrix<T> &dest) const = 0;
  virtual void get_row(numeric_index_type i, std::vector<numeric_index_type> &indices,
                       std::vector<T> &values) const = 0;

  // SparseMatrix::vector_mult (sparse_matrix.C:237-253)
  virtual void vector_mult(NumericVector<T> &dest, const NumericVector<T> &arg) const = 0;
  virtual std::size_t n_nonzeros() const = 0;

  // On-disk formats (sparse_matrix.C: print_matlab :398, read :686, read_matlab
  // :1050, read_petsc_binary :1332)
  virtual void print_matlab(const std::string &name = "") const { libmesh_not_implemented(); }
  virtual void read(const std::string &filename) { libmesh_not_implemented(); }
  virtual void read_matlab(const std::string &filename) { libmesh_not_implemented(); }
  virtual void read_petsc_binary(const std::string &filename) { libmesh_not_implemented(); }

 protected:
  const DofMap *_dof_map = nullptr;
  bool _is_initialized = false;
};

}  // namespace libMesh
