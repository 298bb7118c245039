// sparse_matrix.C -- the SparseMatrix members that the interface implements on
// top of its pure virtuals (src/numerics/sparse_matrix.C): build, block
// insertion, printing, the products through NumericVector::add_vector, and
// the difference norm.
#include <fstream>

#include "libmesh/hip_csr_matrix.h"
#include "libmesh/sparse_matrix.h"

namespace libMesh {

template <typename T>
std::unique_ptr<SparseMatrix<T>> SparseMatrix<T>::build(const Parallel::Communicator &comm,
                                                        const SolverPackage solver_package,
                                                        const MatrixBuildType matrix_build_type) {
  if (solver_package != HIP_SOLVERS || matrix_build_type != MatrixBuildType::AUTOMATIC) libmesh_not_implemented();
  return std::make_unique<HipCsrMatrix<T>>(comm);
}

template <typename T>
void SparseMatrix<T>::zero_rows(std::vector<numeric_index_type> &, T) {
  libmesh_not_implemented();
}

template <typename T>
void SparseMatrix<T>::add_block_matrix(const DenseMatrix<T> &dm, const std::vector<numeric_index_type> &brows,
                                       const std::vector<numeric_index_type> &bcols) {
  if (brows.empty() || bcols.empty() || dm.m() % brows.size() || dm.n() % bcols.size() ||
      dm.m() / brows.size() != dm.n() / bcols.size())
    libmesh_error_msg("add_block_matrix: a " << dm.m() << "x" << dm.n() << " matrix for " << brows.size() << "x"
                                             << bcols.size() << " blocks");
  const numeric_index_type bs = dm.m() / brows.size();
  std::vector<numeric_index_type> rows, cols;
  rows.reserve(dm.m());
  cols.reserve(dm.n());
  for (auto r : brows)
    for (numeric_index_type v = 0; v < bs; ++v) rows.push_back(r * bs + v);
  for (auto c : bcols)
    for (numeric_index_type v = 0; v < bs; ++v) cols.push_back(c * bs + v);
  this->add_matrix(dm, rows, cols);
}

template <typename T>
Real SparseMatrix<T>::l1_norm_diff(const SparseMatrix<T> &other_mat) const {
  auto d = this->clone();
  d->add(T(-1), other_mat);
  return d->l1_norm();
}

// the stored entries, counted row by row
template <typename T>
std::size_t SparseMatrix<T>::n_nonzeros() const {
  std::size_t c = 0;
  std::vector<numeric_index_type> idx;
  std::vector<T> vals;
  for (numeric_index_type i = row_start(); i < row_stop(); ++i) {
    this->get_row(i, idx, vals);
    c += idx.size();
  }
  return c;
}

// dense rows, or "i j value" lines of the nonzero entries
template <typename T>
void SparseMatrix<T>::print(std::ostream &os, const bool sparse) const {
  for (numeric_index_type i = 0; i < this->m(); ++i) {
    for (numeric_index_type j = 0; j < this->n(); ++j) {
      const T c = (*this)(i, j);
      if (!sparse)
        os << c << " ";
      else if (c != T(0))
        os << i << " " << j << " " << c << std::endl;
    }
    if (!sparse) os << std::endl;
  }
}

// the format by the file name: Matlab text for ".m" / ".matlab", PETSc binary
// for ".petsc32" / ".petsc64" (sparse_matrix.C:754-790)
template <typename T>
void SparseMatrix<T>::print(const std::string &filename) const {
  auto ends = [&](const char *suf) {
    const std::string s(suf);
    return filename.size() >= s.size() && filename.compare(filename.size() - s.size(), s.size(), s) == 0;
  };
  if (ends(".matlab") || ends(".m"))
    this->print_matlab(filename);
  else if (ends(".petsc32") || ends(".petsc64"))
    this->print_petsc_binary(filename);
  else
    libmesh_error_msg("Unrecognized matrix file extension on: " << filename);
}

template <typename T>
void SparseMatrix<T>::vector_mult(NumericVector<T> &dest, const NumericVector<T> &arg) const {
  dest.zero();
  this->vector_mult_add(dest, arg);
}

template <typename T>
void SparseMatrix<T>::vector_mult_add(NumericVector<T> &dest, const NumericVector<T> &arg) const {
  dest.add_vector(arg, *this);
}

template <typename T>
void SparseMatrix<T>::scale(const T scale) {
  if (!this->closed()) libmesh_error_msg("SparseMatrix::scale before close()");
  std::vector<numeric_index_type> idx;
  std::vector<T> vals;
  for (numeric_index_type i = row_start(); i < row_stop(); ++i) {
    this->get_row(i, idx, vals);
    for (std::size_t k = 0; k < idx.size(); ++k) this->set(i, idx[k], vals[k] * scale);
  }
  this->close();
}

template class SparseMatrix<double>;

}  // namespace libMesh
