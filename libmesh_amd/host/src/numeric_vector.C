// numeric_vector.C -- the NumericVector members that the interface implements
// on top of its pure virtuals (src/numerics/numeric_vector.C): build, the
// indexed get / add_vector / insert families, subset and difference norms,
// the compare family, printing and the Matlab text files.
#include <algorithm>
#include <cmath>
#include <fstream>
#include <regex>
#include <sstream>

#include "libmesh/numeric_vector.h"
#include "libmesh/shell_matrix.h"

namespace libMesh {

template <typename T>
std::unique_ptr<NumericVector<T>> NumericVector<T>::build(const Parallel::Communicator &comm,
                                                          SolverPackage solver_package, ParallelType parallel_type) {
  if (solver_package != HIP_SOLVERS) libmesh_not_implemented();
  return std::make_unique<HipVector<T>>(comm, parallel_type);
}

template <typename T>
void NumericVector<T>::get(const std::vector<numeric_index_type> &index, T *values) const {
  for (std::size_t k = 0; k < index.size(); ++k) values[k] = (*this)(index[k]);
}

template <typename T>
void NumericVector<T>::get(const std::vector<numeric_index_type> &index, std::vector<T> &values) const {
  values.resize(index.size());
  if (!index.empty()) this->get(index, values.data());
}

template <typename T>
void NumericVector<T>::add_vector(const T *v, const std::vector<numeric_index_type> &dof_indices) {
  for (std::size_t k = 0; k < dof_indices.size(); ++k) this->add(dof_indices[k], v[k]);
}

template <typename T>
void NumericVector<T>::add_vector(const std::vector<T> &v, const std::vector<numeric_index_type> &dof_indices) {
  if (v.size() != dof_indices.size()) libmesh_error_msg("add_vector: " << v.size() << " values, " << dof_indices.size() << " indices");
  if (!v.empty()) this->add_vector(v.data(), dof_indices);
}

template <typename T>
void NumericVector<T>::add_vector(const NumericVector<T> &v, const std::vector<numeric_index_type> &dof_indices) {
  if (v.size() != dof_indices.size()) libmesh_error_msg("add_vector: " << v.size() << " values, " << dof_indices.size() << " indices");
  std::vector<T> vals;
  v.localize(vals);
  this->add_vector(vals, dof_indices);
}

template <typename T>
void NumericVector<T>::add_vector(const DenseVector<T> &v, const std::vector<numeric_index_type> &dof_indices) {
  if (v.size() != dof_indices.size()) libmesh_error_msg("add_vector: " << v.size() << " values, " << dof_indices.size() << " indices");
  if (!dof_indices.empty()) this->add_vector(v.get_values().data(), dof_indices);
}

// this += A v for a shell matrix: A v into a temporary, then added
template <typename T>
void NumericVector<T>::add_vector(const NumericVector<T> &v, const ShellMatrix<T> &A) {
  A.vector_mult_add(*this, v);
}

template <typename T>
void NumericVector<T>::insert(const T *v, const std::vector<numeric_index_type> &dof_indices) {
  for (std::size_t k = 0; k < dof_indices.size(); ++k) this->set(dof_indices[k], v[k]);
}

template <typename T>
void NumericVector<T>::insert(const std::vector<T> &v, const std::vector<numeric_index_type> &dof_indices) {
  if (v.size() != dof_indices.size()) libmesh_error_msg("insert: " << v.size() << " values, " << dof_indices.size() << " indices");
  if (!v.empty()) this->insert(v.data(), dof_indices);
}

template <typename T>
void NumericVector<T>::insert(const NumericVector<T> &v, const std::vector<numeric_index_type> &dof_indices) {
  if (v.size() != dof_indices.size()) libmesh_error_msg("insert: " << v.size() << " values, " << dof_indices.size() << " indices");
  std::vector<T> vals;
  v.localize(vals);
  this->insert(vals, dof_indices);
}

template <typename T>
void NumericVector<T>::insert(const DenseVector<T> &v, const std::vector<numeric_index_type> &dof_indices) {
  if (v.size() != dof_indices.size()) libmesh_error_msg("insert: " << v.size() << " values, " << dof_indices.size() << " indices");
  if (!dof_indices.empty()) this->insert(v.get_values().data(), dof_indices);
}

template <typename T>
void NumericVector<T>::insert(const DenseSubVector<T> &v, const std::vector<numeric_index_type> &dof_indices) {
  if (v.size() != dof_indices.size()) libmesh_error_msg("insert: " << v.size() << " values, " << dof_indices.size() << " indices");
  for (std::size_t k = 0; k < dof_indices.size(); ++k) this->set(dof_indices[k], v((unsigned)k));
}

template <typename T>
Real NumericVector<T>::subset_l1_norm(const std::set<numeric_index_type> &indices) const {
  Real s = 0;
  for (auto i : indices) s += std::abs((*this)(i));
  return s;
}

template <typename T>
Real NumericVector<T>::subset_l2_norm(const std::set<numeric_index_type> &indices) const {
  Real s = 0;
  for (auto i : indices) {
    const Real a = std::abs((*this)(i));
    s += a * a;
  }
  return std::sqrt(s);
}

template <typename T>
Real NumericVector<T>::subset_linfty_norm(const std::set<numeric_index_type> &indices) const {
  Real m = 0;
  for (auto i : indices) m = std::max(m, (Real)std::abs((*this)(i)));
  return m;
}

// |this - v| through a temporary difference (on the device for HipVectors)
template <typename T>
Real NumericVector<T>::l2_norm_diff(const NumericVector<T> &v) const {
  if (this->size() != v.size()) libmesh_error_msg("l2_norm_diff: sizes " << this->size() << " and " << v.size());
  auto d = this->clone();
  *d -= v;
  return d->l2_norm();
}

template <typename T>
Real NumericVector<T>::l1_norm_diff(const NumericVector<T> &v) const {
  if (this->size() != v.size()) libmesh_error_msg("l1_norm_diff: sizes " << this->size() << " and " << v.size());
  auto d = this->clone();
  *d -= v;
  return d->l1_norm();
}

template <typename T>
int NumericVector<T>::compare(const NumericVector<T> &other, const Real threshold) const {
  for (numeric_index_type i = first_local_index(); i < last_local_index(); ++i)
    if (std::abs((*this)(i) - other(i)) > threshold) return (int)i;
  return -1;
}

template <typename T>
int NumericVector<T>::local_relative_compare(const NumericVector<T> &other, const Real threshold) const {
  for (numeric_index_type i = first_local_index(); i < last_local_index(); ++i) {
    const Real a = std::abs((*this)(i)), b = std::abs(other(i));
    if (std::abs((*this)(i) - other(i)) > threshold * std::max(a, b)) return (int)i;
  }
  return -1;
}

template <typename T>
int NumericVector<T>::global_relative_compare(const NumericVector<T> &other, const Real threshold) const {
  const Real tol = std::max(this->linfty_norm(), other.linfty_norm()) * threshold;
  for (numeric_index_type i = first_local_index(); i < last_local_index(); ++i)
    if (std::abs((*this)(i) - other(i)) > tol) return (int)i;
  return -1;
}

template <typename T>
void NumericVector<T>::print(std::ostream &os) const {
  os << "Size\tglobal =  " << this->size() << "\t\tlocal =  " << this->local_size() << std::endl;
  os << "#\tValue" << std::endl;
  for (numeric_index_type i = first_local_index(); i < last_local_index(); ++i) os << i << "\t" << (*this)(i) << std::endl;
}

template <typename T>
void NumericVector<T>::print_global(std::ostream &os) const {
  std::vector<T> v;
  this->localize(v);
  os << "Size\tglobal =  " << this->size() << std::endl;
  os << "#\tValue" << std::endl;
  for (std::size_t i = 0; i < v.size(); ++i) os << i << "\t" << v[i] << std::endl;
}

// The PETSC_VIEWER_ASCII_MATLAB-like text of numeric_vector.C:245-290: a
// two-line header, "Vec = [", one entry per line at the stream's default
// precision, "];".
template <typename T>
void NumericVector<T>::print_matlab(const std::string &filename) const {
  std::vector<T> v;
  this->localize(v);
  std::ofstream file;
  if (!filename.empty()) {
    file.open(filename);
    if (!file) libmesh_error_msg("print_matlab: cannot write " << filename);
  }
  std::ostream &os = filename.empty() ? std::cout : file;
  os << "%Vec Object: () " << this->n_processors() << " MPI processes\n"
     << "%  type: " << (this->n_processors() > 1 ? "mpi" : "seq") << "\n"
     << "Vec = [\n";
  for (const T &x : v) os << x << '\n';
  os << "];\n" << std::endl;
}

// numeric_vector.C:297-440: entries are the lines that parse as a number
// between an assignment "name = [" and a closing "]"; the vector is resized
// to the entry count when it differs.
template <typename T>
void NumericVector<T>::read_matlab(const std::string &filename) {
  std::ifstream in(filename);
  if (!in) libmesh_error_msg("read_matlab: cannot open " << filename);
  const std::regex start_re("\\s*\\w+\\s*=\\s*\\[");
  const std::regex end_re("^[^%]*\\]");
  std::vector<T> entries;
  bool started = false, ended = false;
  for (std::string line; std::getline(in, line);) {
    std::istringstream ls(line);
    T value;
    ls >> value;
    if (!ls.fail()) {
      if (!started) libmesh_error_msg("Confused by premature entries in vector file " << filename);
      entries.push_back(value);
    } else if (std::regex_search(line, start_re)) {
      started = true;
    } else if (std::regex_search(line, end_re)) {
      ended = true;
      break;
    }
  }
  if (!started) libmesh_error_msg("Confused by missing assignment-beginning in vector file " << filename);
  if (!ended) libmesh_error_msg("Confused by missing assignment-ending in vector file " << filename);
  const numeric_index_type n = entries.size();
  if (!this->initialized() || this->size() != n || this->local_size() != n) this->init(n, n);
  std::vector<numeric_index_type> idx(n);
  for (numeric_index_type i = 0; i < n; ++i) idx[i] = i;
  if (n) this->insert(entries.data(), idx);
  this->close();
}

template <typename T>
void NumericVector<T>::swap(NumericVector<T> &v) {
  std::swap(_is_closed, v._is_closed);
  std::swap(_is_initialized, v._is_initialized);
  std::swap(_type, v._type);
}

template class NumericVector<double>;

}  // namespace libMesh
