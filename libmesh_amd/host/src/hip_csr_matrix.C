// hip_csr_matrix.C -- SparseMatrix backend on the GPU CSR of liblmhip.
//
// Two ways to get a pattern, as libMesh's backends do:
//  * init(ParallelType) with an attached, distributed DofMap: the pattern is
//    built on the GPU from the element DoF lists (DofMap::compute_sparsity ->
//    SparsityPattern::Build, need_full_sparsity_pattern() = true);
//  * init(m, n, ...) without a DofMap: entries are collected on the host until
//    the first close(), which freezes their union as the pattern (like
//    EigenSparseMatrix inserting on coeffRef, eigen_sparse_matrix.C:289-300).
// Values live on the device.  Host edits (set/add/add_matrix with separate
// row/column lists) go to a host mirror pushed at close(); square element
// matrices (add_matrix(dm, dofs)) are batched and summed on the device in call
// order by lmh_matrix_add_element_matrices.
#include <algorithm>
#include <cmath>
#include <iomanip>
#include <map>

#include "libmesh/hip_csr_matrix.h"
#include "libmesh/device_dof_data.h"

#include <tuple>

#include <fstream>

#include "libmesh/matrix_io.h"

namespace libMesh {


template <typename T>
HipCsrMatrix<T>::HipCsrMatrix(const Parallel::Communicator &comm, int device)
    : SparseMatrix<T>(comm), _device(device), _ctx(HipDevice::context(device)) {}

template <typename T>
HipCsrMatrix<T>::~HipCsrMatrix() {
  clear();
}

template <typename T>
void HipCsrMatrix<T>::clear() {
  if (_A && _owns_A) lmh_matrix_destroy(_A);
  _owns_A = true;
  if (_dm) lmh_dofmap_destroy(_dm);
  if (_mesh) lmh_mesh_destroy(_mesh);
  _A = nullptr;
  _dm = nullptr;
  _mesh = nullptr;
  _rows = 0;
  _cols = 0;
  _nnz = 0;
  _rp.clear();
  _ci.clear();
  _vals.clear();
  _vals_host_valid = false;
  _vals_host_dirty = false;
  _stage_K.clear();
  _stage_dofs.clear();
  _stage_n = 0;
  _open.clear();
  this->_is_initialized = false;
  _closed = false;
}

template <typename T>
void HipCsrMatrix<T>::update_sparsity_pattern(const SparsityPattern::Graph &sparsity_pattern) {
  _graph_rp.assign(1, 0);
  _graph_ci.clear();
  for (const auto &row : sparsity_pattern) {
    for (std::size_t k = 0; k < row.size(); ++k) {
      if (k && row[k] <= row[k - 1]) libmesh_error_msg("update_sparsity_pattern: row columns must ascend");
      _graph_ci.push_back((int32_t)row[k]);
    }
    _graph_rp.push_back((int32_t)_graph_ci.size());
  }
}

template <typename T>
void HipCsrMatrix<T>::init(ParallelType) {
  const DofMap *dm = this->_dof_map;
  if (!dm) libmesh_error_msg("HipCsrMatrix::init: attach a distributed DofMap first");
  const std::vector<int32_t> graph_rp = _graph_rp, graph_ci = _graph_ci;
  clear();
  _rows = dm->n_dofs();
  _cols = _rows;
  if (_mesh_src) {
    // pattern, scatter maps and constraint rows on the device, from the mesh
    // and DoF map read through their libMesh accessors
    DeviceDofData dev;
    build_device_dofs(_ctx, *_mesh_src, *dm, _asymmetric ? 0 : LMH_CONSTRAINT_SYMMETRIC_ROWS, dev);
    hip_check(lmh_matrix_create(_ctx, dev.dofmap, &_A), "HipCsrMatrix::init (sparsity)");
    _mesh = dev.mesh;
    _dm = dev.dofmap;
    dev.mesh = nullptr;
    dev.dofmap = nullptr;
  } else {
    // libMesh's graph (DofMap::update_sparsity_pattern before init)
    if (graph_rp.size() != _rows + 1)
      libmesh_error_msg("HipCsrMatrix::init: no pattern (attach_mesh, or update_sparsity_pattern with "
                        << _rows << " rows; got " << (graph_rp.empty() ? 0 : graph_rp.size() - 1) << ")");
    hip_check(lmh_matrix_create_csr2(_ctx, _rows, _cols, graph_rp.data(), graph_ci.data(), &_A),
              "HipCsrMatrix::init (graph)");
    _graph_rp = graph_rp;
    _graph_ci = graph_ci;
  }
  _nnz = lmh_matrix_nnz(_A);
  _rp.resize(_rows + 1);
  _ci.resize(_nnz);
  hip_check(lmh_matrix_get_pattern(_A, _rp.data(), _ci.data()), "HipCsrMatrix::init (pattern)");
  _vals.assign(_nnz, T(0));
  _vals_host_valid = true;
  this->_is_initialized = true;
  _closed = true;
}

template <typename T>
void HipCsrMatrix<T>::view_device_matrix(lmh_matrix *A) {
  clear();
  _A = A;
  _owns_A = false;
  _rows = (numeric_index_type)lmh_matrix_rows(A);
  _cols = (numeric_index_type)lmh_matrix_cols(A);
  _nnz = lmh_matrix_nnz(A);
  _rp.resize(_rows + 1);
  _ci.resize(_nnz);
  hip_check(lmh_matrix_get_pattern(_A, _rp.data(), _ci.data()), "HipCsrMatrix::view_device_matrix");
  _vals_host_valid = false;
  this->_is_initialized = true;
  _closed = true;
}

template <typename T>
void HipCsrMatrix<T>::init(numeric_index_type m, numeric_index_type n, numeric_index_type, numeric_index_type,
                           numeric_index_type, numeric_index_type, numeric_index_type) {
  clear();
  _rows = m;
  _cols = n;
  this->_is_initialized = true;
  _closed = false;
}

template <typename T>
int64_t HipCsrMatrix<T>::slot(numeric_index_type i, numeric_index_type j) const {
  if (i >= _rows || _rp.empty()) return -1;
  const int32_t *b = _ci.data() + _rp[i], *e = _ci.data() + _rp[i + 1];
  const int32_t *p = std::lower_bound(b, e, (int32_t)j);
  if (p == e || *p != (int32_t)j) return -1;
  return p - _ci.data();
}

template <typename T>
void HipCsrMatrix<T>::pull_values() const {
  if (_vals_host_valid || !_A) return;
  _vals.resize(_nnz);
  if (_nnz) hip_check(lmh_matrix_get_values(_A, (double *)_vals.data()), "HipCsrMatrix: read values");
  _vals_host_valid = true;
}

template <typename T>
void HipCsrMatrix<T>::push_values() {
  if (!_vals_host_dirty) return;
  hip_check(lmh_matrix_set_values(_A, (const double *)_vals.data()), "HipCsrMatrix: write values");
  _vals_host_dirty = false;
}

template <typename T>
void HipCsrMatrix<T>::flush_elements() {
  if (_stage_n == 0 || _stage_dofs.empty()) return;
  push_values();
  const int n = _stage_n;
  const int64_t ne = (int64_t)(_stage_dofs.size() / n);
  void *Kd = nullptr, *Dd = nullptr;
  hip_check(lmh_malloc(_ctx, sizeof(double) * _stage_K.size(), &Kd), "add_matrix");
  hip_check(lmh_malloc(_ctx, sizeof(int32_t) * _stage_dofs.size(), &Dd), "add_matrix");
  hip_check(lmh_memcpy_h2d(_ctx, Kd, _stage_K.data(), sizeof(double) * _stage_K.size()), "add_matrix");
  hip_check(lmh_memcpy_h2d(_ctx, Dd, _stage_dofs.data(), sizeof(int32_t) * _stage_dofs.size()), "add_matrix");
  const int rc = lmh_matrix_add_element_matrices(_A, ne, n, (const double *)Kd, (const int32_t *)Dd);
  lmh_free(_ctx, Kd);
  lmh_free(_ctx, Dd);
  hip_check(rc, "HipCsrMatrix::add_matrix");
  _stage_K.clear();
  _stage_dofs.clear();
  _stage_n = 0;
  _vals_host_valid = false;
}

template <typename T>
void HipCsrMatrix<T>::flush_entries() {
  // pattern-open mode: freeze the union of inserted entries as the pattern
  if (_A || !this->_is_initialized) return;
  _rp.assign(_rows + 1, 0);
  _ci.clear();
  _vals.clear();
  for (const auto &kv : _open) {
    _rp[kv.first.first + 1]++;
    _ci.push_back((int32_t)kv.first.second);
    _vals.push_back(kv.second);
  }
  for (numeric_index_type r = 0; r < _rows; ++r) _rp[r + 1] += _rp[r];
  _nnz = (int64_t)_ci.size();
  hip_check(lmh_matrix_create_csr2(_ctx, _rows, _cols, _rp.data(), _ci.data(), &_A), "HipCsrMatrix::close (pattern)");
  _vals_host_valid = true;
  _vals_host_dirty = true;
  _open.clear();
}

template <typename T>
void HipCsrMatrix<T>::close() {
  std::lock_guard<std::mutex> lock(_mutex);
  flush_entries();
  flush_elements();
  push_values();
  _closed = true;
}

template <typename T>
void HipCsrMatrix<T>::zero() {
  std::lock_guard<std::mutex> lock(_mutex);
  if (!_A) {
    for (auto &kv : _open) kv.second = T(0);
    return;
  }
  _stage_K.clear();
  _stage_dofs.clear();
  _stage_n = 0;
  hip_check(lmh_matrix_zero(_A), "HipCsrMatrix::zero");
  std::fill(_vals.begin(), _vals.end(), T(0));
  _vals_host_valid = true;
  _vals_host_dirty = false;
}

template <typename T>
void HipCsrMatrix<T>::set(numeric_index_type i, numeric_index_type j, T value) {
  std::lock_guard<std::mutex> lock(_mutex);
  if (!_A) {
    _open[{i, j}] = value;
    _closed = false;
    return;
  }
  flush_elements();
  pull_values();
  const int64_t k = slot(i, j);
  if (k < 0) libmesh_error_msg("HipCsrMatrix::set: (" << i << "," << j << ") is not in the sparsity pattern");
  _vals[k] = value;
  _vals_host_dirty = true;
  _closed = false;
}

template <typename T>
void HipCsrMatrix<T>::add(numeric_index_type i, numeric_index_type j, T value) {
  std::lock_guard<std::mutex> lock(_mutex);
  if (!_A) {
    _open[{i, j}] += value;
    _closed = false;
    return;
  }
  flush_elements();
  pull_values();
  const int64_t k = slot(i, j);
  if (k < 0) libmesh_error_msg("HipCsrMatrix::add: (" << i << "," << j << ") is not in the sparsity pattern");
  _vals[k] += value;
  _vals_host_dirty = true;
  _closed = false;
}

template <typename T>
void HipCsrMatrix<T>::add_matrix(const DenseMatrix<T> &dm, const std::vector<numeric_index_type> &rows,
                                 const std::vector<numeric_index_type> &cols) {
  if (dm.m() != rows.size() || dm.n() != cols.size()) libmesh_error_msg("add_matrix: size mismatch");
  if (rows == cols) {
    add_matrix(dm, rows);
    return;
  }
  for (unsigned a = 0; a < dm.m(); ++a)
    for (unsigned b = 0; b < dm.n(); ++b) add(rows[a], cols[b], dm(a, b));
}

template <typename T>
void HipCsrMatrix<T>::add_matrix(const DenseMatrix<T> &dm, const std::vector<numeric_index_type> &dofs) {
  if (dm.m() != dofs.size() || dm.n() != dofs.size()) libmesh_error_msg("add_matrix: size mismatch");
  std::lock_guard<std::mutex> lock(_mutex);
  _closed = false;
  if (!_A) {
      for (unsigned a = 0; a < dm.m(); ++a)
      for (unsigned b = 0; b < dm.n(); ++b) _open[{dofs[a], dofs[b]}] += dm(a, b);
    return;
  }
  if (_stage_n != 0 && _stage_n != (int)dofs.size()) flush_elements();
  push_values();
  _stage_n = (int)dofs.size();
  for (numeric_index_type d : dofs) {
    if (d >= _rows) libmesh_error_msg("add_matrix: dof " << d << " out of range");
    _stage_dofs.push_back((int32_t)d);
  }
  _stage_K.insert(_stage_K.end(), dm.get_values().begin(), dm.get_values().end());
}

template <typename T>
void HipCsrMatrix<T>::add(T a, const SparseMatrix<T> &X) {
  const auto *Y = dynamic_cast<const HipCsrMatrix<T> *>(&X);
  if (!Y || Y->_rp != _rp || Y->_ci != _ci) libmesh_error_msg("HipCsrMatrix::add(a, X): X must share the pattern");
  close();
  Y->pull_values();
  pull_values();
  for (int64_t k = 0; k < _nnz; ++k) _vals[k] += a * Y->_vals[k];
  _vals_host_dirty = true;
  push_values();
}

template <typename T>
T HipCsrMatrix<T>::operator()(numeric_index_type i, numeric_index_type j) const {
  if (!_closed) libmesh_error_msg("HipCsrMatrix: read before close()");
  pull_values();
  const int64_t k = slot(i, j);
  return k < 0 ? T(0) : _vals[k];
}

template <typename T>
Real HipCsrMatrix<T>::l1_norm() const {
  pull_values();
  std::vector<Real> col(_cols, 0.0);
  for (numeric_index_type r = 0; r < _rows; ++r)
    for (int32_t k = _rp[r]; k < _rp[r + 1]; ++k) col[_ci[k]] += std::abs(_vals[k]);
  return _cols ? *std::max_element(col.begin(), col.end()) : 0.0;
}

template <typename T>
Real HipCsrMatrix<T>::linfty_norm() const {
  pull_values();
  Real m = 0;
  for (numeric_index_type r = 0; r < _rows; ++r) {
    Real s = 0;
    for (int32_t k = _rp[r]; k < _rp[r + 1]; ++k) s += std::abs(_vals[k]);
    m = std::max(m, s);
  }
  return m;
}

template <typename T>
void HipCsrMatrix<T>::print_personal(std::ostream &os) const {
  pull_values();
  os << std::setprecision(17);
  for (numeric_index_type r = 0; r < _rows; ++r)
    for (int32_t k = _rp[r]; k < _rp[r + 1]; ++k) os << r << " " << _ci[k] << " " << _vals[k] << "\n";
}

template <typename T>
void HipCsrMatrix<T>::get_diagonal(NumericVector<T> &dest) const {
  pull_values();
  if (dest.size() != _rows) dest.init(_rows, _rows);
  for (numeric_index_type r = 0; r < _rows; ++r) {
    const int64_t k = slot(r, r);
    dest.set(r, k < 0 ? T(0) : _vals[k]);
  }
  dest.close();
}

template <typename T>
void HipCsrMatrix<T>::get_transpose(SparseMatrix<T> &dest) const {
  auto *D = dynamic_cast<HipCsrMatrix<T> *>(&dest);
  if (!D) libmesh_error_msg("get_transpose: destination must be a HipCsrMatrix");
  pull_values();
  D->clear();
  D->init(_cols, _rows, _cols, _rows);
  for (numeric_index_type r = 0; r < _rows; ++r)
    for (int32_t k = _rp[r]; k < _rp[r + 1]; ++k) D->set((numeric_index_type)_ci[k], r, _vals[k]);
  D->close();
}

template <typename T>
void HipCsrMatrix<T>::get_row(numeric_index_type i, std::vector<numeric_index_type> &indices,
                              std::vector<T> &values) const {
  pull_values();
  indices.clear();
  values.clear();
  for (int32_t k = _rp.at(i); k < _rp.at(i + 1); ++k) {
    indices.push_back((numeric_index_type)_ci[k]);
    values.push_back(_vals[k]);
  }
}

template <typename T>
void HipCsrMatrix<T>::device_vector_mult(HipVector<T> &dest, const HipVector<T> &arg) const {
  if (!_closed) libmesh_error_msg("vector_mult before close()");
  if (arg.size() != _cols) libmesh_error_msg("vector_mult: argument has " << arg.size() << " entries, need " << _cols);
  if (dest.size() != _rows) dest.init(_rows, _rows);
  hip_check(lmh_spmv(_A, (const double *)arg.device_data(), (double *)dest.device_data()), "vector_mult");
  dest.mark_device_modified();
}

template <typename T>
void HipCsrMatrix<T>::vector_mult_transpose(NumericVector<T> &dest, const NumericVector<T> &arg) const {
  if (!_closed) libmesh_error_msg("vector_mult_transpose before close()");
  if (arg.size() != _rows) libmesh_error_msg("vector_mult_transpose: argument has " << arg.size() << " entries, need " << _rows);
  pull_values();
  std::vector<T> a, y(_cols, T(0));
  arg.localize(a);
  for (numeric_index_type r = 0; r < _rows; ++r)
    for (int32_t k = _rp[r]; k < _rp[r + 1]; ++k) y[_ci[k]] += _vals[k] * a[r];
  if (dest.size() != _cols) dest.init(_cols, _cols);
  dest = y;
}

// rows emptied, diag_value on their diagonals (PetscMatrix::zero_rows semantics)
template <typename T>
void HipCsrMatrix<T>::zero_rows(std::vector<numeric_index_type> &rows, T diag_value) {
  close();
  pull_values();
  for (numeric_index_type r : rows) {
    if (r >= _rows) libmesh_error_msg("zero_rows: row " << r << " out of range");
    for (int32_t k = _rp[r]; k < _rp[r + 1]; ++k) _vals[k] = (numeric_index_type)_ci[k] == r ? diag_value : T(0);
  }
  _vals_host_dirty = true;
  push_values();
}

// every stored value times `scale`, on the device
template <typename T>
void HipCsrMatrix<T>::scale(const T scale) {
  close();
  const int32_t *rp = nullptr, *ci = nullptr;
  double *vals = nullptr;
  hip_check(lmh_matrix_device_arrays(_A, &rp, &ci, &vals), "HipCsrMatrix::scale");
  hip_check(lmh_vector_op(_ctx, LMH_VEC_SCALE, _nnz, (double)scale, nullptr, nullptr, vals), "HipCsrMatrix::scale");
  _vals_host_valid = false;
}

template <typename T>
std::unique_ptr<SparseMatrix<T>> HipCsrMatrix<T>::zero_clone() const {
  auto c = std::make_unique<HipCsrMatrix<T>>(this->comm(), _device);
  if (this->_dof_map && (_dm || !_graph_rp.empty())) {
    c->attach_dof_map(*this->_dof_map);
    if (_mesh_src) c->attach_mesh(*_mesh_src);
    c->set_asymmetric_constraint_rows(_asymmetric);
    c->_graph_rp = _graph_rp;
    c->_graph_ci = _graph_ci;
    c->init();
  } else {
    c->init(_rows, _cols, _rows, _cols);
    for (numeric_index_type r = 0; r < _rows; ++r)
      for (int32_t k = _rp[r]; k < _rp[r + 1]; ++k) c->set(r, (numeric_index_type)_ci[k], T(0));
    c->close();
  }
  return c;
}

template <typename T>
std::unique_ptr<SparseMatrix<T>> HipCsrMatrix<T>::clone() const {
  auto c = zero_clone();
  auto *h = static_cast<HipCsrMatrix<T> *>(c.get());
  pull_values();
  h->_vals = _vals;
  h->_vals_host_valid = true;
  h->_vals_host_dirty = true;
  h->close();
  return c;
}

template <typename T>
void HipCsrMatrix<T>::print_matlab(const std::string &name) const {
  if (!_closed) libmesh_error_msg("HipCsrMatrix::print_matlab before close()");
  pull_values();
  std::vector<double> v(_vals.begin(), _vals.end());
  if (name.empty()) {
    MatrixIO::write_matlab(std::cout, _rows, _cols, _rp, _ci, v);
    return;
  }
  std::ofstream os(name);
  if (!os.good()) libmesh_error_msg("HipCsrMatrix::print_matlab: cannot write " << name);
  MatrixIO::write_matlab(os, _rows, _cols, _rp, _ci, v);
}

template <typename T>
void HipCsrMatrix<T>::print_petsc_binary(const std::string &filename) const {
  if (!_closed) libmesh_error_msg("HipCsrMatrix::print_petsc_binary before close()");
  pull_values();
  const bool wide = filename.size() >= 8 && filename.compare(filename.size() - 8, 8, ".petsc64") == 0;
  std::vector<double> v(_vals.begin(), _vals.end());
  MatrixIO::write_petsc_binary(filename, wide ? 8 : 4, _rows, _cols, _rp, _ci, v);
}

// A read matrix takes the pattern of the file's entries (explicit zeros kept),
// as the reference's read_matlab/read_petsc_binary init + set do.
template <typename T>
static void load_entries(HipCsrMatrix<T> &A, const MatrixIO::Entries &E) {
  A.clear();
  A.init(E.m, E.n, E.m, E.n);
  for (const auto &e : E.ijv) A.set(std::get<0>(e), std::get<1>(e), (T)std::get<2>(e));
  A.close();
}

template <typename T>
void HipCsrMatrix<T>::read(const std::string &filename) {
  load_entries(*this, MatrixIO::read(filename));
}

template <typename T>
void HipCsrMatrix<T>::read_matlab(const std::string &filename) {
  load_entries(*this, MatrixIO::read_matlab(filename));
}

template <typename T>
void HipCsrMatrix<T>::read_petsc_binary(const std::string &filename) {
  const bool wide = filename.size() >= 8 && filename.compare(filename.size() - 8, 8, ".petsc64") == 0;
  load_entries(*this, MatrixIO::read_petsc_binary(filename, wide ? 8 : 4));
}

template class HipCsrMatrix<double>;

}  // namespace libMesh
