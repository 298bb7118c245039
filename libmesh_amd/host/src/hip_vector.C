// hip_vector.C -- device-resident NumericVector with a lazily synchronised
// host mirror.  Entry writes (set, add, insert, add_vector) edit the mirror;
// close() pushes them.  Whole-vector operations, norms and dot products run
// on the device after pushing pending writes; reads of single entries pull
// the device values into the mirror first.
#include <cmath>

#include "libmesh/hip_csr_matrix.h"
#include "libmesh/numeric_vector.h"

namespace libMesh {

template <typename T>
HipVector<T>::HipVector(const Parallel::Communicator &comm, const ParallelType type)
    : NumericVector<T>(comm, type), _ctx(HipDevice::context(0)) {}

template <typename T>
HipVector<T>::HipVector(const Parallel::Communicator &comm, const numeric_index_type n, const ParallelType type)
    : NumericVector<T>(comm, type), _ctx(HipDevice::context(0)) {
  init(n, n, false, type);
}

template <typename T>
HipVector<T>::HipVector(const Parallel::Communicator &comm, const numeric_index_type n,
                        const numeric_index_type n_local, const ParallelType type)
    : NumericVector<T>(comm, type), _ctx(HipDevice::context(0)) {
  init(n, n_local, false, type);
}

template <typename T>
HipVector<T>::HipVector(const Parallel::Communicator &comm, const numeric_index_type N,
                        const numeric_index_type n_local, const std::vector<numeric_index_type> &ghost,
                        const ParallelType type)
    : NumericVector<T>(comm, type), _ctx(HipDevice::context(0)) {
  init(N, n_local, ghost, false, type);
}

template <typename T>
HipVector<T>::~HipVector() {
  if (_d) lmh_free(_ctx, _d);
}

template <typename T>
void HipVector<T>::allocate(numeric_index_type n) {
  if (_d) lmh_free(_ctx, _d);
  _d = nullptr;
  _n = n;
  void *p = nullptr;
  hip_check(lmh_malloc(_ctx, sizeof(T) * (n ? n : 1), &p), "HipVector::init");
  _d = (T *)p;
  _h.assign(n, T(0));
  hip_check(lmh_memset(_ctx, _d, 0, sizeof(T) * n), "HipVector::init");
  _host_valid = true;
  _host_dirty = false;
  this->_is_initialized = true;
  this->_is_closed = true;
}

// one process holds the whole vector: the local size is the size
template <typename T>
void HipVector<T>::init(const numeric_index_type n, const numeric_index_type n_local, const bool,
                        const ParallelType ptype) {
  if (n_local != n) libmesh_error_msg("HipVector::init: local size " << n_local << " != size " << n << " on one rank");
  if (ptype != AUTOMATIC) this->_type = ptype;
  allocate(n);
}

template <typename T>
void HipVector<T>::init(const numeric_index_type n, const bool fast, const ParallelType ptype) {
  init(n, n, fast, ptype);
}

template <typename T>
void HipVector<T>::init(const numeric_index_type n, const numeric_index_type n_local,
                        const std::vector<numeric_index_type> &ghost, const bool fast, const ParallelType ptype) {
  for (auto g : ghost)
    if (g >= n) libmesh_error_msg("HipVector::init: ghost index " << g << " beyond size " << n);
  init(n, n_local, fast, ptype);
}

template <typename T>
void HipVector<T>::init(const NumericVector<T> &other, const bool fast) {
  init(other.size(), other.local_size(), fast, other.type());
}

template <typename T>
void HipVector<T>::clear() {
  if (_d) lmh_free(_ctx, _d);
  _d = nullptr;
  _n = 0;
  _h.clear();
  _host_valid = true;
  _host_dirty = false;
  NumericVector<T>::clear();
}

template <typename T>
void HipVector<T>::pull() const {
  if (_host_valid) return;
  if (_n) hip_check(lmh_memcpy_d2h(_ctx, _h.data(), _d, sizeof(T) * _n), "HipVector: device to host");
  _host_valid = true;
}

template <typename T>
void HipVector<T>::zero() {
  hip_check(lmh_memset(_ctx, _d, 0, sizeof(T) * _n), "HipVector::zero");
  std::fill(_h.begin(), _h.end(), T(0));
  _host_valid = true;
  _host_dirty = false;
}

template <typename T>
void HipVector<T>::close() {
  if (_host_dirty) {
    hip_check(lmh_memcpy_h2d(_ctx, _d, _h.data(), sizeof(T) * _n), "HipVector::close");
    _host_dirty = false;
  }
  this->_is_closed = true;
}

template <typename T>
const HipVector<T> &HipVector<T>::same_size(const NumericVector<T> &v, const char *what) const {
  const auto *hv = dynamic_cast<const HipVector<T> *>(&v);
  if (!hv) libmesh_error_msg("HipVector::" << what << ": the other vector must be a HipVector");
  if (hv->size() != _n) libmesh_error_msg("HipVector::" << what << ": sizes " << _n << " and " << hv->size());
  return *hv;
}

template <typename T>
void HipVector<T>::device_op(int op, T a, const HipVector<T> *x, const HipVector<T> *y) {
  T *z = device_data();
  hip_check(lmh_vector_op(_ctx, op, (int64_t)_n, (double)a, x ? (const double *)x->device_data_synced() : nullptr,
                          y ? (const double *)y->device_data_synced() : nullptr, (double *)z),
            "HipVector (device operation)");
  mark_device_modified();
}

template <typename T>
T HipVector<T>::reduce(int op, const HipVector<T> *other) const {
  double out = 0;
  hip_check(lmh_vector_reduce(_ctx, op, (int64_t)_n, (const double *)device_data_synced(),
                              other ? (const double *)other->device_data_synced() : nullptr, &out),
            "HipVector (device reduction)");
  return (T)out;
}

template <typename T>
void HipVector<T>::set(const numeric_index_type i, const T v) {
  if (i >= _n) libmesh_error_msg("HipVector::set: index " << i << " beyond size " << _n);
  pull();
  _h[i] = v;
  _host_dirty = true;
  this->_is_closed = false;
}

template <typename T>
void HipVector<T>::add(const numeric_index_type i, const T v) {
  if (i >= _n) libmesh_error_msg("HipVector::add: index " << i << " beyond size " << _n);
  pull();
  _h[i] += v;
  _host_dirty = true;
  this->_is_closed = false;
}

template <typename T>
T HipVector<T>::operator()(const numeric_index_type i) const {
  if (i >= _n) libmesh_error_msg("HipVector: index " << i << " beyond size " << _n);
  pull();
  return _h[i];
}

template <typename T>
void HipVector<T>::add_vector(const T *v, const std::vector<numeric_index_type> &dof_indices) {
  pull();
  for (std::size_t k = 0; k < dof_indices.size(); ++k) {
    if (dof_indices[k] >= _n) libmesh_error_msg("HipVector::add_vector: index " << dof_indices[k] << " beyond size " << _n);
    _h[dof_indices[k]] += v[k];
  }
  _host_dirty = true;
  this->_is_closed = false;
}

template <typename T>
void HipVector<T>::insert(const T *v, const std::vector<numeric_index_type> &dof_indices) {
  pull();
  for (std::size_t k = 0; k < dof_indices.size(); ++k) {
    if (dof_indices[k] >= _n) libmesh_error_msg("HipVector::insert: index " << dof_indices[k] << " beyond size " << _n);
    _h[dof_indices[k]] = v[k];
  }
  _host_dirty = true;
  this->_is_closed = false;
}

// this += A v (the product on the device into a temporary), this += A^T v
template <typename T>
void HipVector<T>::add_vector(const NumericVector<T> &v, const SparseMatrix<T> &A) {
  auto *H = dynamic_cast<const HipCsrMatrix<T> *>(&A);
  if (!H) libmesh_error_msg("HipVector::add_vector(v, A): a HipCsrMatrix is required");
  const auto *hv = dynamic_cast<const HipVector<T> *>(&v);
  if (!hv) libmesh_error_msg("HipVector::add_vector(v, A): v must be a HipVector");
  HipVector<T> prod(this->comm(), A.m(), A.m());
  H->device_vector_mult(prod, *hv);
  *this += prod;
}

template <typename T>
void HipVector<T>::add_vector_transpose(const NumericVector<T> &v, const SparseMatrix<T> &A) {
  auto *H = dynamic_cast<const HipCsrMatrix<T> *>(&A);
  if (!H) libmesh_error_msg("HipVector::add_vector_transpose: a HipCsrMatrix is required");
  HipVector<T> prod(this->comm(), A.n(), A.n());
  H->vector_mult_transpose(prod, v);
  *this += prod;
}

template <typename T>
NumericVector<T> &HipVector<T>::operator=(const T s) {
  device_op(LMH_VEC_SET, s, nullptr, nullptr);
  return *this;
}

template <typename T>
HipVector<T> &HipVector<T>::operator=(const HipVector<T> &v) {
  if (&v == this) return *this;
  if (v.size() != _n) init(v.size(), v.size());
  device_op(LMH_VEC_COPY, T(0), &v, nullptr);
  return *this;
}

template <typename T>
NumericVector<T> &HipVector<T>::operator=(const NumericVector<T> &v) {
  if (const auto *hv = dynamic_cast<const HipVector<T> *>(&v)) return *this = *hv;
  std::vector<T> vals;
  v.localize(vals);
  return *this = vals;
}

template <typename T>
NumericVector<T> &HipVector<T>::operator=(const std::vector<T> &v) {
  if (v.size() != _n) libmesh_error_msg("HipVector::operator=: " << v.size() << " values for size " << _n);
  _h = v;
  _host_valid = true;
  _host_dirty = true;
  close();
  return *this;
}

template <typename T>
Real HipVector<T>::min() const {
  return (Real)reduce(LMH_VEC_MIN, nullptr);
}

template <typename T>
Real HipVector<T>::max() const {
  return (Real)reduce(LMH_VEC_MAX, nullptr);
}

template <typename T>
T HipVector<T>::sum() const {
  return reduce(LMH_VEC_SUM, nullptr);
}

template <typename T>
Real HipVector<T>::l1_norm() const {
  return (Real)reduce(LMH_VEC_L1, nullptr);
}

template <typename T>
Real HipVector<T>::l2_norm() const {
  return (Real)reduce(LMH_VEC_L2, nullptr);
}

template <typename T>
Real HipVector<T>::linfty_norm() const {
  return (Real)reduce(LMH_VEC_LINF, nullptr);
}

template <typename T>
NumericVector<T> &HipVector<T>::operator+=(const NumericVector<T> &v) {
  device_op(LMH_VEC_AXPY, T(1), &same_size(v, "operator+="), nullptr);
  return *this;
}

template <typename T>
NumericVector<T> &HipVector<T>::operator-=(const NumericVector<T> &v) {
  device_op(LMH_VEC_AXPY, T(-1), &same_size(v, "operator-="), nullptr);
  return *this;
}

template <typename T>
NumericVector<T> &HipVector<T>::operator*=(const NumericVector<T> &v) {
  device_op(LMH_VEC_MUL, T(0), &same_size(v, "operator*="), nullptr);
  return *this;
}

template <typename T>
NumericVector<T> &HipVector<T>::operator/=(const NumericVector<T> &v) {
  device_op(LMH_VEC_DIV, T(0), &same_size(v, "operator/="), nullptr);
  return *this;
}

template <typename T>
void HipVector<T>::reciprocal() {
  device_op(LMH_VEC_RECIPROCAL, T(0), nullptr, nullptr);
}

template <typename T>
void HipVector<T>::add(const T s) {
  device_op(LMH_VEC_ADD_SCALAR, s, nullptr, nullptr);
}

template <typename T>
void HipVector<T>::add(const NumericVector<T> &v) {
  *this += v;
}

template <typename T>
void HipVector<T>::add(const T a, const NumericVector<T> &v) {
  device_op(LMH_VEC_AXPY, a, &same_size(v, "add(a, v)"), nullptr);
}

template <typename T>
void HipVector<T>::scale(const T factor) {
  device_op(LMH_VEC_SCALE, factor, nullptr, nullptr);
}

template <typename T>
void HipVector<T>::abs() {
  device_op(LMH_VEC_ABS, T(0), nullptr, nullptr);
}

template <typename T>
T HipVector<T>::dot(const NumericVector<T> &v) const {
  return reduce(LMH_VEC_DOT, &same_size(v, "dot"));
}

template <typename T>
void HipVector<T>::pointwise_mult(const NumericVector<T> &vec1, const NumericVector<T> &vec2) {
  device_op(LMH_VEC_PMUL, T(0), &same_size(vec1, "pointwise_mult"), &same_size(vec2, "pointwise_mult"));
}

template <typename T>
void HipVector<T>::pointwise_divide(const NumericVector<T> &vec1, const NumericVector<T> &vec2) {
  device_op(LMH_VEC_PDIV, T(0), &same_size(vec1, "pointwise_divide"), &same_size(vec2, "pointwise_divide"));
}

template <typename T>
void HipVector<T>::localize(std::vector<T> &v) const {
  pull();
  v = _h;
}

template <typename T>
void HipVector<T>::localize(NumericVector<T> &v_local) const {
  v_local = *this;
}

template <typename T>
void HipVector<T>::localize(NumericVector<T> &v_local, const std::vector<numeric_index_type> &) const {
  v_local = *this;
}

template <typename T>
void HipVector<T>::localize(std::vector<T> &v_local, const std::vector<numeric_index_type> &indices) const {
  pull();
  v_local.resize(indices.size());
  for (std::size_t k = 0; k < indices.size(); ++k) {
    if (indices[k] >= _n) libmesh_error_msg("HipVector::localize: index " << indices[k] << " beyond size " << _n);
    v_local[k] = _h[indices[k]];
  }
}

// every entry is already local
template <typename T>
void HipVector<T>::localize(const numeric_index_type first_local_idx, const numeric_index_type last_local_idx,
                            const std::vector<numeric_index_type> &) {
  if (first_local_idx != 0 || last_local_idx + 1 != _n)
    libmesh_error_msg("HipVector::localize: local range [" << first_local_idx << ", " << last_local_idx
                                                           << "] is not the whole vector");
}

template <typename T>
void HipVector<T>::localize_to_one(std::vector<T> &v_local, const processor_id_type) const {
  localize(v_local);
}

template <typename T>
void HipVector<T>::swap(NumericVector<T> &other) {
  auto *hv = dynamic_cast<HipVector<T> *>(&other);
  if (!hv) libmesh_error_msg("HipVector::swap: the other vector must be a HipVector");
  NumericVector<T>::swap(other);
  std::swap(_ctx, hv->_ctx);
  std::swap(_n, hv->_n);
  std::swap(_d, hv->_d);
  std::swap(_h, hv->_h);
  std::swap(_host_valid, hv->_host_valid);
  std::swap(_host_dirty, hv->_host_dirty);
}

template <typename T>
std::unique_ptr<NumericVector<T>> HipVector<T>::clone() const {
  auto c = std::make_unique<HipVector<T>>(this->comm(), _n, _n, this->_type);
  *c = *this;
  return c;
}

template <typename T>
std::unique_ptr<NumericVector<T>> HipVector<T>::zero_clone() const {
  return std::make_unique<HipVector<T>>(this->comm(), _n, _n, this->_type);
}

template <typename T>
T *HipVector<T>::device_data() {
  close();
  return _d;
}

template <typename T>
const T *HipVector<T>::device_data() const {
  if (_host_dirty) libmesh_error_msg("HipVector: call close() before using the device data");
  return _d;
}

// pending host writes of a const operand are pushed (the mirror is mutable
// state of the same logical value, as close() would)
template <typename T>
const T *HipVector<T>::device_data_synced() const {
  if (_host_dirty) {
    hip_check(lmh_memcpy_h2d(_ctx, _d, _h.data(), sizeof(T) * _n), "HipVector (push)");
    const_cast<HipVector<T> *>(this)->_host_dirty = false;
  }
  return _d;
}

template class HipVector<double>;

}  // namespace libMesh
