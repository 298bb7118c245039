// hip_vector.C -- device-resident NumericVector with a lazily synchronised host mirror.
#include <cmath>

#include "libmesh/numeric_vector.h"

namespace libMesh {

template <typename T>
HipVector<T>::HipVector(numeric_index_type n, int device) : _ctx(HipDevice::context(device)) {
  if (n) init(n);
}

template <typename T>
HipVector<T>::~HipVector() {
  if (_d) lmh_free(_ctx, _d);
}

template <typename T>
void HipVector<T>::init(numeric_index_type n, bool) {
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
}

template <typename T>
void HipVector<T>::pull() const {
  if (_host_valid) return;
  hip_check(lmh_memcpy_d2h(_ctx, _h.data(), _d, sizeof(T) * _n), "HipVector: device to host");
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
}

template <typename T>
void HipVector<T>::set(numeric_index_type i, T v) {
  pull();
  _h.at(i) = v;
  _host_dirty = true;
}

template <typename T>
void HipVector<T>::add(numeric_index_type i, T v) {
  pull();
  _h.at(i) += v;
  _host_dirty = true;
}

template <typename T>
T HipVector<T>::operator()(numeric_index_type i) const {
  pull();
  return _h.at(i);
}

template <typename T>
void HipVector<T>::add_vector(const DenseVector<T> &v, const std::vector<dof_id_type> &dofs) {
  if (v.size() != dofs.size()) libmesh_error_msg("add_vector: size mismatch");
  pull();
  for (size_t k = 0; k < dofs.size(); ++k) _h.at(dofs[k]) += v(k);
  _host_dirty = true;
}

template <typename T>
Real HipVector<T>::l2_norm() const {
  if (_host_dirty) {
    Real s = 0;
    for (const T &x : _h) s += (Real)x * (Real)x;
    return std::sqrt(s);
  }
  return lmh_norm2(_ctx, (const double *)_d, (int64_t)_n);
}

template <typename T>
void HipVector<T>::localize(std::vector<T> &v) const {
  pull();
  v = _h;
}

template <typename T>
std::unique_ptr<NumericVector<T>> HipVector<T>::clone() const {
  std::unique_ptr<HipVector<T>> c(new HipVector<T>(0));
  c->init(_n);
  pull();
  c->_h = _h;
  c->_host_dirty = true;
  c->close();
  return c;
}

template <typename T>
std::unique_ptr<NumericVector<T>> HipVector<T>::zero_clone() const {
  std::unique_ptr<HipVector<T>> c(new HipVector<T>(0));
  c->init(_n);
  c->zero();
  c->close();
  return c;
}

template <typename T>
void HipVector<T>::add(T a, const NumericVector<T> &v) {
  const auto *hv = dynamic_cast<const HipVector<T> *>(&v);
  if (!hv || hv->size() != _n) libmesh_error_msg("HipVector::add(a, v): HipVector of the same size required");
  close();
  hip_check(lmh_axpy(_ctx, (int64_t)_n, (double)a, (const double *)hv->device_data(), (double *)_d),
            "HipVector::add");
  mark_device_modified();
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

template class HipVector<double>;

}  // namespace libMesh
