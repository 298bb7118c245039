// dense_matrix.h -- DenseMatrix / DenseVector subset used by element assembly
// (include/numerics/dense_matrix.h, dense_vector.h): row-major storage,
// resize-to-zero, element access.
#pragma once

#include <vector>

#include "libmesh/libmesh_common.h"

namespace libMesh {

template <typename T>
class DenseMatrix {
 public:
  DenseMatrix(unsigned m = 0, unsigned n = 0) { resize(m, n); }
  void resize(unsigned m, unsigned n) {
    _m = m;
    _n = n;
    _val.assign((size_t)m * n, T(0));
  }
  void zero() { std::fill(_val.begin(), _val.end(), T(0)); }
  T &operator()(unsigned i, unsigned j) { return _val[(size_t)i * _n + j]; }
  T operator()(unsigned i, unsigned j) const { return _val[(size_t)i * _n + j]; }
  unsigned m() const { return _m; }
  unsigned n() const { return _n; }
  const std::vector<T> &get_values() const { return _val; }
  std::vector<T> &get_values() { return _val; }

 private:
  unsigned _m = 0, _n = 0;
  std::vector<T> _val;
};

template <typename T>
class DenseVector {
 public:
  DenseVector(unsigned n = 0) { resize(n); }
  void resize(unsigned n) { _val.assign(n, T(0)); }
  void zero() { std::fill(_val.begin(), _val.end(), T(0)); }
  T &operator()(unsigned i) { return _val[i]; }
  T operator()(unsigned i) const { return _val[i]; }
  unsigned size() const { return (unsigned)_val.size(); }
  const std::vector<T> &get_values() const { return _val; }
  std::vector<T> &get_values() { return _val; }

 private:
  std::vector<T> _val;
};

// DenseSubVector (include/numerics/dense_subvector.h): a window of
// `n` entries of a parent DenseVector starting at entry `ioff`
template <typename T>
class DenseSubVector {
 public:
  DenseSubVector(DenseVector<T> &parent, unsigned ioff = 0, unsigned n = 0) : _parent(parent), _i(ioff), _n(n) {}
  void reposition(unsigned ioff, unsigned n) {
    _i = ioff;
    _n = n;
  }
  T &operator()(unsigned i) { return _parent(_i + i); }
  T operator()(unsigned i) const { return _parent(_i + i); }
  unsigned size() const { return _n; }
  unsigned i_off() const { return _i; }

 private:
  DenseVector<T> &_parent;
  unsigned _i, _n;
};

}  // namespace libMesh
