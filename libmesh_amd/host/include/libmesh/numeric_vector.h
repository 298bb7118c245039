// numeric_vector.h -- NumericVector<T> subset (include/numerics/numeric_vector.h)
// and HipVector<T>, a serial vector whose values live in device memory with a
// host mirror synchronised on demand (close() pushes host edits to the device).
#pragma once

#include <memory>
#include <vector>

#include "libmesh/dense_matrix.h"
#include "libmesh/hip_device.h"

namespace libMesh {

template <typename T>
class NumericVector {
 public:
  virtual ~NumericVector() = default;
  virtual void init(numeric_index_type n, bool fast = false) = 0;
  virtual numeric_index_type size() const = 0;
  virtual void zero() = 0;
  virtual void close() = 0;
  virtual void set(numeric_index_type i, T v) = 0;
  virtual void add(numeric_index_type i, T v) = 0;
  virtual T operator()(numeric_index_type i) const = 0;
  // NumericVector::add_vector(const T*, dofs) (src/numerics/numeric_vector.C:690-697)
  virtual void add_vector(const DenseVector<T> &v, const std::vector<dof_id_type> &dofs) = 0;
  virtual Real l2_norm() const = 0;
  virtual void localize(std::vector<T> &v) const = 0;
  virtual std::unique_ptr<NumericVector<T>> clone() const = 0;
  // NumericVector::zero_clone: same size, zero entries
  virtual std::unique_ptr<NumericVector<T>> zero_clone() const = 0;
  // NumericVector::add(a, v): this += a v
  virtual void add(T a, const NumericVector<T> &v) = 0;
};

template <typename T>
class HipVector : public NumericVector<T> {
 public:
  explicit HipVector(numeric_index_type n = 0, int device = 0);
  ~HipVector() override;
  HipVector(const HipVector &) = delete;
  HipVector &operator=(const HipVector &) = delete;

  void init(numeric_index_type n, bool fast = false) override;
  numeric_index_type size() const override { return _n; }
  void zero() override;
  void close() override;
  void set(numeric_index_type i, T v) override;
  void add(numeric_index_type i, T v) override;
  T operator()(numeric_index_type i) const override;
  void add_vector(const DenseVector<T> &v, const std::vector<dof_id_type> &dofs) override;
  Real l2_norm() const override;
  void localize(std::vector<T> &v) const override;
  std::unique_ptr<NumericVector<T>> clone() const override;
  std::unique_ptr<NumericVector<T>> zero_clone() const override;
  void add(T a, const NumericVector<T> &v) override;

  // device storage (valid after close())
  T *device_data();
  const T *device_data() const;
  // mark device values as newer than the host mirror (after a kernel wrote them)
  void mark_device_modified() { _host_valid = false; }
  lmh_context *context() const { return _ctx; }

 private:
  void pull() const;  // device -> host mirror if stale
  lmh_context *_ctx;
  numeric_index_type _n = 0;
  T *_d = nullptr;
  mutable std::vector<T> _h;
  mutable bool _host_valid = true;
  bool _host_dirty = false;
};

}  // namespace libMesh
