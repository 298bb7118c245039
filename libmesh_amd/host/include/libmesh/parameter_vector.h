// parameter_vector.h -- ParameterAccessor, ParameterPointer and
// ParameterVector (include/systems/parameter_accessor.h, parameter_pointer.h,
// parameter_vector.h): the parameters a sensitivity computation
// differentiates with respect to, each reached through an accessor that can
// read and write it.
#pragma once

#include <memory>
#include <vector>

#include "libmesh/libmesh_common.h"

namespace libMesh {

template <typename T>
class ParameterAccessor;

// ParameterProxy (parameter_accessor.h): *accessor reads and assigns the parameter
template <typename T>
class ParameterProxy {
 public:
  ParameterProxy(ParameterAccessor<T> &accessor) : _accessor(accessor) {}
  ParameterProxy &operator=(const ParameterProxy<T> &new_value) {
    _accessor.set(new_value.get());
    return *this;
  }
  ParameterProxy &operator=(const T &new_value) {
    _accessor.set(new_value);
    return *this;
  }
  ParameterProxy &operator+=(const T &v) { return *this = this->get() + v; }
  ParameterProxy &operator-=(const T &v) { return *this = this->get() - v; }
  ParameterProxy &operator*=(const T &v) { return *this = this->get() * v; }
  operator T() const { return _accessor.get(); }
  T get() const { return _accessor.get(); }

 private:
  ParameterAccessor<T> &_accessor;
};

template <typename T>
class ParameterAccessor {
 public:
  virtual ~ParameterAccessor() = default;
  virtual void set(const T &new_value) = 0;
  virtual const T &get() const = 0;
  ParameterProxy<T> operator*() { return ParameterProxy<T>(*this); }
  T operator*() const { return this->get(); }
  virtual std::unique_ptr<ParameterAccessor<T>> clone() const = 0;
};

template <typename T>
class ParameterPointer : public ParameterAccessor<T> {
 public:
  ParameterPointer(T *param_ptr) : _ptr(param_ptr) {}
  void set(const T &new_value) override { *_ptr = new_value; }
  const T &get() const override { return *_ptr; }
  std::unique_ptr<ParameterAccessor<T>> clone() const override {
    return std::make_unique<ParameterPointer<T>>(_ptr);
  }

 private:
  T *_ptr;
};

class ParameterVector {
 public:
  ParameterVector() = default;
  explicit ParameterVector(const std::vector<Number *> &params) {
    for (Number *p : params) _params.push_back(std::make_unique<ParameterPointer<Number>>(p));
  }
  std::size_t size() const { return _params.size(); }
  void clear() { _params.clear(); }
  void push_back(std::unique_ptr<ParameterAccessor<Number>> new_accessor) {
    if (!new_accessor) libmesh_error_msg("ParameterVector::push_back: null accessor");
    _params.push_back(std::move(new_accessor));
  }
  const ParameterAccessor<Number> &operator[](std::size_t i) const { return *_params.at(i); }
  ParameterAccessor<Number> &operator[](std::size_t i) { return *_params.at(i); }

 private:
  std::vector<std::unique_ptr<ParameterAccessor<Number>>> _params;
};

}  // namespace libMesh
