// numeric_vector.h -- NumericVector<T> (include/numerics/numeric_vector.h:69-880):
// the reference's pure virtuals with their signatures, and the non-virtual
// helpers it implements on top of them (src/numerics/numeric_vector.C);
// HipVector<T>, a serial vector whose values live in device memory with a
// host mirror synchronised on demand: entry writes (set, add, insert,
// add_vector) go to the mirror and close() pushes them; whole-vector
// operations and reductions run on the device (lmh_vector_op /
// lmh_vector_reduce).
#pragma once

#include <iostream>
#include <memory>
#include <set>
#include <string>
#include <vector>

#include "libmesh/dense_matrix.h"
#include "libmesh/hip_device.h"
#include "libmesh/parallel_object.h"

namespace libMesh {

template <typename T>
class SparseMatrix;
template <typename T>
class ShellMatrix;

template <typename T>
class NumericVector : public ParallelObject {
 public:
  explicit NumericVector(const Parallel::Communicator &comm_in, const ParallelType ptype = AUTOMATIC)
      : ParallelObject(comm_in), _type(ptype) {}
  explicit NumericVector(const Parallel::Communicator &comm_in, const numeric_index_type n,
                         const ParallelType ptype = AUTOMATIC)
      : ParallelObject(comm_in), _type(ptype) {
    (void)n;
  }
  NumericVector(const Parallel::Communicator &comm_in, const numeric_index_type n, const numeric_index_type n_local,
                const ParallelType ptype = AUTOMATIC)
      : ParallelObject(comm_in), _type(ptype) {
    (void)n;
    (void)n_local;
  }
  NumericVector(const Parallel::Communicator &comm_in, const numeric_index_type N, const numeric_index_type n_local,
                const std::vector<numeric_index_type> &ghost, const ParallelType ptype = AUTOMATIC)
      : ParallelObject(comm_in), _type(ptype) {
    (void)N;
    (void)n_local;
    (void)ghost;
  }
  virtual NumericVector<T> &operator=(const NumericVector<T> &v) = 0;
  virtual ~NumericVector() = default;

  // NumericVector::build (numeric_vector.C:47-90): the GPU vector for
  // HIP_SOLVERS (the only package this library provides)
  static std::unique_ptr<NumericVector<T>> build(const Parallel::Communicator &comm,
                                                 SolverPackage solver_package = HIP_SOLVERS,
                                                 ParallelType parallel_type = AUTOMATIC);

  virtual bool initialized() const { return _is_initialized; }
  ParallelType type() const { return _type; }
  void set_type(ParallelType t) { _type = t; }
  bool is_effectively_serial() const { return this->n_processors() == 1 || _type == SERIAL; }
  virtual bool closed() const { return _is_closed; }
  virtual void close() = 0;
  virtual void clear() {
    _is_closed = false;
    _is_initialized = false;
  }
  virtual void zero() = 0;
  virtual std::unique_ptr<NumericVector<T>> zero_clone() const = 0;
  virtual std::unique_ptr<NumericVector<T>> clone() const = 0;
  virtual void init(const numeric_index_type n, const numeric_index_type n_local, const bool fast = false,
                    const ParallelType ptype = AUTOMATIC) = 0;
  virtual void init(const numeric_index_type n, const bool fast = false, const ParallelType ptype = AUTOMATIC) = 0;
  virtual void init(const numeric_index_type n, const numeric_index_type n_local,
                    const std::vector<numeric_index_type> &ghost, const bool fast = false,
                    const ParallelType ptype = AUTOMATIC) = 0;
  virtual void init(const NumericVector<T> &other, const bool fast = false) = 0;
  virtual NumericVector<T> &operator=(const T s) = 0;
  virtual NumericVector<T> &operator=(const std::vector<T> &v) = 0;
  virtual Real min() const = 0;
  virtual Real max() const = 0;
  virtual T sum() const = 0;
  virtual Real l1_norm() const = 0;
  virtual Real l2_norm() const = 0;
  virtual Real linfty_norm() const = 0;
  // subset norms (numeric_vector.C:322-375)
  virtual Real subset_l1_norm(const std::set<numeric_index_type> &indices) const;
  virtual Real subset_l2_norm(const std::set<numeric_index_type> &indices) const;
  virtual Real subset_linfty_norm(const std::set<numeric_index_type> &indices) const;
  // |this - other| (numeric_vector.C:379-420)
  Real l2_norm_diff(const NumericVector<T> &other_vec) const;
  Real l1_norm_diff(const NumericVector<T> &other_vec) const;
  virtual numeric_index_type size() const = 0;
  virtual numeric_index_type local_size() const = 0;
  virtual numeric_index_type first_local_index() const = 0;
  virtual numeric_index_type last_local_index() const = 0;
  virtual T operator()(const numeric_index_type i) const = 0;
  virtual T el(const numeric_index_type i) const { return (*this)(i); }
  virtual void get(const std::vector<numeric_index_type> &index, T *values) const;
  void get(const std::vector<numeric_index_type> &index, std::vector<T> &values) const;
  virtual NumericVector<T> &operator+=(const NumericVector<T> &v) = 0;
  virtual NumericVector<T> &operator-=(const NumericVector<T> &v) = 0;
  NumericVector<T> &operator*=(const T a) {
    this->scale(a);
    return *this;
  }
  virtual NumericVector<T> &operator*=(const NumericVector<T> &v) = 0;
  NumericVector<T> &operator/=(const T a) {
    this->scale(1. / a);
    return *this;
  }
  virtual NumericVector<T> &operator/=(const NumericVector<T> &v) = 0;
  virtual void reciprocal() = 0;
  virtual void conjugate() = 0;
  virtual void set(const numeric_index_type i, const T value) = 0;
  virtual void add(const numeric_index_type i, const T value) = 0;
  virtual void add(const T s) = 0;
  virtual void add(const NumericVector<T> &v) = 0;
  virtual void add(const T a, const NumericVector<T> &v) = 0;
  // add_vector / insert families (numeric_vector.C:640-760)
  virtual void add_vector(const T *v, const std::vector<numeric_index_type> &dof_indices);
  void add_vector(const std::vector<T> &v, const std::vector<numeric_index_type> &dof_indices);
  void add_vector(const NumericVector<T> &v, const std::vector<numeric_index_type> &dof_indices);
  void add_vector(const DenseVector<T> &v, const std::vector<numeric_index_type> &dof_indices);
  virtual void add_vector(const NumericVector<T> &v, const SparseMatrix<T> &A) = 0;
  void add_vector(const NumericVector<T> &v, const ShellMatrix<T> &A);
  virtual void add_vector_transpose(const NumericVector<T> &v, const SparseMatrix<T> &A) = 0;
  virtual void insert(const T *v, const std::vector<numeric_index_type> &dof_indices);
  void insert(const std::vector<T> &v, const std::vector<numeric_index_type> &dof_indices);
  void insert(const NumericVector<T> &v, const std::vector<numeric_index_type> &dof_indices);
  void insert(const DenseVector<T> &v, const std::vector<numeric_index_type> &dof_indices);
  void insert(const DenseSubVector<T> &v, const std::vector<numeric_index_type> &dof_indices);
  virtual void scale(const T factor) = 0;
  virtual void abs() = 0;
  virtual T dot(const NumericVector<T> &v) const = 0;
  virtual void localize(std::vector<T> &v_local) const = 0;
  virtual void localize(NumericVector<T> &v_local) const = 0;
  virtual void localize(NumericVector<T> &v_local, const std::vector<numeric_index_type> &send_list) const = 0;
  virtual void localize(std::vector<T> &v_local, const std::vector<numeric_index_type> &indices) const = 0;
  virtual void localize(const numeric_index_type first_local_idx, const numeric_index_type last_local_idx,
                        const std::vector<numeric_index_type> &send_list) = 0;
  virtual void localize_to_one(std::vector<T> &v_local, const processor_id_type proc_id = 0) const = 0;
  // numeric_vector.C:110-200: index of the first differing entry, -1 if none
  virtual int compare(const NumericVector<T> &other_vector, const Real threshold = TOLERANCE) const;
  virtual int local_relative_compare(const NumericVector<T> &other_vector, const Real threshold = TOLERANCE) const;
  virtual int global_relative_compare(const NumericVector<T> &other_vector, const Real threshold = TOLERANCE) const;
  virtual void pointwise_mult(const NumericVector<T> &vec1, const NumericVector<T> &vec2) = 0;
  virtual void pointwise_divide(const NumericVector<T> &vec1, const NumericVector<T> &vec2) = 0;
  virtual void print(std::ostream &os = std::cout) const;
  virtual void print_global(std::ostream &os = std::cout) const;
  // Matlab text (numeric_vector.C:218-318): "%Vector of size N" then
  // "Vec_0 = [" and one entry per line
  virtual void print_matlab(const std::string &filename = "") const;
  virtual void read_matlab(const std::string &filename);
  virtual void create_subvector(NumericVector<T> &, const std::vector<numeric_index_type> &,
                                bool = true) const {
    libmesh_not_implemented();
  }
  virtual std::unique_ptr<NumericVector<T>> get_subvector(const std::vector<numeric_index_type> &) {
    libmesh_not_implemented();
  }
  virtual void restore_subvector(std::unique_ptr<NumericVector<T>>, const std::vector<numeric_index_type> &) {
    libmesh_not_implemented();
  }
  virtual void swap(NumericVector<T> &v);
  virtual std::size_t max_allowed_id() const = 0;
  bool readable() const { return this->initialized() && this->closed(); }
  bool compatible(const NumericVector<T> &v) const {
    return this->readable() && v.readable() && this->size() == v.size();
  }

 protected:
  bool _is_closed = false;
  bool _is_initialized = false;
  ParallelType _type;
};

// relative_fuzzy_equals (include/base/libmesh_common.h / fuzzy_equals.h) on vectors
template <typename T>
bool relative_fuzzy_equals(const NumericVector<T> &a, const NumericVector<T> &b, Real tol = TOLERANCE * TOLERANCE) {
  return a.l1_norm_diff(b) <= tol * std::max(a.l1_norm(), b.l1_norm());
}

template <typename T>
class HipVector : public NumericVector<T> {
 public:
  explicit HipVector(const Parallel::Communicator &comm, const ParallelType type = AUTOMATIC);
  explicit HipVector(const Parallel::Communicator &comm, const numeric_index_type n,
                     const ParallelType type = AUTOMATIC);
  HipVector(const Parallel::Communicator &comm, const numeric_index_type n, const numeric_index_type n_local,
            const ParallelType type = AUTOMATIC);
  HipVector(const Parallel::Communicator &comm, const numeric_index_type N, const numeric_index_type n_local,
            const std::vector<numeric_index_type> &ghost, const ParallelType type = AUTOMATIC);
  ~HipVector() override;
  HipVector(const HipVector &) = delete;
  HipVector &operator=(const HipVector &v);

  void close() override;
  void clear() override;
  void zero() override;
  std::unique_ptr<NumericVector<T>> zero_clone() const override;
  std::unique_ptr<NumericVector<T>> clone() const override;
  void init(const numeric_index_type n, const numeric_index_type n_local, const bool fast = false,
            const ParallelType ptype = AUTOMATIC) override;
  void init(const numeric_index_type n, const bool fast = false, const ParallelType ptype = AUTOMATIC) override;
  void init(const numeric_index_type n, const numeric_index_type n_local,
            const std::vector<numeric_index_type> &ghost, const bool fast = false,
            const ParallelType ptype = AUTOMATIC) override;
  void init(const NumericVector<T> &other, const bool fast = false) override;
  NumericVector<T> &operator=(const T s) override;
  NumericVector<T> &operator=(const NumericVector<T> &v) override;
  NumericVector<T> &operator=(const std::vector<T> &v) override;
  Real min() const override;
  Real max() const override;
  T sum() const override;
  Real l1_norm() const override;
  Real l2_norm() const override;
  Real linfty_norm() const override;
  numeric_index_type size() const override { return _n; }
  numeric_index_type local_size() const override { return _n; }
  numeric_index_type first_local_index() const override { return 0; }
  numeric_index_type last_local_index() const override { return _n; }
  T operator()(const numeric_index_type i) const override;
  NumericVector<T> &operator+=(const NumericVector<T> &v) override;
  NumericVector<T> &operator-=(const NumericVector<T> &v) override;
  NumericVector<T> &operator*=(const NumericVector<T> &v) override;
  NumericVector<T> &operator/=(const NumericVector<T> &v) override;
  using NumericVector<T>::operator*=;
  using NumericVector<T>::operator/=;
  void reciprocal() override;
  void conjugate() override {}  // real numbers
  void set(const numeric_index_type i, const T value) override;
  void add(const numeric_index_type i, const T value) override;
  void add(const T s) override;
  void add(const NumericVector<T> &v) override;
  void add(const T a, const NumericVector<T> &v) override;
  using NumericVector<T>::add_vector;
  void add_vector(const T *v, const std::vector<numeric_index_type> &dof_indices) override;
  // this += A v and this += A^T v, on the device for a HipCsrMatrix
  void add_vector(const NumericVector<T> &v, const SparseMatrix<T> &A) override;
  void add_vector_transpose(const NumericVector<T> &v, const SparseMatrix<T> &A) override;
  using NumericVector<T>::insert;
  void insert(const T *v, const std::vector<numeric_index_type> &dof_indices) override;
  void scale(const T factor) override;
  void abs() override;
  T dot(const NumericVector<T> &v) const override;
  void localize(std::vector<T> &v_local) const override;
  void localize(NumericVector<T> &v_local) const override;
  void localize(NumericVector<T> &v_local, const std::vector<numeric_index_type> &send_list) const override;
  void localize(std::vector<T> &v_local, const std::vector<numeric_index_type> &indices) const override;
  void localize(const numeric_index_type first_local_idx, const numeric_index_type last_local_idx,
                const std::vector<numeric_index_type> &send_list) override;
  void localize_to_one(std::vector<T> &v_local, const processor_id_type proc_id = 0) const override;
  void pointwise_mult(const NumericVector<T> &vec1, const NumericVector<T> &vec2) override;
  void pointwise_divide(const NumericVector<T> &vec1, const NumericVector<T> &vec2) override;
  void swap(NumericVector<T> &v) override;
  std::size_t max_allowed_id() const override { return (std::size_t)INT32_MAX; }

  // device storage (pushes pending host writes first)
  T *device_data();
  const T *device_data() const;
  // mark device values as newer than the host mirror (after a kernel wrote them)
  void mark_device_modified() { _host_valid = false; }
  lmh_context *context() const { return _ctx; }

 private:
  void allocate(numeric_index_type n);
  void pull() const;  // device -> host mirror if stale
  const T *device_data_synced() const;  // pushes pending host writes of a const operand
  const HipVector<T> &same_size(const NumericVector<T> &v, const char *what) const;
  T reduce(int op, const HipVector<T> *other) const;
  void device_op(int op, T a, const HipVector<T> *x, const HipVector<T> *y);
  lmh_context *_ctx;
  numeric_index_type _n = 0;
  T *_d = nullptr;
  mutable std::vector<T> _h;
  mutable bool _host_valid = true;
  bool _host_dirty = false;
};

}  // namespace libMesh
