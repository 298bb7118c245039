// shell_matrix.h -- ShellMatrix<T> (include/numerics/shell_matrix.h:50-110):
// an operator known only by its action on vectors.  The GPU solver takes
// assembled HipCsrMatrix operators; its ShellMatrix solve overloads are not
// implemented, as in EigenSparseLinearSolver (eigen_sparse_linear_solver.C:355-390).
#pragma once

#include "libmesh/parallel_object.h"
#include "libmesh/libmesh_common.h"

namespace libMesh {

template <typename T>
class NumericVector;

template <typename T>
class ShellMatrix : public ParallelObject {
 public:
  ShellMatrix(const Parallel::Communicator &comm_in) : ParallelObject(comm_in) {}
  virtual ~ShellMatrix() = default;
  virtual numeric_index_type m() const = 0;
  virtual numeric_index_type n() const = 0;
  virtual void vector_mult(NumericVector<T> &dest, const NumericVector<T> &arg) const = 0;
  virtual void vector_mult_add(NumericVector<T> &dest, const NumericVector<T> &arg) const = 0;
  virtual void get_diagonal(NumericVector<T> &dest) const = 0;
  virtual void clear() { libmesh_error_msg("Not implemented yet"); }
  virtual void init() { libmesh_error_msg("Not implemented yet"); }
};

}  // namespace libMesh
