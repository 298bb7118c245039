// libmesh_common.h -- the subset of libMesh's basic types, enums and error
// conventions that the GPU drop-in classes use (include/base/libmesh_common.h,
// include/enums/*.h, include/base/libmesh_exceptions.h).  Values of the enums
// match libMesh so code written against libMesh compiles unchanged.
#pragma once

#include <algorithm>
#include <cstdint>
#include <sstream>
#include <stdexcept>
#include <string>

namespace libMesh {

typedef double Real;
typedef double Number;
typedef uint32_t dof_id_type;
typedef int16_t boundary_id_type;
typedef uint64_t numeric_index_type;

// libmesh_common.h:172 (double precision build)
static constexpr Real TOLERANCE = 1.e-6;

// include/enums/enum_elem_type.h
enum ElemType : int { TET4 = 8, TET10 = 9, HEX8 = 10, HEX20 = 11, HEX27 = 12, INVALID_ELEM = 99 };
// include/enums/enum_order.h
enum Order : int { CONSTANT = 0, FIRST = 1, SECOND = 2, THIRD = 3, FOURTH = 4, FIFTH = 5 };
// include/enums/enum_fe_family.h (only LAGRANGE is on this path)
enum FEFamily : int { LAGRANGE = 0 };
// include/enums/enum_solver_type.h
enum SolverType : int { CG = 0, BICGSTAB = 8, INVALID_SOLVER = 99 };
// include/enums/enum_preconditioner_type.h
enum PreconditionerType : int { IDENTITY_PRECOND = 0, JACOBI_PRECOND = 1, INVALID_PRECONDITIONER = 99 };
// include/enums/enum_convergence_flags.h:33
enum LinearConvergenceReason : int {
  CONVERGED_RTOL = 2,
  CONVERGED_ATOL = 3,
  CONVERGED_ITS = 4,
  DIVERGED_NULL = -2,
  DIVERGED_ITS = -3,
  DIVERGED_DTOL = -4,
  DIVERGED_BREAKDOWN = -5,
  DIVERGED_NANORINF = -9,
  UNKNOWN_FLAG = -128
};
// include/enums/enum_solver_package.h:34 has no GPU package; the drop-in
// classes report this value and are injected explicitly.
enum SolverPackage : int { PETSC_SOLVERS = 0, TRILINOS_SOLVERS, LASPACK_SOLVERS, SLEPC_SOLVERS, EIGEN_SOLVERS,
                           NLOPT_SOLVERS, INVALID_SOLVER_PACKAGE, HIP_SOLVERS = 100 };
enum ParallelType : int { AUTOMATIC = 0, SERIAL, PARALLEL, GHOSTED, INVALID_PARALLELIZATION };
// include/enums/enum_subset_solve_mode.h
enum SubsetSolveMode : int { SUBSET_ZERO = 0, SUBSET_COPY_RHS, SUBSET_DONT_TOUCH };

// include/base/libmesh_exceptions.h:65, 116
class LogicError : public std::logic_error {
 public:
  explicit LogicError(const std::string &msg = "Error in libMesh internal logic") : std::logic_error(msg) {}
};
class DegenerateMap : public LogicError {
 public:
  explicit DegenerateMap(const std::string &msg = "Problem with the mapping") : LogicError(msg) {}
};
class NotImplemented : public std::logic_error {
 public:
  NotImplemented() : std::logic_error("Error: functionality not implemented!") {}
};

#define libmesh_error_msg(msg)                                                 \
  do {                                                                         \
    std::ostringstream _s;                                                     \
    _s << msg;                                                                 \
    throw ::libMesh::LogicError(_s.str());                                     \
  } while (0)
#define libmesh_not_implemented() throw ::libMesh::NotImplemented()

}  // namespace libMesh
