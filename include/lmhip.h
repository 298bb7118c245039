/*
 * lmhip.h -- C ABI of the MI355X-native libMesh assembly + Jacobi-CG hot path.
 *
 * Every entry point takes plain pointers, sizes and opaque handles and returns
 * an int status (LMH_OK = 0).  The libMesh-facing C++ classes (HipCsrMatrix,
 * HipLinearSolver, HipAssembler) call only these functions; the comment on each
 * names the reference interface it replaces (paths relative to the libMesh
 * source tree).
 *
 * Pointer conventions: arguments suffixed _h are host pointers, _d are device
 * pointers (allocated with lmh_malloc or by any HIP allocator on the context's
 * device).  All work is enqueued on the context's stream; results in device
 * memory are ready after lmh_synchronize or any call that returns host data.
 */
#ifndef LMHIP_H
#define LMHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (mapped by the C++ shim to libMesh's exceptions:
 * LMH_ERR_INVALID -> LogicError (libmesh_common.h:441),
 * LMH_ERR_DEGENERATE -> DegenerateMap (fe_map.C:467-505),
 * LMH_ERR_UNSUPPORTED -> libmesh_not_implemented) ---- */
enum {
  LMH_OK = 0,
  LMH_ERR_INVALID = 1,
  LMH_ERR_HIP = 2,
  LMH_ERR_DEGENERATE = 3,
  LMH_ERR_ALLOC = 4,
  LMH_ERR_UNSUPPORTED = 5
};

/* libMesh ElemType values (include/enums/enum_elem_type.h). */
enum { LMH_TET4 = 8, LMH_TET10 = 9, LMH_HEX8 = 10, LMH_HEX27 = 12 };

/* Arithmetic of a call. */
enum { LMH_FP64 = 0, LMH_FP32 = 1 };

/* Scatter strategy for Ke/Fe -> CSR. */
enum { LMH_SCATTER_COLORED = 0, LMH_SCATTER_GATHER = 1 };

/* Converged reasons (subset of LinearConvergenceReason,
 * include/enums/enum_convergence_flags.h:33). */
enum { LMH_CONVERGED_RTOL = 2, LMH_DIVERGED_ITS = -3, LMH_DIVERGED_NANORINF = -9, LMH_DIVERGED_BREAKDOWN = -5 };

typedef struct lmh_context lmh_context;
typedef struct lmh_mesh lmh_mesh;
typedef struct lmh_dofmap lmh_dofmap;
typedef struct lmh_matrix lmh_matrix;
typedef struct lmh_batch lmh_batch;
typedef struct lmh_condensation lmh_condensation;

/* Last error message of the calling thread ("" if none). */
const char *lmh_last_error(void);

/* ---- context: one device + one stream (LibMeshInit / SolverPackage selection,
 * src/base/libmesh.C:1064-1122) ---- */
int lmh_context_create(int device, lmh_context **out);
int lmh_context_destroy(lmh_context *ctx);
int lmh_synchronize(lmh_context *ctx);
/* hipStream_t of the context, as void* */
void *lmh_context_stream(lmh_context *ctx);

/* ---- device memory (NumericVector storage, numeric_vector.h) ---- */
int lmh_malloc(lmh_context *ctx, size_t bytes, void **d_out);
int lmh_free(lmh_context *ctx, void *d);
int lmh_memcpy_h2d(lmh_context *ctx, void *d_dst, const void *h_src, size_t bytes);
int lmh_memcpy_d2h(lmh_context *ctx, void *h_dst, const void *d_src, size_t bytes);
int lmh_memcpy_d2d(lmh_context *ctx, void *d_dst, const void *d_src, size_t bytes);
int lmh_memset(lmh_context *ctx, void *d, int value, size_t bytes);

/* ---- mesh: element connectivity + nodal coordinates, the data FE::reinit reads
 * through Elem::point (src/fe/fe.C:197-368, fe_map.C:958-1090).
 * coords_h: n_nodes*3 FP64; conn_h: n_elem*nodes_per_elem (libMesh local node
 * order, 0-based node ids). ---- */
int lmh_mesh_create(lmh_context *ctx, int elem_type, int64_t n_nodes, const double *coords_h, int64_t n_elem,
                    const int32_t *conn_h, lmh_mesh **out);
int lmh_mesh_destroy(lmh_mesh *m);
/* Replace the nodal coordinates (e.g. after MeshTools::Modification). */
int lmh_mesh_set_coords(lmh_mesh *m, const double *coords_h);

/* ---- DoF map: DofMap::dof_indices per element (dof_map.C:2201-2316) and the
 * Dirichlet-constrained set (DofMap::is_constrained_dof after
 * create_dof_constraints, dof_map_constraints.C:592-740).
 * elem_dofs_h: n_elem * (n_vars*nodes_per_elem), variable-major per element;
 * constrained_h: n_dofs bytes, nonzero = constrained (may be NULL). ---- */
int lmh_dofmap_create(lmh_context *ctx, const lmh_mesh *m, int n_vars, int64_t n_dofs, const int32_t *elem_dofs_h,
                      const uint8_t *constrained_h, lmh_dofmap **out);

/* ---- general constraint rows (DofMap::add_constraint_row,
 * dof_map_constraints.C:2180-2208, then DofMap::process_constraints,
 * :4341-4497).  Replaces the constraint set by the rows given at creation
 * (empty rows, value 0) plus these n_rows rows, added in order: DoF dof_h[k]
 * = rhs_h[k] + sum of coefs_h[t] * x[cols_h[t]] over t in
 * [row_ptr_h[k], row_ptr_h[k+1]) (rhs_h NULL: 0).  Rows are then substituted
 * until they refer to unconstrained DoFs only.  Matrices of this dofmap must be
 * created after the call.  Assembly applies the rows with
 * DofMap::heterogeneously_constrain_element_matrix_and_vector
 * (dof_map_constraints.C:2518-2616) and the pattern couples each element's DoFs
 * with the DoFs they are constrained by (find_connected_dofs, dof_map.C:2915).
 * flags: LMH_CONSTRAINT_FORBID_OVERWRITE (error if a DoF is constrained
 * already, the forbid_constraint_overwrite argument), LMH_CONSTRAINT_ERROR_ON_LOOP
 * (DofMap::set_error_on_constraint_loop: check_for_constraint_loops,
 * :4509-4597), LMH_CONSTRAINT_SYMMETRIC_ROWS (asymmetric_constraint_rows =
 * false: a constrained row keeps only its diagonal and a zero rhs; use with
 * CG and lmh_enforce_constraints).  A loop that survives substitution is
 * always an error here.  Elements with general rows: at most 32 DoFs of their
 * own and 64 once expanded. */
enum {
  LMH_CONSTRAINT_FORBID_OVERWRITE = 1,
  LMH_CONSTRAINT_ERROR_ON_LOOP = 2,
  LMH_CONSTRAINT_SYMMETRIC_ROWS = 4
};
int lmh_dofmap_set_constraints(lmh_dofmap *dm, int64_t n_rows, const int32_t *dof_h, const int32_t *row_ptr_h,
                               const int32_t *cols_h, const double *coefs_h, const double *rhs_h, int flags);
/* Processed rows: counts (any pointer may be NULL), then the rows ascending by
 * DoF with entries ascending (cols_h, coefs_h, rhs_h may be NULL). */
int lmh_dofmap_constraint_counts(const lmh_dofmap *dm, int64_t *n_rows, int64_t *n_entries, int64_t *n_general_elems);
int lmh_dofmap_get_constraints(const lmh_dofmap *dm, int32_t *dof_h, int32_t *row_ptr_h, int32_t *cols_h,
                               double *coefs_h, double *rhs_h);
/* DofMap::enforce_constraints_exactly (dof_map_constraints.C:3088-3170):
 * x_d = value_d + sum_j c_dj x_j on every constrained DoF (value omitted when
 * homogeneous != 0). */
int lmh_enforce_constraints(const lmh_dofmap *dm, double *x_d, int homogeneous);
int lmh_dofmap_destroy(lmh_dofmap *dm);

/* ---- sparsity + matrix: DofMap::compute_sparsity / SparsityPattern::Build
 * (dof_map.C:62-133, 1960-1978; sparsity_pattern.C:96-600) built on the GPU,
 * handed to SparseMatrix::init/update_sparsity_pattern (sparse_matrix.h,
 * eigen_sparse_matrix.C:64-115).  Also builds the element scatter map and the
 * element colouring used by lmh_assemble_*. ---- */
int lmh_matrix_create(lmh_context *ctx, const lmh_dofmap *dm, lmh_matrix **out);
/* A matrix with a caller-given CSR pattern (SparseMatrix::update_sparsity_pattern
 * from a SparsityPattern::Graph, sparse_matrix.h:175; sorted rows, square, FP64
 * values zeroed).  It has no DoF map: element assembly calls reject it. */
int lmh_matrix_create_csr(lmh_context *ctx, int64_t rows, const int32_t *row_ptr_h, const int32_t *col_idx_h,
                          lmh_matrix **out);
/* Rectangular variant (rows x cols), for matrices read from files
 * (SparseMatrix::init(m, n, ...) with m != n, sparse_matrix.h); SpMV takes x of
 * length cols.  The solvers require rows == cols. */
int lmh_matrix_create_csr2(lmh_context *ctx, int64_t rows, int64_t cols, const int32_t *row_ptr_h,
                           const int32_t *col_idx_h, lmh_matrix **out);
int lmh_matrix_destroy(lmh_matrix *A);
int64_t lmh_matrix_rows(const lmh_matrix *A);
int64_t lmh_matrix_cols(const lmh_matrix *A);
int64_t lmh_matrix_nnz(const lmh_matrix *A);
int lmh_matrix_n_colors(const lmh_matrix *A);
/* Download the pattern (row_ptr_h: rows+1, col_idx_h: nnz). */
int lmh_matrix_get_pattern(const lmh_matrix *A, int32_t *row_ptr_h, int32_t *col_idx_h);
/* Device pointers of the CSR arrays (row_ptr int32, col_idx int32, FP64 values). */
int lmh_matrix_device_arrays(lmh_matrix *A, const int32_t **row_ptr_d, const int32_t **col_idx_d, double **vals_d);
/* SparseMatrix::zero (eigen_sparse_matrix.C:186-198). */
int lmh_matrix_zero(lmh_matrix *A);
/* Values in/out, FP64, nnz entries in pattern order. */
int lmh_matrix_get_values(const lmh_matrix *A, double *vals_h);
int lmh_matrix_set_values(lmh_matrix *A, const double *vals_h);
/* FP32 operator values (after FP32 assembly or lmh_matrix_make_fp32). */
int lmh_matrix_get_values_fp32(const lmh_matrix *A, float *vals_h);
/* SparseMatrix::add_matrix(DenseMatrix, dofs) for a batch of dense element
 * matrices held on the device (eigen_sparse_matrix.C:121-135; laspack_matrix.C:626-640):
 * Ke_d: n_elem*n*n row-major, dofs_d: n_elem*n.  Entries must exist in the
 * pattern.  Deterministic: each entry accumulates in ascending batch order. */
int lmh_matrix_add_element_matrices(lmh_matrix *A, int64_t n_elem, int n, const double *Ke_d, const int32_t *dofs_d);
/* NumericVector::add_vector(values, dofs) for a batch (numeric_vector.C:690-697). */
int lmh_vector_add_element_vectors(lmh_context *ctx, double *v_d, int64_t n_elem, int n, const double *Fe_d,
                                   const int32_t *dofs_d);

/* ---- assembly: System::assemble for -div(kappa grad u) = f with QGauss(npts^3)
 * and Lagrange FE of the element's own order (introduction_ex4.C:313-560;
 * kappa-weighted form adjoints_ex2/L-shaped.C:101), element constraints
 * (DofMap::heterogeneously_constrain_element_matrix_and_vector,
 * dof_map_constraints.C:2518-2616) and the CSR/rhs scatter (eigen_sparse_matrix.C:121-135,
 * numeric_vector.C:690-697).  Overwrites the matrix values and rhs_d.
 * kappa_d: per-element coefficient (NULL -> 1); f_kind 0: f = 1, 1: f =
 * 3 pi^2 sin(pi x) sin(pi y) sin(pi z).  precision LMH_FP64 builds the FP64
 * CSR; LMH_FP32 runs the element contraction in FP32 on MFMA and writes the
 * FP32 operator used by the mixed-precision inner solver (geometry stays FP64).
 * scatter: LMH_SCATTER_COLORED or LMH_SCATTER_GATHER.  rhs_d may be NULL
 * (matrix only).  The rhs of an FP32 assembly equals the FP64 one except for
 * the K H terms of nonzero constraint values, formed from the FP32 element
 * matrices there. */
int lmh_assemble_poisson(lmh_matrix *A, const lmh_mesh *m, const lmh_dofmap *dm, int npts, const double *kappa_d,
                         int f_kind, double *rhs_d, int precision, int scatter);
/* Residual of the same physics at u (FEMSystem::assembly(get_residual = true,
 * get_jacobian = false), fem_system.C:875-1010 with the residual of
 * element_time_derivative, R_e = K_e u_e - F_e): element residuals from the
 * unconstrained element matrices, C^T applied for the constraint rows
 * (constrain_element_vector / constrain_element_matrix_and_vector,
 * dof_map_constraints.C:2439-2516: constrained entries 0), added in element
 * order into R_d (overwritten).  Uses the matrix's element buffers.  With
 * general constraint rows, u_d is the vector with its constraints enforced
 * (C u + H, lmh_enforce_constraints) and R_d = C^T (K u - F), constrained
 * entries 0. */
int lmh_assemble_residual_poisson(lmh_matrix *A, const lmh_mesh *m, const lmh_dofmap *dm, int npts,
                                  const double *kappa_d, int f_kind, const double *u_d, double *R_d);
/* Element matrices/vectors only (no scatter), constrained if constrain != 0:
 * Ke_d n_elem*dpe*dpe, Fe_d n_elem*dpe. */
int lmh_element_poisson(lmh_context *ctx, const lmh_mesh *m, const lmh_dofmap *dm, int npts, const double *kappa_d,
                        int f_kind, int constrain, double *Ke_d, double *Fe_d);

/* ---- user physics: the element loop of FEMSystem::assembly
 * (src/systems/fem_system.C:875-1200) with element_time_derivative
 * (include/physics/diff_physics.h:125) written as device code.  The template
 * element loop is in lmh_fem_physics.h (compiled with the caller's hipcc);
 * these calls give it the device arrays of a mesh and DoF map and put its
 * element results into the system matrix and a vector. ---- */
typedef struct lmh_fem_data {
  int elem_type;              /* LMH_HEX8 ... */
  int n_nodes;                /* nodes per element */
  int n_vars;                 /* variables (one Lagrange type) */
  int n_dofs;                 /* DoFs per element, n_vars * n_nodes, variable-major */
  int n_qp;                   /* quadrature points */
  int64_t n_elem, n_dofs_total;
  const double *xyz;          /* node coordinates, 3 per node */
  const int32_t *conn;        /* n_elem * n_nodes, libMesh local node order */
  const int32_t *elem_dofs;   /* n_elem * n_dofs */
  const uint8_t *constrained; /* n_dofs_total: nonzero = homogeneous Dirichlet DoF */
  const double *w;            /* QGauss weights, n_qp */
  const double *phi;          /* reference shapes, phi[a*n_qp + q] */
  const double *dref;         /* reference derivatives, dref[(a*n_qp + q)*3 + r] */
  int *flag;                  /* device int the element loop sets to 1 on a degenerate map */
  void *stream;               /* the context's hipStream_t */
} lmh_fem_data;
/* Fills *out for the mesh and DoF map with the QGauss rule of `npts` points
 * per direction (hex) or that order (tet), and clears *flag.  Only
 * homogeneous Dirichlet rows (the element loop's mask): a DoF map with general
 * constraint rows gives LMH_ERR_UNSUPPORTED. */
int lmh_fem_element_data(const lmh_mesh *m, const lmh_dofmap *dm, int npts, lmh_fem_data *out);
/* LMH_ERR_DEGENERATE when the element loop set *flag (then clears it). */
int lmh_fem_check_flag(const lmh_fem_data *d);
/* SparseMatrix::add_matrix of every element matrix (n_elem * n_dofs^2 in
 * element order, dofs of the matrix's DoF map) and NumericVector::add_vector
 * of every element vector, summed per entry in ascending element order (the
 * order of the reference's element loop): A's values and R_d (if given) are
 * overwritten by the sums.  Ke_d may be NULL (R_d only) or Fe_d (A only). */
int lmh_matrix_scatter_elements(lmh_matrix *A, const double *Ke_d, const double *Fe_d, double *R_d);

/* ---- linear algebra ---- */
/* y = A x (SparseMatrix::vector_mult, sparse_matrix.C:237-253). */
int lmh_spmv(lmh_matrix *A, const double *x_d, double *y_d);
/* y = A32 x on the FP32 operator (after FP32 assembly or lmh_matrix_make_fp32). */
int lmh_spmv_fp32(lmh_matrix *A, const float *x_d, float *y_d);
/* r = b - A x in FP64 (LinearImplicitSystem::assembly residual, linear_implicit_system.C:367-381). */
int lmh_residual(lmh_matrix *A, const double *b_d, const double *x_d, double *r_d);
double lmh_norm2(lmh_context *ctx, const double *x_d, int64_t n);
/* y += a x (NumericVector::add(a, v), numeric_vector.h). */
int lmh_axpy(lmh_context *ctx, int64_t n, double a, const double *x_d, double *y_d);

/* Whole-vector updates of the device-resident NumericVector
 * (include/numerics/numeric_vector.h:288 operator=(T), :503 add(T), :618
 * scale, :517 add(a, v), :457 operator*=(v), :474 operator/=(v), :480
 * reciprocal, :623 abs, :730 pointwise_mult, :738 pointwise_divide, :119
 * operator=(v)), on n entries of z:
 *   SET z = a; ADD_SCALAR z += a; SCALE z *= a; AXPY z += a x; MUL z *= x;
 *   DIV z /= x; RECIPROCAL z = 1/z; ABS z = |z|; PMUL z = x y; PDIV z = x / y;
 *   COPY z = x.  Operands that an operation does not read may be NULL. */
enum {
  LMH_VEC_SET = 0, LMH_VEC_ADD_SCALAR, LMH_VEC_SCALE, LMH_VEC_AXPY, LMH_VEC_MUL, LMH_VEC_DIV,
  LMH_VEC_RECIPROCAL, LMH_VEC_ABS, LMH_VEC_PMUL, LMH_VEC_PDIV, LMH_VEC_COPY
};
int lmh_vector_op(lmh_context *ctx, int op, int64_t n, double a, const double *x_d, const double *y_d, double *z_d);
/* Reductions of the same vector (numeric_vector.h:312 sum, :318 l1_norm, :324
 * l2_norm, :330 linfty_norm, :301 min, :307 max, :631 dot, and
 * l1_norm_diff / l2_norm_diff) over n entries of x (and y), deterministic
 * (fixed slices per block, fixed fold order); the value is written to *out_h.
 * n = 0 gives 0 (MIN: +inf, MAX: -inf). */
enum {
  LMH_VEC_SUM = 0, LMH_VEC_L1, LMH_VEC_L2, LMH_VEC_LINF, LMH_VEC_MIN, LMH_VEC_MAX, LMH_VEC_DOT,
  LMH_VEC_L1_DIFF, LMH_VEC_L2_DIFF
};
int lmh_vector_reduce(lmh_context *ctx, int op, int64_t n, const double *x_d, const double *y_d, double *out_h);

/* Jacobi-preconditioned CG, FP64 (EigenSparseLinearSolver::solve with CG +
 * JACOBI_PRECOND -> Eigen::internal::conjugate_gradient,
 * eigen_sparse_linear_solver.C:82-170, ConjugateGradient.h:28-90).  x_d holds
 * the initial guess.  Stops when |r|^2 < max(rtol^2 |b|^2, DBL_MIN); returns
 * the iteration count and sqrt(|r|^2/|b|^2) exactly as Eigen reports them. */
int lmh_cg_solve(lmh_matrix *A, const double *b_d, double *x_d, double rtol, int64_t maxits, int64_t *iters,
                 double *rel_res, int *reason);
/* Allocation-free lmh_cg_solve.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * stopping rule and results as lmh_cg_solve; it calls no allocator.  The work
 * vectors (r, p, q, 1/a_ii) and the diagonal slots live in one fixed device
 * array of the library sized for LMH_CG_NOALLOC_MAX_ROWS rows (4 doubles and
 * one int32 per row: 144 MiB), which every GPU the library runs on holds from
 * its first context until the process ends.  A matrix with more rows:
 * LMH_ERR_INVALID.  A's compressed column windows are used when an earlier
 * lmh_cg_solve or lmh_mixed_solve on A built them (the iterates are the same
 * either way).  All calls on one GPU share the array and keep the last call's
 * contents: calls that overlap on one GPU (from several contexts or threads)
 * race.  Not recorded by lmh_profile_enable. */
#define LMH_CG_NOALLOC_MAX_ROWS 4194304
int lmh_cg_solve_noalloc(lmh_matrix *A, const double *b_d, double *x_d, double rtol, int64_t maxits,
                         int64_t *iters, double *rel_res, int *reason);

/* Jacobi-preconditioned BiCGSTAB, FP64 (EigenSparseLinearSolver::solve with
 * BICGSTAB + JACOBI_PRECOND -> Eigen::BiCGSTAB, eigen_sparse_linear_solver.C:175-205,
 * BiCGSTAB.h:28-112), for nonsymmetric matrices such as those with asymmetric
 * constraint rows.  Same contract as lmh_cg_solve: stops when
 * |r|^2 <= rtol^2 |b|^2; b = 0 gives x = 0.  Eigen's restart when (r0, r)
 * vanishes relative to |r0|^2 is reported as LMH_DIVERGED_BREAKDOWN. */
int lmh_bicgstab_solve(lmh_matrix *A, const double *b_d, double *x_d, double rtol, int64_t maxits, int64_t *iters,
                       double *rel_res, int *reason);
/* The same BiCGSTAB on the transpose, A^T x = b: what LinearSolver::
 * adjoint_solve does for BICGSTAB (src/solvers/linear_solver.C:147-171, the
 * explicit transpose of eigen_sparse_linear_solver.C:331-350).  The pattern
 * must be structurally symmetric (every (i, j) stored with its (j, i), as the
 * constraint-coupled finite-element patterns are; otherwise
 * LMH_ERR_UNSUPPORTED); A^T then has A's pattern with mirrored values, which
 * are regathered on the device at each call. */
int lmh_bicgstab_solve_transpose(lmh_matrix *A, const double *b_d, double *x_d, double rtol, int64_t maxits,
                                 int64_t *iters, double *rel_res, int *reason);

/* Mixed precision: FP32 Jacobi-CG on the FP32 operator inside FP64 refinement
 * of the FP64 residual r = b - A x.  mode 0: reliable updates (the FP32
 * recurrence continues; r and x are refreshed in FP64 whenever the FP32
 * residual falls below inner_rtol times the last true residual); mode 1:
 * classic iterative refinement (each correction solve restarts from zero and
 * runs to inner_rtol).  Stops when |b - A x| / |b| <= rtol.  Requires both the
 * FP64 and FP32 operators (lmh_assemble_poisson with each precision, or
 * lmh_matrix_make_fp32).  stats_h (may be NULL) receives
 * {inner iterations, outer refreshes, final true relative residual}. */
int lmh_mixed_solve(lmh_matrix *A, const double *b_d, double *x_d, double rtol, double inner_rtol, int mode,
                    int64_t maxits, int64_t *iters, double *rel_res, int *reason, double *stats_h);
/* Build the FP32 operator by rounding the FP64 values. */
int lmh_matrix_make_fp32(lmh_matrix *A);
/* Column storage of the solvers' SpMV: on (default) stores, per row block,
 * the block's distinct columns once (16-bit offsets against two bases) and
 * per entry the index of its column in that list: 8-bit on blocks of at most
 * 1016 entries and 256 distinct columns (5 bytes per entry with the FP32
 * value instead of 8, 9 with the FP64 value instead of 12), else 16-bit on
 * the 2048-entry blocks, whenever every block fits; x is gathered once per
 * distinct column.  Off keeps the int32 CSR columns.  Built with the matrix;
 * used by the FP32 inner iterations of lmh_mixed_solve and by the iterations
 * of lmh_cg_solve.  Row sums are bit-identical either way; the CG's (p, Ap)
 * partials follow the block partition, so solves agree to rounding.
 * lmh_matrix_fp32_compression reports 1 when the compressed columns are in
 * use, 0 otherwise; lmh_matrix_window_bits their index width (8 or 16, 0 when
 * not in use). */
int lmh_matrix_set_fp32_compression(lmh_matrix *A, int on);
int lmh_matrix_fp32_compression(lmh_matrix *A);
int lmh_matrix_window_bits(lmh_matrix *A);
/* Preconditioner of lmh_cg_solve, lmh_bicgstab_solve and lmh_mixed_solve on
 * this matrix, with libMesh's PreconditionerType values: LMH_PRECOND_JACOBI
 * (default; Eigen DiagonalPreconditioner, 1/a_ii or 1 where the diagonal is
 * absent or zero) or LMH_PRECOND_IDENTITY (Eigen IdentityPreconditioner),
 * the two EigenSparseLinearSolver::solve runs without a factorisation
 * (src/solvers/eigen_sparse_linear_solver.C:136-170, 177-205).  Other values:
 * LMH_ERR_UNSUPPORTED. */
enum { LMH_PRECOND_IDENTITY = 0, LMH_PRECOND_JACOBI = 1 };
int lmh_matrix_set_preconditioner(lmh_matrix *A, int type);

/* ---- adjoint sensitivity (ImplicitSystem::adjoint_qoi_parameter_sensitivity,
 * implicit_system.C:517-623, for per-element kappa parameters) ---- */
/* dJ/du for J = integral of u over [lo,hi]^3 (per-qp inclusion), with the
 * constraints applied as System::assemble_qoi_derivative(apply_constraints =
 * true) does: C^T dq, constrained entries zero (constrain_element_vector,
 * dof_map_constraints.C:2981-3020); for Dirichlet-only maps that is the
 * constrained entries zeroed. */
int lmh_qoi_box_derivative(const lmh_mesh *m, const lmh_dofmap *dm, int npts, double lo, double hi, double *dq_d);
/* dJ/dkappa_e = -(1/kappa_e) lambda_e^T Kc_e u_e for every element, FP64.
 * With general constraint rows, u is the solution with its constraints
 * enforced (lmh_enforce_constraints, heterogeneous) and lambda the adjoint
 * with them enforced homogeneously (enforce_adjoint_constraints_exactly,
 * implicit_system.C:230-240); the adjoint itself solves A^T lambda = dq
 * (lmh_cg_solve for symmetric rows, lmh_bicgstab_solve_transpose for the
 * default asymmetric rows). */
int lmh_adjoint_kappa_gradient(const lmh_mesh *m, const lmh_dofmap *dm, int npts, const double *kappa_d,
                               const double *u_d, const double *lambda_d, double *grad_d);

/* ---- batched FP32 linear-elasticity assembly of independent problems ----
 * Each problem is one LinearElasticity system of systems_of_equations_ex6.C
 * (assemble_elasticity, :127-236: C_ijkl = l1 d_ij d_kl + l2 (d_ik d_jl +
 * d_il d_jk), body force (0,0,fz), QGauss at the FE type's default order,
 * element constraints, add_matrix / add_vector).  libMesh assembles one system
 * per System::assemble call; here a whole batch of problems of uneven size and
 * mixed element type (HEX8, HEX27, TET4, TET10; 3 variables per node) is
 * assembled by one hipGraph launch of a persistent element kernel followed by a
 * persistent row-gather scatter.  Values are FP32 (lmh_matrix_get_values_fp32,
 * lmh_spmv_fp32), right-hand sides FP64. */
int lmh_batch_create(lmh_context *ctx, int n_prob, const lmh_dofmap *const *dms, lmh_batch **out);
int lmh_batch_destroy(lmh_batch *b);
int lmh_batch_size(const lmh_batch *b);
/* The batch owns each problem's matrix and right-hand side. */
int lmh_batch_matrix(const lmh_batch *b, int p, lmh_matrix **A);
int lmh_batch_rhs(const lmh_batch *b, int p, double **rhs_d);
/* Tile shape per element type, indexed HEX8, HEX27, TET4, TET10: lanes per
 * element (16..256) and passes (groups of concurrently assembled elements) per
 * tile.  The defaults are the smallest lane count the LDS budget allows. */
int lmh_batch_set_tiles(lmh_batch *b, const int *lanes_per_elem, const int *passes_per_tile);
int lmh_batch_get_tiles(const lmh_batch *b, int *lanes_per_elem, int *passes_per_tile);
/* Time every admissible tile shape of each element type present on the batch
 * itself (best of `reps` runs after one warm-up), keep the fastest, and
 * capture the assembly graph for (E, nu, fz).  best_ms[4] (may be NULL)
 * receives the element-kernel time of the chosen shape per type (-1: absent). */
int lmh_batch_autotune(lmh_batch *b, double E, double nu, double fz, int reps, double *best_ms);
/* Assemble every problem: one hipGraphLaunch (the graph is captured on first
 * use and again when E, nu or fz change).  LMH_ERR_DEGENERATE if any element
 * map has det J <= 0. */
int lmh_batch_assemble_elasticity(lmh_batch *b, double E, double nu, double fz);

/* ---- per-kernel timing with HIP events on the context stream ---- */
int lmh_profile_enable(lmh_context *ctx, int on);
/* Average duration in ms and launch count of a named kernel since enable. */
int lmh_profile_query(lmh_context *ctx, const char *kernel, double *avg_ms, int64_t *count);
int lmh_profile_reset(lmh_context *ctx);

/* ---- static condensation (StaticCondensation, include/numerics/static_condensation.h:64;
 * src/numerics/static_condensation.C) ----
 * The DoFs of element-internal nodes (static_condensation_dof_map.C:187-199,
 * elem.is_internal) are eliminated: the reduced matrix holds A's pattern
 * without the condensed rows and columns (reduced DoFs keep the full order)
 * and the values S = A_uu - A_uc A_cc^-1 A_cu, which for DoFs that each live
 * on one element equals the sum of the element Schur complements
 * (static_condensation.C:205-211).  The condensed block must be diagonal (one
 * condensed DoF per coupled block, e.g. a scalar variable on HEX27; exact
 * zeros between condensed DoFs are admitted): otherwise LMH_ERR_UNSUPPORTED;
 * a zero pivot gives LMH_ERR_DEGENERATE.  Every column of a condensed row c
 * must lie, structurally, in every reduced row that holds column c (true for
 * element-internal DoFs), whatever the values at creation: otherwise
 * LMH_ERR_UNSUPPORTED.  A keeps its values and must outlive every call on the
 * condensation except lmh_condensation_destroy, which may follow A's
 * lmh_matrix_destroy (the context must still exist). */

/* DoFs of the internal nodes of dm's elements (HEX27: local node 26, every
 * variable), element by element.  *n receives the count; dofs_h (may be NULL
 * to query the count) receives the DoFs. */
int lmh_dofmap_condensed_dofs(const lmh_dofmap *dm, int64_t *n, int32_t *dofs_h);
/* Builds the reduced pattern and values from A's current values (StaticCondensation::init + close).
 * Also builds the device position map that later refreshes use: 2 bytes per
 * entry of each condensed row, counted once per reduced row that couples to it
 * (27 x 26 x 2 B per condensed DoF for HEX27, 368 MB for 64^3); without room
 * for it (or for rows over 384 entries) refreshes search the pattern instead. */
int lmh_condensation_create(lmh_matrix *A, int64_t n_cond, const int32_t *cond_dofs_h, lmh_condensation **out);
/* Recomputes the reduced values after A's values changed (StaticCondensation::setup/close). */
int lmh_condensation_setup(lmh_condensation *sc);
/* The reduced operator (get_condensed_mat); owned by sc; every solver and lmh_matrix_make_fp32 apply. */
lmh_matrix *lmh_condensation_reduced_matrix(lmh_condensation *sc);
int64_t lmh_condensation_n_condensed(const lmh_condensation *sc);
/* full_index_h[r] = full DoF of reduced DoF r (rows of the reduced matrix). */
int lmh_condensation_reduced_dofs(const lmh_condensation *sc, int32_t *full_index_h);
/* forward_elimination (static_condensation.C:383-417): g = b_u - A_uc A_cc^-1 b_c. */
int lmh_condensation_forward(lmh_condensation *sc, const double *full_rhs_d, double *reduced_rhs_d);
/* backwards_substitution (:419-456): x_u = y, x_c = A_cc^-1 (b_c - A_cu x_u). */
int lmh_condensation_backward(lmh_condensation *sc, const double *full_rhs_d, const double *reduced_sol_d,
                              double *full_sol_d);
/* apply (:458-473): forward elimination, Jacobi-CG on the reduced matrix from
 * x's uncondensed values (rtol on the reduced residual), back substitution into x. */
int lmh_condensation_solve(lmh_condensation *sc, const double *b_d, double *x_d, double rtol, int64_t maxits,
                           int64_t *iters, double *rel_res, int *reason);
int lmh_condensation_destroy(lmh_condensation *sc);

#ifdef __cplusplus
}
#endif
#endif
