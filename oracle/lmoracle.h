/*
 * lmoracle.h -- CPU oracle for the libMesh assembly + Jacobi-CG hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain C++ restatement of libMesh's
 * FP64 path (mesh generation, first-touch renumbering, var-major DoF
 * numbering, Dirichlet constraint set, sparsity, QGauss, Lagrange shapes,
 * FEMap, element assembly, element constraint application, CSR scatter,
 * Eigen/Laspack Jacobi-CG, adjoint parameter sensitivity).  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * only as the checker: nothing in the product links or calls it.
 *
 * Each function cites the reference file:line it restates (paths relative to
 * the libMesh source tree).  Pinning of each part (golden vectors, closed
 * forms, vendored Eigen 3.3.9 via oracle/_ref) is done by tests/test_oracle_*.
 */
#ifndef LMORACLE_H
#define LMORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* libMesh ElemType values (include/enums/enum_elem_type.h). */
enum { ORC_TET4 = 8, ORC_TET10 = 9, ORC_HEX8 = 10, ORC_HEX27 = 12 };

typedef struct orc_mesh orc_mesh;
typedef struct orc_system orc_system;

/* --- mesh (src/mesh/mesh_generation.C:317, 1160-1940; replicated_mesh.C:700-830) --- */
orc_mesh *orc_build_cube(int nx, int ny, int nz, int elem_type);
void orc_mesh_free(orc_mesh *m);
int64_t orc_mesh_n_nodes(const orc_mesh *m);
int64_t orc_mesh_n_elem(const orc_mesh *m);
int orc_mesh_nodes_per_elem(const orc_mesh *m);
int orc_mesh_elem_type(const orc_mesh *m);
void orc_mesh_coords(const orc_mesh *m, double *xyz);          /* n_nodes*3 */
void orc_mesh_set_coords(orc_mesh *m, const double *xyz);
void orc_mesh_conn(const orc_mesh *m, int32_t *conn);          /* n_elem*npe */
void orc_mesh_node_bmask(const orc_mesh *m, uint8_t *mask);    /* nodeset ids, bit b = id b */
void orc_mesh_side_bid(const orc_mesh *m, int8_t *bid);        /* n_elem*n_sides, -1 = none */
/* Seeded uniform perturbation of interior nodes by U[-frac*h_c, frac*h_c] per
 * coordinate c, h_c = node-lattice spacing along axis c. */
void orc_perturb_interior(orc_mesh *m, uint64_t seed, double frac);
/* kappa_e = exp(z_e), z_e ~ N(0,1) by Box-Muller over mt19937_64(seed). */
void orc_kappa_lognormal(int64_t n_elem, uint64_t seed, double *kappa);
/* min det J over all quadrature points of all elements (npts per direction). */
double orc_min_detj(const orc_mesh *m, int npts);

/* --- DoF map + constraints + sparsity (dof_map.C:949-1530, 2201-2316;
 *     dof_map_constraints.C:592-740; sparsity_pattern.C:96-600) --- */
orc_system *orc_system_create(const orc_mesh *m, int n_vars, int order, unsigned dirichlet_bid_mask);
void orc_system_free(orc_system *s);
int64_t orc_n_dofs(const orc_system *s);
int64_t orc_nnz(const orc_system *s);
int orc_dofs_per_elem(const orc_system *s);
void orc_elem_dofs(const orc_system *s, int32_t *dofs);       /* n_elem*dpe */
void orc_constrained(const orc_system *s, uint8_t *flags);     /* n_dofs */
void orc_pattern(const orc_system *s, int32_t *row_ptr, int32_t *col_idx);

/* --- general constraint rows (dof_map_constraints.C:2180-2208 add_constraint_row,
 *     4341-4497 process_constraints, 4509-4597 check_for_constraint_loops,
 *     3545-3660 build_constraint_matrix_and_vector, 2518-2616 heterogeneous
 *     element constraint; dof_map.C:2915-2972 find_connected_dofs) ---
 * Dirichlet DoFs of orc_system_create hold empty rows.  add returns -1 when
 * forbid is set and the DoF is constrained; process returns -1 on a loop when
 * error_on_loop is set, and rebuilds the pattern (connected DoFs included). */
int orc_add_constraint_row(orc_system *s, int32_t dof, int n, const int32_t *cols, const double *coefs, double rhs,
                           int forbid);
int orc_process_constraints(orc_system *s, int error_on_loop);
/* asymmetric_constraint_rows of the element constraint (default 1). */
void orc_set_asymmetric_constraint_rows(orc_system *s, int on);
void orc_constraint_counts(const orc_system *s, int64_t *n_rows, int64_t *n_entries);
void orc_constraint_rows(const orc_system *s, int32_t *dofs, int32_t *row_ptr, int32_t *cols, double *coefs,
                         double *rhs);
/* FEMSystem residual of the Poisson physics at u (Dirichlet rows only):
 * R = sum_e C^T (K_e u_e - F_e), constrained entries 0.  Returns 0 or -1. */
int orc_residual_poisson(const orc_system *s, int npts, const double *kappa, int f_kind, const double *u, double *R);
/* Constrained element matrix/vector of element e from its unconstrained K, F:
 * returns the expanded size m (<= cap) with dofs_out[m], K_out[m*m], F_out[m];
 * -1 if m > cap. */
int orc_constrain_element(const orc_system *s, int64_t e, const double *K, const double *F, int32_t *dofs_out,
                          double *K_out, double *F_out, int cap);
/* x_i = value_i (0 if homogeneous) + sum_j c_ij x_j for every constrained DoF
 * (enforce_constraints_exactly). */
void orc_enforce_constraints(const orc_system *s, double *x, int homogeneous);

/* --- assembly (introduction_ex4.C:313-560; fe.C:197-368; fe_map.C:958-1090,
 *     1167-1246; dof_map_constraints.C:2518-2616; eigen_sparse_matrix.C:121-135) ---
 * f_kind: 0 -> f = 1; 1 -> f = 3 pi^2 sin(pi x) sin(pi y) sin(pi z).
 * kappa may be NULL (kappa = 1).  Returns 0, or -1 on a degenerate map. */
int orc_assemble_poisson(const orc_system *s, int npts, const double *kappa, int f_kind,
                         double *vals, double *rhs);
/* Element matrices/vectors for elements [e0, e0+count): Ke row-major dpe*dpe.
 * constrain != 0 applies the element constraint step. */
int orc_element_poisson(const orc_system *s, int npts, const double *kappa, int f_kind,
                        int64_t e0, int64_t count, int constrain, double *Ke, double *Fe);
/* FEMSystem::assembly of the user physics -div(kappa grad u) + sigma u^3 = f at
 * the solution u (one variable): constrained element residual Fe and Jacobian Ke
 * (Ke may be NULL), or their element-order sums into vals / R (either may be
 * NULL).  variant 0; 1 and 2 plant mistakes for the pin tests. */
int orc_fem_element_reaction_diffusion(const orc_system *s, int npts, const double *kappa, double sigma,
                                       int f_kind, const double *u, int64_t e0, int64_t count, int constrain,
                                       int variant, double *Ke, double *Fe);
int orc_fem_reaction_diffusion(const orc_system *s, int npts, const double *kappa, double sigma, int f_kind,
                               const double *u, int variant, double *vals, double *R);
/* Linear elasticity (systems_of_equations_ex6.C), body force (0,0,fz). */
int orc_assemble_elasticity(const orc_system *s, int npts, double E, double nu, double fz,
                            double *vals, double *rhs);
int orc_element_elasticity(const orc_system *s, int npts, double E, double nu, double fz,
                           int64_t e0, int64_t count, int constrain, double *Ke, double *Fe);

/* --- solvers --- */
/* Eigen 3.3.9 conjugate_gradient + DiagonalPreconditioner restated
 * (contrib/eigen/3.3.9/Eigen/src/IterativeLinearSolvers/ConjugateGradient.h:28-90,
 * BasicPreconditioners.h:66-76).  x holds the initial guess on entry. */
void orc_cg_eigen(int64_t n, const int32_t *row_ptr, const int32_t *col_idx, const double *vals,
                  const double *b, double *x, double tol, int64_t maxits, int64_t *iters,
                  double *rel_err);
/* Eigen BiCGSTAB + DiagonalPreconditioner (BiCGSTAB.h:28-112), restated. */
void orc_bicgstab_eigen(int64_t n, const int32_t *row_ptr, const int32_t *col_idx, const double *vals,
                        const double *b, double *x, double tol, int64_t maxits, int64_t *iters, double *rel_err);
/* The same two with the preconditioner chosen: precond 1 JACOBI_PRECOND,
 * 0 IDENTITY_PRECOND (Eigen IdentityPreconditioner: z = r). */
void orc_cg_eigen_pc(int64_t n, const int32_t *row_ptr, const int32_t *col_idx, const double *vals,
                     const double *b, double *x, double tol, int64_t maxits, int precond, int64_t *iters,
                     double *rel_err);
void orc_bicgstab_eigen_pc(int64_t n, const int32_t *row_ptr, const int32_t *col_idx, const double *vals,
                           const double *b, double *x, double tol, int64_t maxits, int precond, int64_t *iters,
                           double *rel_err);
/* Laspack CGIter + JacobiPrecond restated (contrib/laspack/itersolv.C:377-476,
 * precond.C:26-46, rtc.C:48-72); x zeroed first as laspack_linear_solver.C:130. */
void orc_cg_laspack(int64_t n, const int32_t *row_ptr, const int32_t *col_idx, const double *vals,
                    const double *b, double *x, double tol, int64_t maxits, int64_t *iters,
                    double *rel_res);
void orc_spmv(int64_t n, const int32_t *row_ptr, const int32_t *col_idx, const double *vals,
              const double *x, double *y);

/* --- adjoint (implicit_system.C:196-240, 474-623; adjoints_ex2/L-qoi.C) --- */
/* dJ/du for J = integral of u over the box [lo,hi]^3 (per-qp inclusion), with
 * constrained entries zeroed. */
int orc_qoi_box_derivative(const orc_system *s, int npts, double lo, double hi, double *dq);
double orc_qoi_box(const orc_system *s, int npts, double lo, double hi, const double *u);
/* Restated adjoint_qoi_parameter_sensitivity for per-element kappa parameters:
 * sens_e = dq/dkappa_e + [(R(k-d) - R(k+d)) / (2d)] . z, R = A u - b,
 * d = 1e-6 * max(|kappa_e|, 1e-3).  Only element e's terms of R depend on kappa_e. */
int orc_adjoint_sensitivity(const orc_system *s, int npts, const double *kappa, int f_kind,
                            const double *u, const double *z, const int64_t *elems,
                            int64_t n_elems, double *sens);
/* Closed form used by the GPU kernel, for comparison: -(1/kappa_e) z_e^T Kc_e u_e. */
int orc_adjoint_exact(const orc_system *s, int npts, const double *kappa, const double *u,
                      const double *z, double *sens);

/* --- reference quantities for pinning --- */
void orc_qgauss_1d(int npts, double *x, double *w);
void orc_qgauss_hex(int npts, double *xyz, double *w);
int orc_qgauss_tet(int order, double *xyz, double *w); /* returns n points */
/* reference shapes: phi[i*nqp+q], dphi[(i*nqp+q)*3+c] at given points */
int orc_shapes(int elem_type, int order, int nq, const double *pts, double *phi, double *dphi);

#ifdef __cplusplus
}
#endif
#endif
