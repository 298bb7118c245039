/* lmh_fem_physics.h -- the element loop of FEMSystem::assembly
 * (src/systems/fem_system.C:875-1200) for physics written as device code.
 *
 * A physics is a struct with a device member
 *
 *   __device__ void element_time_derivative(bool request_jacobian, lmh_fem::Context &c) const;
 *
 * the counterpart of DifferentiablePhysics::element_time_derivative
 * (include/physics/diff_physics.h:125): it adds the element residual F and,
 * when asked, the Jacobian dF/du through c.get_elem_residual(var, i) and
 * c.get_elem_jacobian(var_i, var_j, i, j), reading the FE values the way
 * libMesh physics read FEMContext / FEBase (include/systems/fem_context.h):
 * JxW, xyz, phi, dphi, interior_value, interior_gradient.  One wavefront
 * evaluates one element: the code runs on every lane, and each lane owns the
 * test functions i = lane(), lane() + 64, ... (the F(i) and K(i, j) it
 * writes), e.g.
 *
 *   for (int i = c.lane(); i < c.n_dofs(0); i += c.width()) { ... F(i) += ... }
 *
 * element_loop<Physics>() runs it on every element of a mesh (arrays from
 * lmh_fem_element_data), then applies the homogeneous Dirichlet rows as
 * DofMap::constrain_element_matrix_and_vector does for empty rows
 * (dof_map_constraints.C:2518-2616: constrained rows and columns cleared,
 * 1 on the constrained diagonal, constrained residual entries 0), leaving
 * the element results in Fe (n_elem * n_dofs) and Ke (n_elem * n_dofs^2).
 * lmh_matrix_scatter_elements then sums them into the system matrix and
 * residual in element order.  Compile with hipcc --offload-arch=gfx950. */
#ifndef LMH_FEM_PHYSICS_H
#define LMH_FEM_PHYSICS_H

#include <hip/hip_runtime.h>

#include "lmh_fe_map.h"
#include "lmhip.h"

namespace lmh_fem {

constexpr int kWave = 64;
constexpr int kMaxNodes = 27, kMaxQp = 64, kMaxVars = 3;

// FEMContext of one element on the device.  The element's FE values live in
// LDS (JxW, xyz, dphi, u and grad u at the points), phi in the reference
// table in global memory, the outputs in the element's slots of Fe / Ke.
struct Context {
  int64_t elem;
  int nodes, vars, qps, dofs, lane_id;
  double coef;
  const double *jxw, *xq, *phiq, *dphiq, *uq, *guq;
  double *F, *K;

  __device__ int lane() const { return lane_id; }
  __device__ int width() const { return kWave; }
  __device__ int64_t elem_id() const { return elem; }
  __device__ unsigned int n_vars() const { return vars; }
  // FEMContext::n_dof_indices(var): one Lagrange type, so the element's nodes
  __device__ int n_dofs(int /*var*/) const { return nodes; }
  __device__ int n_qpoints() const { return qps; }
  // the element's coefficient (the per-element parameter array given to
  // element_loop; 1 when there is none)
  __device__ double coefficient() const { return coef; }
  // FEBase::get_JxW / get_xyz / get_phi / get_dphi at quadrature point qp
  __device__ double JxW(int qp) const { return jxw[qp]; }
  __device__ double xyz(int qp, int c) const { return xq[3 * qp + c]; }
  __device__ double phi(int i, int qp) const { return phiq[i * qps + qp]; }
  __device__ double dphi(int i, int qp, int c) const { return dphiq[(i * qps + qp) * 3 + c]; }
  // FEMContext::interior_value / interior_gradient of the current solution
  __device__ double interior_value(int var, int qp) const { return uq[var * qps + qp]; }
  __device__ double interior_gradient(int var, int qp, int c) const { return guq[(var * qps + qp) * 3 + c]; }
  // FEMContext::get_elem_residual(var)(i), get_elem_jacobian(vi, vj)(i, j)
  __device__ double &get_elem_residual(int var, int i) const { return F[var * nodes + i]; }
  __device__ double &get_elem_jacobian(int var_i, int var_j, int i, int j) const {
    return K[(var_i * nodes + i) * dofs + var_j * nodes + j];
  }
};

// LDS of one block for n nodes, nq points, nv variables: the element's own
// data only (the reference tables w, phi and dref are read from global memory,
// where every element's reads hit the cache), so several workgroups fit a CU
inline size_t lds_bytes(int n, int nq, int nv) {
  const size_t nd = (size_t)nv * n;
  const size_t doubles = (size_t)n * 3 + nd                         // X, u_e
                         + nq + (size_t)nq * 3 + (size_t)nq * 9     // JxW, xyz, inverse map
                         + (size_t)n * nq * 3                       // physical dphi
                         + (size_t)nv * nq * 4;                     // u, grad u at the points
  return doubles * sizeof(double) + nd;                            // + constrained flags
}

template <class Physics>
__global__ __launch_bounds__(kWave) void k_element_loop(Physics phys, lmh_fem_data d, const double *__restrict__ u,
                                                        const double *__restrict__ coef, int request_jacobian,
                                                        double *__restrict__ Fe, double *__restrict__ Ke) {
  extern __shared__ double sm[];
  const int lane = threadIdx.x;
  const int n = d.n_nodes, nq = d.n_qp, nv = d.n_vars, nd = d.n_dofs;
  const double *__restrict__ w = d.w, *__restrict__ phi = d.phi, *__restrict__ dref = d.dref;
  double *X = sm;
  double *ue = X + n * 3;
  double *jxw = ue + nd;
  double *xq = jxw + nq;
  double *G = xq + nq * 3;
  double *dphi = G + nq * 9;
  double *uq = dphi + n * nq * 3;
  double *gq = uq + nv * nq;
  unsigned char *cons = (unsigned char *)(gq + nv * nq * 3);
  for (int64_t e = blockIdx.x; e < d.n_elem; e += gridDim.x) {
    for (int a = lane; a < n; a += kWave) {
      const int64_t nd3 = 3 * (int64_t)d.conn[e * n + a];
      X[a * 3] = d.xyz[nd3];
      X[a * 3 + 1] = d.xyz[nd3 + 1];
      X[a * 3 + 2] = d.xyz[nd3 + 2];
    }
    for (int k = lane; k < nd; k += kWave) {
      const int32_t dof = d.elem_dofs[e * nd + k];
      ue[k] = u ? u[dof] : 0.0;
      cons[k] = d.constrained[dof] != 0;
    }
    __syncthreads();
    for (int q = lane; q < nq; q += kWave) {
      double Gq[3][3];
      const double det = lmh_qp_map(X, dref, n, nq, q, Gq);
      if (!(det > 0.0)) {
        *d.flag = 1;  // degenerate map: the host raises DegenerateMap
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
          for (int c = 0; c < 3; ++c) Gq[r][c] = 0.0;
      }
      jxw[q] = det * w[q];
      double x = 0, y = 0, z = 0;
      for (int a = 0; a < n; ++a) {
        const double p = phi[a * nq + q];
        x += X[a * 3] * p;
        y += X[a * 3 + 1] * p;
        z += X[a * 3 + 2] * p;
      }
      xq[q * 3] = x;
      xq[q * 3 + 1] = y;
      xq[q * 3 + 2] = z;
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) G[q * 9 + r * 3 + c] = Gq[r][c];
    }
    __syncthreads();
    // physical gradients: dphi/dx_c = sum_r dref[r] dxi_r/dx_c
    for (int t = lane; t < n * nq; t += kWave) {
      const int q = t % nq;
      const double *dr = &dref[t * 3];
      const double *Gq = &G[q * 9];
#pragma unroll
      for (int c = 0; c < 3; ++c) dphi[t * 3 + c] = dr[0] * Gq[c] + dr[1] * Gq[3 + c] + dr[2] * Gq[6 + c];
    }
    __syncthreads();
    for (int t = lane; t < nv * nq; t += kWave) {
      const int v = t / nq, q = t % nq;
      double s = 0, g0 = 0, g1 = 0, g2 = 0;
      for (int a = 0; a < n; ++a) {
        const double ua = ue[v * n + a];
        s += ua * phi[a * nq + q];
        const double *dp = &dphi[(a * nq + q) * 3];
        g0 += ua * dp[0];
        g1 += ua * dp[1];
        g2 += ua * dp[2];
      }
      uq[t] = s;
      gq[t * 3] = g0;
      gq[t * 3 + 1] = g1;
      gq[t * 3 + 2] = g2;
    }
    double *Fel = Fe + e * nd;
    double *Kel = request_jacobian ? Ke + e * (int64_t)nd * nd : nullptr;
    for (int k = lane; k < nd; k += kWave) Fel[k] = 0.0;
    if (Kel)
      for (int k = lane; k < nd * nd; k += kWave) Kel[k] = 0.0;
    __syncthreads();
    Context c{e, n, nv, nq, nd, lane, coef ? coef[e] : 1.0, jxw, xq, phi, dphi, uq, gq, Fel, Kel};
    phys.element_time_derivative(request_jacobian != 0, c);
    __syncthreads();
    for (int k = lane; k < nd; k += kWave)
      if (cons[k]) Fel[k] = 0.0;
    if (Kel)
      for (int t = lane; t < nd * nd; t += kWave) {
        const int i = t / nd, j = t % nd;
        if (cons[i] || cons[j]) Kel[t] = (i == j) ? 1.0 : 0.0;
      }
    __syncthreads();
  }
}

// Runs phys on every element of d at the solution u_d (n_dofs_total entries)
// with per-element coefficients coef_d (may be NULL): Fe_d (n_elem * n_dofs)
// and, with request_jacobian, Ke_d (n_elem * n_dofs^2) receive the
// constrained element results.  Stream-ordered on d.stream; returns
// LMH_ERR_DEGENERATE for a non-positive Jacobian determinant.
template <class Physics>
inline int element_loop(const lmh_fem_data &d, const Physics &phys, const double *u_d, const double *coef_d,
                        bool request_jacobian, double *Fe_d, double *Ke_d) {
  if (d.n_nodes > kMaxNodes || d.n_qp > kMaxQp || d.n_vars > kMaxVars || d.n_vars < 1 || !Fe_d ||
      (request_jacobian && !Ke_d))
    return LMH_ERR_INVALID;
  if (d.n_elem == 0) return LMH_OK;
  hipStream_t s = (hipStream_t)d.stream;
  const int grid = (int)(d.n_elem < 65536 ? d.n_elem : 65536);
  hipLaunchKernelGGL(k_element_loop<Physics>, dim3(grid), dim3(kWave), lds_bytes(d.n_nodes, d.n_qp, d.n_vars), s,
                     phys, d, u_d, coef_d, request_jacobian ? 1 : 0, Fe_d, Ke_d);
  if (hipGetLastError() != hipSuccess) return LMH_ERR_HIP;
  return lmh_fem_check_flag(&d);
}

}  // namespace lmh_fem

#endif
