/* lmh_fe_map.h -- the per-quadrature-point element map of the device
 * kernels (FEMap::compute_single_point_map, src/fe/fe_map.C:958-1090, 3D):
 *   J[r][c] = sum_a X_a[c] dphi_a/dxi_r(q),  det = det J,
 *   G[r][c] = dxi_r/dx_c (the inverse map), dphi_a/dx_c = sum_r dref_a[r] G[r][c].
 * Inputs: element node coordinates X[a*3+c] and the reference derivative
 * table dref[(a*nq+q)*3+r] of the element's Lagrange shapes.  Returns det J;
 * G is written only when det J > 0 (a non-positive det J is a degenerate map,
 * fe_map.C:467-505).  Shared by the library kernels and the user-physics
 * element loop of lmh_fem_physics.h.  HIP device code. */
#ifndef LMH_FE_MAP_H
#define LMH_FE_MAP_H

#include <hip/hip_runtime.h>

__device__ inline double lmh_qp_map(const double *X, const double *dref, int n, int nq, int q, double G[3][3]) {
  double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};  // J[r][c] = d x_c / d xi_r
#pragma unroll 4
  for (int a = 0; a < n; ++a) {
    const double *d = &dref[(a * nq + q) * 3];
    const double x = X[a * 3], y = X[a * 3 + 1], z = X[a * 3 + 2];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      J[r][0] += x * d[r];
      J[r][1] += y * d[r];
      J[r][2] += z * d[r];
    }
  }
  const double dx_dxi = J[0][0], dy_dxi = J[0][1], dz_dxi = J[0][2];
  const double dx_deta = J[1][0], dy_deta = J[1][1], dz_deta = J[1][2];
  const double dx_dzeta = J[2][0], dy_dzeta = J[2][1], dz_dzeta = J[2][2];
  const double det = (dx_dxi * (dy_deta * dz_dzeta - dz_deta * dy_dzeta) +
                      dy_dxi * (dz_deta * dx_dzeta - dx_deta * dz_dzeta) +
                      dz_dxi * (dx_deta * dy_dzeta - dy_deta * dx_dzeta));
  if (!(det > 0.0)) return det;
  const double inv = 1.0 / det;
  G[0][0] = (dy_deta * dz_dzeta - dz_deta * dy_dzeta) * inv;
  G[0][1] = (dz_deta * dx_dzeta - dx_deta * dz_dzeta) * inv;
  G[0][2] = (dx_deta * dy_dzeta - dy_deta * dx_dzeta) * inv;
  G[1][0] = (dz_dxi * dy_dzeta - dy_dxi * dz_dzeta) * inv;
  G[1][1] = (dx_dxi * dz_dzeta - dz_dxi * dx_dzeta) * inv;
  G[1][2] = (dy_dxi * dx_dzeta - dx_dxi * dy_dzeta) * inv;
  G[2][0] = (dy_dxi * dz_deta - dz_dxi * dy_deta) * inv;
  G[2][1] = (dz_dxi * dx_deta - dx_dxi * dz_deta) * inv;
  G[2][2] = (dx_dxi * dy_deta - dy_dxi * dx_deta) * inv;
  return det;
}

#endif
