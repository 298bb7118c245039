// device_physics.h -- DifferentiablePhysics (include/physics/diff_physics.h:76)
// whose element_time_derivative runs on the GPU.  FEMSystem::assembly calls
// element_loop() once per assembly for all elements; the element code itself
// is a device functor (lmh_fem_physics.h) wrapped by DevicePhysicsFor<P> in
// device_physics_hip.h, which a hipcc-compiled translation unit instantiates.
// This header is plain C++, so host code compiled by g++ can hold and attach
// a DevicePhysics.
#pragma once

#include <vector>

#include "libmesh/libmesh_common.h"
#include "../../../../include/lmhip.h"

namespace libMesh {

class DevicePhysics {
 public:
  virtual ~DevicePhysics() = default;

  // element_time_derivative (diff_physics.h:125) of every element of d at the
  // device solution u_d, with per-element coefficients coef_d (NULL: 1):
  // constrained element residuals into Fe_d (n_elem * n_dofs) and, with
  // request_jacobian, Jacobians into Ke_d (n_elem * n_dofs^2).  Returns an
  // lmh status code.
  virtual int element_loop(const lmh_fem_data &d, const double *u_d, const double *coef_d, bool request_jacobian,
                           double *Fe_d, double *Ke_d) const = 0;

  // per-element coefficient read by Context::coefficient() (empty: 1),
  // uploaded by FEMSystem at every assembly
  std::vector<Real> &coefficient() { return _coef; }
  const std::vector<Real> &coefficient() const { return _coef; }

 private:
  std::vector<Real> _coef;
};

}  // namespace libMesh
