// device_physics_hip.h -- DevicePhysicsFor<P>: a DevicePhysics whose
// element_time_derivative is the device member of P (see lmh_fem_physics.h
// for the Context it receives).  Include from a translation unit compiled with
// hipcc --offload-arch=gfx950:
//
//   struct MyPhysics {
//     __device__ void element_time_derivative(bool request_jacobian, lmh_fem::Context &c) const;
//   };
//   libMesh::DevicePhysicsFor<MyPhysics> physics;
//   fem_system.attach_physics(&physics);
#pragma once

#include "libmesh/device_physics.h"
#include "../../../../include/lmh_fem_physics.h"

namespace libMesh {

template <class P>
class DevicePhysicsFor : public DevicePhysics {
 public:
  explicit DevicePhysicsFor(const P &p = P()) : _p(p) {}
  P &physics() { return _p; }
  const P &physics() const { return _p; }

  int element_loop(const lmh_fem_data &d, const double *u_d, const double *coef_d, bool request_jacobian,
                   double *Fe_d, double *Ke_d) const override {
    return lmh_fem::element_loop(d, _p, u_d, coef_d, request_jacobian, Fe_d, Ke_d);
  }

 private:
  P _p;
};

}  // namespace libMesh
