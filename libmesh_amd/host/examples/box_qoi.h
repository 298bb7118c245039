// box_qoi.h -- the example programs' System::QOIDerivative: dJ/du of
// J = integral of u over [0.25, 0.75]^3 into the adjoint right-hand side
// (the QoI derivative of adjoints_ex2/L-qoi.C:29-95), computed on the device.
#pragma once

#include "libmesh/hip_assembler.h"
#include "libmesh/system.h"

namespace libMesh {

class BoxQoIDerivative : public System::QOIDerivative {
 public:
  BoxQoIDerivative(System &sys, HipAssembler &assembler) : _sys(sys), _assembler(assembler) {}
  void qoi_derivative(const QoISet &qoi_indices, bool, bool) override {
    if (qoi_indices.has_index(0)) _assembler.qoi_box_derivative(0.25, 0.75, _sys.get_adjoint_rhs(0));
  }

 private:
  System &_sys;
  HipAssembler &_assembler;
};

}  // namespace libMesh
