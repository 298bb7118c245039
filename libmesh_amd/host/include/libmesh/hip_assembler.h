// hip_assembler.h -- HipAssembler: a System::Assembly that performs the whole
// element loop of a Poisson-type assemble function on the GPU
// (introduction_ex4.C:313-560 with a per-element coefficient as
// adjoints_ex2/L-shaped.C:101): FE::reinit + FEMap for every element,
// Ke = sum_q JxW kappa dphi_i . dphi_j, Fe = sum_q JxW f phi_i, the element
// constraint step (dof_map_constraints.C:2518-2616) and the scatter into the
// system's HipCsrMatrix and rhs (eigen_sparse_matrix.C:121-135,
// numeric_vector.C:690-697).
#pragma once

#include <vector>

#include "libmesh/hip_csr_matrix.h"
#include "libmesh/quadrature_gauss.h"
#include "libmesh/system.h"

namespace libMesh {

class HipAssembler : public System::Assembly {
 public:
  enum Forcing { UNIT = 0, SINE = 1 };  // f = 1, or f = 3 pi^2 sin(pi x) sin(pi y) sin(pi z)

  HipAssembler(ImplicitSystem &system, const MeshBase &mesh, const DofMap &dof_map, const FEType &fe_type,
               const QGauss &qrule, Forcing f = UNIT);
  ~HipAssembler() override;

  // per-element coefficient kappa_e (empty: kappa = 1).  The assembler keeps
  // the values on the host and uploads them at every assembly, so parameters
  // that point into coefficient() (ParameterPointer) act on the next one.
  void set_coefficient(const std::vector<Real> &kappa);
  std::vector<Number> &coefficient() { return _kappa_h; }
  // the element whose coefficient *p is, or (dof_id_type)-1
  dof_id_type coefficient_index(const Number *p) const {
    if (_kappa_h.empty() || p < _kappa_h.data() || p >= _kappa_h.data() + _kappa_h.size()) return (dof_id_type)-1;
    return (dof_id_type)(p - _kappa_h.data());
  }
  // FP32 element contraction + FP32 operator for the mixed-precision solver
  void set_fp32_operator(bool on) { _fp32 = on; }
  void set_scatter(int mode) { _scatter = mode; }

  void assemble() override;

  // The same physics as the FEMSystem residual/Jacobian pair (fem_system.C:875-1010):
  // R(u) = sum_e C^T (K_e u_e - F_e) into R, and the constrained Jacobian K into
  // the system matrix (its rhs is not written).
  void assemble_residual(const NumericVector<Number> &u, NumericVector<Number> &R);
  void assemble_jacobian();

  // dJ/dkappa_e for J = integral of u over [lo,hi]^3, from the forward
  // solution u and the adjoint solution z (adjoint_qoi_parameter_sensitivity,
  // src/systems/implicit_system.C:517-623, for per-element parameters)
  void qoi_box_derivative(Real lo, Real hi, NumericVector<Number> &dq);
  std::vector<Real> kappa_sensitivities(const NumericVector<Number> &u, const NumericVector<Number> &z);

 private:
  ImplicitSystem &_system;
  const MeshBase &_mesh;
  const DofMap &_dof_map;
  FEType _fe_type;
  QGauss _qrule;
  Forcing _f;
  bool _fp32 = false;
  int _scatter = 1;
  void upload_coefficient();
  std::vector<Number> _kappa_h;
  double *_kappa_d = nullptr;
};

}  // namespace libMesh
