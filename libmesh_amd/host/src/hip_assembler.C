// hip_assembler.C -- System::Assembly object that runs the element loop on the GPU.
#include "libmesh/hip_assembler.h"

namespace libMesh {

HipAssembler::HipAssembler(ImplicitSystem &system, const MeshBase &mesh, const DofMap &dof_map,
                           const FEType &fe_type, const QGauss &qrule, Forcing f)
    : _system(system), _mesh(mesh), _dof_map(dof_map), _fe_type(fe_type), _qrule(qrule), _f(f) {}

HipAssembler::~HipAssembler() {
  if (_kappa_d) lmh_free(HipDevice::context(), _kappa_d);
}

void HipAssembler::set_coefficient(const std::vector<Real> &kappa) {
  if (!kappa.empty() && kappa.size() != _mesh.n_elem()) libmesh_error_msg("coefficient needs one value per element");
  _kappa_h.assign(kappa.begin(), kappa.end());
  upload_coefficient();
}

// the host coefficients to the device (2 MB at 262,144 elements)
void HipAssembler::upload_coefficient() {
  lmh_context *ctx = HipDevice::context();
  if (_kappa_h.empty()) {
    if (_kappa_d) lmh_free(ctx, _kappa_d);
    _kappa_d = nullptr;
    return;
  }
  if (_kappa_h.size() != _mesh.n_elem()) libmesh_error_msg("coefficient needs one value per element");
  if (!_kappa_d) {
    void *p = nullptr;
    hip_check(lmh_malloc(ctx, sizeof(double) * _kappa_h.size(), &p), "HipAssembler (coefficient)");
    _kappa_d = (double *)p;
  }
  hip_check(lmh_memcpy_h2d(ctx, _kappa_d, _kappa_h.data(), sizeof(double) * _kappa_h.size()),
            "HipAssembler (coefficient)");
}

void HipAssembler::assemble() {
  auto *A = dynamic_cast<HipCsrMatrix<Number> *>(&_system.get_system_matrix());
  auto *b = dynamic_cast<HipVector<Number> *>(_system.rhs);
  if (!A || !b) libmesh_error_msg("HipAssembler needs a HipCsrMatrix system matrix and HipVector rhs");
  if (!A->device_matrix() || !A->device_dofmap()) libmesh_error_msg("HipAssembler: matrix not initialised from the DofMap");
  if (&_dof_map != A->dof_map()) libmesh_error_msg("HipAssembler: DofMap differs from the matrix's");
  if (_fe_type.order != _dof_map.variable_type(0).order) libmesh_error_msg("HipAssembler: FEType mismatch");
  upload_coefficient();
  const int arg = _qrule.kernel_arg(_mesh.elem_type());
  hip_check(lmh_assemble_poisson(A->device_matrix(), A->device_mesh(), A->device_dofmap(), arg, _kappa_d, (int)_f,
                                 (double *)b->device_data(), LMH_FP64, _scatter),
            "HipAssembler::assemble");
  if (_fp32) hip_check(lmh_matrix_make_fp32(A->device_matrix()), "HipAssembler::assemble (FP32 operator)");
  A->mark_device_modified();
  b->mark_device_modified();
}

void HipAssembler::assemble_residual(const NumericVector<Number> &u_in, NumericVector<Number> &R_in) {
  auto *A = dynamic_cast<HipCsrMatrix<Number> *>(&_system.get_system_matrix());
  const auto *u = dynamic_cast<const HipVector<Number> *>(&u_in);
  auto *R = dynamic_cast<HipVector<Number> *>(&R_in);
  if (!A || !u || !R) libmesh_error_msg("assemble_residual: HipCsrMatrix / HipVector required");
  if (!A->device_matrix()) libmesh_error_msg("assemble_residual: matrix not initialised from the DofMap");
  upload_coefficient();
  const int arg = _qrule.kernel_arg(_mesh.elem_type());
  hip_check(lmh_assemble_residual_poisson(A->device_matrix(), A->device_mesh(), A->device_dofmap(), arg, _kappa_d,
                                          (int)_f, (const double *)u->device_data(), (double *)R->device_data()),
            "HipAssembler::assemble_residual");
  R->mark_device_modified();
}

void HipAssembler::assemble_jacobian() {
  auto *A = dynamic_cast<HipCsrMatrix<Number> *>(&_system.get_system_matrix());
  if (!A || !A->device_matrix()) libmesh_error_msg("assemble_jacobian: HipCsrMatrix initialised from the DofMap required");
  upload_coefficient();
  const int arg = _qrule.kernel_arg(_mesh.elem_type());
  hip_check(lmh_assemble_poisson(A->device_matrix(), A->device_mesh(), A->device_dofmap(), arg, _kappa_d, (int)_f,
                                 nullptr, LMH_FP64, _scatter),
            "HipAssembler::assemble_jacobian");
  if (_fp32) hip_check(lmh_matrix_make_fp32(A->device_matrix()), "HipAssembler::assemble_jacobian (FP32 operator)");
  A->mark_device_modified();
}

void HipAssembler::qoi_box_derivative(Real lo, Real hi, NumericVector<Number> &dq_in) {
  auto *A = dynamic_cast<HipCsrMatrix<Number> *>(&_system.get_system_matrix());
  auto *dq = dynamic_cast<HipVector<Number> *>(&dq_in);
  if (!A || !dq) libmesh_error_msg("qoi_box_derivative: HipCsrMatrix / HipVector required");
  const numeric_index_type nd = _dof_map.n_dofs();
  if (dq->size() != nd) dq->init(nd, nd);
  const int arg = _qrule.kernel_arg(_mesh.elem_type());
  hip_check(lmh_qoi_box_derivative(A->device_mesh(), A->device_dofmap(), arg, lo, hi, (double *)dq->device_data()),
            "qoi_box_derivative");
  dq->mark_device_modified();
}

std::vector<Real> HipAssembler::kappa_sensitivities(const NumericVector<Number> &u_in,
                                                    const NumericVector<Number> &z_in) {
  auto *A = dynamic_cast<HipCsrMatrix<Number> *>(&_system.get_system_matrix());
  const auto *u = dynamic_cast<const HipVector<Number> *>(&u_in);
  const auto *z = dynamic_cast<const HipVector<Number> *>(&z_in);
  if (!A || !u || !z) libmesh_error_msg("kappa_sensitivities: HipCsrMatrix / HipVector required");
  lmh_context *ctx = HipDevice::context();
  upload_coefficient();
  void *g = nullptr;
  hip_check(lmh_malloc(ctx, sizeof(double) * _mesh.n_elem(), &g), "kappa_sensitivities");
  const int arg = _qrule.kernel_arg(_mesh.elem_type());
  const int rc = lmh_adjoint_kappa_gradient(A->device_mesh(), A->device_dofmap(), arg, _kappa_d,
                                            (const double *)u->device_data(), (const double *)z->device_data(),
                                            (double *)g);
  std::vector<Real> out(_mesh.n_elem());
  if (rc == LMH_OK) lmh_memcpy_d2h(ctx, out.data(), g, sizeof(double) * out.size());
  lmh_free(ctx, g);
  hip_check(rc, "kappa_sensitivities");
  return out;
}

}  // namespace libMesh
