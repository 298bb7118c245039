// static_condensation.C -- StaticCondensation over the device condensation
// (csrc/lmh_condense.hip): the reference's element Schur complements
// (static_condensation.C:185-227), forward elimination (:383-417), back
// substitution (:419-456) and apply (:458-473).
#include "libmesh/static_condensation.h"

#include "libmesh/hip_device.h"
#include "libmesh/hip_linear_solver.h"
#include "libmesh/system.h"

namespace libMesh {

namespace {
// a device copy of a NumericVector that is not a HipVector (or the vector itself)
struct DeviceCopy {
  HipVector<Number> *hv;
  std::unique_ptr<HipVector<Number>> tmp;
  explicit DeviceCopy(const NumericVector<Number> &v)
      : hv(dynamic_cast<HipVector<Number> *>(const_cast<NumericVector<Number> *>(&v))) {
    if (!hv) {
      tmp = std::make_unique<HipVector<Number>>(v.comm(), v.size(), v.size());
      std::vector<Number> vals;
      v.localize(vals);
      *tmp = vals;
      hv = tmp.get();
    }
    hv->close();
  }
};
}  // namespace

StaticCondensation::StaticCondensation(const MeshBase &mesh, System &system, const DofMap &full_dof_map,
                                       StaticCondensationDofMap &reduced_dof_map)
    : HipCsrMatrix<Number>(mesh.comm()),
      _sc_mesh(mesh),
      _system(system),
      _reduced_dof_map(reduced_dof_map) {}

StaticCondensation::~StaticCondensation() { clear(); }

void StaticCondensation::clear() {
  _reduced_view.reset();
  if (_sc) lmh_condensation_destroy(_sc);
  _sc = nullptr;
  _sc_stale = true;
  HipCsrMatrix<Number>::clear();
}

void StaticCondensation::init(numeric_index_type m, numeric_index_type n, numeric_index_type m_l,
                              numeric_index_type n_l, numeric_index_type nnz, numeric_index_type noz,
                              numeric_index_type blocksize) {
  HipCsrMatrix<Number>::init(m, n, m_l, n_l, nnz, noz, blocksize);
  init();
}

void StaticCondensation::init(ParallelType type) {
  HipCsrMatrix<Number>::init(type);
  init();
}

void StaticCondensation::init() {
  if (!_reduced_dof_map.initialized()) _reduced_dof_map.reinit();
  if (!_reduced_solver) {
    _reduced_solver = std::make_unique<HipLinearSolver<Number>>(this->comm());
    _reduced_solver->init((_system.name() + "_condensed_").c_str());
  }
}

// The full values are final; the reduced matrix follows them when it is next
// needed (HipCsrMatrix operations close the matrix on their way, and values
// between those steps are not meant to be condensed).
void StaticCondensation::close() {
  HipCsrMatrix<Number>::close();
  _sc_stale = true;
}

// the reduced pattern (first time) and the Schur complement values from the
// current full values
void StaticCondensation::condense() const {
  if (!device_matrix()) libmesh_error_msg("StaticCondensation: the matrix is not initialised");
  if (!_sc) {
    const std::vector<dof_id_type> &c = _reduced_dof_map.condensed_dofs();
    std::vector<int32_t> cond(c.begin(), c.end());
    hip_check(lmh_condensation_create(device_matrix(), (int64_t)cond.size(), cond.data(), &_sc),
              "StaticCondensation (reduced pattern)");
  } else if (_sc_stale) {
    hip_check(lmh_condensation_setup(_sc), "StaticCondensation (Schur complements)");
  }
  _sc_stale = false;
  if (!_reduced_view) _reduced_view = std::make_unique<HipCsrMatrix<Number>>(this->comm());
  _reduced_view->view_device_matrix(lmh_condensation_reduced_matrix(_sc));
}

void StaticCondensation::setup() {
  if (!closed()) libmesh_error_msg("StaticCondensation::setup: close the matrix first");
  condense();
}

const SparseMatrix<Number> &StaticCondensation::get_condensed_mat() const {
  if (!_sc || _sc_stale) condense();
  return *_reduced_view;
}

LinearSolver<Number> &StaticCondensation::reduced_system_solver() {
  init();
  return *_reduced_solver;
}

void StaticCondensation::dont_condense_vars(const std::unordered_set<unsigned int> &vars) {
  if (_sc) libmesh_error_msg("StaticCondensation::dont_condense_vars: the condensation is already built");
  _reduced_dof_map.dont_condense_vars(vars);
  _reduced_dof_map.clear();
}

LinearConvergenceReason StaticCondensation::reduced_converged_reason() const {
  return _reduced_solver ? _reduced_solver->get_converged_reason() : UNKNOWN_FLAG;
}

std::pair<unsigned int, Real> StaticCondensation::condensed_solve(const NumericVector<Number> &full_rhs,
                                                                  NumericVector<Number> &full_sol, double rtol,
                                                                  unsigned int max_its) {
  init();
  close();  // values the device assembly wrote count as final
  condense();
  const numeric_index_type nr = (numeric_index_type)lmh_matrix_rows(lmh_condensation_reduced_matrix(_sc));
  DeviceCopy b(full_rhs), x(full_sol);
  HipVector<Number> g(this->comm(), nr, nr), y(this->comm(), nr, nr);
  // forward elimination: g = b_u - A_uc A_cc^-1 b_c
  hip_check(lmh_condensation_forward(_sc, b.hv->device_data(), g.device_data()),
            "StaticCondensation (forward elimination)");
  g.mark_device_modified();
  // the reduced solve from zero, as apply zeroes the solution it is handed
  y.zero();
  const auto r = _reduced_solver->solve(*_reduced_view, y, g, rtol, max_its);
  // back substitution: x_u = y, x_c = A_cc^-1 (b_c - A_cu x_u)
  y.close();
  hip_check(lmh_condensation_backward(_sc, b.hv->device_data(), y.device_data(), x.hv->device_data()),
            "StaticCondensation (back substitution)");
  x.hv->mark_device_modified();
  if (x.tmp) {
    std::vector<Number> vals;
    x.tmp->localize(vals);
    full_sol = vals;
  }
  full_sol.close();
  return r;
}

void StaticCondensation::apply(const NumericVector<Number> &full_rhs, NumericVector<Number> &full_sol) {
  condensed_solve(full_rhs, full_sol, 1e-5, 300);
}

}  // namespace libMesh
