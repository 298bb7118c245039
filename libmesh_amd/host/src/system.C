// system.C -- System / ImplicitSystem / LinearImplicitSystem control flow of
// the assembly + solve path (src/systems/system.C, implicit_system.C,
// linear_implicit_system.C), the QoI hooks and the adjoint entry points.
#include "libmesh/system.h"
#include "libmesh/static_condensation.h"

#include <cmath>

#include "libmesh/hip_assembler.h"
#include "libmesh/hip_csr_matrix.h"
#include "libmesh/hip_linear_solver.h"

namespace libMesh {

System::System(EquationSystems &es, const std::string &name, unsigned number)
    : _es(es), _name(name), _number(number), _dof_map(number) {
  _dof_map._system = this;
}

SparseMatrix<Number> &System::add_matrix(std::string_view mat_name, std::unique_ptr<SparseMatrix<Number>> matrix,
                                         ParallelType) {
  if (have_matrix(mat_name)) libmesh_error_msg("matrix " << mat_name << " already exists");
  SparseMatrix<Number> &ref = *matrix;
  _matrices[std::string(mat_name)] = std::move(matrix);
  return ref;
}

SparseMatrix<Number> &System::get_matrix(std::string_view mat_name) {
  auto it = _matrices.find(mat_name);
  if (it == _matrices.end()) libmesh_error_msg("no matrix named " << mat_name);
  return *it->second;
}

const SparseMatrix<Number> &System::get_matrix(std::string_view mat_name) const {
  auto it = _matrices.find(mat_name);
  if (it == _matrices.end()) libmesh_error_msg("no matrix named " << mat_name);
  return *it->second;
}

void System::attach_assemble_function(void fptr(EquationSystems &, const std::string &)) {
  _assemble_fptr = fptr;
  _assemble_obj = nullptr;
}

void System::attach_assemble_object(Assembly &a) {
  _assemble_obj = &a;
  _assemble_fptr = nullptr;
}

void System::attach_QOI_function(void fptr(EquationSystems &, const std::string &, const QoISet &)) {
  _qoi_fptr = fptr;
  _qoi_obj = nullptr;
}

void System::attach_QOI_object(QOI &qoi) {
  _qoi_obj = &qoi;
  _qoi_fptr = nullptr;
}

void System::attach_QOI_derivative(void fptr(EquationSystems &, const std::string &, const QoISet &, bool, bool)) {
  _qoi_derivative_fptr = fptr;
  _qoi_derivative_obj = nullptr;
}

void System::attach_QOI_derivative_object(QOIDerivative &qoi_derivative) {
  _qoi_derivative_obj = &qoi_derivative;
  _qoi_derivative_fptr = nullptr;
}

// System::user_QOI (system.C:2140-2152): nothing attached leaves the values
void System::assemble_qoi(const QoISet &qoi_indices) {
  if (_qoi_obj)
    _qoi_obj->qoi(qoi_indices);
  else if (_qoi_fptr)
    _qoi_fptr(_es, _name, qoi_indices);
}

void System::assemble_qoi_derivative(const QoISet &qoi_indices, bool include_liftfunc, bool apply_constraints) {
  if (_qoi_derivative_obj)
    _qoi_derivative_obj->qoi_derivative(qoi_indices, include_liftfunc, apply_constraints);
  else if (_qoi_derivative_fptr)
    _qoi_derivative_fptr(_es, _name, qoi_indices, include_liftfunc, apply_constraints);
}

namespace {
std::string indexed(const char *base, unsigned int i) { return std::string(base) + std::to_string(i); }
}  // namespace

NumericVector<Number> &System::add_adjoint_solution(unsigned int i) {
  return add_vector(indexed("adjoint_solution", i), false);
}
NumericVector<Number> &System::get_adjoint_solution(unsigned int i) { return get_vector(indexed("adjoint_solution", i)); }
const NumericVector<Number> &System::get_adjoint_solution(unsigned int i) const {
  return get_vector(indexed("adjoint_solution", i));
}
NumericVector<Number> &System::add_adjoint_rhs(unsigned int i) { return add_vector(indexed("adjoint_rhs", i), false); }
NumericVector<Number> &System::get_adjoint_rhs(unsigned int i) { return get_vector(indexed("adjoint_rhs", i)); }
const NumericVector<Number> &System::get_adjoint_rhs(unsigned int i) const {
  return get_vector(indexed("adjoint_rhs", i));
}
NumericVector<Number> &System::add_sensitivity_rhs(unsigned int i) {
  return add_vector(indexed("sensitivity_rhs", i), false);
}
NumericVector<Number> &System::get_sensitivity_rhs(unsigned int i) { return get_vector(indexed("sensitivity_rhs", i)); }

std::pair<unsigned int, Real> System::adjoint_solve(const QoISet &) {
  libmesh_not_implemented();
  return std::make_pair(0, 0.0);
}

void System::adjoint_qoi_parameter_sensitivity(const QoISet &, const ParameterVector &, SensitivityData &) {
  libmesh_not_implemented();
}

void SensitivityData::allocate_data(const QoISet &qoi_indices, const System &sys,
                                    const ParameterVector &parameter_vector) {
  const std::size_t Np = parameter_vector.size();
  const unsigned int Nq = sys.n_qois();
  if (_grad_data.size() < Nq) _grad_data.resize(Nq);
  for (unsigned int i = 0; i != Nq; ++i)
    if (qoi_indices.has_index(i)) _grad_data[i].assign(Np, Number(0));
}

// System::init_data: distribute DoFs, add and process constraint rows, size
// the vectors; then init_matrices
void System::init() {
  MeshBase &mesh = _es.get_mesh();
  _dof_map.distribute_dofs(mesh);
  if (_constrain_obj) _constrain_obj->constrain();
  _dof_map.process_constraints(mesh);
  if (_dof_map.has_static_condensation()) _dof_map.get_static_condensation().reinit();
  solution = std::make_unique<HipVector<Number>>(_es.comm(), _dof_map.n_dofs());
  _rhs = std::make_unique<HipVector<Number>>(_es.comm(), _dof_map.n_dofs());
  rhs = _rhs.get();
  for (auto &kv : _vectors)
    if (kv.second.vec->size() != _dof_map.n_dofs()) kv.second.vec->init(_dof_map.n_dofs());
  init_matrices();
}

NumericVector<Number> &System::add_vector(std::string_view vec_name, const bool projections, const ParallelType type) {
  auto it = _vectors.find(vec_name);
  if (it != _vectors.end()) return *it->second.vec;  // libMesh returns the existing vector
  NamedVector &nv = _vectors[std::string(vec_name)];
  nv.projections = projections;
  nv.type = type;
  // sized now when the DoFs are distributed, else by init()
  nv.vec = std::make_unique<HipVector<Number>>(_es.comm(), solution ? _dof_map.n_dofs() : 0);
  return *nv.vec;
}

NumericVector<Number> &System::get_vector(std::string_view vec_name) {
  auto it = _vectors.find(vec_name);
  if (it == _vectors.end()) libmesh_error_msg("System::get_vector: no vector named " << vec_name);
  return *it->second.vec;
}

const NumericVector<Number> &System::get_vector(std::string_view vec_name) const {
  auto it = _vectors.find(vec_name);
  if (it == _vectors.end()) libmesh_error_msg("System::get_vector: no vector named " << vec_name);
  return *it->second.vec;
}

// Matrices get the DoF map (SparseMatrix::attach_dof_map) and, for the GPU
// matrix, the mesh its pattern is built from.
void System::init_matrices() {
  for (auto &kv : _matrices) {
    kv.second->attach_dof_map(_dof_map);
    if (auto *h = dynamic_cast<HipCsrMatrix<Number> *>(kv.second.get())) h->attach_mesh(_es.get_mesh());
    kv.second->init();
  }
}

// System::user_assembly (system.C:2109-2119)
void System::assemble() {
  if (_assemble_obj)
    _assemble_obj->assemble();
  else if (_assemble_fptr)
    _assemble_fptr(_es, _name);
  else
    libmesh_error_msg("System::assemble: no assembly attached");
}

// System::create_static_condensation (system.C): the DoF map of the
// condensation, filled when the DoFs are distributed
void System::create_static_condensation() { _dof_map.create_static_condensation(_es.get_mesh(), *this); }

// ImplicitSystem::create_static_condensation (implicit_system.C:57-67): the
// system matrix is the StaticCondensation (built by init_matrices)
void ImplicitSystem::create_static_condensation() {
  System::create_static_condensation();
  if (have_matrix("System Matrix"))
    libmesh_error_msg("create_static_condensation: the system already has a System Matrix");
}

StaticCondensation &ImplicitSystem::get_static_condensation() {
  auto *sc = have_matrix("System Matrix") ? dynamic_cast<StaticCondensation *>(&get_system_matrix()) : nullptr;
  if (!sc) libmesh_error_msg("ImplicitSystem::get_static_condensation: no static condensation");
  return *sc;
}

// ImplicitSystem::add_matrices: the "System Matrix" a caller injected before
// init() is adopted; otherwise one is built (implicit_system.C:113-128), a
// StaticCondensation when the system condenses (implicit_system.C:51-60).
void ImplicitSystem::init_matrices() {
  if (!have_matrix("System Matrix") && has_static_condensation())
    add_matrix("System Matrix", std::make_unique<StaticCondensation>(_es.get_mesh(), *this, _dof_map,
                                                                     _dof_map.get_static_condensation()));
  if (!have_matrix("System Matrix")) add_matrix("System Matrix", std::make_unique<HipCsrMatrix<Number>>(_es.comm()));
  if (!linear_solver) linear_solver = std::make_unique<HipLinearSolver<Number>>(_es.comm());
  System::init_matrices();
}

HipAssembler *ImplicitSystem::device_assembler() const { return dynamic_cast<HipAssembler *>(_assemble_obj); }

void ImplicitSystem::assemble() {
  get_system_matrix().zero();
  rhs->zero();
  System::assemble();
}

// LinearImplicitSystem::assembly (linear_implicit_system.C:367-382): the
// residual of the linear system, R = A u - b
void ImplicitSystem::assembly(bool get_residual, bool, bool, bool) {
  assemble();
  rhs->close();
  get_system_matrix().close();
  if (get_residual) {
    *rhs *= -1.0;
    rhs->add_vector(*solution, get_system_matrix());
  }
}

std::pair<unsigned int, Real> ImplicitSystem::adjoint_solve(const QoISet &qoi_indices) {
  if (assemble_before_solve) assembly(false, true);
  LinearSolver<Number> *solver = get_linear_solver();
  if (!solver) libmesh_error_msg("ImplicitSystem::adjoint_solve: no linear solver");
  for (unsigned int i = 0; i < n_qois(); ++i)
    if (qoi_indices.has_index(i)) {
      add_adjoint_rhs(i).zero();
      add_adjoint_solution(i);
    }
  assemble_qoi_derivative(qoi_indices, false, true);
  std::pair<unsigned int, Real> total(0, 0.0);
  for (unsigned int i = 0; i < n_qois(); ++i)
    if (qoi_indices.has_index(i)) {
      const auto r = solver->adjoint_solve(get_system_matrix(), add_adjoint_solution(i), get_adjoint_rhs(i),
                                           _es.linear_solver_tolerance, _es.linear_solver_maximum_iterations);
      total.first += r.first;
      total.second += r.second;
    }
  // the linear solver may not have fit the constraints exactly
  for (unsigned int i = 0; i < n_qois(); ++i)
    if (qoi_indices.has_index(i)) _dof_map.enforce_adjoint_constraints_exactly(get_adjoint_solution(i), i);
  return total;
}

void ImplicitSystem::assemble_residual_derivatives(const ParameterVector &parameters_in) {
  ParameterVector &params = const_cast<ParameterVector &>(parameters_in);
  for (unsigned int p = 0; p != params.size(); ++p) {
    NumericVector<Number> &sensitivity_rhs = add_sensitivity_rhs(p);
    const Number old_parameter = *params[p];
    const Real delta_p = TOLERANCE * std::max(std::abs(old_parameter), 1e-3);
    *params[p] = old_parameter - delta_p;
    assembly(true, false, false);
    rhs->close();
    sensitivity_rhs = *rhs;
    *params[p] = old_parameter + delta_p;
    assembly(true, false, false);
    rhs->close();
    sensitivity_rhs -= *rhs;
    sensitivity_rhs /= (2 * delta_p);
    sensitivity_rhs.close();
    *params[p] = old_parameter;
  }
}

void ImplicitSystem::adjoint_qoi_parameter_sensitivity(const QoISet &qoi_indices, const ParameterVector &parameters_in,
                                                       SensitivityData &sensitivities) {
  ParameterVector &params = const_cast<ParameterVector &>(parameters_in);
  const unsigned int Np = (unsigned int)params.size(), Nq = n_qois();
  if (!is_adjoint_already_solved()) adjoint_solve(qoi_indices);
  // device path: every parameter is a per-element coefficient of the
  // attached HipAssembler
  HipAssembler *hip = device_assembler();
  std::vector<dof_id_type> elem_of;
  if (hip && Np) {
    for (unsigned int j = 0; j < Np; ++j) {
      const dof_id_type e = hip->coefficient_index(&params[j].get());
      if (e == (dof_id_type)-1) {
        elem_of.clear();
        break;
      }
      elem_of.push_back(e);
    }
  }
  if (elem_of.empty()) assemble_residual_derivatives(parameters_in);
  sensitivities.allocate_data(qoi_indices, *this, params);
  for (unsigned int i = 0; i != Nq; ++i) {
    if (!qoi_indices.has_index(i)) continue;
    std::vector<Real> dev;
    // -z_i . partial R / partial kappa_e for every element at once
    if (!elem_of.empty()) dev = hip->kappa_sensitivities(*solution, get_adjoint_solution(i));
    for (unsigned int j = 0; j != Np; ++j)
      sensitivities[i][j] =
          elem_of.empty() ? get_sensitivity_rhs(j).dot(get_adjoint_solution(i)) : (Number)dev[elem_of[j]];
  }
  // partial q / partial p by central differences of the attached QoI
  if (_qoi_obj || _qoi_fptr) {
    for (unsigned int j = 0; j != Np; ++j) {
      const Number old_parameter = *params[j];
      const Real delta_p = TOLERANCE * std::max(std::abs(old_parameter), 1e-3);
      *params[j] = old_parameter - delta_p;
      assemble_qoi(qoi_indices);
      const std::vector<Number> qoi_minus = get_qoi_values();
      *params[j] = old_parameter + delta_p;
      assemble_qoi(qoi_indices);
      const std::vector<Number> qoi_plus = get_qoi_values();
      *params[j] = old_parameter;
      for (unsigned int i = 0; i != Nq; ++i)
        if (qoi_indices.has_index(i)) sensitivities[i][j] += (qoi_plus[i] - qoi_minus[i]) / (2. * delta_p);
    }
    assemble_qoi(qoi_indices);
  }
}

// LinearImplicitSystem::solve (linear_implicit_system.C:122-161): the solve
// with the "Preconditioner" matrix when the system has one
void LinearImplicitSystem::solve() {
  if (assemble_before_solve) assemble();
  linear_solver->init();
  linear_solver->init_systems(*this);
  const auto r = linear_solver->solve(get_system_matrix(), request_matrix("Preconditioner"), *solution, *rhs,
                                      _es.linear_solver_tolerance, _es.linear_solver_maximum_iterations);
  _n_linear_iterations = r.first;
  _final_linear_residual = r.second;
}

}  // namespace libMesh
