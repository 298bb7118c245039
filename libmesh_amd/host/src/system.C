// system.C -- System / ImplicitSystem / LinearImplicitSystem control flow of
// the assembly + solve path (src/systems/system.C, implicit_system.C,
// linear_implicit_system.C).
#include "libmesh/system.h"

#include "libmesh/hip_csr_matrix.h"
#include "libmesh/hip_linear_solver.h"

namespace libMesh {

System::System(EquationSystems &es, const std::string &name, unsigned number)
    : _es(es), _name(name), _number(number), _dof_map(number) {}

SparseMatrix<Number> &System::add_matrix(const std::string &mat_name, std::unique_ptr<SparseMatrix<Number>> matrix) {
  if (_matrices.count(mat_name)) libmesh_error_msg("matrix " << mat_name << " already exists");
  SparseMatrix<Number> &ref = *matrix;
  _matrices[mat_name] = std::move(matrix);
  return ref;
}

SparseMatrix<Number> &System::get_matrix(const std::string &mat_name) {
  auto it = _matrices.find(mat_name);
  if (it == _matrices.end()) libmesh_error_msg("no matrix named " << mat_name);
  return *it->second;
}

void System::attach_assemble_function(void (*fptr)(EquationSystems &, const std::string &)) {
  _assemble_fptr = fptr;
  _assemble_obj = nullptr;
}

void System::attach_assemble_object(Assembly &a) {
  _assemble_obj = &a;
  _assemble_fptr = nullptr;
}

// System::init_data: distribute DoFs, size the vectors; then init_matrices
void System::init() {
  _dof_map.distribute_dofs(_es.get_mesh());
  if (_constrain_obj) _constrain_obj->constrain();
  auto sol = std::make_unique<HipVector<Number>>(_dof_map.n_dofs());
  solution = std::move(sol);
  _rhs = std::make_unique<HipVector<Number>>(_dof_map.n_dofs());
  rhs = _rhs.get();
  for (auto &kv : _vectors)
    if (kv.second.vec->size() != _dof_map.n_dofs()) kv.second.vec->init(_dof_map.n_dofs());
  init_matrices();
}

NumericVector<Number> &System::add_vector(const std::string &vec_name, bool projections, ParallelType type) {
  auto it = _vectors.find(vec_name);
  if (it != _vectors.end()) return *it->second.vec;  // libMesh returns the existing vector
  NamedVector &nv = _vectors[vec_name];
  nv.projections = projections;
  nv.type = type;
  // sized now when the DoFs are distributed, else by init()
  nv.vec = std::make_unique<HipVector<Number>>(solution ? _dof_map.n_dofs() : 0);
  return *nv.vec;
}

NumericVector<Number> &System::get_vector(const std::string &vec_name) {
  auto it = _vectors.find(vec_name);
  if (it == _vectors.end()) libmesh_error_msg("System::get_vector: no vector named " << vec_name);
  return *it->second.vec;
}

const NumericVector<Number> &System::get_vector(const std::string &vec_name) const {
  auto it = _vectors.find(vec_name);
  if (it == _vectors.end()) libmesh_error_msg("System::get_vector: no vector named " << vec_name);
  return *it->second.vec;
}

void System::init_matrices() {
  for (auto &kv : _matrices) {
    kv.second->attach_dof_map(_dof_map);
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

// ImplicitSystem::add_matrices: the "System Matrix" a caller injected before
// init() is adopted; otherwise one is built (implicit_system.C:113-128).
void ImplicitSystem::init_matrices() {
  if (!have_matrix("System Matrix")) add_matrix("System Matrix", std::make_unique<HipCsrMatrix<Number>>());
  if (!linear_solver) linear_solver = std::make_unique<HipLinearSolver<Number>>();
  System::init_matrices();
}

void ImplicitSystem::assemble() {
  get_system_matrix().zero();
  rhs->zero();
  System::assemble();
}

std::pair<unsigned int, Real> ImplicitSystem::adjoint_solve(NumericVector<Number> &z, NumericVector<Number> &dq) {
  const auto r = linear_solver->adjoint_solve(get_system_matrix(), z, dq, _es.linear_solver_tolerance,
                                              _es.linear_solver_maximum_iterations);
  // the linear solver may not have fit the constraints exactly:
  // enforce_adjoint_constraints_exactly (implicit_system.C:230-240), the
  // homogeneous form of the constraint rows
  if (!_dof_map.added_constraint_rows().empty()) _dof_map.enforce_constraints_exactly(*this, &z, true);
  return r;
}

void LinearImplicitSystem::solve() {
  if (assemble_before_solve) assemble();
  linear_solver->init();
  linear_solver->init_systems(*this);
  const auto r = linear_solver->solve(get_system_matrix(), *solution, *rhs, _es.linear_solver_tolerance,
                                      _es.linear_solver_maximum_iterations);
  _n_linear_iterations = r.first;
  _final_linear_residual = r.second;
}

}  // namespace libMesh
