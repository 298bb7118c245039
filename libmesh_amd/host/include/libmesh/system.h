// system.h -- the System / ImplicitSystem / LinearImplicitSystem /
// EquationSystems surface the assembly + solve path goes through
// (include/systems/*.h).  Injection points are libMesh's own:
//   System::add_matrix("System Matrix", unique_ptr<SparseMatrix>)  (system.C:1041-1061)
//   ImplicitSystem::linear_solver, a public unique_ptr            (implicit_system.h:328)
//   System::attach_assemble_object / attach_assemble_function     (system.C:1959-1987)
#pragma once

#include <functional>
#include <map>
#include <memory>
#include <string>
#include <string_view>

#include "libmesh/dof_map.h"
#include "libmesh/linear_solver.h"
#include "libmesh/parameter_vector.h"
#include "libmesh/qoi_set.h"
#include "libmesh/replicated_mesh.h"
#include "libmesh/sensitivity_data.h"
#include "libmesh/xdr_cxx.h"

namespace libMesh {

class EquationSystems;
class HipAssembler;
class StaticCondensation;

class System {
 public:
  // System::Assembly (include/systems/system.h:170)
  class Assembly {
   public:
    virtual ~Assembly() = default;
    virtual void assemble() = 0;
  };

  System(EquationSystems &es, const std::string &name, unsigned number);
  virtual ~System() = default;

  const std::string &name() const { return _name; }
  unsigned number() const { return _number; }
  // System::system_type (include/systems/system.h:510 and overrides)
  virtual std::string system_type() const { return "Basic"; }
  // DoFs in the order of EquationSystems files: nodes by id, each node's
  // variables in order (System::write_serialized_blocked_dof_objects,
  // src/systems/system_io.C:1626-1910)
  std::vector<dof_id_type> node_major_dofs() const;
  EquationSystems &get_equation_systems() { return _es; }
  DofMap &get_dof_map() { return _dof_map; }
  const DofMap &get_dof_map() const { return _dof_map; }
  unsigned add_variable(const std::string &var, Order order = FIRST, FEFamily family = LAGRANGE) {
    return _dof_map.add_variable(var, FEType(order, family));
  }

  // System::add_matrix with a caller-built matrix (system.C:1041-1061)
  SparseMatrix<Number> &add_matrix(std::string_view mat_name, std::unique_ptr<SparseMatrix<Number>> matrix,
                                   ParallelType type = PARALLEL);
  const SparseMatrix<Number> &get_matrix(std::string_view mat_name) const;
  SparseMatrix<Number> &get_matrix(std::string_view mat_name);
  bool have_matrix(std::string_view mat_name) const { return _matrices.find(mat_name) != _matrices.end(); }

  void attach_assemble_function(void fptr(EquationSystems &es, const std::string &name));
  void attach_assemble_object(Assembly &assemble);

  // System::Constraint (include/systems/system.h:232): user constraint rows,
  // added with DofMap::add_constraint_row when the DoFs are distributed
  // (DofMap::create_dof_constraints calls it, dof_map_constraints.C:1736-1800)
  class Constraint {
   public:
    virtual ~Constraint() = default;
    virtual void constrain() = 0;
  };
  void attach_constraint_object(Constraint &constrain) { _constrain_obj = &constrain; }

  // System::QOI and System::QOIDerivative (include/systems/system.h:205-245):
  // user quantities of interest, set with set_qoi, and their derivatives with
  // respect to the solution, written to get_adjoint_rhs(i)
  class QOI {
   public:
    virtual ~QOI() = default;
    virtual void qoi(const QoISet &qoi_indices) = 0;
  };
  class QOIDerivative {
   public:
    virtual ~QOIDerivative() = default;
    virtual void qoi_derivative(const QoISet &qoi_indices, bool include_liftfunc, bool apply_constraints) = 0;
  };
  void attach_QOI_function(void fptr(EquationSystems &es, const std::string &name, const QoISet &qoi_indices));
  void attach_QOI_object(QOI &qoi);
  void attach_QOI_derivative(void fptr(EquationSystems &es, const std::string &name, const QoISet &qoi_indices,
                                       bool include_liftfunc, bool apply_constraints));
  void attach_QOI_derivative_object(QOIDerivative &qoi_derivative);
  // System::assemble_qoi / assemble_qoi_derivative (system.C:575-600): the
  // attached user functions
  virtual void assemble_qoi(const QoISet &qoi_indices = QoISet());
  virtual void assemble_qoi_derivative(const QoISet &qoi_indices = QoISet(), bool include_liftfunc = true,
                                       bool apply_constraints = true);
  // QoI values (system.h:1682-1720)
  unsigned int n_qois() const { return (unsigned int)_qoi.size(); }
  void init_qois(unsigned int n_qois) { _qoi.resize(n_qois, Number(0)); }
  void set_qoi(unsigned int qoi_index, Number qoi_value) { _qoi.at(qoi_index) = qoi_value; }
  Number get_qoi_value(unsigned int qoi_index) const { return _qoi.at(qoi_index); }
  std::vector<Number> get_qoi_values() const { return _qoi; }
  // adjoint, adjoint-rhs and sensitivity-rhs vectors, named as libMesh names
  // them ("adjoint_solution0", "adjoint_rhs0", "sensitivity_rhs0";
  // system.C:1057-1180)
  NumericVector<Number> &add_adjoint_solution(unsigned int i = 0);
  NumericVector<Number> &get_adjoint_solution(unsigned int i = 0);
  const NumericVector<Number> &get_adjoint_solution(unsigned int i = 0) const;
  NumericVector<Number> &add_adjoint_rhs(unsigned int i = 0);
  NumericVector<Number> &get_adjoint_rhs(unsigned int i = 0);
  const NumericVector<Number> &get_adjoint_rhs(unsigned int i = 0) const;
  NumericVector<Number> &add_sensitivity_rhs(unsigned int i = 0);
  NumericVector<Number> &get_sensitivity_rhs(unsigned int i = 0);
  void set_adjoint_already_solved(bool setting) { _adjoint_already_solved = setting; }
  bool is_adjoint_already_solved() const { return _adjoint_already_solved; }
  // System::adjoint_solve / adjoint_qoi_parameter_sensitivity (system.h:388,
  // 447): not implemented by the base class
  virtual std::pair<unsigned int, Real> adjoint_solve(const QoISet &qoi_indices = QoISet());
  virtual void adjoint_qoi_parameter_sensitivity(const QoISet &qoi_indices, const ParameterVector &parameters,
                                                 SensitivityData &sensitivities);

  // System::init -> init_data (dof distribution, vectors) + init_matrices (system.C:196-371)
  virtual void init();
  // System::assemble -> user_assembly (system.C:554-561, 2109-2119)
  virtual void assemble();
  virtual void solve() {}

  std::unique_ptr<NumericVector<Number>> solution;
  NumericVector<Number> *rhs = nullptr;
  // quadrature order added to the FE type's default rule (system.h:1640)
  int extra_quadrature_order = 0;

  // System::create_static_condensation (system.h:1985): condense the DoFs of
  // element-internal nodes; call before EquationSystems::init
  virtual void create_static_condensation();
  bool has_static_condensation() const { return _dof_map.has_static_condensation(); }

  // System::add_vector / get_vector / have_vector / n_vectors
  // (src/systems/system.C:751-800, 893-960): named vectors sized like the
  // solution (created now when the system is initialised, else by init());
  // listed, with their projection flag and parallel type, in EquationSystems
  // files written with WRITE_ADDITIONAL_DATA.
  NumericVector<Number> &add_vector(std::string_view vec_name, const bool projections = true,
                                    const ParallelType type = PARALLEL);
  const NumericVector<Number> &get_vector(std::string_view vec_name) const;
  NumericVector<Number> &get_vector(std::string_view vec_name);
  bool have_vector(std::string_view vec_name) const { return _vectors.find(vec_name) != _vectors.end(); }
  unsigned int n_vectors() const { return (unsigned int)_vectors.size(); }
  struct NamedVector {
    std::unique_ptr<NumericVector<Number>> vec;
    bool projections = true;
    ParallelType type = PARALLEL;
  };
  const std::map<std::string, NamedVector, std::less<>> &vectors() const { return _vectors; }

 protected:
  virtual void init_matrices();
  EquationSystems &_es;
  std::string _name;
  unsigned _number;
  DofMap _dof_map;
  std::map<std::string, std::unique_ptr<SparseMatrix<Number>>, std::less<>> _matrices;
  std::map<std::string, NamedVector, std::less<>> _vectors;  // name order, as libMesh's std::map
  std::unique_ptr<NumericVector<Number>> _rhs;
  void (*_assemble_fptr)(EquationSystems &, const std::string &) = nullptr;
  Assembly *_assemble_obj = nullptr;
  Constraint *_constrain_obj = nullptr;
  void (*_qoi_fptr)(EquationSystems &, const std::string &, const QoISet &) = nullptr;
  QOI *_qoi_obj = nullptr;
  void (*_qoi_derivative_fptr)(EquationSystems &, const std::string &, const QoISet &, bool, bool) = nullptr;
  QOIDerivative *_qoi_derivative_obj = nullptr;
  std::vector<Number> _qoi;
  bool _adjoint_already_solved = false;
};

class ImplicitSystem : public System {
 public:
  using System::System;
  std::string system_type() const override { return "Implicit"; }
  SparseMatrix<Number> &get_system_matrix() { return get_matrix("System Matrix"); }
  // ImplicitSystem::assemble zeroes matrix and rhs first (implicit_system.C:94-108)
  void assemble() override;
  // ImplicitSystem::assembly (implicit_system.h:120; LinearImplicitSystem's,
  // linear_implicit_system.C:367-382): assemble, then with get_residual the
  // residual R = A u - b into rhs
  virtual void assembly(bool get_residual, bool get_jacobian, bool apply_heterogeneous_constraints = false,
                        bool apply_no_constraints = false);
  // ImplicitSystem::adjoint_solve (implicit_system.C:196-240): assembles the
  // matrix, the QoI derivatives into the adjoint right-hand sides, solves
  // K^T z_i = dq_i with LinearSolver::adjoint_solve and enforces the
  // constraint rows homogeneously on each z_i
  std::pair<unsigned int, Real> adjoint_solve(const QoISet &qoi_indices = QoISet()) override;
  // ImplicitSystem::adjoint_qoi_parameter_sensitivity (implicit_system.C:517-623):
  // dq_i/dp_j = partial q_i/partial p_j - z_i . partial R/partial p_j, with
  // the partial derivatives by central differences.  When the attached
  // assembly is a HipAssembler and every parameter points at one of its
  // per-element coefficients, the residual term of all of them comes from
  // one launch of the device adjoint kernel (exact, as R is linear in each
  // coefficient) instead of two assemblies per parameter.
  void adjoint_qoi_parameter_sensitivity(const QoISet &qoi_indices, const ParameterVector &parameters,
                                         SensitivityData &sensitivities) override;
  // ImplicitSystem::assemble_residual_derivatives (implicit_system.C:474-515):
  // -(partial R/partial p_j) by central differences into get_sensitivity_rhs(j)
  virtual void assemble_residual_derivatives(const ParameterVector &parameters);
  // System::request_matrix (system.C): the named matrix, or null
  SparseMatrix<Number> *request_matrix(std::string_view mat_name) {
    auto it = _matrices.find(mat_name);
    return it == _matrices.end() ? nullptr : it->second.get();
  }
  LinearSolver<Number> *get_linear_solver() const { return linear_solver.get(); }
  // ImplicitSystem::linear_solver (public, include/systems/implicit_system.h:328)
  mutable std::unique_ptr<LinearSolver<Number>> linear_solver;
  // ImplicitSystem::create_static_condensation / get_static_condensation
  // (implicit_system.h:330-335, implicit_system.C:51-67): the system matrix
  // becomes a StaticCondensation, which the HipLinearSolver solves through
  // the reduced system
  void create_static_condensation() override;
  StaticCondensation &get_static_condensation();
  bool assemble_before_solve = true;

 protected:
  void init_matrices() override;
  // the device assembly whose per-element coefficients parameters may point
  // at (the attached HipAssembler), or null
  virtual HipAssembler *device_assembler() const;
};

class LinearImplicitSystem : public ImplicitSystem {
 public:
  using ImplicitSystem::ImplicitSystem;
  std::string system_type() const override { return "LinearImplicit"; }
  // LinearImplicitSystem::solve (src/systems/linear_implicit_system.C:122-161)
  void solve() override;
  unsigned n_linear_iterations() const { return _n_linear_iterations; }
  Real final_linear_residual() const { return _final_linear_residual; }

 private:
  unsigned _n_linear_iterations = 0;
  Real _final_linear_residual = 0;
};

class EquationSystems : public ParallelObject {
 public:
  explicit EquationSystems(MeshBase &mesh) : ParallelObject(mesh.comm()), _mesh(mesh) {}
  template <typename T>
  T &add_system(const std::string &name) {
    auto s = std::make_unique<T>(*this, name, (unsigned)_systems.size());
    T &ref = *s;
    _systems[name] = std::move(s);
    return ref;
  }
  template <typename T>
  T &get_system(const std::string &name) {
    auto it = _systems.find(name);
    if (it == _systems.end()) libmesh_error_msg("no system named " << name);
    return dynamic_cast<T &>(*it->second);
  }
  MeshBase &get_mesh() { return _mesh; }
  void init() {
    for (auto &kv : _systems) kv.second->init();
  }
  // include/systems/equation_systems.h:78-91
  enum ReadFlags {
    READ_HEADER = 1,
    READ_DATA = 2,
    READ_ADDITIONAL_DATA = 4,
    READ_LEGACY_FORMAT = 8,
    TRY_READ_IFEMS = 16,
    READ_BASIC_ONLY = 32
  };
  enum WriteFlags { WRITE_DATA = 1, WRITE_ADDITIONAL_DATA = 2, WRITE_PARALLEL_FILES = 4, WRITE_SERIAL_FILES = 8 };
  // EquationSystems::write (src/systems/equation_systems_io.C:335-560): the
  // serial file, ASCII or XDR by name (".xdr") or mode, nodes in id order (a
  // libMesh build without libHilbert writes the same order; with libHilbert
  // and MPI, partition_agnostic renumbers along a Hilbert curve first)
  void write(const std::string &name, unsigned int write_flags = WRITE_DATA) const;
  void write(const std::string &name, XdrMODE mode, unsigned int write_flags = WRITE_DATA) const;
  // EquationSystems::read (equation_systems_io.C:59-330) into systems that
  // are already added and initialised: the file's header must describe them
  // (names, variables, orders, families); READ_DATA fills their solutions
  void read(const std::string &name, unsigned int read_flags = READ_HEADER | READ_DATA);
  void read(const std::string &name, XdrMODE mode, unsigned int read_flags = READ_HEADER | READ_DATA);
  // EquationSystems parameters (equation_systems.C:59-60)
  Real linear_solver_tolerance = 1e-12;
  unsigned linear_solver_maximum_iterations = 5000;

 private:
  MeshBase &_mesh;
  std::map<std::string, std::unique_ptr<System>> _systems;
};

}  // namespace libMesh
