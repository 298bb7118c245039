// periodic_x.h -- the example programs' System::Constraint object: constraint
// rows u(x = 1) = u(x = 0) on the unit cube, the way libMesh applications add
// rows from a System::Constraint (DofMap::add_constraint_row in constrain(),
// called by create_dof_constraints).
#pragma once

#include <cmath>
#include <map>
#include <utility>
#include <vector>

#include "libmesh/system.h"

namespace libMesh {

// Periodic-style rows on the unit cube: every unconstrained DoF on x = 1 equals
// the DoF on x = 0 at the same (y, z) (boundary nodes sit exactly on the
// lattice).  A first row gives the first interior DoF of the x = 0 face the
// value 0.3, so the rows of its partner chain through it.
class PeriodicX : public System::Constraint {
 public:
  PeriodicX(System &sys, const MeshBase &mesh, unsigned cells) : _sys(sys), _mesh(mesh), _cells(cells) {}
  void constrain() override {
    DofMap &dm = _sys.get_dof_map();
    std::vector<dof_id_type> node_dof(_mesh.n_nodes()), di;
    for (dof_id_type e = 0; e < _mesh.n_elem(); ++e) {
      const Elem &elem = _mesh.elem_ref(e);
      dm.dof_indices(&elem, di);
      for (unsigned a = 0; a < elem.n_nodes(); ++a) node_dof[elem.node_id(a)] = di[a];
    }
    auto key = [&](const Point &p) {
      return std::make_pair((long)std::lround(p.y * 2 * _cells), (long)std::lround(p.z * 2 * _cells));
    };
    std::map<std::pair<long, long>, dof_id_type> near;
    bool valued = false;
    for (dof_id_type k = 0; k < _mesh.n_nodes(); ++k) {
      const Point p = _mesh.point(k);
      if (std::abs(p.x) > 1e-12) continue;
      near[key(p)] = node_dof[k];
      const bool interior = p.y > 1e-9 && p.y < 1 - 1e-9 && p.z > 1e-9 && p.z < 1 - 1e-9;
      if (!valued && interior && !dm.is_constrained_dof(node_dof[k])) {
        dm.add_constraint_row(node_dof[k], DofConstraintRow(), 0.3, true);
        valued = true;
      }
    }
    for (dof_id_type k = 0; k < _mesh.n_nodes(); ++k) {
      const Point p = _mesh.point(k);
      if (std::abs(p.x - 1) > 1e-12 || dm.is_constrained_dof(node_dof[k])) continue;
      DofConstraintRow row;
      row[near.at(key(p))] = 1.0;
      dm.add_constraint_row(node_dof[k], row, true);
    }
  }

 private:
  System &_sys;
  const MeshBase &_mesh;
  unsigned _cells;
};

}  // namespace libMesh
