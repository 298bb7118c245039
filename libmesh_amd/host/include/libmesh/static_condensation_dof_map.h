// static_condensation_dof_map.h -- StaticCondensationDofMap
// (include/numerics/static_condensation_dof_map.h): the DoFs static
// condensation eliminates and the reduced numbering of the others.  The
// eliminated DoFs are those of element-internal nodes (Elem::is_internal;
// static_condensation_dof_map.C:187-199), for every variable not excluded with
// dont_condense_vars; for the Lagrange elements here that is the centre node of
// a HEX27 (HEX8, TET4 and TET10 have none).  The reduced DoFs keep the full
// order.
#pragma once

#include <unordered_set>
#include <vector>

#include "libmesh/replicated_mesh.h"

namespace libMesh {

class System;
class DofMap;

class StaticCondensationDofMap {
 public:
  StaticCondensationDofMap(MeshBase &mesh, System &system, const DofMap &dof_map);
  virtual ~StaticCondensationDofMap() = default;

  // StaticCondensationDofMap::reinit: after the full DoFs are distributed
  void reinit();
  void dont_condense_vars(const std::unordered_set<unsigned int> &vars) {
    _uncondensed_vars.insert(vars.begin(), vars.end());
  }
  const std::unordered_set<unsigned int> &uncondensed_vars() const { return _uncondensed_vars; }
  bool initialized() const { return _init; }
  virtual void clear();
  const MeshBase &mesh() const { return _mesh; }

  // DofMapBase::n_dofs: the reduced system's size
  dof_id_type n_dofs() const { return _n_reduced; }
  // the eliminated DoFs, element by element (ascending element id)
  const std::vector<dof_id_type> &condensed_dofs() const { return _condensed; }
  // the element's uncondensed DoFs of variable vn in the reduced numbering
  // (the reduced system's dof_indices; p_level is accepted and unused, as
  // there is no p-refinement here)
  void dof_indices(const Elem *const elem, std::vector<dof_id_type> &di, const unsigned int vn,
                   int p_level = -12345) const;
  // full DoF -> reduced DoF, or (dof_id_type)-1 for an eliminated DoF
  dof_id_type reduced_dof(dof_id_type full_dof) const { return _full_to_reduced.at(full_dof); }

 private:
  MeshBase &_mesh;
  System &_system;
  const DofMap &_dof_map;
  std::unordered_set<unsigned int> _uncondensed_vars;
  std::vector<dof_id_type> _condensed;
  std::vector<dof_id_type> _full_to_reduced;
  dof_id_type _n_reduced = 0;
  bool _init = false;
};

}  // namespace libMesh
