// static_condensation_dof_map.C -- the condensed DoF set and the reduced
// numbering (static_condensation_dof_map.C:187-199 of libMesh).
#include "libmesh/static_condensation_dof_map.h"

#include "libmesh/dof_map.h"

namespace libMesh {

namespace {
// Elem::is_internal for the Lagrange elements here: HEX27's local node 26
// (the centre) touches no side, edge or vertex; no other type has such a node
bool is_internal_node(ElemType type, unsigned int local_node) { return type == HEX27 && local_node == 26; }
}  // namespace

StaticCondensationDofMap::StaticCondensationDofMap(MeshBase &mesh, System &system, const DofMap &dof_map)
    : _mesh(mesh), _system(system), _dof_map(dof_map) {}

void StaticCondensationDofMap::clear() {
  _condensed.clear();
  _full_to_reduced.clear();
  _n_reduced = 0;
  _init = false;
}

void StaticCondensationDofMap::reinit() {
  clear();
  const dof_id_type n_full = _dof_map.n_dofs();
  std::vector<char> cond(n_full, 0);
  std::vector<dof_id_type> di;
  for (dof_id_type e = 0; e < _mesh.n_elem(); ++e) {
    const Elem &elem = _mesh.elem_ref(e);
    for (unsigned int v = 0; v < _dof_map.n_variables(); ++v) {
      if (_uncondensed_vars.count(v)) continue;
      _dof_map.dof_indices(&elem, di, v);  // one DoF per node for Lagrange variables
      for (unsigned int a = 0; a < elem.n_nodes() && a < di.size(); ++a)
        if (is_internal_node(elem.type(), a) && !cond[di[a]]) {
          cond[di[a]] = 1;
          _condensed.push_back(di[a]);
        }
    }
  }
  _full_to_reduced.assign(n_full, (dof_id_type)-1);
  for (dof_id_type d = 0; d < n_full; ++d)
    if (!cond[d]) _full_to_reduced[d] = _n_reduced++;
  _init = true;
}

void StaticCondensationDofMap::dof_indices(const Elem *const elem, std::vector<dof_id_type> &di,
                                           const unsigned int vn, int) const {
  std::vector<dof_id_type> full;
  _dof_map.dof_indices(elem, full, vn);
  di.clear();
  for (dof_id_type d : full)
    if (_full_to_reduced.at(d) != (dof_id_type)-1) di.push_back(_full_to_reduced[d]);
}

}  // namespace libMesh
