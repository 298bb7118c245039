// dof_map.C -- DofMap numbering and Dirichlet constraint set for one group of
// Lagrange variables (dof_map.C:949-1530, dof_map_constraints.C:592-740).
#include "libmesh/dof_map.h"

namespace libMesh {
namespace {
// Local nodes on each side (side_nodes_map of Hex27/Hex8 in cell_hex27.C:37-45
// and Tet10/Tet4 in cell_tet10.C:37-43; linear elements use the leading
// vertices of each list).
const int kHexSide[6][9] = {{0, 3, 2, 1, 11, 10, 9, 8, 20},  {0, 1, 5, 4, 8, 13, 16, 12, 21},
                            {1, 2, 6, 5, 9, 14, 17, 13, 22}, {2, 3, 7, 6, 10, 15, 18, 14, 23},
                            {3, 0, 4, 7, 11, 12, 19, 15, 24}, {4, 5, 6, 7, 16, 17, 18, 19, 25}};
const int kTetSide[4][6] = {{0, 2, 1, 6, 5, 4}, {0, 1, 3, 4, 8, 7}, {1, 2, 3, 5, 9, 8}, {2, 0, 3, 6, 7, 9}};
}  // namespace

unsigned DofMap::add_variable(const std::string &name, const FEType &type) {
  if (!_types.empty() && (_types[0].order != type.order || _types[0].family != type.family))
    libmesh_error_msg("all variables must share one Lagrange FEType on this path");
  _names.push_back(name);
  _types.push_back(type);
  return (unsigned)_names.size() - 1;
}

void DofMap::distribute_dofs(const MeshBase &mesh) {
  if (_types.empty()) libmesh_error_msg("DofMap::distribute_dofs: no variables");
  const ElemType t = mesh.elem_type();
  const Order elem_order = (t == HEX8 || t == TET4) ? FIRST : SECOND;
  if (_types[0].order != elem_order)
    libmesh_error_msg("FE order must equal the element's own Lagrange order on this path");
  _mesh = &mesh;
  const unsigned nv = n_variables(), npe = mesh.nodes_per_elem();
  const dof_id_type ne = mesh.n_elem(), nn = mesh.n_nodes();
  _dpe = nv * npe;
  // first-touch variable-group bases in element order, local node order
  std::vector<int64_t> base(nn, -1);
  int64_t next = 0;
  for (dof_id_type e = 0; e < ne; ++e)
    for (unsigned a = 0; a < npe; ++a) {
      int64_t &b = base[mesh.node_id(e, a)];
      if (b < 0) {
        b = next;
        next += nv;
      }
    }
  _n_dofs = (dof_id_type)next;
  _edofs.resize((size_t)ne * _dpe);
  for (dof_id_type e = 0; e < ne; ++e)
    for (unsigned v = 0; v < nv; ++v)
      for (unsigned a = 0; a < npe; ++a) _edofs[(size_t)e * _dpe + v * npe + a] = (int32_t)(base[mesh.node_id(e, a)] + v);
  // Dirichlet: every listed variable's DoF at every node of a side whose
  // boundary id is in the boundary's set
  _cons.assign(_n_dofs, 0);
  const bool hex = (t == HEX8 || t == HEX27);
  const int side_n = t == HEX8 ? 4 : t == HEX27 ? 9 : t == TET4 ? 3 : 6;
  for (const auto &d : _dirichlet)
    for (dof_id_type e = 0; e < ne; ++e)
      for (unsigned s = 0; s < mesh.sides_per_elem(); ++s) {
        const int b = mesh.side_boundary_id(e, s);
        if (b < 0 || !d.b.count((boundary_id_type)b)) continue;
        for (int k = 0; k < side_n; ++k) {
          const int a = hex ? kHexSide[s][k] : kTetSide[s][k];
          for (unsigned v : d.variables) _cons[_edofs[(size_t)e * _dpe + v * npe + a]] = 1;
        }
      }
  _dirichlet_cons = _cons;
  _rows.clear();
}

void DofMap::add_constraint_row(dof_id_type dof_number, const DofConstraintRow &constraint_row, Number constraint_rhs,
                                bool forbid_constraint_overwrite) {
  if (dof_number >= _n_dofs) libmesh_error_msg("add_constraint_row: DoF " << dof_number << " out of range");
  if (forbid_constraint_overwrite && is_constrained_dof(dof_number))
    libmesh_error_msg("ERROR: DOF " << dof_number << " was already constrained!");
  if (constraint_row.count(dof_number))
    libmesh_error_msg("Error: constraint_row for dof " << dof_number << " should not contain an entry for dof "
                                                       << dof_number);
  for (const auto &it : constraint_row)
    if (it.first >= _n_dofs) libmesh_error_msg("add_constraint_row: DoF " << it.first << " out of range");
  _rows.emplace_back(dof_number, constraint_row, constraint_rhs);
  _cons[dof_number] = 1;
}

void DofMap::dof_indices(dof_id_type e, std::vector<dof_id_type> &di) const {
  di.assign(_edofs.begin() + (size_t)e * _dpe, _edofs.begin() + (size_t)(e + 1) * _dpe);
}

dof_id_type DofMap::n_constrained_dofs() const {
  dof_id_type c = 0;
  for (uint8_t x : _cons) c += x;
  return c;
}

}  // namespace libMesh
