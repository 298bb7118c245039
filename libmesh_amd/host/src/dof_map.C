// dof_map.C -- DofMap numbering and Dirichlet constraint set for one group of
// Lagrange variables (dof_map.C:949-1530, dof_map_constraints.C:592-740).
#include "libmesh/dof_map.h"
#include "libmesh/device_dof_data.h"

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

void DofMap::distribute_dofs(MeshBase &mesh) {
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
  // boundary id is in the boundary's set gets an empty row (u = 0)
  _dof_constraints.clear();
  _primal_constraint_values.clear();
  const bool hex = (t == HEX8 || t == HEX27);
  const int side_n = t == HEX8 ? 4 : t == HEX27 ? 9 : t == TET4 ? 3 : 6;
  for (const auto &d : _dirichlet)
    for (dof_id_type e = 0; e < ne; ++e)
      for (unsigned s = 0; s < mesh.sides_per_elem(); ++s) {
        const int b = mesh.side_boundary_id(e, s);
        if (b < 0 || !d.b.count((boundary_id_type)b)) continue;
        for (int k = 0; k < side_n; ++k) {
          const int a = hex ? kHexSide[s][k] : kTetSide[s][k];
          for (unsigned v : d.variables) _dof_constraints[(dof_id_type)_edofs[(size_t)e * _dpe + v * npe + a]];
        }
      }
}

void DofMap::add_constraint_row(const dof_id_type dof_number, const DofConstraintRow &constraint_row,
                                const Number constraint_rhs, const bool forbid_constraint_overwrite) {
  if (dof_number >= _n_dofs) libmesh_error_msg("add_constraint_row: DoF " << dof_number << " out of range");
  if (forbid_constraint_overwrite && is_constrained_dof(dof_number))
    libmesh_error_msg("ERROR: DOF " << dof_number << " was already constrained!");
  if (constraint_row.count(dof_number))
    libmesh_error_msg("Error: constraint_row for dof " << dof_number << " should not contain an entry for dof "
                                                       << dof_number);
  for (const auto &it : constraint_row)
    if (it.first >= _n_dofs) libmesh_error_msg("add_constraint_row: DoF " << it.first << " out of range");
  _dof_constraints[dof_number] = constraint_row;
  if (constraint_rhs != Number(0))
    _primal_constraint_values[dof_number] = constraint_rhs;
  else
    _primal_constraint_values.erase(dof_number);
}

void DofMap::dof_indices(const Elem *const elem, std::vector<dof_id_type> &di) const {
  const dof_id_type e = elem->id();
  if (!_dpe || (size_t)(e + 1) * _dpe > _edofs.size()) libmesh_error_msg("dof_indices: element " << e << " unknown");
  di.assign(_edofs.begin() + (size_t)e * _dpe, _edofs.begin() + (size_t)(e + 1) * _dpe);
}

void DofMap::dof_indices(const Elem *const elem, std::vector<dof_id_type> &di, const unsigned int vn) const {
  const dof_id_type e = elem->id();
  const unsigned nv = n_variables(), npe = nv ? _dpe / nv : 0;
  if (vn >= nv) libmesh_error_msg("dof_indices: variable " << vn << " of " << nv);
  if (!_dpe || (size_t)(e + 1) * _dpe > _edofs.size()) libmesh_error_msg("dof_indices: element " << e << " unknown");
  const size_t o = (size_t)e * _dpe + (size_t)vn * npe;
  di.assign(_edofs.begin() + o, _edofs.begin() + o + npe);
}

std::vector<uint8_t> DofMap::constrained() const {
  std::vector<uint8_t> c(_n_dofs, 0);
  for (const auto &r : _dof_constraints) c[r.first] = 1;
  return c;
}

bool DofMap::has_general_constraints() const {
  if (!_primal_constraint_values.empty()) return true;
  for (const auto &r : _dof_constraints)
    if (!r.second.empty()) return true;
  return false;
}

// Processing on the device (lmh_dofmap_set_constraints substitutes the rows)
// and the processed rows read back, so that the constraint accessors return
// what libMesh's DofMap holds after process_constraints.
void DofMap::process_constraints(MeshBase &mesh) {
  if (!has_general_constraints()) return;
  DeviceDofData dev;
  build_device_dofs(HipDevice::context(0), mesh, *this, _error_on_loop ? LMH_CONSTRAINT_ERROR_ON_LOOP : 0, dev);
  int64_t nr = 0, ne = 0;
  hip_check(lmh_dofmap_constraint_counts(dev.dofmap, &nr, &ne, nullptr), "DofMap::process_constraints");
  std::vector<int32_t> dof(nr), ptr(nr + 1), col(ne);
  std::vector<double> coef(ne), rhs(nr);
  hip_check(lmh_dofmap_get_constraints(dev.dofmap, dof.data(), ptr.data(), col.data(), coef.data(), rhs.data()),
            "DofMap::process_constraints");
  _dof_constraints.clear();
  _primal_constraint_values.clear();
  for (int64_t k = 0; k < nr; ++k) {
    DofConstraintRow &row = _dof_constraints[(dof_id_type)dof[k]];
    for (int32_t t = ptr[k]; t < ptr[k + 1]; ++t) row[(dof_id_type)col[t]] = coef[t];
    if (rhs[k] != 0.0) _primal_constraint_values[(dof_id_type)dof[k]] = rhs[k];
  }
}

}  // namespace libMesh
