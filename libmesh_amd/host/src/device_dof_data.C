// device_dof_data.C -- mesh and DoF numbering to the device through libMesh's
// accessors.
#include "libmesh/device_dof_data.h"

namespace libMesh {

void DeviceDofData::release() {
  if (dofmap) lmh_dofmap_destroy(dofmap);
  if (mesh) lmh_mesh_destroy(mesh);
  dofmap = nullptr;
  mesh = nullptr;
}

void build_device_dofs(lmh_context *ctx, const MeshBase &mesh, const DofMap &dof_map, int constraint_flags,
                       DeviceDofData &out) {
  out.release();
  const dof_id_type ne = mesh.n_elem(), nn = mesh.n_nodes(), nd = dof_map.n_dofs();
  if (ne == 0) libmesh_error_msg("build_device_dofs: the mesh has no elements");
  const Elem &e0 = mesh.elem_ref(0);
  const ElemType type = e0.type();
  const unsigned npe = e0.n_nodes(), nv = dof_map.n_variables(), dpe = nv * npe;
  std::vector<double> xyz((size_t)3 * nn);
  for (dof_id_type n = 0; n < nn; ++n) {
    const Point &p = mesh.point(n);
    xyz[3 * n] = p(0);
    xyz[3 * n + 1] = p(1);
    xyz[3 * n + 2] = p(2);
  }
  std::vector<int32_t> conn((size_t)ne * npe), edofs((size_t)ne * dpe);
  std::vector<dof_id_type> di;
  for (dof_id_type e = 0; e < ne; ++e) {
    const Elem &el = mesh.elem_ref(e);
    if (el.type() != type || el.n_nodes() != npe)
      libmesh_error_msg("build_device_dofs: element " << e << " differs in type from element 0");
    for (unsigned a = 0; a < npe; ++a) conn[(size_t)e * npe + a] = (int32_t)el.node_id(a);
    dof_map.dof_indices(&el, di);
    if (di.size() != dpe) libmesh_error_msg("build_device_dofs: element " << e << " has " << di.size() << " DoFs, expected " << dpe);
    for (unsigned k = 0; k < dpe; ++k) edofs[(size_t)e * dpe + k] = (int32_t)di[k];
  }
  // homogeneous Dirichlet rows as the mask, the rest as general rows in DoF order
  std::vector<uint8_t> mask(nd, 0);
  std::vector<int32_t> dof, ptr{0}, col;
  std::vector<double> coef, val;
  // libMesh offers the right-hand sides only through the non-const accessor; they are read here
  const DofConstraintValueMap &values = const_cast<DofMap &>(dof_map).get_primal_constraint_values();
  for (auto it = dof_map.constraint_rows_begin(); it != dof_map.constraint_rows_end(); ++it) {
    const auto v = values.find(it->first);
    const double rhs = v == values.end() ? 0.0 : (double)v->second;
    if (it->second.empty() && rhs == 0.0) {
      mask[it->first] = 1;
      continue;
    }
    dof.push_back((int32_t)it->first);
    for (const auto &c : it->second) {
      col.push_back((int32_t)c.first);
      coef.push_back(c.second);
    }
    ptr.push_back((int32_t)col.size());
    val.push_back(rhs);
  }
  hip_check(lmh_mesh_create(ctx, (int)type, nn, xyz.data(), ne, conn.data(), &out.mesh), "device mesh");
  hip_check(lmh_dofmap_create(ctx, out.mesh, (int)nv, nd, edofs.data(), mask.data(), &out.dofmap), "device DoF map");
  if (!dof.empty() || constraint_flags)
    hip_check(lmh_dofmap_set_constraints(out.dofmap, (int64_t)dof.size(), dof.data(), ptr.data(), col.data(),
                                         coef.data(), val.data(), constraint_flags),
              "device constraint rows");
}

}  // namespace libMesh
