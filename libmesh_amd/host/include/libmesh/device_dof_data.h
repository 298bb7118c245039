// device_dof_data.h -- the device copy of a mesh and its DoF numbering, read
// only through libMesh's own accessors (MeshBase::n_nodes / point / n_elem /
// elem_ref, Elem::type / n_nodes / node_id, DofMap::n_variables / n_dofs /
// dof_indices / constraint_rows_begin / constraint_rows_end /
// get_primal_constraint_values), so the drop-in classes work on any mesh and
// DofMap that offer them.
#pragma once

#include "libmesh/dof_map.h"
#include "libmesh/hip_device.h"

namespace libMesh {

struct DeviceDofData {
  lmh_mesh *mesh = nullptr;
  lmh_dofmap *dofmap = nullptr;
  DeviceDofData() = default;
  DeviceDofData(const DeviceDofData &) = delete;
  DeviceDofData &operator=(const DeviceDofData &) = delete;
  ~DeviceDofData() { release(); }
  void release();
};

// Uploads coordinates, connectivity and element DoF lists; constraint rows
// that are empty with a zero right-hand side (homogeneous Dirichlet) become
// the element kernels' mask, every other row goes to
// lmh_dofmap_set_constraints with `constraint_flags` (LMH_CONSTRAINT_*).
// One element type per mesh.
void build_device_dofs(lmh_context *ctx, const MeshBase &mesh, const DofMap &dof_map, int constraint_flags,
                       DeviceDofData &out);

}  // namespace libMesh
