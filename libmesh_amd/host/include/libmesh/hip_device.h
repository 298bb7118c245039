// hip_device.h -- process-wide device context of the drop-in classes (the GPU
// counterpart of the solver-package selection in LibMeshInit,
// src/base/libmesh.C:1064-1122), and the mapping of C-ABI status codes onto
// libMesh's exceptions (libmesh_common.h:441, fe_map.C:467-505).
#pragma once

#include "../../../../include/lmhip.h"
#include "libmesh/libmesh_common.h"

namespace libMesh {

class HipDevice {
 public:
  // The context of `device` (created on first use, destroyed at exit).
  static lmh_context *context(int device = 0);
};

// Throws the libMesh exception matching an lmhip status code.
void hip_check(int rc, const char *what);

}  // namespace libMesh
