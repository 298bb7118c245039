// libmesh.h -- LibMeshInit (include/base/libmesh.h:90-130): owns the world
// communicator a program's meshes, matrices, vectors and solvers are built on.
// Here the serial communicator; the GPU context is created on first use by
// the drop-in classes (HipDevice).
#pragma once

#include "libmesh/parallel_object.h"

namespace libMesh {

class LibMeshInit {
 public:
  LibMeshInit(int argc, const char *const *argv, int COMM_WORLD_IN = 0, int n_threads = -1);
  virtual ~LibMeshInit() = default;
  const Parallel::Communicator &comm() const { return _comm; }
  Parallel::Communicator &comm() { return _comm; }

 private:
  Parallel::Communicator _comm;
};

}  // namespace libMesh
