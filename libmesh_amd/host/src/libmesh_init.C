// libmesh_init.C -- LibMeshInit: the serial world communicator (the GPU
// context is created on first use by the drop-in classes).
#include "libmesh/libmesh.h"

namespace libMesh {

LibMeshInit::LibMeshInit(int, const char *const *, int, int) {}

}  // namespace libMesh
