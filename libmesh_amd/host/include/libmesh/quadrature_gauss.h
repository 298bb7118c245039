// quadrature_gauss.h -- QGauss (include/quadrature/quadrature_gauss.h:39): the
// rule is identified by dimension and order; for hexes the points per
// direction are (order + 2) / 2 (THIRD -> 2, FIFTH -> 3); for tets the order
// selects the rule of quadrature_gauss_3D.C:55-260.
#pragma once

#include "libmesh/libmesh_common.h"

namespace libMesh {

class QGauss {
 public:
  QGauss(unsigned dim, Order order) : _dim(dim), _order(order) {
    if (dim != 3) libmesh_error_msg("QGauss: only 3D rules are on this path");
  }
  Order get_order() const { return _order; }
  unsigned get_dim() const { return _dim; }
  // kernel argument for an element type: hex points/direction, tet order
  int kernel_arg(ElemType t) const {
    if (t == HEX8 || t == HEX27) return ((int)_order + 2) / 2;
    return (int)_order;
  }

 private:
  unsigned _dim;
  Order _order;
};

}  // namespace libMesh
