// sparsity_pattern.h -- SparsityPattern::Row and Graph
// (include/base/sparsity_pattern.h:52-53): the sorted column lists of each
// row that DofMap hands to matrices whose need_full_sparsity_pattern() is
// true (DofMap::update_sparsity_pattern, src/base/dof_map.C:269-290).
#pragma once

#include <vector>

#include "libmesh/libmesh_common.h"

namespace libMesh {
namespace SparsityPattern {

typedef std::vector<dof_id_type> Row;
class Graph : public std::vector<Row> {};

}  // namespace SparsityPattern
}  // namespace libMesh
