// qoi_set.h -- QoISet (include/systems/qoi_set.h:45-190): which quantities of
// interest a call computes and their weights.  An empty index list means
// "all QoIs"; a missing weight means 1.
#pragma once

#include <vector>

#include "libmesh/libmesh_common.h"

namespace libMesh {

class System;

class QoISet {
 public:
  QoISet() = default;
  // every QoI of the system, weight 1
  QoISet(const System &) {}
  QoISet(std::vector<bool> indices) : _indices(std::move(indices)) {}
  QoISet(const std::vector<unsigned int> &indices) { add_indices(indices); }
  void clear() {
    _indices.clear();
    _weights.clear();
  }
  void add_indices(const std::vector<unsigned int> &indices) {
    unsigned int mx = 0;
    for (auto i : indices) mx = std::max(mx, i + 1);
    // listing indices selects those only
    if (_indices.empty()) _indices.assign(mx, false);
    for (auto i : indices) add_index(i);
  }
  void add_index(std::size_t i) {
    if (i >= _indices.size()) _indices.resize(i + 1, true);
    _indices[i] = true;
  }
  void remove_index(std::size_t i) {
    if (i >= _indices.size()) _indices.resize(i + 1, true);
    _indices[i] = false;
  }
  void set_weight(std::size_t i, Real w) {
    if (_weights.size() <= i) _weights.resize(i + 1, 1.0);
    _weights[i] = w;
  }
  Real weight(std::size_t i) const { return _weights.size() <= i ? 1.0 : _weights[i]; }
  bool has_index(std::size_t i) const { return _indices.size() <= i || _indices[i]; }

 private:
  std::vector<bool> _indices;
  std::vector<Real> _weights;
};

}  // namespace libMesh
