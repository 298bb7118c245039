// parallel_object.h -- Parallel::Communicator and ParallelObject
// (include/parallel/parallel_object.h:47-120, the TIMPI Communicator behind
// Parallel::Communicator).  The drop-in classes run one process per GPU with
// one rank per communicator, so the communicator is the serial one: rank 0 of
// 1, collectives are identities.  Every class that libMesh derives from
// ParallelObject (meshes, matrices, vectors, solvers) takes it in its
// constructor, as there.
#pragma once

#include <cstdint>
#include <vector>

namespace libMesh {

// include/base/id_types.h:108 (default build)
typedef uint32_t processor_id_type;

namespace Parallel {

class Communicator {
 public:
  Communicator() = default;
  Communicator(const Communicator &) = delete;
  Communicator &operator=(const Communicator &) = delete;
  processor_id_type rank() const { return 0; }
  processor_id_type size() const { return 1; }
  void barrier() const {}
  template <typename T>
  void max(T &) const {}
  template <typename T>
  void min(T &) const {}
  template <typename T>
  void sum(T &) const {}
  template <typename T>
  void allgather(std::vector<T> &) const {}
  template <typename T>
  void broadcast(T &, processor_id_type = 0) const {}
};

}  // namespace Parallel

class ParallelObject {
 public:
  ParallelObject(const Parallel::Communicator &comm_in) : _communicator(comm_in) {}
  ParallelObject(const ParallelObject &other) : _communicator(other._communicator) {}
  ParallelObject &operator=(const ParallelObject &) { return *this; }
  virtual ~ParallelObject() = default;
  const Parallel::Communicator &comm() const { return _communicator; }
  processor_id_type n_processors() const { return _communicator.size(); }
  processor_id_type processor_id() const { return _communicator.rank(); }

 protected:
  const Parallel::Communicator &_communicator;
};

}  // namespace libMesh
