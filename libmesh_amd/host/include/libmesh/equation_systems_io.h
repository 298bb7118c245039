// equation_systems_io.h -- the layout of libMesh's EquationSystems files
// (serial, non-legacy: "libMesh-0.7.4" and later), held as data so that the
// same code reads and writes both encodings:
//   version string; number of systems; per system (in name order): name,
//   type, number of variables, per variable: name, subdomains, approximation
//   order, FE family; number of additional vectors, per vector: name,
//   projection flag, parallel type (the last two from 1.7.0 on); then, per
//   system: the solution and each listed additional vector as a length and
//   that many values in node-major order (nodes by id, each node's variables,
//   each variable's components; then elements the same way).
// src/systems/equation_systems_io.C:335-560 (write), 59-330 (read);
// src/systems/system_io.C:97-320 (read_header), 533-600 (read_serialized_data),
// 1117-1300 (write_header), 1518-1580 (write_serialized_data), 1968-2010
// (write_serialized_vector), 1626-1910 (blocked node-major layout).
#pragma once

#include <string>
#include <vector>

#include "libmesh/xdr_cxx.h"

namespace libMesh {

struct XdaVariable {
  std::string name;
  std::vector<uint16_t> subdomains;  // empty: every subdomain
  int order = 1, family = 0;
};

struct XdaVector {
  std::string name;
  int projection = 0, parallel_type = 0;
  std::vector<Number> values;  // node-major, as stored
};

struct XdaSystem {
  std::string name, type;
  // the system number appears only in comments ("# Name, System No. N"); the
  // ASCII reader recovers it from there, the XDR reader numbers in file order
  unsigned number = 0;
  std::vector<XdaVariable> vars;
  std::vector<Number> solution;  // node-major, as stored
  std::vector<XdaVector> vectors;
};

struct XdaFile {
  std::string version = "libMesh-1.7.0";
  std::vector<XdaSystem> systems;
};

// ".xdr" anywhere in the name selects XDR, otherwise ASCII (equation_systems_io.C:337-346, 81-90)
XdrMODE xdr_mode_for(const std::string &name, bool writing);
// header, then (write_data) the vectors
void write_equation_systems_file(Xdr &io, XdaFile &f, bool write_data);
// header, then (read_data) the vectors; throws on a legacy, parallel or
// infinite-element file and on truncated or malformed input
void read_equation_systems_file(Xdr &io, XdaFile &f, bool read_data);

}  // namespace libMesh
