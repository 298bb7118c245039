// equation_systems_io.C -- EquationSystems::write / read in libMesh's serial
// file layout (src/systems/equation_systems_io.C, src/systems/system_io.C;
// layout summarised in equation_systems_io.h).
#include "libmesh/equation_systems_io.h"

#include <cstdio>

#include "libmesh/system.h"

namespace libMesh {

namespace {

// LIBMESH_VERSION_ID (include/base/libmesh_version.h)
constexpr int version_id(int major, int minor, int micro) { return major * 10000 + minor * 100 + micro; }

std::string quoted(const std::string &s) { return "\"" + s + "\""; }

void write_vector(Xdr &io, std::vector<Number> &v, const std::string &comment) {
  // System::write_serialized_vector: length, node-major values, then the
  // vector's comment line (system_io.C:1968-2010, 1518-1575)
  unsigned int n = (unsigned int)v.size();
  io.data(n, "# vector length");
  io.data_stream(v.data(), n);
  io.comment(comment);
}

void read_vector(Xdr &io, std::vector<Number> &v) {
  unsigned int n = 0;
  io.data(n, "# vector length");
  if (n > (1u << 31)) libmesh_error_msg("EquationSystems file: implausible vector length " << n);
  v.resize(n);
  io.data_stream(v.data(), n);
  io.comment("");
}

}  // namespace

XdrMODE xdr_mode_for(const std::string &name, bool writing) {
  const bool xdr = name.find(".xdr") != std::string::npos;
  return writing ? (xdr ? ENCODE : WRITE) : (xdr ? DECODE : READ);
}

void write_equation_systems_file(Xdr &io, XdaFile &f, bool write_data) {
  if (!io.writing()) libmesh_error_msg("write_equation_systems_file: stream is not open for writing");
  // 1.) - 2.)
  io.data(f.version, "# File Format Identifier");
  unsigned int n_sys = (unsigned int)f.systems.size();
  io.data(n_sys, "# No. of Equation Systems");
  for (auto &s : f.systems) {
    const std::string num = std::to_string(s.number), sn = quoted(s.name);
    // 3.) - 4.)
    io.data(s.name, "# Name, System No. " + num);
    io.data(s.type, "# Type, System No. " + num);
    // 5.) - 9.): System::write_header (system_io.C:1117-1300)
    unsigned int nv = (unsigned int)s.vars.size();
    io.data(nv, "# No. of Variables in System " + sn);
    for (unsigned int v = 0; v < nv; ++v) {
      XdaVariable &var = s.vars[v];
      const std::string vn = quoted(var.name);
      io.data(var.name, "#   Name, Variable No. " + std::to_string(v) + ", System " + sn);
      io.data(var.subdomains, "#     Subdomains, Variable " + vn + ", System " + sn);
      io.data(var.order, "#     Approximation Order, Variable " + vn + ", System " + sn);
      io.data(var.family, "#     FE Family, Variable " + vn + ", System " + sn);
    }
    unsigned int nvec = (unsigned int)s.vectors.size();
    io.data(nvec, "# No. of Additional Vectors, System " + sn);
    for (unsigned int k = 0; k < nvec; ++k) {
      const std::string dth = std::to_string(k) + "th vector";
      io.data(s.vectors[k].name, "# Name of " + dth);
      io.data(s.vectors[k].projection, "# Whether to do projections for " + dth);
      io.data(s.vectors[k].parallel_type, "# Parallel type of " + dth);
    }
  }
  if (!write_data) return;
  // 10.) - 11.): System::write_serialized_data, systems in the same order
  for (auto &s : f.systems) {
    write_vector(io, s.solution, "# System " + quoted(s.name) + " Solution Vector");
    for (auto &vec : s.vectors)
      write_vector(io, vec.values, "# System " + quoted(s.name) + " Additional Vector " + quoted(vec.name));
  }
}

void read_equation_systems_file(Xdr &io, XdaFile &f, bool read_data) {
  if (!io.reading()) libmesh_error_msg("read_equation_systems_file: stream is not open for reading");
  io.data(f.version, "# File Format Identifier");
  int ma = 0, mi = 0, mc = 0;
  if (std::sscanf(f.version.c_str(), "libMesh-%d.%d.%d", &ma, &mi, &mc) != 3)
    libmesh_error_msg("EquationSystems file: unrecognised version header '" << f.version
                                                                              << "' (legacy files are not read)");
  const int ver = version_id(ma, mi, mc);
  if (ver < version_id(0, 7, 4))
    libmesh_error_msg("EquationSystems file: version " << f.version << " stores variables var-major; not supported");
  if (f.version.find(" parallel") != std::string::npos)
    libmesh_error_msg("EquationSystems file: parallel-data files are not supported");
  if (f.version.find(" with infinite elements") != std::string::npos)
    libmesh_error_msg("EquationSystems file: infinite-element files are not supported");
  unsigned int n_sys = 0;
  io.data(n_sys, "# No. of Equation Systems");
  if (n_sys > 4096) libmesh_error_msg("EquationSystems file: implausible system count " << n_sys);
  f.systems.assign(n_sys, XdaSystem());
  for (unsigned int i = 0; i < n_sys; ++i) {
    XdaSystem &s = f.systems[i];
    io.data(s.name, "");
    s.number = i;
    const std::string &c = io.last_comment();
    const size_t at = c.find("System No. ");
    if (at != std::string::npos) s.number = (unsigned)std::strtoul(c.c_str() + at + 11, nullptr, 10);
    io.data(s.type, "");
    unsigned int nv = 0;
    io.data(nv, "");
    if (nv > 4096) libmesh_error_msg("EquationSystems file: implausible variable count " << nv);
    s.vars.assign(nv, XdaVariable());
    for (auto &var : s.vars) {
      io.data(var.name, "");
      if (ver >= version_id(0, 7, 2)) io.data(var.subdomains, "");
      io.data(var.order, "");
      io.data(var.family, "");
    }
    unsigned int nvec = 0;
    io.data(nvec, "");
    if (nvec > 4096) libmesh_error_msg("EquationSystems file: implausible vector count " << nvec);
    s.vectors.assign(nvec, XdaVector());
    for (auto &vec : s.vectors) {
      io.data(vec.name, "");
      if (ver >= version_id(1, 7, 0)) {
        io.data(vec.projection, "");
        io.data(vec.parallel_type, "");
      }
    }
  }
  if (!read_data) return;
  for (auto &s : f.systems) {
    read_vector(io, s.solution);
    for (auto &vec : s.vectors) read_vector(io, vec.values);
  }
}

std::vector<dof_id_type> System::node_major_dofs() const {
  const MeshBase *mesh = _dof_map.mesh();
  if (!mesh) libmesh_error_msg("System::node_major_dofs: DoFs are not distributed (call init())");
  const unsigned nv = _dof_map.n_variables(), npe = mesh->nodes_per_elem(), dpe = _dof_map.dofs_per_elem();
  const dof_id_type invalid = (dof_id_type)-1;
  std::vector<dof_id_type> slot((size_t)mesh->n_nodes() * nv, invalid);
  const std::vector<int32_t> &ed = _dof_map.elem_dofs();
  for (dof_id_type e = 0; e < mesh->n_elem(); ++e)
    for (unsigned v = 0; v < nv; ++v)
      for (unsigned a = 0; a < npe; ++a)
        slot[(size_t)mesh->node_id(e, a) * nv + v] = (dof_id_type)ed[(size_t)e * dpe + v * npe + a];
  // nodes of no element hold no DoFs and write nothing (n_comp = 0)
  std::vector<dof_id_type> out;
  out.reserve(slot.size());
  for (dof_id_type d : slot)
    if (d != invalid) out.push_back(d);
  if (out.size() != _dof_map.n_dofs()) libmesh_error_msg("System::node_major_dofs: DoFs off the nodes");
  return out;
}

void EquationSystems::write(const std::string &name, unsigned int write_flags) const {
  write(name, xdr_mode_for(name, true), write_flags);
}

void EquationSystems::write(const std::string &name, XdrMODE mode, unsigned int write_flags) const {
  if (write_flags & WRITE_PARALLEL_FILES)
    libmesh_error_msg("EquationSystems::write: parallel data files are not supported (one process)");
  if (mode != WRITE && mode != ENCODE) libmesh_error_msg("EquationSystems::write: mode must be WRITE or ENCODE");
  const bool write_data = write_flags & WRITE_DATA;
  XdaFile f;
  for (const auto &kv : _systems) {
    const System &sys = *kv.second;
    XdaSystem s;
    s.name = kv.first;
    s.type = sys.system_type();
    s.number = sys.number();
    const DofMap &dm = sys.get_dof_map();
    for (unsigned v = 0; v < dm.n_variables(); ++v) {
      XdaVariable var;
      var.name = dm.variable_name(v);
      var.order = (int)dm.variable_type(v).order;
      var.family = (int)dm.variable_type(v).family;
      s.vars.push_back(var);
    }
    // additional vectors, name order, listed only with WRITE_ADDITIONAL_DATA (system_io.C:1277-1300)
    if (write_flags & WRITE_ADDITIONAL_DATA)
      for (const auto &v : sys.vectors()) {
        XdaVector xv;
        xv.name = v.first;
        xv.projection = v.second.projections ? 1 : 0;
        xv.parallel_type = (int)v.second.type;
        s.vectors.push_back(std::move(xv));
      }
    if (write_data) {
      if (!sys.solution) libmesh_error_msg("EquationSystems::write: system " << s.name << " is not initialised");
      const std::vector<dof_id_type> order = sys.node_major_dofs();
      auto node_major = [&](const NumericVector<Number> &vec, std::vector<Number> &out) {
        std::vector<Number> x;
        vec.localize(x);
        out.resize(order.size());
        for (size_t i = 0; i < order.size(); ++i) out[i] = x.at(order[i]);
      };
      node_major(*sys.solution, s.solution);
      for (XdaVector &xv : s.vectors) node_major(sys.get_vector(xv.name), xv.values);
    }
    f.systems.push_back(std::move(s));
  }
  Xdr io(name, mode);
  write_equation_systems_file(io, f, write_data);
  io.close();
}

void EquationSystems::read(const std::string &name, unsigned int read_flags) {
  read(name, xdr_mode_for(name, false), read_flags);
}

void EquationSystems::read(const std::string &name, XdrMODE mode, unsigned int read_flags) {
  if (read_flags & READ_LEGACY_FORMAT) libmesh_error_msg("EquationSystems::read: legacy files are not supported");
  if (mode != READ && mode != DECODE) libmesh_error_msg("EquationSystems::read: mode must be READ or DECODE");
  const bool read_data = read_flags & READ_DATA;
  XdaFile f;
  {
    Xdr io(name, mode);
    read_equation_systems_file(io, f, read_data);
    io.close();
  }
  // systems are declared (variables added, init() called) before reading:
  // the header must describe them.  Additional vectors follow
  // System::read_serialized_data (system_io.C:533-600): with
  // READ_ADDITIONAL_DATA a system that holds vectors must hold as many as the
  // file and receives them in order (both in name order); a system without
  // vectors, or a read without the flag, drops them.
  for (const XdaSystem &s : f.systems) {
    auto it = _systems.find(s.name);
    if (it == _systems.end())
      libmesh_error_msg("EquationSystems::read: file system '" << s.name << "' is not in this EquationSystems");
    System &sys = *it->second;
    const DofMap &dm = sys.get_dof_map();
    if (s.vars.size() != dm.n_variables())
      libmesh_error_msg("EquationSystems::read: system '" << s.name << "' has " << s.vars.size()
                                                          << " variables in the file, " << dm.n_variables()
                                                          << " here");
    for (unsigned v = 0; v < dm.n_variables(); ++v)
      if (s.vars[v].name != dm.variable_name(v) || s.vars[v].order != (int)dm.variable_type(v).order ||
          s.vars[v].family != (int)dm.variable_type(v).family)
        libmesh_error_msg("EquationSystems::read: variable " << v << " of system '" << s.name
                                                             << "' differs from the file's '" << s.vars[v].name
                                                             << "'");
    if (!read_data) continue;
    if (!sys.solution) libmesh_error_msg("EquationSystems::read: system " << s.name << " is not initialised");
    const std::vector<dof_id_type> order = sys.node_major_dofs();
    auto fill = [&](NumericVector<Number> &vec, const std::vector<Number> &vals, const std::string &what) {
      if (vals.size() != order.size())
        libmesh_error_msg("EquationSystems::read: system '" << s.name << "' " << what << " has " << vals.size()
                                                            << " values in the file, " << order.size()
                                                            << " DoFs here");
      for (size_t i = 0; i < order.size(); ++i) vec.set(order[i], vals[i]);
      vec.close();
    };
    fill(*sys.solution, s.solution, "solution");
    if ((read_flags & READ_ADDITIONAL_DATA) && sys.n_vectors()) {
      if (sys.n_vectors() != s.vectors.size())
        libmesh_error_msg("EquationSystems::read: additional vectors in file do not match system '" << s.name
                                                                                                      << "'");
      auto it = sys.vectors().begin();
      for (const XdaVector &xv : s.vectors) {
        fill(*it->second.vec, xv.values, "vector " + xv.name);
        ++it;
      }
    }
  }
}

}  // namespace libMesh
