// xdr_cxx.C -- ASCII and XDR streams of EquationSystems files
// (src/utils/xdr_cxx.C of libMesh; encodings described in xdr_cxx.h).
#include "libmesh/xdr_cxx.h"

#include <cstring>
#include <iomanip>
#include <limits>

namespace libMesh {

Xdr::Xdr(const std::string &name, XdrMODE mode) : _mode(mode), _name(name) {
  switch (mode) {
    case WRITE:
      _out = std::make_unique<std::ofstream>(name);
      if (!_out->good()) libmesh_error_msg("Xdr: cannot open " << name << " for writing");
      break;
    case READ:
      _in = std::make_unique<std::ifstream>(name);
      if (!_in->good()) libmesh_error_msg("Xdr: cannot open " << name << " for reading");
      break;
    case ENCODE:
    case DECODE:
      _fp = std::fopen(name.c_str(), mode == ENCODE ? "wb" : "rb");
      if (!_fp) libmesh_error_msg("Xdr: cannot open " << name);
      break;
    default:
      libmesh_error_msg("Xdr: invalid mode " << (int)mode);
  }
}

Xdr::~Xdr() {
  if (_fp) std::fclose(_fp);
}

void Xdr::close() {
  if (_out) {
    _out->flush();
    if (!_out->good()) libmesh_error_msg("Xdr: write to " << _name << " failed");
    _out->close();
    _out.reset();
  }
  if (_in) _in.reset();
  if (_fp) {
    const bool bad = std::fclose(_fp) != 0;
    _fp = nullptr;
    if (bad && _mode == ENCODE) libmesh_error_msg("Xdr: write to " << _name << " failed");
  }
}

void Xdr::check_stream(const char *what) {
  if ((_in && !_in->good()) || (_out && !_out->good()))
    libmesh_error_msg("Xdr: " << what << " failed on " << _name << " (truncated or malformed file)");
}

void Xdr::put_u32(uint32_t v) {
  const unsigned char b[4] = {(unsigned char)(v >> 24), (unsigned char)(v >> 16), (unsigned char)(v >> 8),
                              (unsigned char)v};
  if (std::fwrite(b, 1, 4, _fp) != 4) libmesh_error_msg("Xdr: write to " << _name << " failed");
}

uint32_t Xdr::get_u32() {
  unsigned char b[4];
  if (std::fread(b, 1, 4, _fp) != 4) libmesh_error_msg("Xdr: unexpected end of " << _name);
  return ((uint32_t)b[0] << 24) | ((uint32_t)b[1] << 16) | ((uint32_t)b[2] << 8) | (uint32_t)b[3];
}

template <typename T>
void Xdr::ascii_write(const T &a, const std::string &comment) {
  *_out << a;
  if (!comment.empty()) *_out << "\t " << comment;
  *_out << '\n';
  check_stream("write");
}

template <typename T>
void Xdr::ascii_read(T &a) {
  *_in >> a;
  check_stream("read");
  std::getline(*_in, _comm);
}

void Xdr::data(int &a, const std::string &comment) {
  switch (_mode) {
    case WRITE: return ascii_write(a, comment);
    case READ: return ascii_read(a);
    case ENCODE: return put_u32((uint32_t)a);
    default: a = (int)get_u32();
  }
}

void Xdr::data(unsigned int &a, const std::string &comment) {
  switch (_mode) {
    case WRITE: return ascii_write(a, comment);
    case READ: return ascii_read(a);
    case ENCODE: return put_u32(a);
    default: a = get_u32();
  }
}

void Xdr::data(std::string &a, const std::string &comment) {
  switch (_mode) {
    case WRITE: return ascii_write(a, comment);
    case READ: {
      // the line up to its first tab; the rest is the comment
      std::string line;
      std::getline(*_in, line);
      check_stream("read");
      const size_t tab = line.find('\t');
      a = line.substr(0, tab);
      _comm = tab == std::string::npos ? std::string() : line.substr(tab);
      return;
    }
    case ENCODE: {
      put_u32((uint32_t)a.size());
      static const char zeros[4] = {0, 0, 0, 0};
      const size_t pad = (4 - a.size() % 4) % 4;
      if (std::fwrite(a.data(), 1, a.size(), _fp) != a.size() || std::fwrite(zeros, 1, pad, _fp) != pad)
        libmesh_error_msg("Xdr: write to " << _name << " failed");
      return;
    }
    default: {
      const uint32_t n = get_u32();
      if (n > (1u << 20)) libmesh_error_msg("Xdr: implausible string length " << n << " in " << _name);
      std::vector<char> buf(n + (4 - n % 4) % 4);
      if (std::fread(buf.data(), 1, buf.size(), _fp) != buf.size())
        libmesh_error_msg("Xdr: unexpected end of " << _name);
      a.assign(buf.data(), n);
    }
  }
}

void Xdr::data(std::vector<uint16_t> &a, const std::string &comment) {
  switch (_mode) {
    case WRITE: {
      // Xdr::do_write(std::vector<T>) inside Xdr::data (xdr_cxx.C:820-835, 893-915)
      *_out << a.size() << "\t # vector length\n";
      for (uint16_t v : a) *_out << v << "\t ";
      if (!comment.empty()) *_out << "\t " << comment;
      *_out << '\n';
      check_stream("write");
      return;
    }
    case READ: {
      unsigned int n = 0;
      ascii_read(n);
      a.resize(n);
      for (auto &v : a) *_in >> v;
      check_stream("read");
      std::getline(*_in, _comm);
      return;
    }
    case ENCODE:
      put_u32((uint32_t)a.size());
      for (uint16_t v : a) put_u32(v);
      return;
    default: {
      const uint32_t n = get_u32();
      if (n > (1u << 16)) libmesh_error_msg("Xdr: implausible vector length " << n << " in " << _name);
      a.resize(n);
      for (auto &v : a) v = (uint16_t)get_u32();
    }
  }
}

void Xdr::data_stream(double *val, unsigned int len) {
  switch (_mode) {
    case WRITE: {
      const std::ios_base::fmtflags f = _out->flags();
      const std::streamsize p = _out->precision();
      *_out << std::scientific << std::setprecision(std::numeric_limits<double>::max_digits10);
      for (unsigned int i = 0; i < len; ++i) *_out << val[i] << ' ';
      _out->flags(f);
      _out->precision(p);
      check_stream("write");
      return;
    }
    case READ:
      for (unsigned int i = 0; i < len; ++i) *_in >> val[i];
      check_stream("read");
      return;
    case ENCODE:
      for (unsigned int i = 0; i < len; ++i) {
        uint64_t u;
        std::memcpy(&u, &val[i], 8);
        put_u32((uint32_t)(u >> 32));
        put_u32((uint32_t)u);
      }
      return;
    default:
      for (unsigned int i = 0; i < len; ++i) {
        const uint64_t hi = get_u32(), lo = get_u32();
        const uint64_t u = (hi << 32) | lo;
        std::memcpy(&val[i], &u, 8);
      }
  }
}

void Xdr::comment(const std::string &c) {
  if (_mode == WRITE) {
    *_out << "\t " << c << '\n';
    check_stream("write");
  } else if (_mode == READ) {
    std::getline(*_in, _comm);
  }
}

}  // namespace libMesh
