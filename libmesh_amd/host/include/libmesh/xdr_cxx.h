// xdr_cxx.h -- Xdr (include/utils/xdr_cxx.h): the stream that EquationSystems
// files are written through, in libMesh's two encodings.
//   WRITE / READ (".xda"): ASCII.  Every datum on a line of its own, followed
//     by a tab, a space and its comment; floating point in scientific notation
//     with max_digits10 significant digits after the first; a stream of values
//     is written value-space-value-space on one line (src/utils/xdr_cxx.C:
//     860-920 data, 1175-1225 data_stream, 1380-1410 comment; reading takes the
//     value with operator>> and drops the rest of the line, 744-790).
//   ENCODE / DECODE (".xdr"): XDR, RFC 4506.  Big-endian 4-byte integers
//     (unsigned short widened to 4 bytes), 8-byte IEEE doubles, strings as a
//     4-byte length, the bytes and zero padding to a multiple of 4, vectors as a
//     4-byte length and the elements; value streams carry no length; comments
//     are not stored (xdr_translate, xdr_cxx.C:484-560).
#pragma once

#include <cstdint>
#include <cstdio>
#include <fstream>
#include <memory>
#include <string>
#include <vector>

#include "libmesh/libmesh_common.h"

namespace libMesh {

// include/enums/enum_xdr_mode.h
enum XdrMODE : int { UNKNOWN = -1, ENCODE = 0, DECODE, WRITE, READ };

class Xdr {
 public:
  Xdr(const std::string &name, XdrMODE mode);
  ~Xdr();
  Xdr(const Xdr &) = delete;
  Xdr &operator=(const Xdr &) = delete;

  bool reading() const { return _mode == DECODE || _mode == READ; }
  bool writing() const { return _mode == ENCODE || _mode == WRITE; }
  bool binary() const { return _mode == ENCODE || _mode == DECODE; }
  XdrMODE access_mode() const { return _mode; }

  void data(int &a, const std::string &comment = "");
  void data(unsigned int &a, const std::string &comment = "");
  void data(std::string &a, const std::string &comment = "");
  void data(std::vector<uint16_t> &a, const std::string &comment = "");
  // len values, no length field
  void data_stream(double *val, unsigned int len);
  // WRITE: a line holding only the comment; READ: drops the rest of the line
  void comment(const std::string &c);
  // READ: what followed the last datum on its line (its comment), "" otherwise
  const std::string &last_comment() const { return _comm; }
  void close();

 private:
  void put_u32(uint32_t v);
  uint32_t get_u32();
  void check_stream(const char *what);
  template <typename T>
  void ascii_write(const T &a, const std::string &comment);
  template <typename T>
  void ascii_read(T &a);

  XdrMODE _mode;
  std::string _name;
  std::unique_ptr<std::ofstream> _out;
  std::unique_ptr<std::ifstream> _in;
  FILE *_fp = nullptr;
  std::string _comm;
};

}  // namespace libMesh
