// matrix_io.C -- Matlab text and PETSc binary sparse-matrix files
// (SparseMatrix::read / read_matlab / print_matlab, sparse_matrix.C:398-491,
// 686-745, 1050-1330).
#include "libmesh/matrix_io.h"

#include <zlib.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <regex>
#include <sstream>

namespace libMesh {
namespace MatrixIO {

namespace {

bool ends_with(const std::string &s, const std::string &suffix) {
  return s.size() >= suffix.size() && s.compare(s.size() - suffix.size(), suffix.size(), suffix) == 0;
}

// whole file as text; gzip-compressed files are inflated
std::string slurp(const std::string &filename) {
  gzFile f = gzopen(filename.c_str(), "rb");  // reads plain files unchanged
  if (!f) libmesh_error_msg("ERROR: cannot read file:\n\t" << filename);
  std::string out;
  char buf[1 << 16];
  int got;
  while ((got = gzread(f, buf, sizeof(buf))) > 0) out.append(buf, (size_t)got);
  const bool bad = got < 0;
  gzclose(f);
  if (bad) libmesh_error_msg("ERROR: cannot inflate file:\n\t" << filename);
  return out;
}

uint64_t read_be(const unsigned char *p, int bytes) {
  uint64_t v = 0;
  for (int k = 0; k < bytes; ++k) v = (v << 8) | p[k];
  return v;
}

}  // namespace

Entries read_matlab(const std::string &filename) {
  const std::string text = slurp(filename);
  std::istringstream file(text);
  // the regular expressions of read_matlab (sparse_matrix.C:1097-1135)
  const std::regex start_regex("\\s*\\w+\\s*=\\s*\\[");
  const std::regex end_regex("^[^%]*\\]");
  const std::regex size_regex("%\\s*[Ss][Ii][Zz][Ee]\\s*=\\s*(\\d+)\\s+(\\d+)");
  Entries E;
  bool started = false, ended = false;
  std::size_t cur_row = 1, max_i = 0, max_j = 0;
  for (std::string line; std::getline(file, line);) {
    std::istringstream l(line);
    std::size_t i, j;
    double v;
    l >> i >> j >> v;
    if (!l.fail()) {
      if (!started) libmesh_error_msg("Confused by premature entries in matrix file " << filename);
      if (!i || !j) libmesh_error_msg("Expected 1-based indexing in matrix file " << filename);
      if (i < cur_row) libmesh_error_msg("Can't handle out-of-order entries in matrix file " << filename);
      cur_row = i;
      E.ijv.emplace_back(i - 1, j - 1, v);
      max_i = std::max(max_i, i);
      max_j = std::max(max_j, j);
      continue;
    }
    std::smatch sm;
    if (std::regex_search(line, sm, size_regex)) {
      E.m = std::stoull(sm[1]);
      E.n = std::stoull(sm[2]);
    } else if (std::regex_search(line, start_regex)) {
      started = true;
    } else if (std::regex_search(line, end_regex)) {
      ended = true;
      break;
    }
  }
  if (!started || !ended) libmesh_error_msg("Incomplete matrix in file " << filename);
  if (E.m == 0 && E.n == 0) {
    E.m = max_i;
    E.n = max_j;
  }
  if (max_i > E.m || max_j > E.n) libmesh_error_msg("Entries outside the declared size in matrix file " << filename);
  return E;
}

Entries read_petsc_binary(const std::string &filename, int ib) {
  if (ib != 4 && ib != 8) libmesh_error_msg("PETSc binary: index width must be 4 or 8 bytes");
  std::ifstream in(filename, std::ios::binary);
  if (!in.good()) libmesh_error_msg("ERROR: cannot read file:\n\t" << filename);
  std::vector<unsigned char> b((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
  // the class id is a 32-bit int in the 32-bit format, index-width in the 64-bit one
  const std::size_t hdr = 4 * (std::size_t)ib;
  if (b.size() < hdr) libmesh_error_msg("PETSc binary: truncated header in " << filename);
  const uint64_t cid = read_be(b.data(), ib), M = read_be(b.data() + ib, ib), N = read_be(b.data() + 2 * ib, ib),
                 nz = read_be(b.data() + 3 * ib, ib);
  if (cid != 1211216) libmesh_error_msg("PETSc binary: not a matrix file (class id " << cid << ") " << filename);
  const std::size_t need = hdr + (std::size_t)M * ib + (std::size_t)nz * ib + (std::size_t)nz * 8;
  if (b.size() != need) libmesh_error_msg("PETSc binary: size mismatch in " << filename);
  Entries E;
  E.m = M;
  E.n = N;
  const unsigned char *rl = b.data() + hdr, *ci = rl + M * ib, *va = ci + nz * ib;
  std::size_t k = 0;
  for (std::size_t r = 0; r < M; ++r) {
    const uint64_t len = read_be(rl + r * ib, ib);
    for (uint64_t t = 0; t < len; ++t, ++k) {
      if (k >= nz) libmesh_error_msg("PETSc binary: row lengths exceed nz in " << filename);
      const uint64_t c = read_be(ci + k * ib, ib);
      if (c >= N) libmesh_error_msg("PETSc binary: column out of range in " << filename);
      const uint64_t bits = read_be(va + k * 8, 8);
      double v;
      std::memcpy(&v, &bits, 8);
      E.ijv.emplace_back(r, (std::size_t)c, v);
    }
  }
  if (k != nz) libmesh_error_msg("PETSc binary: row lengths do not sum to nz in " << filename);
  return E;
}

Entries read(const std::string &filename) {
  std::string base = filename;
  const bool gz = ends_with(base, ".gz");
  if (gz) base.resize(base.size() - 3);
  if (ends_with(base, ".m") || ends_with(base, ".matlab")) return read_matlab(filename);
  if (gz) libmesh_error_msg("Gzipped " << base << " matrices are not supported");
  if (ends_with(base, ".petsc64")) return read_petsc_binary(filename, 8);
  if (ends_with(base, ".petsc32")) return read_petsc_binary(filename, 4);
  libmesh_error_msg("ERROR: Unrecognized matrix file extension on: " << filename
                    << " (understood: .m, .matlab, .m.gz, .matlab.gz, .petsc32, .petsc64)");
  return Entries{};
}

void write_matlab(std::ostream &os, std::size_t m, std::size_t n, const std::vector<int32_t> &rp,
                  const std::vector<int32_t> &ci, const std::vector<double> &vals) {
  std::size_t nonzeros = 0;
  for (double v : vals) nonzeros += v != 0.0;
  // header of SparseMatrix::print_matlab (one process)
  os << "%Mat Object: () 1 MPI processes\n"
     << "%  type: seqaij\n"
     << "% Size = " << m << ' ' << n << '\n'
     << "% Nonzeros = " << nonzeros << '\n'
     << "zzz = zeros(" << nonzeros << ",3);\n"
     << "zzz = [\n";
  char buf[64];
  for (std::size_t r = 0; r < m; ++r)
    for (int32_t k = rp[r]; k < rp[r + 1]; ++k) {
      if (vals[k] == 0.0) continue;
      std::snprintf(buf, sizeof(buf), "%.16e", vals[k]);
      os << (r + 1) << ' ' << (ci[k] + 1) << "  " << buf << '\n';
    }
  os << "];\n"
     << "Mat_sparse = spconvert(zzz);" << std::endl;
}

// Big-endian words of `bytes` bytes (the layout read_petsc_binary reads).
static void put_be(std::vector<unsigned char> &out, uint64_t v, int bytes) {
  for (int k = bytes - 1; k >= 0; --k) out.push_back((unsigned char)(v >> (8 * k)));
}

void write_petsc_binary(const std::string &filename, int ib, std::size_t m, std::size_t n,
                        const std::vector<int32_t> &rp, const std::vector<int32_t> &ci,
                        const std::vector<double> &vals) {
  if (ib != 4 && ib != 8) libmesh_error_msg("PETSc binary: index width must be 4 or 8 bytes");
  const std::size_t nz = vals.size();
  std::vector<unsigned char> b;
  b.reserve(4 * ib + m * ib + nz * (ib + 8));
  put_be(b, 1211216, ib);
  put_be(b, m, ib);
  put_be(b, n, ib);
  put_be(b, nz, ib);
  for (std::size_t r = 0; r < m; ++r) put_be(b, (uint64_t)(rp[r + 1] - rp[r]), ib);
  for (std::size_t k = 0; k < nz; ++k) put_be(b, (uint64_t)ci[k], ib);
  for (std::size_t k = 0; k < nz; ++k) {
    uint64_t bits;
    std::memcpy(&bits, &vals[k], 8);
    put_be(b, bits, 8);
  }
  std::ofstream os(filename, std::ios::binary);
  if (!os.good()) libmesh_error_msg("ERROR: cannot write to file:\n\t" << filename);
  os.write(reinterpret_cast<const char *>(b.data()), (std::streamsize)b.size());
}

}  // namespace MatrixIO
}  // namespace libMesh
