// matrix_io.h -- on-disk sparse-matrix formats of SparseMatrix::read / print
// (src/numerics/sparse_matrix.C: read :686-745, print_matlab :398-491,
// read_matlab :1050-1330): the Matlab "spconvert" text format, gzipped or
// not, and PETSc's binary matrix format with 32- or 64-bit indices.
#pragma once

#include <cstdint>
#include <iosfwd>
#include <string>
#include <tuple>
#include <vector>

#include "libmesh/libmesh_common.h"

namespace libMesh {
namespace MatrixIO {

// A matrix read from disk: sizes and 0-based (row, column, value) entries in
// file order (rows non-decreasing).
struct Entries {
  std::size_t m = 0, n = 0;
  std::vector<std::tuple<std::size_t, std::size_t, double>> ijv;
};

// Matlab text (`.m`, `.matlab`, optionally `.gz`): entries "i j value" with
// 1-based indices between a "name = [" line and the closing "]"; the size comes
// from a "% Size = m n" comment, else from the largest indices seen.  Entries
// must come in non-decreasing row order (as read_matlab requires).
Entries read_matlab(const std::string &filename);

// PETSc binary (`.petsc32` / `.petsc64`): big-endian class id 1211216, M, N,
// nz, then M row lengths, nz column indices, nz values (FP64); index width
// 4 or 8 bytes.
Entries read_petsc_binary(const std::string &filename, int index_bytes);

// Dispatch on the extension as SparseMatrix::read does (.m/.matlab[.gz],
// .petsc32, .petsc64); HDF5 is not supported (LogicError).
Entries read(const std::string &filename);

// Matlab text of a CSR matrix with the header of SparseMatrix::print_matlab;
// values that are exactly zero are skipped, values are printed with 17
// significant digits so that read_matlab returns them exactly.
void write_matlab(std::ostream &os, std::size_t m, std::size_t n, const std::vector<int32_t> &row_ptr,
                  const std::vector<int32_t> &col_idx, const std::vector<double> &vals);

// PETSc binary of a CSR matrix, the layout read_petsc_binary reads (every
// stored entry, explicit zeros included; index width 4 or 8 bytes)
void write_petsc_binary(const std::string &filename, int index_bytes, std::size_t m, std::size_t n,
                        const std::vector<int32_t> &row_ptr, const std::vector<int32_t> &col_idx,
                        const std::vector<double> &vals);

}  // namespace MatrixIO
}  // namespace libMesh
