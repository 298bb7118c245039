// lmh_internal.h -- shared internals of liblmhip (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/lmhip.h"
#include "../../include/lmh_fe_map.h"

namespace lmh {

// Tail padding (entries) of CSR value/column arrays (kept for aligned vector
// reads past the last stored entry).
constexpr int64_t LMH_PAD = 8;

// SpMV row blocks hold at most this many stored entries
#ifndef LMH_RB_CAP
#define LMH_RB_CAP 2048
#endif
// workgroups per CU of the solvers' SpMV-type kernels and vector kernels
#ifndef LMH_SPMV_GRID_PER_CU
#define LMH_SPMV_GRID_PER_CU 16
#endif
// devices whose scratch addresses the library keeps
#define LMH_MAX_DEVICES 64

#ifndef LMH_GRID_PER_CU
#define LMH_GRID_PER_CU 4
#endif

// ---- errors ---------------------------------------------------------------
int set_error(int code, const char *fmt, ...);

#define LMH_HIP(call)                                                                                  \
  do {                                                                                                 \
    hipError_t _e = (call);                                                                            \
    if (_e != hipSuccess)                                                                              \
      return ::lmh::set_error(LMH_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(_e),     \
                              __FILE__, __LINE__);                                                     \
  } while (0)

#define LMH_CHECK(cond, code, ...)                                                                     \
  do {                                                                                                 \
    if (!(cond)) return ::lmh::set_error(code, __VA_ARGS__);                                           \
  } while (0)

#define LMH_TRY(call)                                                                                  \
  do {                                                                                                 \
    int _rc = (call);                                                                                  \
    if (_rc != LMH_OK) return _rc;                                                                     \
  } while (0)

// ---- FE reference tables (QGauss + Lagrange shapes) -------------------------
// Layout (all FP64): w[nq], phi[n*nq] (phi[a*nq+q]), dref[n*nq*3]
// (dref[(a*nq+q)*3+r]).
struct FETable {
  int type = 0, n = 0, nq = 0, arg = 0;
  std::vector<double> w, phi, dref, qp;
  double *d_data = nullptr;  // device copy: w | phi | dref
  const double *d_w() const { return d_data; }
  const double *d_phi() const { return d_data + nq; }
  const double *d_dref() const { return d_data + nq + (size_t)n * nq; }
};
// Build the table for an element type; arg = points per direction (hex) or
// QGauss order (tet).  Returns false for unsupported combinations.
bool build_fe_table(int type, int arg, FETable &t);
int nodes_per_elem(int type);

// ---- profiling: HIP events around each tracked launch on the context stream ---
struct ProfEntry {
  double total_ms = 0;
  int64_t count = 0;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
};

}  // namespace lmh

struct lmh_context {
  int device = 0;
  hipStream_t stream = nullptr;
  std::map<std::pair<int, int>, std::unique_ptr<lmh::FETable>> tables;
  bool prof_on = false;
  std::map<std::string, lmh::ProfEntry> prof;
  std::vector<hipEvent_t> event_pool;
  // small device scratch for reductions / solver scalars
  double *d_scal = nullptr;     // 64 doubles
  double *d_part = nullptr;     // reduction partials
  unsigned *d_count = nullptr;  // last-block counters (zeroed, self-resetting)
  int *d_flags = nullptr;       // error flags
  double *h_pinned = nullptr;   // pinned host scalars
  hipEvent_t ev_poll[2] = {nullptr, nullptr};  // solver convergence polls
  void *cg_scratch = nullptr;   // this device's address of the allocation-free CG's scratch
  int n_part = 0;
  int cu_count = 256;
};

struct lmh_mesh {
  lmh_context *ctx = nullptr;
  int type = 0, npe = 0;
  int64_t n_nodes = 0, n_elem = 0;
  double *xyz = nullptr;     // n_nodes*3
  int32_t *conn = nullptr;   // n_elem*npe
};

// General constraint rows after processing (DofMap::_dof_constraints with
// _primal_constraint_values).  "Empty" rows (no entries, value 0: homogeneous
// Dirichlet) are applied by the element kernels' mask; every element holding a
// DoF with a non-empty row or a nonzero value is a "general element" whose
// matrix goes through the dense constraint product C^T K C on its expanded DoF
// list (its own DoFs, then the DoFs its rows refer to).
struct lmh_constraints {
  int64_t version = 0;          // bumped by every lmh_dofmap_set_constraints
  bool asymmetric = true;       // asymmetric_constraint_rows
  int64_t n_general = 0;        // rows with entries or a nonzero value
  // all rows, ascending DoF, entries ascending (host mirror for queries)
  std::vector<int32_t> dof, ptr, col;
  std::vector<double> coef, val;
  // the same rows on the device (enforce_constraints_exactly)
  int32_t *d_dof = nullptr, *d_ptr = nullptr, *d_col = nullptr;
  double *d_coef = nullptr, *d_val = nullptr;
  // general elements
  int64_t n_ge = 0;
  int mx = 0;                   // expanded DoF count bound (<= 64)
  int32_t *ge_elem = nullptr;   // n_ge element ids, ascending
  int32_t *ge_m = nullptr;      // n_ge expanded sizes
  int32_t *ge_dofs = nullptr;   // n_ge*mx expanded DoF lists, -1 padded
  uint8_t *ge_flag = nullptr;   // n_ge*dpe: 1 = the element's DoF i is constrained
  double *ge_val = nullptr;     // n_ge*dpe: its constraint value (H)
  int32_t *ge_cptr = nullptr;   // n_ge*(mx+1): C by expanded column, absolute offsets
  int32_t *ge_ci = nullptr;     // local row i of each C entry
  double *ge_cv = nullptr;      // C entry
  // DoF -> sorted list of (g*mx + a): expanded position a of general element g
  std::vector<int32_t> h_xip;   // n_dofs+1 (host: pattern candidate bound)
  int32_t *xip = nullptr, *xinc = nullptr;
  int32_t *xrows = nullptr;     // rows with at least one such entry
  int64_t n_xrows = 0;
};

struct lmh_dofmap {
  lmh_context *ctx = nullptr;
  const lmh_mesh *mesh = nullptr;
  int nv = 1, dpe = 0;
  int64_t n_dofs = 0;
  int32_t *edofs = nullptr;   // n_elem*dpe
  uint8_t *cons = nullptr;    // n_dofs: 1 = constrained by an empty row (masked in the element kernels)
  std::vector<uint8_t> h_cons0;  // the constrained set given at creation
  // incidence: dof -> sorted list of (e*dpe + i)
  int32_t *inc_ptr = nullptr;  // n_dofs+1
  int32_t *inc = nullptr;      // n_elem*dpe
  lmh_constraints cr;
};

struct lmh_matrix {
  lmh_context *ctx = nullptr;
  const lmh_dofmap *dm = nullptr;
  int64_t rows = 0, cols = 0, nnz = 0;  // cols == rows except for rectangular lmh_matrix_create_csr2
  int max_row = 0;
  int32_t *rp = nullptr, *ci = nullptr;
  int32_t *rb = nullptr;  // SpMV row-block boundaries (nrb+1 rows)
  int64_t nrb = 0;
  double *vals = nullptr;
  float *vals32 = nullptr;
  bool have32 = false;
  int precond = LMH_PRECOND_JACOBI;  // preconditioner of the solvers on this matrix
  // transpose solves of a structurally symmetric pattern: mirror position of
  // each entry (0 not built, 1 built, -1 pattern not symmetric) and the
  // mirrored values
  int tpos_state = 0;
  int32_t *tpos = nullptr;
  double *vals_t = nullptr;
  // Column compression of the SpMV operators (x windows): each row block of
  // the window partition rbw lists its distinct columns once, sorted, as
  // 16-bit offsets woff16[w0 .. w0+nu), col = off + (off <= s0 ? b0 : base1);
  // every entry stores the index of its column in that list, 8-bit (blocks of
  // at most 1024 entries and 256 distinct columns) or else 16-bit (the
  // 2048-entry SpMV blocks, up to 768 columns).  rbmeta[2bk] = {r0, r1, k0,
  // cnt} (the block's rows and entries), rbmeta[2bk+1] = {b0, base1, s0 | nu << 16, w0}.
  int fp32_index = 1;          // 1: use the compressed columns when they fit
  int idx16_state = 0;         // 0 not built, 1 built, -1 does not fit
  int wbits = 0;               // 8 or 16 once built
  int32_t *rbw = nullptr;      // window partition (nrbw + 1 row boundaries)
  int64_t nrbw = 0;
  uint8_t *widx8 = nullptr;    // nnz: 8-bit window index of each entry's column
  uint16_t *widx16 = nullptr;  // nnz: 16-bit window index of each entry's column
  uint16_t *woff16 = nullptr;  // sum of nu over the row blocks
  int4 *rbmeta = nullptr;      // 2 nrbw
  uint16_t *off = nullptr;  // n_elem*dpe*dpe slot offsets within rows
  int32_t *diag = nullptr;  // rows: position of the diagonal entry, -1 if absent
  int ncolors = 0;
  std::vector<int64_t> color_ptr;  // host, ncolors+1
  int32_t *color_elems = nullptr;  // device, n_elem (grouped by colour, ascending id)
  // element buffers
  double *Ke = nullptr, *Fe = nullptr;
  float *Ke32 = nullptr;
  // general-element matrices (constraint version they were built for, slot
  // offsets of expanded entries, C^T K C, C^T (F - K H))
  int64_t cons_version = 0;
  uint16_t *offx = nullptr;   // n_ge*mx*mx
  double *Kx = nullptr, *Fx = nullptr;
  float *Kx32 = nullptr;
  // solver work vectors (allocated on first solve)
  double *w64 = nullptr;  // 4*rows
  double *wbcg = nullptr;  // 9*rows (BiCGSTAB)
  float *w32 = nullptr;   // 5*rows
};

namespace lmh {

struct ProfScope {
  lmh_context *ctx;
  const char *name;
  hipEvent_t a = nullptr, b = nullptr;
  ProfScope(lmh_context *c, const char *n);
  ~ProfScope();
};

hipEvent_t get_event(lmh_context *ctx);

// launch helper: records HIP events on the context stream when profiling is on
#define LMH_LAUNCH(ctx, name, kernel, grid, block, shmem, ...)                                          \
  do {                                                                                                 \
    ::lmh::ProfScope _ps((ctx), (name));                                                               \
    hipLaunchKernelGGL(kernel, grid, block, shmem, (ctx)->stream, __VA_ARGS__);                        \
  } while (0)

int check_launch(const char *what);
// per-device address of the allocation-free CG's scratch into ctx (context creation)
int cg_scratch_setup(lmh_context *ctx);
// x windows of a new matrix's pattern (csrc/lmh_solve.hip)
int matrix_build_windows(lmh_matrix *A);

// device helpers shared across translation units
int exclusive_scan_i32(lmh_context *ctx, const int32_t *in, int32_t *out, int64_t n, int64_t *total);
int build_incidence(lmh_context *ctx, const int32_t *edofs, int64_t n_elem, int dpe, int64_t n_dofs,
                    int32_t **inc_ptr, int32_t **inc);
int fe_table(lmh_context *ctx, int type, int arg, const FETable **out);
// CSR pattern, scatter map, (if colour) element colouring and (if fp64) zeroed
// FP64 values for a DoF map
int matrix_create_impl(lmh_context *ctx, const lmh_dofmap *dm, bool colour, bool fp64, lmh_matrix **out);
// general constraints around the scatter: before it, the general elements'
// matrices are replaced by their constrained, expanded forms (and zeroed in
// Ke/Fe); after it, those are added into their rows
template <typename T>
int constrain_general_impl(lmh_matrix *A, T *Ke, double *Fe);
template <typename T>
int scatter_extra_impl(lmh_matrix *A, T *vals, double *rhs);
int release_constraints(lmh_constraints &c);

// ---- element map at one quadrature point ----------------------------------
// From LDS coordinates X[a*3+c] and the reference derivative table
// dref[(a*nq+q)*3+r]: J[r][c] = sum_a X_a[c] dphi_a/dxi_r(q), G = J^{-1} by
// cofactors as FEMap::compute_single_point_map (fe_map.C:1046-1068),
// G[r][c] = d xi_r / d x_c.  Returns det J; G is only meaningful when det > 0.
__device__ inline double qp_map(const double *X, const double *dref, int n, int nq, int q, double G[3][3]) {
  return lmh_qp_map(X, dref, n, nq, q, G);
}

// MI355X dispatches workgroups round-robin over its 8 XCDs; remap so each XCD
// owns one contiguous range of logical blocks (its L2 then holds the data its
// neighbouring rows share).
constexpr int NXCD = 8;
__device__ inline int logical_block() {
  const int G = gridDim.x, b = blockIdx.x;
  if (G % NXCD) return b;
  return (b % NXCD) * (G / NXCD) + b / NXCD;
}
// this workgroup's contiguous share [lo, hi) of n items
__device__ inline void block_range(int64_t n, int64_t &lo, int64_t &hi) {
  const int lb = logical_block();
  const int64_t per = (n + gridDim.x - 1) / gridDim.x;
  lo = (int64_t)lb * per;
  hi = lo + per < n ? lo + per : n;
}

// the same with lo a multiple of W (the share rounded up to W items)
template <int W>
__device__ inline void block_range_aligned(int64_t n, int64_t &lo, int64_t &hi) {
  const int lb = logical_block();
  int64_t per = (n + gridDim.x - 1) / gridDim.x;
  per = (per + W - 1) / W * W;
  lo = (int64_t)lb * per;
  if (lo > n) lo = n;
  hi = lo + per < n ? lo + per : n;
}

// grid for an XCD-aware 1-D launch
inline int grid_for(int64_t work, int per_block, int max_blocks) {
  int64_t g = (work + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > max_blocks) g = max_blocks;
  return (int)g;
}

}  // namespace lmh
