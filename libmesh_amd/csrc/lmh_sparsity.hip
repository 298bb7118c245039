// lmh_sparsity.hip -- GPU sparsity pattern (SparsityPattern::Build), element
// scatter map, element colouring and the CSR matrix object.
//
// Pattern rule (src/base/sparsity_pattern.C:96-365, DefaultCoupling with no
// neighbour layers): row r is the sorted, unique union of the DoFs of every
// element that contains DoF r.  The GPU builds it from the DoF->element
// incidence lists: one wavefront per row gathers the candidate DoFs of its
// incident elements into LDS, bitonic-sorts them and counts (pass 1) or writes
// (pass 2) the unique values.
#include <algorithm>
#include <climits>
#include <cstring>

#include "lmh_internal.h"

namespace lmh {

// one wave (64 lanes) per row, rows grid-strided; cap = power of two >= max
// candidates.  With general constraints (xip non-null) a row also takes the
// expanded DoF lists (xd, mx per general element, -1 padded) of the general
// elements it appears in (find_connected_dofs, dof_map.C:2915-2972).
template <bool FILL>
__global__ __launch_bounds__(64) void k_row_union(const int32_t *__restrict__ inc_ptr, const int32_t *__restrict__ inc,
                                                  const int32_t *__restrict__ edofs, int dpe, int64_t n_rows, int cap,
                                                  const int32_t *__restrict__ xip, const int32_t *__restrict__ xinc,
                                                  const int32_t *__restrict__ xd, int mx,
                                                  int32_t *__restrict__ rowlen, const int32_t *__restrict__ rp,
                                                  int32_t *__restrict__ ci) {
  extern __shared__ int32_t keys[];
  const int lane = threadIdx.x;
  for (int64_t r = blockIdx.x; r < n_rows; r += gridDim.x) {
    const int32_t b = inc_ptr[r], e = inc_ptr[r + 1];
    const int32_t xb = xip ? xip[r] : 0, xe = xip ? xip[r + 1] : 0;
    const int nbase = (e - b) * dpe;
    const int ncand = nbase + (xe - xb) * mx;
    // gather candidates: incidence k covers element inc[k]/dpe; expanded
    // incidence k covers general element xinc[k]/mx
    for (int t = lane; t < cap; t += 64) {
      int32_t v = INT_MAX;
      if (t < nbase) {
        const int k = t / dpe, j = t - k * dpe;
        const int32_t el = inc[b + k] / dpe;
        v = edofs[(int64_t)el * dpe + j];
      } else if (t < ncand) {
        const int tt = t - nbase;
        const int k = tt / mx, j = tt - k * mx;
        const int32_t g = xinc[xb + k] / mx;
        v = xd[(int64_t)g * mx + j];
        if (v < 0) v = INT_MAX;
      }
      keys[t] = v;
    }
    __syncthreads();
    // bitonic sort ascending
    for (int k = 2; k <= cap; k <<= 1) {
      for (int j = k >> 1; j > 0; j >>= 1) {
        for (int i = lane; i < cap; i += 64) {
          const int p = i ^ j;
          if (p > i) {
            const int32_t a = keys[i], c = keys[p];
            const bool up = (i & k) == 0;
            if ((a > c) == up) {
              keys[i] = c;
              keys[p] = a;
            }
          }
        }
        __syncthreads();
      }
    }
    // unique
    int base = 0;
    for (int t0 = 0; t0 < cap; t0 += 64) {
      const int t = t0 + lane;
      const int32_t v = keys[t];
      const bool uniq = (v != INT_MAX) && (t == 0 || keys[t - 1] != v);
      const unsigned long long m = __ballot(uniq);
      if (FILL) {
        if (uniq) {
          const int pos = base + __popcll(m & ((1ull << lane) - 1ull));
          ci[rp[r] + pos] = v;
        }
      }
      base += __popcll(m);
    }
    if (!FILL && lane == 0) rowlen[r] = base;
    __syncthreads();
  }
}

// slot offset of every (e, i, j) within row edofs[e][i]: binary search of edofs[e][j]
__global__ void k_scatter_offsets(const int32_t *__restrict__ edofs, int64_t n_elem, int dpe,
                                  const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                  uint16_t *__restrict__ off, int *__restrict__ err) {
  const int64_t total = n_elem * dpe * dpe;
  const int64_t dd = (int64_t)dpe * dpe;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t e = t / dd;
    const int rem = (int)(t - e * dd);
    const int i = rem / dpe, j = rem - i * dpe;
    const int32_t r = edofs[e * dpe + i], c = edofs[e * dpe + j];
    int32_t lo = rp[r], hi = rp[r + 1];
    const int32_t base = lo;
    while (lo < hi) {
      const int32_t mid = (lo + hi) >> 1;
      if (ci[mid] < c)
        lo = mid + 1;
      else
        hi = mid;
    }
    if (lo >= rp[r + 1] || ci[lo] != c) atomicOr(err, 1);
    off[t] = (uint16_t)(lo - base);
  }
}

// slot offset of every (g, a, b) of the general elements' expanded matrices
// within row X_g[a]; 0 in the padding
__global__ void k_xoffsets(const int32_t *__restrict__ xd, const int32_t *__restrict__ gm, int64_t n_ge, int mx,
                           const int32_t *__restrict__ rp, const int32_t *__restrict__ ci, uint16_t *__restrict__ off,
                           int *__restrict__ err) {
  const int64_t mm = (int64_t)mx * mx, total = n_ge * mm;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t g = t / mm;
    const int rem = (int)(t - g * mm);
    const int a = rem / mx, bb = rem - a * mx;
    const int m = gm[g];
    uint16_t o = 0;
    if (a < m && bb < m) {
      const int32_t r = xd[g * mx + a], c = xd[g * mx + bb];
      int32_t lo = rp[r], hi = rp[r + 1];
      const int32_t base = lo;
      while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (ci[mid] < c)
          lo = mid + 1;
        else
          hi = mid;
      }
      if (lo >= rp[r + 1] || ci[lo] != c) atomicOr(err, 1);
      o = (uint16_t)(lo - base);
    }
    off[t] = o;
  }
}

// Greedy first-fit colouring in element order on the host: an element takes the
// smallest colour not used by any element sharing a DoF with it.  On a
// structured hex lattice this gives the 8 parity colours; the scatter within a
// colour is race-free, and the colour order fixes the summation order.
static int color_elements(const std::vector<int32_t> &edofs, int64_t n_elem, int dpe, int64_t n_dofs,
                          std::vector<int64_t> &cptr, std::vector<int32_t> &celems, int &ncolors) {
  std::vector<int32_t> ip(n_dofs + 1, 0);
  for (int64_t t = 0; t < n_elem * dpe; ++t) ip[edofs[t] + 1]++;
  for (int64_t d = 0; d < n_dofs; ++d) ip[d + 1] += ip[d];
  std::vector<int32_t> lst(ip[n_dofs]);
  std::vector<int32_t> fill(ip.begin(), ip.end() - 1);
  for (int64_t e = 0; e < n_elem; ++e)
    for (int i = 0; i < dpe; ++i) lst[fill[edofs[e * dpe + i]]++] = (int32_t)e;
  std::vector<int32_t> color(n_elem, -1);
  ncolors = 0;
  for (int64_t e = 0; e < n_elem; ++e) {
    uint64_t used = 0;
    for (int i = 0; i < dpe; ++i) {
      const int32_t d = edofs[e * dpe + i];
      for (int32_t k = ip[d]; k < ip[d + 1]; ++k) {
        const int32_t f = lst[k];
        if (f != e && color[f] >= 0 && color[f] < 64) used |= 1ull << color[f];
      }
    }
    int c = 0;
    while (c < 64 && (used >> c) & 1ull) ++c;
    if (c >= 64) return set_error(LMH_ERR_UNSUPPORTED, "element colouring needs more than 64 colours");
    color[e] = c;
    ncolors = std::max(ncolors, c + 1);
  }
  cptr.assign(ncolors + 1, 0);
  for (int64_t e = 0; e < n_elem; ++e) cptr[color[e] + 1]++;
  for (int c = 0; c < ncolors; ++c) cptr[c + 1] += cptr[c];
  celems.resize(n_elem);
  std::vector<int64_t> pos(cptr.begin(), cptr.end() - 1);
  for (int64_t e = 0; e < n_elem; ++e) celems[pos[color[e]]++] = (int32_t)e;
  return LMH_OK;
}

static int build_row_blocks(lmh_matrix *A, const std::vector<int32_t> &rp) {
  // SpMV row blocks: consecutive rows holding at most kRowBlockCap stored
  // entries (a longer row forms a block of its own).  The cap leaves room for
  // the FP32 SpMV's alignment of a block's start down to 8 entries.
  const int kRowBlockCap = LMH_RB_CAP - 8;
  std::vector<int32_t> rb{0};
  const int64_t n = A->rows;
  for (int64_t r = 0; r < n; ++r)
    if (r > rb.back() && rp[r + 1] - rp[rb.back()] > kRowBlockCap) rb.push_back((int32_t)r);
  rb.push_back((int32_t)n);
  A->nrb = (int64_t)rb.size() - 1;
  LMH_HIP(hipMalloc((void **)&A->rb, sizeof(int32_t) * rb.size()));
  LMH_HIP(hipMemcpy(A->rb, rb.data(), sizeof(int32_t) * rb.size(), hipMemcpyHostToDevice));
  return LMH_OK;
}

}  // namespace lmh

using namespace lmh;

namespace lmh {

int matrix_create_impl(lmh_context *ctx, const lmh_dofmap *dm, bool colour, bool fp64, lmh_matrix **out) {
  LMH_CHECK(ctx && dm && out, LMH_ERR_INVALID, "null argument");
  LMH_HIP(hipSetDevice(ctx->device));
  const int64_t n = dm->n_dofs, ne = dm->mesh->n_elem;
  const int dpe = dm->dpe;
  std::unique_ptr<lmh_matrix> A(new lmh_matrix);
  A->ctx = ctx;
  A->dm = dm;
  A->rows = n;
  A->cols = n;

  // max candidates per row
  std::vector<int32_t> ip(n + 1);
  LMH_HIP(hipMemcpy(ip.data(), dm->inc_ptr, sizeof(int32_t) * (n + 1), hipMemcpyDeviceToHost));
  const lmh_constraints &cr = dm->cr;
  const bool gen = cr.n_ge > 0;
  A->cons_version = cr.version;
  int64_t maxcand = 0;
  for (int64_t d = 0; d < n; ++d)
    maxcand = std::max<int64_t>(maxcand, (int64_t)(ip[d + 1] - ip[d]) * dpe +
                                             (gen ? (int64_t)(cr.h_xip[d + 1] - cr.h_xip[d]) * cr.mx : 0));
  int cap = 64;
  while (cap < maxcand) cap <<= 1;
  LMH_CHECK(cap <= 16384, LMH_ERR_UNSUPPORTED, "row candidate count %lld too large", (long long)maxcand);
  const size_t shm = sizeof(int32_t) * cap;
  const int32_t *xip = gen ? cr.xip : nullptr;

  int32_t *rowlen = nullptr;
  LMH_HIP(hipMalloc((void **)&rowlen, sizeof(int32_t) * (n + 1)));
  LMH_HIP(hipMemsetAsync(rowlen, 0, sizeof(int32_t) * (n + 1), ctx->stream));
  const int g = (int)std::min<int64_t>(n, 65536);
  LMH_LAUNCH(ctx, "sparsity_count", k_row_union<false>, dim3(g), dim3(64), shm, dm->inc_ptr, dm->inc, dm->edofs, dpe,
             n, cap, xip, cr.xinc, cr.ge_dofs, cr.mx, rowlen, (const int32_t *)nullptr, (int32_t *)nullptr);
  LMH_TRY(check_launch("k_row_union<count>"));
  LMH_HIP(hipMalloc((void **)&A->rp, sizeof(int32_t) * (n + 1)));
  int64_t nnz = 0;
  LMH_TRY(exclusive_scan_i32(ctx, rowlen, A->rp, n, &nnz));
  A->nnz = nnz;
  LMH_HIP(hipMalloc((void **)&A->ci, sizeof(int32_t) * (nnz + LMH_PAD)));
  LMH_HIP(hipMemsetAsync(A->ci + nnz, 0, sizeof(int32_t) * LMH_PAD, ctx->stream));
  LMH_LAUNCH(ctx, "sparsity_fill", k_row_union<true>, dim3(g), dim3(64), shm, dm->inc_ptr, dm->inc, dm->edofs, dpe,
             n, cap, xip, cr.xinc, cr.ge_dofs, cr.mx, (int32_t *)nullptr, (const int32_t *)A->rp, A->ci);
  LMH_TRY(check_launch("k_row_union<fill>"));

  // row length bound for the 16-bit scatter offsets
  std::vector<int32_t> rp(n + 1);
  LMH_HIP(hipMemcpyAsync(rp.data(), A->rp, sizeof(int32_t) * (n + 1), hipMemcpyDeviceToHost, ctx->stream));
  LMH_HIP(hipStreamSynchronize(ctx->stream));
  for (int64_t r = 0; r < n; ++r) A->max_row = std::max<int>(A->max_row, rp[r + 1] - rp[r]);
  LMH_TRY(build_row_blocks(A.get(), rp));
  LMH_HIP(hipFree(rowlen));
  LMH_CHECK(A->max_row <= 65535, LMH_ERR_UNSUPPORTED, "row length %d exceeds 16-bit scatter offsets", A->max_row);

  LMH_HIP(hipMalloc((void **)&A->off, sizeof(uint16_t) * ne * dpe * dpe));
  LMH_HIP(hipMemsetAsync(ctx->d_flags, 0, sizeof(int), ctx->stream));
  LMH_LAUNCH(ctx, "scatter_offsets", k_scatter_offsets, dim3(grid_for(ne * dpe * dpe, 256, 16384)), dim3(256), 0,
             dm->edofs, ne, dpe, A->rp, A->ci, A->off, ctx->d_flags);
  LMH_TRY(check_launch("k_scatter_offsets"));
  int bad = 0;
  LMH_HIP(hipMemcpyAsync(&bad, ctx->d_flags, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  LMH_HIP(hipStreamSynchronize(ctx->stream));
  LMH_CHECK(!bad, LMH_ERR_INVALID, "internal: element DoF pair missing from the sparsity pattern");
  if (gen) {
    const int64_t tot = cr.n_ge * cr.mx * cr.mx;
    LMH_HIP(hipMalloc((void **)&A->offx, sizeof(uint16_t) * tot));
    LMH_LAUNCH(ctx, "scatter_offsets_expanded", k_xoffsets, dim3(grid_for(tot, 256, 16384)), dim3(256), 0,
               cr.ge_dofs, cr.ge_m, cr.n_ge, cr.mx, A->rp, A->ci, A->offx, ctx->d_flags);
    LMH_TRY(check_launch("k_xoffsets"));
    LMH_HIP(hipMemcpyAsync(&bad, ctx->d_flags, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    LMH_HIP(hipStreamSynchronize(ctx->stream));
    LMH_CHECK(!bad, LMH_ERR_INVALID, "internal: expanded DoF pair missing from the sparsity pattern");
  }

  if (fp64) {
    LMH_HIP(hipMalloc((void **)&A->vals, sizeof(double) * (nnz + LMH_PAD)));
    LMH_HIP(hipMemsetAsync(A->vals, 0, sizeof(double) * (nnz + LMH_PAD), ctx->stream));
  }

  // element colouring (host greedy); matrices assembled only by the gather
  // scatter skip it
  if (colour) {
    std::vector<int32_t> ed((size_t)ne * dpe);
    LMH_HIP(hipMemcpy(ed.data(), dm->edofs, sizeof(int32_t) * ed.size(), hipMemcpyDeviceToHost));
    std::vector<int32_t> celems;
    LMH_TRY(color_elements(ed, ne, dpe, n, A->color_ptr, celems, A->ncolors));
    LMH_HIP(hipMalloc((void **)&A->color_elems, sizeof(int32_t) * ne));
    LMH_HIP(hipMemcpy(A->color_elems, celems.data(), sizeof(int32_t) * ne, hipMemcpyHostToDevice));
  }
  LMH_HIP(hipStreamSynchronize(ctx->stream));
  LMH_TRY(matrix_build_windows(A.get()));
  *out = A.release();
  return LMH_OK;
}

}  // namespace lmh

extern "C" {

int lmh_matrix_create(lmh_context *ctx, const lmh_dofmap *dm, lmh_matrix **out) {
  return matrix_create_impl(ctx, dm, true, true, out);
}

int lmh_matrix_create_csr2(lmh_context *ctx, int64_t rows, int64_t cols, const int32_t *row_ptr_h,
                           const int32_t *col_idx_h, lmh_matrix **out) {
  LMH_CHECK(ctx && row_ptr_h && out && rows > 0 && cols > 0 && cols <= INT32_MAX, LMH_ERR_INVALID, "bad argument");
  LMH_CHECK(row_ptr_h[0] == 0, LMH_ERR_INVALID, "row_ptr[0] must be 0");
  const int64_t nnz = row_ptr_h[rows];
  LMH_CHECK(nnz >= 0 && (nnz == 0 || col_idx_h), LMH_ERR_INVALID, "bad nnz");
  std::vector<int32_t> rp(row_ptr_h, row_ptr_h + rows + 1);
  int max_row = 0;
  for (int64_t r = 0; r < rows; ++r) {
    LMH_CHECK(rp[r + 1] >= rp[r], LMH_ERR_INVALID, "row_ptr not monotone at row %lld", (long long)r);
    max_row = std::max<int>(max_row, rp[r + 1] - rp[r]);
    for (int32_t k = rp[r]; k < rp[r + 1]; ++k) {
      LMH_CHECK(col_idx_h[k] >= 0 && col_idx_h[k] < cols, LMH_ERR_INVALID, "column index out of range");
      LMH_CHECK(k == rp[r] || col_idx_h[k] > col_idx_h[k - 1], LMH_ERR_INVALID,
                "columns of row %lld not strictly increasing", (long long)r);
    }
  }
  LMH_HIP(hipSetDevice(ctx->device));
  std::unique_ptr<lmh_matrix> A(new lmh_matrix);
  A->ctx = ctx;
  A->rows = rows;
  A->cols = cols;
  A->nnz = nnz;
  A->max_row = max_row;
  LMH_HIP(hipMalloc((void **)&A->rp, sizeof(int32_t) * (rows + 1)));
  LMH_HIP(hipMalloc((void **)&A->ci, sizeof(int32_t) * (nnz + LMH_PAD)));
  LMH_HIP(hipMalloc((void **)&A->vals, sizeof(double) * (nnz + LMH_PAD)));
  LMH_HIP(hipMemset(A->ci, 0, sizeof(int32_t) * (nnz + LMH_PAD)));
  LMH_HIP(hipMemcpy(A->rp, rp.data(), sizeof(int32_t) * (rows + 1), hipMemcpyHostToDevice));
  if (nnz) LMH_HIP(hipMemcpy(A->ci, col_idx_h, sizeof(int32_t) * nnz, hipMemcpyHostToDevice));
  LMH_HIP(hipMemset(A->vals, 0, sizeof(double) * (nnz + LMH_PAD)));
  LMH_TRY(build_row_blocks(A.get(), rp));
  LMH_TRY(matrix_build_windows(A.get()));
  *out = A.release();
  return LMH_OK;
}

int lmh_matrix_create_csr(lmh_context *ctx, int64_t rows, const int32_t *row_ptr_h, const int32_t *col_idx_h,
                          lmh_matrix **out) {
  return lmh_matrix_create_csr2(ctx, rows, rows, row_ptr_h, col_idx_h, out);
}

int lmh_matrix_destroy(lmh_matrix *A) {
  if (!A) return LMH_OK;
  (void)hipStreamSynchronize(A->ctx->stream);
  (void)hipFree(A->rp);
  (void)hipFree(A->ci);
  (void)hipFree(A->rb);
  (void)hipFree(A->widx16);
  (void)hipFree(A->widx8);
  (void)hipFree(A->rbw);
  (void)hipFree(A->woff16);
  (void)hipFree(A->rbmeta);
  (void)hipFree(A->vals);
  (void)hipFree(A->vals32);
  (void)hipFree(A->tpos);
  (void)hipFree(A->vals_t);
  (void)hipFree(A->off);
  (void)hipFree(A->diag);
  (void)hipFree(A->color_elems);
  (void)hipFree(A->Ke);
  (void)hipFree(A->Fe);
  (void)hipFree(A->Ke32);
  (void)hipFree(A->offx);
  (void)hipFree(A->Kx);
  (void)hipFree(A->Fx);
  (void)hipFree(A->Kx32);
  (void)hipFree(A->w64);
  (void)hipFree(A->wbcg);
  (void)hipFree(A->w32);
  delete A;
  return LMH_OK;
}

int64_t lmh_matrix_rows(const lmh_matrix *A) { return A ? A->rows : -1; }
int64_t lmh_matrix_cols(const lmh_matrix *A) { return A ? A->cols : -1; }
int64_t lmh_matrix_nnz(const lmh_matrix *A) { return A ? A->nnz : -1; }
int lmh_matrix_n_colors(const lmh_matrix *A) { return A ? A->ncolors : -1; }

int lmh_matrix_get_pattern(const lmh_matrix *A, int32_t *row_ptr_h, int32_t *col_idx_h) {
  LMH_CHECK(A && row_ptr_h && col_idx_h, LMH_ERR_INVALID, "null argument");
  LMH_HIP(hipMemcpyAsync(row_ptr_h, A->rp, sizeof(int32_t) * (A->rows + 1), hipMemcpyDeviceToHost, A->ctx->stream));
  LMH_HIP(hipMemcpyAsync(col_idx_h, A->ci, sizeof(int32_t) * A->nnz, hipMemcpyDeviceToHost, A->ctx->stream));
  LMH_HIP(hipStreamSynchronize(A->ctx->stream));
  return LMH_OK;
}

int lmh_matrix_device_arrays(lmh_matrix *A, const int32_t **row_ptr_d, const int32_t **col_idx_d, double **vals_d) {
  LMH_CHECK(A, LMH_ERR_INVALID, "null matrix");
  if (row_ptr_d) *row_ptr_d = A->rp;
  if (col_idx_d) *col_idx_d = A->ci;
  if (vals_d) *vals_d = A->vals;
  return LMH_OK;
}

int lmh_matrix_zero(lmh_matrix *A) {
  LMH_CHECK(A, LMH_ERR_INVALID, "null matrix");
  if (!A->vals) {  // FP32-only matrix
    LMH_HIP(hipMemsetAsync(A->vals32, 0, sizeof(float) * A->nnz, A->ctx->stream));
    return LMH_OK;
  }
  LMH_HIP(hipMemsetAsync(A->vals, 0, sizeof(double) * A->nnz, A->ctx->stream));
  A->have32 = false;
  return LMH_OK;
}

int lmh_matrix_get_values(const lmh_matrix *A, double *vals_h) {
  LMH_CHECK(A && vals_h, LMH_ERR_INVALID, "null argument");
  LMH_CHECK(A->vals, LMH_ERR_UNSUPPORTED, "matrix holds FP32 values only");
  LMH_HIP(hipMemcpyAsync(vals_h, A->vals, sizeof(double) * A->nnz, hipMemcpyDeviceToHost, A->ctx->stream));
  LMH_HIP(hipStreamSynchronize(A->ctx->stream));
  return LMH_OK;
}

int lmh_matrix_set_values(lmh_matrix *A, const double *vals_h) {
  LMH_CHECK(A && vals_h, LMH_ERR_INVALID, "null argument");
  LMH_CHECK(A->vals, LMH_ERR_UNSUPPORTED, "matrix holds FP32 values only");
  LMH_HIP(hipMemcpyAsync(A->vals, vals_h, sizeof(double) * A->nnz, hipMemcpyHostToDevice, A->ctx->stream));
  LMH_HIP(hipStreamSynchronize(A->ctx->stream));
  A->have32 = false;
  return LMH_OK;
}

}  // extern "C"
