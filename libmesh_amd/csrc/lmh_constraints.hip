// lmh_constraints.hip -- general constraint rows on the assembly path.
//
// DofMap::add_constraint_row (src/base/dof_map_constraints.C:2180-2208) stores
// x_d = sum_j c_dj x_j + value_d; process_constraints (:4341-4497) substitutes
// constrained DoFs inside rows until rows refer to unconstrained DoFs only;
// assembly applies DofMap::heterogeneously_constrain_element_matrix_and_vector
// (:2518-2616) with the element constraint matrix of
// build_constraint_matrix_and_vector (:3545-3660); the sparsity pattern couples
// each element's DoFs with the DoFs they are constrained by
// (find_connected_dofs, dof_map.C:2915-2972, used by sparsity_pattern.C:139);
// enforce_constraints_exactly (:3088-3170) sets constrained entries of a vector.
//
// Split of the work: rows are processed on the host once per constraint set
// (map algebra over a few thousand rows).  Empty rows with value 0 (homogeneous
// Dirichlet) stay with the element kernels' mask.  Every element holding a DoF
// with a non-empty row or a nonzero value is a general element: one workgroup
// per general element forms C^T K C and C^T (F - K H) on its expanded DoF list
// in FP64 from the element kernel's output, writes them to an expanded buffer
// and zeroes the element's own block, so that the ordinary scatter adds exact
// zeros for it; a row-gather pass then adds the expanded matrices into their
// rows in general-element order (deterministic).
#include <algorithm>
#include <climits>
#include <map>
#include <set>

#include "lmh_internal.h"

namespace lmh {

constexpr int kMaxExpanded = 64;  // expanded DoFs per general element (one lane each)
constexpr int kMaxGeneralDpe = 32;

using Row = std::map<int32_t, double>;

// DofMap::check_for_constraint_loops (dof_map_constraints.C:4509-4597):
// process_constraints' substitution on a copy, failing when a row being
// substituted refers to the DoF it constrains.
static bool has_constraint_loop(std::map<int32_t, Row> rows) {
  std::set<int32_t> todo;
  for (const auto &r : rows) todo.insert(r.first);
  while (!todo.empty())
    for (auto i = todo.begin(); i != todo.end();) {
      Row &row = rows[*i];
      std::vector<int32_t> expand;
      for (const auto &it : row)
        if (it.first != *i && rows.count(it.first)) {
          todo.insert(it.first);
          expand.push_back(it.first);
        }
      for (int32_t x : expand) {
        const double c = row[x];
        for (const auto &it : rows.at(x)) {
          if (it.first == x) return true;
          row[it.first] += it.second * c;
        }
        row.erase(x);
      }
      if (expand.empty())
        i = todo.erase(i);
      else
        ++i;
    }
  return false;
}

// DofMap::process_constraints (dof_map_constraints.C:4341-4497), same
// substitution order (ascending DoF sweeps until no row refers to a
// constrained DoF other than itself); values that become 0 are dropped.
static void process_rows(std::map<int32_t, Row> &rows, std::map<int32_t, double> &vals) {
  std::set<int32_t> todo;
  for (const auto &r : rows) todo.insert(r.first);
  while (!todo.empty())
    for (auto i = todo.begin(); i != todo.end();) {
      Row &row = rows[*i];
      auto vit = vals.find(*i);
      double rhs = vit == vals.end() ? 0.0 : vit->second;
      std::vector<int32_t> expand;
      for (const auto &it : row)
        if (it.first != *i && rows.count(it.first)) {
          todo.insert(it.first);
          expand.push_back(it.first);
        }
      for (int32_t x : expand) {
        const double c = row[x];
        for (const auto &it : rows.at(x)) row[it.first] += it.second * c;
        auto sub = vals.find(x);
        if (sub != vals.end()) rhs += sub->second * c;
        row.erase(x);
      }
      if (rhs != 0.0)
        vals[*i] = rhs;
      else
        vals.erase(*i);
      if (expand.empty())
        i = todo.erase(i);
      else
        ++i;
    }
}

template <typename V>
static int upload(lmh_context *ctx, V **dst, const std::vector<V> &src) {
  LMH_HIP(hipMalloc((void **)dst, sizeof(V) * std::max<size_t>(1, src.size())));
  if (!src.empty()) LMH_HIP(hipMemcpyAsync(*dst, src.data(), sizeof(V) * src.size(), hipMemcpyHostToDevice, ctx->stream));
  return LMH_OK;
}

int release_constraints(lmh_constraints &c) {
  for (void *p : {(void *)c.d_dof, (void *)c.d_ptr, (void *)c.d_col, (void *)c.d_coef, (void *)c.d_val,
                  (void *)c.ge_elem, (void *)c.ge_m, (void *)c.ge_dofs, (void *)c.ge_flag, (void *)c.ge_val,
                  (void *)c.ge_cptr, (void *)c.ge_ci, (void *)c.ge_cv, (void *)c.xip, (void *)c.xinc,
                  (void *)c.xrows})
    (void)hipFree(p);
  const int64_t version = c.version;
  c = lmh_constraints();
  c.version = version;
  return LMH_OK;
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------
struct GeView {
  const int32_t *elem, *m;
  const uint8_t *flag;
  const double *val;
  const int32_t *cptr, *ci;
  const double *cv;
  int64_t n_ge;
  int n, mx, asym;
};

// One workgroup per general element g (element e, n own DoFs, m expanded):
//   G = F - K H, KC = K C, K' = C^T (K C), F' = C^T G;
//   a constrained row a of K' becomes e_a (and -c_aj at the expanded positions
//   of its row's DoFs when asymmetric), F'(a) its value (0 when symmetric).
// K' and F' go to Kx[g], Fx[g] (mx-strided); the element's block of Ke/Fe is
// zeroed afterwards.  C is stored by expanded column: entries (i, c_ia).
template <typename T>
__global__ __launch_bounds__(256) void k_constrain_general(GeView v, T *__restrict__ Ke, double *__restrict__ Fe,
                                                           T *__restrict__ Kx, double *__restrict__ Fx) {
  extern __shared__ double csm[];
  const int n = v.n, mx = v.mx, tid = threadIdx.x;
  double *K = csm, *KC = csm + n * n, *G = KC + n * mx;
  for (int64_t g = blockIdx.x; g < v.n_ge; g += gridDim.x) {
    const int64_t e = v.elem[g];
    const int m = v.m[g];
    T *Kel = Ke + e * n * n;
    double *Fel = Fe + e * n;
    const uint8_t *flag = v.flag + g * n;
    const double *H = v.val + g * n;
    const int32_t *cp = v.cptr + g * (mx + 1);
    for (int t = tid; t < n * n; t += blockDim.x) K[t] = (double)Kel[t];
    __syncthreads();
    for (int i = tid; i < n; i += blockDim.x) {
      double kh = 0.0;
      for (int j = 0; j < n; ++j) kh += K[i * n + j] * H[j];
      G[i] = Fel[i] - kh;
    }
    for (int t = tid; t < n * m; t += blockDim.x) {
      const int i = t / m, b = t - i * m;
      double acc = 0.0;
      for (int32_t k = cp[b]; k < cp[b + 1]; ++k) acc += K[i * n + v.ci[k]] * v.cv[k];
      KC[i * mx + b] = acc;
    }
    __syncthreads();
    for (int t = tid; t < n * n; t += blockDim.x) Kel[t] = (T)0;
    for (int t = tid; t < n; t += blockDim.x) Fel[t] = 0.0;
    T *Kg = Kx + g * mx * mx;
    for (int t = tid; t < m * m; t += blockDim.x) {
      const int a = t / m, b = t - a * m;
      double acc = 0.0;
      if (a < n && flag[a]) {
        acc = (a == b) ? 1.0 : 0.0;
        if (v.asym)
          for (int32_t k = cp[b]; k < cp[b + 1]; ++k)
            if (v.ci[k] == a) acc = -v.cv[k];
      } else {
        for (int32_t k = cp[a]; k < cp[a + 1]; ++k) acc += v.cv[k] * KC[v.ci[k] * mx + b];
      }
      Kg[a * mx + b] = (T)acc;
    }
    for (int a = tid; a < m; a += blockDim.x) {
      double acc = 0.0;
      if (a < n && flag[a]) {
        acc = v.asym ? H[a] : 0.0;
      } else {
        for (int32_t k = cp[a]; k < cp[a + 1]; ++k) acc += v.cv[k] * G[v.ci[k]];
      }
      Fx[g * mx + a] = acc;
    }
    __syncthreads();
  }
}

// Rows touched by general elements: one wave per row adds the expanded
// entries of each (general element, position) in ascending order into the row
// (staged in LDS), and their F' into rhs.
template <typename T>
__global__ __launch_bounds__(64) void k_scatter_extra(const int32_t *__restrict__ xrows, int64_t n_xrows,
                                                      const int32_t *__restrict__ xip, const int32_t *__restrict__ xinc,
                                                      const int32_t *__restrict__ gm, int mx,
                                                      const int32_t *__restrict__ rp, const uint16_t *__restrict__ offx,
                                                      const T *__restrict__ Kx, const double *__restrict__ Fx,
                                                      T *__restrict__ vals, double *__restrict__ rhs) {
  extern __shared__ __align__(16) unsigned char xsm[];
  T *acc = (T *)xsm;
  const int lane = threadIdx.x;
  for (int64_t k = blockIdx.x; k < n_xrows; k += gridDim.x) {
    const int32_t r = xrows[k];
    const int32_t b0 = rp[r], len = rp[r + 1] - b0;
    for (int t = lane; t < len; t += 64) acc[t] = vals[b0 + t];
    __syncthreads();
    double f = 0.0;
    for (int32_t q = xip[r]; q < xip[r + 1]; ++q) {
      const int32_t ga = xinc[q];
      const int m = gm[ga / mx];
      const int64_t base = (int64_t)ga * mx;
      if (lane < m) acc[offx[base + lane]] += Kx[base + lane];
      f += Fx[ga];
      __syncthreads();
    }
    for (int t = lane; t < len; t += 64) vals[b0 + t] = acc[t];
    if (rhs && lane == 0) rhs[r] += f;
    __syncthreads();
  }
}

// x_d = value_d (unless homogeneous) + sum_j c_dj x_j, entries ascending.
// Rows refer to unconstrained DoFs only, so updating in place reads the
// vector as it was on entry.
__global__ void k_enforce(const int32_t *__restrict__ dof, const int32_t *__restrict__ ptr,
                          const int32_t *__restrict__ col, const double *__restrict__ coef,
                          const double *__restrict__ val, int64_t n_rows, int homogeneous, double *__restrict__ x) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n_rows; r += (int64_t)gridDim.x * blockDim.x) {
    double t = homogeneous ? 0.0 : val[r];
    for (int32_t k = ptr[r]; k < ptr[r + 1]; ++k) t += coef[k] * x[col[k]];
    x[dof[r]] = t;
  }
}

// ---------------------------------------------------------------------------
// Assembly hooks
// ---------------------------------------------------------------------------
template <typename T>
int constrain_general_impl(lmh_matrix *A, T *Ke, double *Fe) {
  const lmh_dofmap *dm = A->dm;
  const lmh_constraints &c = dm->cr;
  if (c.n_ge == 0) return LMH_OK;
  LMH_CHECK(A->cons_version == c.version && A->offx, LMH_ERR_INVALID,
            "the constraints changed after the matrix was created: create the matrix again");
  lmh_context *ctx = A->ctx;
  const int64_t kx = c.n_ge * c.mx * c.mx;
  T *Kx;
  if (sizeof(T) == 8) {
    if (!A->Kx) LMH_HIP(hipMalloc((void **)&A->Kx, sizeof(double) * kx));
    Kx = (T *)A->Kx;
  } else {
    if (!A->Kx32) LMH_HIP(hipMalloc((void **)&A->Kx32, sizeof(float) * kx));
    Kx = (T *)A->Kx32;
  }
  if (!A->Fx) LMH_HIP(hipMalloc((void **)&A->Fx, sizeof(double) * c.n_ge * c.mx));
  GeView v{c.ge_elem, c.ge_m, c.ge_flag, c.ge_val, c.ge_cptr, c.ge_ci, c.ge_cv, c.n_ge, dm->dpe, c.mx,
           c.asymmetric ? 1 : 0};
  const size_t shm = sizeof(double) * ((size_t)dm->dpe * dm->dpe + (size_t)dm->dpe * c.mx + dm->dpe);
  const int g = (int)std::min<int64_t>(c.n_ge, (int64_t)ctx->cu_count * 8);
  LMH_LAUNCH(ctx, sizeof(T) == 8 ? "constrain_general_fp64" : "constrain_general_fp32", k_constrain_general<T>,
             dim3(g), dim3(256), shm, v, Ke, Fe, Kx, A->Fx);
  return check_launch("k_constrain_general");
}

template <typename T>
int scatter_extra_impl(lmh_matrix *A, T *vals, double *rhs) {
  const lmh_constraints &c = A->dm->cr;
  if (c.n_ge == 0) return LMH_OK;
  lmh_context *ctx = A->ctx;
  const T *Kx = sizeof(T) == 8 ? (const T *)A->Kx : (const T *)A->Kx32;
  const int g = (int)std::min<int64_t>(c.n_xrows, (int64_t)ctx->cu_count * 64);
  LMH_LAUNCH(ctx, sizeof(T) == 8 ? "scatter_expanded_fp64" : "scatter_expanded_fp32", k_scatter_extra<T>, dim3(g),
             dim3(64), sizeof(T) * A->max_row, c.xrows, c.n_xrows, c.xip, c.xinc, c.ge_m, c.mx, A->rp, A->offx, Kx,
             A->Fx, vals, rhs);
  return check_launch("k_scatter_extra");
}

template int constrain_general_impl<double>(lmh_matrix *, double *, double *);
template int constrain_general_impl<float>(lmh_matrix *, float *, double *);
template int scatter_extra_impl<double>(lmh_matrix *, double *, double *);
template int scatter_extra_impl<float>(lmh_matrix *, float *, double *);

}  // namespace lmh

using namespace lmh;

extern "C" {

int lmh_dofmap_set_constraints(lmh_dofmap *dm, int64_t n_rows, const int32_t *dof_h, const int32_t *row_ptr_h,
                               const int32_t *cols_h, const double *coefs_h, const double *rhs_h, int flags) {
  LMH_CHECK(dm && n_rows >= 0, LMH_ERR_INVALID, "bad argument");
  LMH_CHECK(n_rows == 0 || (dof_h && row_ptr_h), LMH_ERR_INVALID, "null argument");
  LMH_CHECK((flags & ~(LMH_CONSTRAINT_FORBID_OVERWRITE | LMH_CONSTRAINT_ERROR_ON_LOOP |
                       LMH_CONSTRAINT_SYMMETRIC_ROWS)) == 0,
            LMH_ERR_INVALID, "unknown constraint flags 0x%x", flags);
  lmh_context *ctx = dm->ctx;
  const int64_t n = dm->n_dofs, ne = dm->mesh->n_elem;
  const int dpe = dm->dpe;
  if (n_rows == 0) {  // the rows of creation alone: empty, value 0, nothing to process
    LMH_HIP(hipStreamSynchronize(ctx->stream));
    LMH_TRY(release_constraints(dm->cr));
    lmh_constraints &c = dm->cr;
    c.version++;
    c.asymmetric = !(flags & LMH_CONSTRAINT_SYMMETRIC_ROWS);
    for (int64_t d = 0; d < n; ++d)
      if (dm->h_cons0[d]) c.dof.push_back((int32_t)d);
    c.ptr.assign(c.dof.size() + 1, 0);
    c.val.assign(c.dof.size(), 0.0);
    LMH_TRY(upload(ctx, &c.d_dof, c.dof));
    LMH_TRY(upload(ctx, &c.d_ptr, c.ptr));
    LMH_TRY(upload(ctx, &c.d_col, c.col));
    LMH_TRY(upload(ctx, &c.d_coef, c.coef));
    LMH_TRY(upload(ctx, &c.d_val, c.val));
    LMH_HIP(hipMemcpyAsync(dm->cons, dm->h_cons0.data(), n, hipMemcpyHostToDevice, ctx->stream));
    LMH_HIP(hipStreamSynchronize(ctx->stream));
    return LMH_OK;
  }
  // DofMap::add_constraint_row on top of the Dirichlet rows of creation
  std::map<int32_t, Row> rows;
  std::map<int32_t, double> vals;
  for (int64_t d = 0; d < n; ++d)
    if (dm->h_cons0[d]) rows[(int32_t)d];
  for (int64_t k = 0; k < n_rows; ++k) {
    const int32_t d = dof_h[k];
    LMH_CHECK(d >= 0 && d < n, LMH_ERR_INVALID, "constrained DoF %d out of range", d);
    LMH_CHECK(!(flags & LMH_CONSTRAINT_FORBID_OVERWRITE) || !rows.count(d), LMH_ERR_INVALID,
              "DOF %d was already constrained", d);
    LMH_CHECK(row_ptr_h[k + 1] >= row_ptr_h[k] && (row_ptr_h[k + 1] == row_ptr_h[k] || (cols_h && coefs_h)),
              LMH_ERR_INVALID, "bad constraint row %lld", (long long)k);
    Row row;
    for (int32_t t = row_ptr_h[k]; t < row_ptr_h[k + 1]; ++t) {
      LMH_CHECK(cols_h[t] >= 0 && cols_h[t] < n, LMH_ERR_INVALID, "constraint entry DoF %d out of range", cols_h[t]);
      LMH_CHECK(cols_h[t] != d, LMH_ERR_INVALID, "constraint row for DoF %d must not contain an entry for itself", d);
      row[cols_h[t]] = coefs_h[t];
    }
    rows[d] = row;
    vals[d] = rhs_h ? rhs_h[k] : 0.0;
  }
  LMH_CHECK(!(flags & LMH_CONSTRAINT_ERROR_ON_LOOP) || !has_constraint_loop(rows), LMH_ERR_INVALID,
            "Constraint loop detected");
  process_rows(rows, vals);
  for (const auto &r : rows)
    for (const auto &it : r.second)
      LMH_CHECK(!rows.count(it.first), LMH_ERR_UNSUPPORTED,
                "Constraint loop detected: DoF %d stays constrained in terms of constrained DoF %d", r.first, it.first);

  LMH_HIP(hipStreamSynchronize(ctx->stream));
  LMH_TRY(release_constraints(dm->cr));
  lmh_constraints &c = dm->cr;
  c.version++;
  c.asymmetric = !(flags & LMH_CONSTRAINT_SYMMETRIC_ROWS);
  std::vector<uint8_t> general(n, 0), mask(n, 0);
  c.ptr.push_back(0);
  for (const auto &r : rows) {
    auto v = vals.find(r.first);
    const double value = v == vals.end() ? 0.0 : v->second;
    c.dof.push_back(r.first);
    for (const auto &it : r.second) {
      c.col.push_back(it.first);
      c.coef.push_back(it.second);
    }
    c.ptr.push_back((int32_t)c.col.size());
    c.val.push_back(value);
    if (!r.second.empty() || value != 0.0) {
      general[r.first] = 1;
      c.n_general++;
    } else {
      mask[r.first] = 1;
    }
  }
  LMH_TRY(upload(ctx, &c.d_dof, c.dof));
  LMH_TRY(upload(ctx, &c.d_ptr, c.ptr));
  LMH_TRY(upload(ctx, &c.d_col, c.col));
  LMH_TRY(upload(ctx, &c.d_coef, c.coef));
  LMH_TRY(upload(ctx, &c.d_val, c.val));
  LMH_HIP(hipMemcpyAsync(dm->cons, mask.data(), n, hipMemcpyHostToDevice, ctx->stream));

  if (c.n_general > 0) {
    LMH_CHECK(dpe <= kMaxGeneralDpe, LMH_ERR_UNSUPPORTED,
              "general constraint rows need at most %d DoFs per element (have %d)", kMaxGeneralDpe, dpe);
    std::vector<int32_t> ed((size_t)ne * dpe);
    LMH_HIP(hipMemcpy(ed.data(), dm->edofs, sizeof(int32_t) * ed.size(), hipMemcpyDeviceToHost));
    // build_constraint_matrix_and_vector: own DoFs, then the DoFs their rows
    // refer to that are not among them, ascending
    std::vector<int32_t> gelem;
    std::vector<std::vector<int32_t>> X;
    for (int64_t e = 0; e < ne; ++e) {
      const int32_t *d = &ed[(size_t)e * dpe];
      bool any = false;
      for (int i = 0; i < dpe; ++i) any |= general[d[i]] != 0;
      if (!any) continue;
      std::set<int32_t> add;
      for (int i = 0; i < dpe; ++i)
        if (general[d[i]])
          for (const auto &it : rows.at(d[i])) add.insert(it.first);
      for (int i = 0; i < dpe; ++i) add.erase(d[i]);
      std::vector<int32_t> x(d, d + dpe);
      x.insert(x.end(), add.begin(), add.end());
      LMH_CHECK((int)x.size() <= kMaxExpanded, LMH_ERR_UNSUPPORTED,
                "element %lld couples %d DoFs through its constraints (at most %d)", (long long)e, (int)x.size(),
                kMaxExpanded);
      gelem.push_back((int32_t)e);
      X.push_back(std::move(x));
    }
    c.n_ge = (int64_t)gelem.size();
    int mx = dpe;
    for (const auto &x : X) mx = std::max<int>(mx, (int)x.size());
    c.mx = mx;
    LMH_CHECK(c.n_ge * mx * mx < INT32_MAX, LMH_ERR_UNSUPPORTED, "too many general elements");
    std::vector<int32_t> gm(c.n_ge), gd((size_t)c.n_ge * mx, -1), cptr((size_t)c.n_ge * (mx + 1)), ci;
    std::vector<uint8_t> gf((size_t)c.n_ge * dpe);
    std::vector<double> gv((size_t)c.n_ge * dpe), cv;
    std::vector<int32_t> xcnt(n + 1, 0);
    for (int64_t g = 0; g < c.n_ge; ++g) {
      const std::vector<int32_t> &x = X[g];
      const int m = (int)x.size();
      gm[g] = m;
      std::copy(x.begin(), x.end(), gd.begin() + g * mx);
      std::map<int32_t, int> pos;
      for (int a = 0; a < m; ++a) pos[x[a]] = a;
      std::vector<std::vector<std::pair<int, double>>> cols(m);
      for (int i = 0; i < dpe; ++i) {
        const int32_t d = x[i];
        auto r = rows.find(d);
        gf[g * dpe + i] = r != rows.end();
        auto v = vals.find(d);
        gv[g * dpe + i] = v == vals.end() ? 0.0 : v->second;
        if (r == rows.end())
          cols[i].push_back({i, 1.0});
        else
          for (const auto &it : r->second) cols[pos.at(it.first)].push_back({i, it.second});
      }
      for (int b = 0; b < m; ++b) {
        cptr[g * (mx + 1) + b] = (int32_t)ci.size();
        for (const auto &p : cols[b]) {
          ci.push_back(p.first);
          cv.push_back(p.second);
        }
      }
      for (int b = m; b <= mx; ++b) cptr[g * (mx + 1) + b] = (int32_t)ci.size();
      for (int a = 0; a < m; ++a) xcnt[x[a] + 1]++;
    }
    for (int64_t d = 0; d < n; ++d) xcnt[d + 1] += xcnt[d];
    std::vector<int32_t> xinc(xcnt[n]), fill(xcnt.begin(), xcnt.end() - 1), xrows;
    for (int64_t g = 0; g < c.n_ge; ++g)
      for (int a = 0; a < gm[g]; ++a) xinc[fill[X[g][a]]++] = (int32_t)(g * mx + a);
    for (int64_t d = 0; d < n; ++d)
      if (xcnt[d + 1] > xcnt[d]) xrows.push_back((int32_t)d);
    c.h_xip = xcnt;
    c.n_xrows = (int64_t)xrows.size();
    LMH_TRY(upload(ctx, &c.ge_elem, gelem));
    LMH_TRY(upload(ctx, &c.ge_m, gm));
    LMH_TRY(upload(ctx, &c.ge_dofs, gd));
    LMH_TRY(upload(ctx, &c.ge_flag, gf));
    LMH_TRY(upload(ctx, &c.ge_val, gv));
    LMH_TRY(upload(ctx, &c.ge_cptr, cptr));
    LMH_TRY(upload(ctx, &c.ge_ci, ci));
    LMH_TRY(upload(ctx, &c.ge_cv, cv));
    LMH_TRY(upload(ctx, &c.xip, xcnt));
    LMH_TRY(upload(ctx, &c.xinc, xinc));
    LMH_TRY(upload(ctx, &c.xrows, xrows));
  }
  LMH_HIP(hipStreamSynchronize(ctx->stream));
  return LMH_OK;
}

int lmh_dofmap_constraint_counts(const lmh_dofmap *dm, int64_t *n_rows, int64_t *n_entries, int64_t *n_general_elems) {
  LMH_CHECK(dm, LMH_ERR_INVALID, "null argument");
  const lmh_constraints &c = dm->cr;
  if (n_rows) *n_rows = (int64_t)c.dof.size();
  if (n_entries) *n_entries = (int64_t)c.col.size();
  if (n_general_elems) *n_general_elems = c.n_ge;
  return LMH_OK;
}

int lmh_dofmap_get_constraints(const lmh_dofmap *dm, int32_t *dof_h, int32_t *row_ptr_h, int32_t *cols_h,
                               double *coefs_h, double *rhs_h) {
  LMH_CHECK(dm && dof_h && row_ptr_h, LMH_ERR_INVALID, "null argument");
  const lmh_constraints &c = dm->cr;
  std::copy(c.dof.begin(), c.dof.end(), dof_h);
  std::copy(c.ptr.begin(), c.ptr.end(), row_ptr_h);
  if (cols_h) std::copy(c.col.begin(), c.col.end(), cols_h);
  if (coefs_h) std::copy(c.coef.begin(), c.coef.end(), coefs_h);
  if (rhs_h) std::copy(c.val.begin(), c.val.end(), rhs_h);
  return LMH_OK;
}

int lmh_enforce_constraints(const lmh_dofmap *dm, double *x_d, int homogeneous) {
  LMH_CHECK(dm && x_d, LMH_ERR_INVALID, "null argument");
  const lmh_constraints &c = dm->cr;
  lmh_context *ctx = dm->ctx;
  const int64_t nr = (int64_t)c.dof.size();
  if (nr == 0) return LMH_OK;
  LMH_LAUNCH(ctx, "enforce_constraints", k_enforce, dim3(grid_for(nr, 256, 4096)), dim3(256), 0, c.d_dof, c.d_ptr,
             c.d_col, c.d_coef, c.d_val, nr, homogeneous ? 1 : 0, x_d);
  LMH_TRY(check_launch("k_enforce"));
  LMH_HIP(hipStreamSynchronize(ctx->stream));
  return LMH_OK;
}

}  // extern "C"
