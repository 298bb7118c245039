// lmh_core.hip -- context, memory, errors, profiling, FE tables, mesh and DoF map.
#include <cstring>
#include <rocprim/rocprim.hpp>

#include <cmath>
#include <cstdio>
#include <cstring>

#include "lmh_internal.h"

namespace lmh {

static thread_local std::string g_err;

int set_error(int code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

int check_launch(const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(LMH_ERR_HIP, "launch of %s failed: %s", what, hipGetErrorString(e));
  return LMH_OK;
}

hipEvent_t get_event(lmh_context *ctx) {
  if (!ctx->event_pool.empty()) {
    hipEvent_t e = ctx->event_pool.back();
    ctx->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

ProfScope::ProfScope(lmh_context *c, const char *n) : ctx(c), name(n) {
  if (ctx && ctx->prof_on) {
    a = get_event(ctx);
    b = get_event(ctx);
    (void)hipEventRecord(a, ctx->stream);
  }
}
ProfScope::~ProfScope() {
  if (a) {
    (void)hipEventRecord(b, ctx->stream);
    ctx->prof[name].pending.emplace_back(a, b);
  }
}

// ---------------------------------------------------------------------------
// FE tables: QGauss points/weights and Lagrange reference shapes.
// QGauss 1D: 2 points +-sqrt(3)/3 (w 1), 3 points +-0.7745966692414834, 0
// (w 5/9, 8/9); hex rules are tensor products with the first coordinate
// fastest (quadrature_gauss_1D.C:30-80, quadrature.C:283-305).  Tet rules:
// order 3 -> 5 points with a negative centre weight, order 5 -> 14 points
// (quadrature_gauss_3D.C:55-260).  Shapes: 1D linear/quadratic Lagrange
// tensor products with the libMesh node-to-index tables for HEX8/HEX27
// (fe_lagrange_shape_3D.C:277-279, 348-350) and barycentric TET4/TET10 shapes
// (fe_lagrange_shape_3D.C:715-740, 957-1005).
// ---------------------------------------------------------------------------
int nodes_per_elem(int type) {
  switch (type) {
    case LMH_HEX8: return 8;
    case LMH_HEX27: return 27;
    case LMH_TET4: return 4;
    case LMH_TET10: return 10;
  }
  return 0;
}

namespace {
const int kIdx8[3][8] = {{0, 1, 1, 0, 0, 1, 1, 0}, {0, 0, 1, 1, 0, 0, 1, 1}, {0, 0, 0, 0, 1, 1, 1, 1}};
const int kIdx27[3][27] = {
    {0, 1, 1, 0, 0, 1, 1, 0, 2, 1, 2, 0, 0, 1, 1, 0, 2, 1, 2, 0, 2, 2, 1, 2, 0, 2, 2},
    {0, 0, 1, 1, 0, 0, 1, 1, 0, 2, 1, 2, 0, 0, 1, 1, 0, 2, 1, 2, 2, 0, 2, 1, 2, 2, 2},
    {0, 0, 0, 0, 1, 1, 1, 1, 0, 0, 0, 0, 2, 2, 2, 2, 1, 1, 1, 1, 0, 2, 2, 2, 2, 1, 2}};
const int kTetEdge[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3}, {2, 3}};

double lin(int i, double x) { return i == 0 ? 0.5 * (1.0 - x) : 0.5 * (1.0 + x); }
double dlin(int i) { return i == 0 ? -0.5 : 0.5; }
double quad(int i, double x) { return i == 0 ? 0.5 * x * (x - 1.0) : (i == 1 ? 0.5 * x * (x + 1.0) : 1.0 - x * x); }
double dquad(int i, double x) { return i == 0 ? x - 0.5 : (i == 1 ? x + 0.5 : -2.0 * x); }

void gauss1d(int np, double *x, double *w) {
  if (np == 2) {
    x[0] = -std::sqrt(3.0) / 3.0;
    x[1] = -x[0];
    w[0] = w[1] = 1.0;
  } else if (np == 3) {
    x[0] = -7.7459666924148337703585307995648e-01;
    x[1] = 0.0;
    x[2] = -x[0];
    w[0] = 5.0 / 9.0;
    w[1] = 8.0 / 9.0;
    w[2] = w[0];
  } else {
    x[0] = 0.0;
    w[0] = 2.0;
  }
}
}  // namespace

bool build_fe_table(int type, int arg, FETable &t) {
  t.type = type;
  t.arg = arg;
  t.n = nodes_per_elem(type);
  t.qp.clear();
  t.w.clear();
  if (type == LMH_HEX8 || type == LMH_HEX27) {
    if (arg < 1 || arg > 3) return false;
    double x[3], w[3];
    gauss1d(arg, x, w);
    for (int k = 0; k < arg; ++k)
      for (int j = 0; j < arg; ++j)
        for (int i = 0; i < arg; ++i) {
          t.qp.insert(t.qp.end(), {x[i], x[j], x[k]});
          t.w.push_back(w[i] * w[j] * w[k]);
        }
  } else if (type == LMH_TET4 || type == LMH_TET10) {
    auto add = [&](double a, double b, double c, double w) {
      t.qp.insert(t.qp.end(), {a, b, c});
      t.w.push_back(w);
    };
    if (arg <= 1) {
      add(0.25, 0.25, 0.25, 1.0 / 6.0);
    } else if (arg == 2) {
      const double b = 0.25 * (1.0 - std::sqrt(5.0) / 5.0), a = 1.0 - 3.0 * b;
      add(a, b, b, 1.0 / 24);
      add(b, a, b, 1.0 / 24);
      add(b, b, a, 1.0 / 24);
      add(b, b, b, 1.0 / 24);
    } else if (arg == 3) {
      add(0.25, 0.25, 0.25, -2.0 / 15.0);
      add(0.5, 1.0 / 6, 1.0 / 6, 0.075);
      add(1.0 / 6, 0.5, 1.0 / 6, 0.075);
      add(1.0 / 6, 1.0 / 6, 0.5, 0.075);
      add(1.0 / 6, 1.0 / 6, 1.0 / 6, 0.075);
    } else {
      const double a[3] = {0.31088591926330060980, 0.092735250310891226402, 0.045503704125649649492};
      const double wt[3] = {0.018781320953002641800, 0.012248840519393658257, 0.0070910034628469110730};
      for (int i = 0; i < 2; ++i) {
        const double b = 1.0 - 3.0 * a[i];
        add(a[i], a[i], a[i], wt[i]);
        add(a[i], b, a[i], wt[i]);
        add(b, a[i], a[i], wt[i]);
        add(a[i], a[i], b, wt[i]);
      }
      const double b = 0.5 * (1.0 - 2.0 * a[2]);
      add(b, b, a[2], wt[2]);
      add(b, a[2], a[2], wt[2]);
      add(a[2], a[2], b, wt[2]);
      add(a[2], b, a[2], wt[2]);
      add(b, a[2], b, wt[2]);
      add(a[2], b, b, wt[2]);
    }
  } else {
    return false;
  }
  t.nq = (int)t.w.size();
  t.phi.assign((size_t)t.n * t.nq, 0.0);
  t.dref.assign((size_t)t.n * t.nq * 3, 0.0);
  for (int q = 0; q < t.nq; ++q) {
    const double xi = t.qp[3 * q], eta = t.qp[3 * q + 1], zeta = t.qp[3 * q + 2];
    for (int a = 0; a < t.n; ++a) {
      double *d = &t.dref[((size_t)a * t.nq + q) * 3];
      double &p = t.phi[(size_t)a * t.nq + q];
      if (type == LMH_HEX8) {
        const int i = kIdx8[0][a], j = kIdx8[1][a], k = kIdx8[2][a];
        p = lin(i, xi) * lin(j, eta) * lin(k, zeta);
        d[0] = dlin(i) * lin(j, eta) * lin(k, zeta);
        d[1] = lin(i, xi) * dlin(j) * lin(k, zeta);
        d[2] = lin(i, xi) * lin(j, eta) * dlin(k);
      } else if (type == LMH_HEX27) {
        const int i = kIdx27[0][a], j = kIdx27[1][a], k = kIdx27[2][a];
        p = quad(i, xi) * quad(j, eta) * quad(k, zeta);
        d[0] = dquad(i, xi) * quad(j, eta) * quad(k, zeta);
        d[1] = quad(i, xi) * dquad(j, eta) * quad(k, zeta);
        d[2] = quad(i, xi) * quad(j, eta) * dquad(k, zeta);
      } else {
        const double L[4] = {1.0 - xi - eta - zeta, xi, eta, zeta};
        const double G[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
        if (type == LMH_TET4) {
          p = L[a];
          for (int c = 0; c < 3; ++c) d[c] = G[a][c];
        } else if (a < 4) {
          p = L[a] * (2.0 * L[a] - 1.0);
          for (int c = 0; c < 3; ++c) d[c] = (4.0 * L[a] - 1.0) * G[a][c];
        } else {
          const int u = kTetEdge[a - 4][0], v = kTetEdge[a - 4][1];
          p = 4.0 * L[u] * L[v];
          for (int c = 0; c < 3; ++c) d[c] = 4.0 * (G[u][c] * L[v] + L[u] * G[v][c]);
        }
      }
    }
  }
  return true;
}

int fe_table(lmh_context *ctx, int type, int arg, const FETable **out) {
  auto key = std::make_pair(type, arg);
  auto it = ctx->tables.find(key);
  if (it == ctx->tables.end()) {
    std::unique_ptr<FETable> t(new FETable);
    LMH_CHECK(build_fe_table(type, arg, *t), LMH_ERR_UNSUPPORTED, "no quadrature/shape table for type %d arg %d",
              type, arg);
    const size_t cnt = t->w.size() + t->phi.size() + t->dref.size();
    LMH_HIP(hipMalloc(&t->d_data, cnt * sizeof(double)));
    std::vector<double> flat;
    flat.reserve(cnt);
    flat.insert(flat.end(), t->w.begin(), t->w.end());
    flat.insert(flat.end(), t->phi.begin(), t->phi.end());
    flat.insert(flat.end(), t->dref.begin(), t->dref.end());
    LMH_HIP(hipMemcpy(t->d_data, flat.data(), cnt * sizeof(double), hipMemcpyHostToDevice));
    it = ctx->tables.emplace(key, std::move(t)).first;
  }
  *out = it->second.get();
  return LMH_OK;
}

// ---------------------------------------------------------------------------
// scans and incidence lists
// ---------------------------------------------------------------------------
int exclusive_scan_i32(lmh_context *ctx, const int32_t *in, int32_t *out, int64_t n, int64_t *total) {
  // 64-bit accumulation so overflow of the int32 result can be detected
  int64_t *tmp64 = nullptr;
  LMH_HIP(hipMallocAsync((void **)&tmp64, sizeof(int64_t) * (n + 1), ctx->stream));
  size_t bytes = 0;
  auto conv = rocprim::make_transform_iterator(in, [] __device__(int32_t v) { return (int64_t)v; });
  LMH_HIP(rocprim::exclusive_scan(nullptr, bytes, conv, tmp64, (int64_t)0, (size_t)n + 1, rocprim::plus<int64_t>(),
                                  ctx->stream));
  void *scratch = nullptr;
  LMH_HIP(hipMallocAsync(&scratch, bytes, ctx->stream));
  // the element at index n of `in` must be readable: callers pass n+1 entries with in[n] = 0
  LMH_HIP(rocprim::exclusive_scan(scratch, bytes, conv, tmp64, (int64_t)0, (size_t)n + 1, rocprim::plus<int64_t>(),
                                  ctx->stream));
  int64_t tot = 0;
  LMH_HIP(hipMemcpyAsync(&tot, tmp64 + n, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
  LMH_HIP(hipStreamSynchronize(ctx->stream));
  LMH_HIP(hipFreeAsync(scratch, ctx->stream));
  if (tot >= (int64_t)INT32_MAX) {
    (void)hipFreeAsync(tmp64, ctx->stream);
    return set_error(LMH_ERR_UNSUPPORTED, "count %lld exceeds the int32 index range", (long long)tot);
  }
  // narrow to int32
  auto narrow = rocprim::make_transform_iterator(tmp64, [] __device__(int64_t v) { return (int32_t)v; });
  LMH_HIP(rocprim::transform(narrow, out, (size_t)n + 1, [] __device__(int32_t v) { return v; }, ctx->stream));
  LMH_HIP(hipFreeAsync(tmp64, ctx->stream));
  *total = tot;
  return LMH_OK;
}

__global__ void k_inc_count(const int32_t *__restrict__ edofs, int64_t total, int32_t *__restrict__ cnt) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x)
    atomicAdd(&cnt[edofs[t]], 1);
}

__global__ void k_inc_fill(const int32_t *__restrict__ edofs, int64_t total, const int32_t *__restrict__ ptr,
                           int32_t *__restrict__ fillc, int32_t *__restrict__ inc) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int32_t d = edofs[t];
    const int32_t p = atomicAdd(&fillc[d], 1);
    inc[ptr[d] + p] = (int32_t)t;
  }
}

// Sort each dof's incidence list ascending (element order, then local index):
// small lists, insertion sort per thread.
__global__ void k_inc_sort(const int32_t *__restrict__ ptr, int64_t n, int32_t *__restrict__ inc) {
  for (int64_t d = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; d < n; d += (int64_t)gridDim.x * blockDim.x) {
    const int32_t b = ptr[d], e = ptr[d + 1];
    for (int32_t i = b + 1; i < e; ++i) {
      const int32_t v = inc[i];
      int32_t j = i - 1;
      while (j >= b && inc[j] > v) {
        inc[j + 1] = inc[j];
        --j;
      }
      inc[j + 1] = v;
    }
  }
}

int build_incidence(lmh_context *ctx, const int32_t *edofs, int64_t n_elem, int dpe, int64_t n_dofs,
                    int32_t **inc_ptr, int32_t **inc) {
  const int64_t total = n_elem * dpe;
  int32_t *cnt = nullptr, *ptr = nullptr, *lst = nullptr;
  LMH_HIP(hipMallocAsync((void **)&cnt, sizeof(int32_t) * (n_dofs + 1), ctx->stream));
  LMH_HIP(hipMemsetAsync(cnt, 0, sizeof(int32_t) * (n_dofs + 1), ctx->stream));
  LMH_HIP(hipMalloc((void **)&ptr, sizeof(int32_t) * (n_dofs + 1)));
  LMH_HIP(hipMalloc((void **)&lst, sizeof(int32_t) * (total > 0 ? total : 1)));
  const int g = grid_for(total, 256, 8192);
  hipLaunchKernelGGL(k_inc_count, dim3(g), dim3(256), 0, ctx->stream, edofs, total, cnt);
  LMH_TRY(check_launch("k_inc_count"));
  int64_t tot = 0;
  LMH_TRY(exclusive_scan_i32(ctx, cnt, ptr, n_dofs, &tot));
  LMH_HIP(hipMemsetAsync(cnt, 0, sizeof(int32_t) * (n_dofs + 1), ctx->stream));
  hipLaunchKernelGGL(k_inc_fill, dim3(g), dim3(256), 0, ctx->stream, edofs, total, ptr, cnt, lst);
  LMH_TRY(check_launch("k_inc_fill"));
  hipLaunchKernelGGL(k_inc_sort, dim3(grid_for(n_dofs, 256, 8192)), dim3(256), 0, ctx->stream, ptr, n_dofs, lst);
  LMH_TRY(check_launch("k_inc_sort"));
  LMH_HIP(hipFreeAsync(cnt, ctx->stream));
  *inc_ptr = ptr;
  *inc = lst;
  return LMH_OK;
}

// Plain device copy, 16 bytes per thread per step.  Also the reference kernel
// of known traffic (reads and writes `n16*16` bytes) for hardware counters.
__global__ void k_copy16(const int4 *__restrict__ src, int4 *__restrict__ dst, int64_t n16) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n16; t += (int64_t)gridDim.x * blockDim.x)
    dst[t] = src[t];
}

}  // namespace lmh

using namespace lmh;

// ===========================================================================
// C ABI: context, memory, profiling
// ===========================================================================
extern "C" {

const char *lmh_last_error(void) { return lmh::g_err.c_str(); }

int lmh_context_create(int device, lmh_context **out) {
  LMH_CHECK(out, LMH_ERR_INVALID, "null output pointer");
  int ndev = 0;
  LMH_HIP(hipGetDeviceCount(&ndev));
  LMH_CHECK(device >= 0 && device < ndev, LMH_ERR_INVALID, "device %d not present (%d devices)", device, ndev);
  LMH_HIP(hipSetDevice(device));
  std::unique_ptr<lmh_context> c(new lmh_context);
  c->device = device;
  LMH_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  hipDeviceProp_t prop;
  LMH_HIP(hipGetDeviceProperties(&prop, device));
  c->cu_count = prop.multiProcessorCount;
  c->n_part = 16384;
  LMH_HIP(hipMalloc((void **)&c->d_scal, 64 * sizeof(double)));
  LMH_HIP(hipMemset(c->d_scal, 0, 64 * sizeof(double)));
  LMH_HIP(hipMalloc((void **)&c->d_part, (size_t)c->n_part * 4 * sizeof(double)));
  LMH_HIP(hipMalloc((void **)&c->d_count, 16 * sizeof(unsigned)));
  LMH_HIP(hipMemset(c->d_count, 0, 16 * sizeof(unsigned)));
  LMH_HIP(hipMalloc((void **)&c->d_flags, 16 * sizeof(int)));
  LMH_HIP(hipMemset(c->d_flags, 0, 16 * sizeof(int)));
  LMH_HIP(hipHostMalloc((void **)&c->h_pinned, 64 * sizeof(double), hipHostMallocDefault));
  LMH_HIP(hipEventCreateWithFlags(&c->ev_poll[0], hipEventDisableTiming));
  LMH_HIP(hipEventCreateWithFlags(&c->ev_poll[1], hipEventDisableTiming));
  LMH_TRY(cg_scratch_setup(c.get()));
  *out = c.release();
  return LMH_OK;
}

int lmh_context_destroy(lmh_context *ctx) {
  if (!ctx) return LMH_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (auto &kv : ctx->tables) (void)hipFree(kv.second->d_data);
  for (auto &kv : ctx->prof)
    for (auto &pr : kv.second.pending) {
      (void)hipEventDestroy(pr.first);
      (void)hipEventDestroy(pr.second);
    }
  for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
  for (auto e : ctx->ev_poll)
    if (e) (void)hipEventDestroy(e);
  (void)hipFree(ctx->d_scal);
  (void)hipFree(ctx->d_part);
  (void)hipFree(ctx->d_count);
  (void)hipFree(ctx->d_flags);
  (void)hipHostFree(ctx->h_pinned);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return LMH_OK;
}

int lmh_synchronize(lmh_context *ctx) {
  LMH_CHECK(ctx, LMH_ERR_INVALID, "null context");
  LMH_HIP(hipStreamSynchronize(ctx->stream));
  return LMH_OK;
}

void *lmh_context_stream(lmh_context *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int lmh_malloc(lmh_context *ctx, size_t bytes, void **d_out) {
  LMH_CHECK(ctx && d_out, LMH_ERR_INVALID, "null argument");
  LMH_HIP(hipSetDevice(ctx->device));
  hipError_t e = hipMalloc(d_out, bytes ? bytes : 1);
  if (e != hipSuccess) return set_error(LMH_ERR_ALLOC, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  return LMH_OK;
}
int lmh_free(lmh_context *ctx, void *d) {
  LMH_CHECK(ctx, LMH_ERR_INVALID, "null context");
  if (d) {
    LMH_HIP(hipStreamSynchronize(ctx->stream));
    LMH_HIP(hipFree(d));
  }
  return LMH_OK;
}
int lmh_memcpy_h2d(lmh_context *ctx, void *d_dst, const void *h_src, size_t bytes) {
  LMH_CHECK(ctx, LMH_ERR_INVALID, "null context");
  LMH_HIP(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, ctx->stream));
  LMH_HIP(hipStreamSynchronize(ctx->stream));
  return LMH_OK;
}
int lmh_memcpy_d2h(lmh_context *ctx, void *h_dst, const void *d_src, size_t bytes) {
  LMH_CHECK(ctx, LMH_ERR_INVALID, "null context");
  LMH_HIP(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  LMH_HIP(hipStreamSynchronize(ctx->stream));
  return LMH_OK;
}
int lmh_memcpy_d2d(lmh_context *ctx, void *d_dst, const void *d_src, size_t bytes) {
  LMH_CHECK(ctx, LMH_ERR_INVALID, "null context");
  if (bytes == 0) return LMH_OK;
  if (((uintptr_t)d_dst | (uintptr_t)d_src | bytes) % 16 == 0) {
    const int64_t n16 = (int64_t)(bytes / 16);
    LMH_LAUNCH(ctx, "copy", k_copy16, dim3(grid_for(n16, 256, ctx->cu_count * 16)), dim3(256), 0,
               (const int4 *)d_src, (int4 *)d_dst, n16);
    return LMH_OK;
  }
  LMH_HIP(hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  return LMH_OK;
}
int lmh_memset(lmh_context *ctx, void *d, int value, size_t bytes) {
  LMH_CHECK(ctx, LMH_ERR_INVALID, "null context");
  LMH_HIP(hipMemsetAsync(d, value, bytes, ctx->stream));
  return LMH_OK;
}

int lmh_profile_enable(lmh_context *ctx, int on) {
  LMH_CHECK(ctx, LMH_ERR_INVALID, "null context");
  ctx->prof_on = on != 0;
  return LMH_OK;
}

static int drain_profile(lmh_context *ctx) {
  LMH_HIP(hipStreamSynchronize(ctx->stream));
  for (auto &kv : ctx->prof) {
    for (auto &pr : kv.second.pending) {
      float ms = 0;
      LMH_HIP(hipEventElapsedTime(&ms, pr.first, pr.second));
      kv.second.total_ms += ms;
      kv.second.count += 1;
      ctx->event_pool.push_back(pr.first);
      ctx->event_pool.push_back(pr.second);
    }
    kv.second.pending.clear();
  }
  return LMH_OK;
}

int lmh_profile_query(lmh_context *ctx, const char *kernel, double *avg_ms, int64_t *count) {
  LMH_CHECK(ctx && kernel, LMH_ERR_INVALID, "null argument");
  LMH_TRY(drain_profile(ctx));
  auto it = ctx->prof.find(kernel);
  if (it == ctx->prof.end() || it->second.count == 0) {
    if (avg_ms) *avg_ms = 0;
    if (count) *count = 0;
    return LMH_OK;
  }
  if (avg_ms) *avg_ms = it->second.total_ms / it->second.count;
  if (count) *count = it->second.count;
  return LMH_OK;
}

int lmh_profile_reset(lmh_context *ctx) {
  LMH_CHECK(ctx, LMH_ERR_INVALID, "null context");
  LMH_TRY(drain_profile(ctx));
  ctx->prof.clear();
  return LMH_OK;
}

// ===========================================================================
// C ABI: mesh and DoF map
// ===========================================================================
int lmh_mesh_create(lmh_context *ctx, int elem_type, int64_t n_nodes, const double *coords_h, int64_t n_elem,
                    const int32_t *conn_h, lmh_mesh **out) {
  LMH_CHECK(ctx && out && coords_h && conn_h, LMH_ERR_INVALID, "null argument");
  const int npe = nodes_per_elem(elem_type);
  LMH_CHECK(npe > 0, LMH_ERR_UNSUPPORTED, "unsupported element type %d", elem_type);
  LMH_CHECK(n_nodes > 0 && n_elem > 0, LMH_ERR_INVALID, "empty mesh");
  LMH_CHECK(n_nodes < INT32_MAX, LMH_ERR_UNSUPPORTED, "too many nodes");
  for (int64_t t = 0; t < n_elem * npe; ++t)
    LMH_CHECK(conn_h[t] >= 0 && conn_h[t] < n_nodes, LMH_ERR_INVALID, "connectivity entry %lld out of range",
              (long long)t);
  LMH_HIP(hipSetDevice(ctx->device));
  std::unique_ptr<lmh_mesh> m(new lmh_mesh);
  m->ctx = ctx;
  m->type = elem_type;
  m->npe = npe;
  m->n_nodes = n_nodes;
  m->n_elem = n_elem;
  LMH_HIP(hipMalloc((void **)&m->xyz, sizeof(double) * 3 * n_nodes));
  LMH_HIP(hipMalloc((void **)&m->conn, sizeof(int32_t) * npe * n_elem));
  LMH_HIP(hipMemcpy(m->xyz, coords_h, sizeof(double) * 3 * n_nodes, hipMemcpyHostToDevice));
  LMH_HIP(hipMemcpy(m->conn, conn_h, sizeof(int32_t) * npe * n_elem, hipMemcpyHostToDevice));
  *out = m.release();
  return LMH_OK;
}

int lmh_mesh_destroy(lmh_mesh *m) {
  if (!m) return LMH_OK;
  (void)hipStreamSynchronize(m->ctx->stream);
  (void)hipFree(m->xyz);
  (void)hipFree(m->conn);
  delete m;
  return LMH_OK;
}

int lmh_mesh_set_coords(lmh_mesh *m, const double *coords_h) {
  LMH_CHECK(m && coords_h, LMH_ERR_INVALID, "null argument");
  LMH_HIP(hipMemcpyAsync(m->xyz, coords_h, sizeof(double) * 3 * m->n_nodes, hipMemcpyHostToDevice, m->ctx->stream));
  LMH_HIP(hipStreamSynchronize(m->ctx->stream));
  return LMH_OK;
}

int lmh_dofmap_create(lmh_context *ctx, const lmh_mesh *m, int n_vars, int64_t n_dofs, const int32_t *elem_dofs_h,
                      const uint8_t *constrained_h, lmh_dofmap **out) {
  LMH_CHECK(ctx && m && out && elem_dofs_h, LMH_ERR_INVALID, "null argument");
  LMH_CHECK(n_vars >= 1 && n_vars <= 3, LMH_ERR_UNSUPPORTED, "n_vars %d not supported", n_vars);
  LMH_CHECK(n_dofs > 0 && n_dofs < INT32_MAX, LMH_ERR_INVALID, "bad n_dofs");
  const int dpe = n_vars * m->npe;
  for (int64_t t = 0; t < m->n_elem * dpe; ++t)
    LMH_CHECK(elem_dofs_h[t] >= 0 && elem_dofs_h[t] < n_dofs, LMH_ERR_INVALID, "dof index %lld out of range",
              (long long)t);
  std::unique_ptr<lmh_dofmap> d(new lmh_dofmap);
  d->ctx = ctx;
  d->mesh = m;
  d->nv = n_vars;
  d->dpe = dpe;
  d->n_dofs = n_dofs;
  LMH_HIP(hipMalloc((void **)&d->edofs, sizeof(int32_t) * dpe * m->n_elem));
  LMH_HIP(hipMemcpy(d->edofs, elem_dofs_h, sizeof(int32_t) * dpe * m->n_elem, hipMemcpyHostToDevice));
  LMH_HIP(hipMalloc((void **)&d->cons, n_dofs));
  if (constrained_h)
    LMH_HIP(hipMemcpy(d->cons, constrained_h, n_dofs, hipMemcpyHostToDevice));
  else
    LMH_HIP(hipMemset(d->cons, 0, n_dofs));
  d->h_cons0.assign(n_dofs, 0);
  if (constrained_h)
    for (int64_t t = 0; t < n_dofs; ++t) d->h_cons0[t] = constrained_h[t] != 0;
  LMH_TRY(build_incidence(ctx, d->edofs, m->n_elem, dpe, n_dofs, &d->inc_ptr, &d->inc));
  // the constrained set as empty constraint rows
  LMH_TRY(lmh_dofmap_set_constraints(d.get(), 0, nullptr, nullptr, nullptr, nullptr, nullptr, 0));
  LMH_HIP(hipStreamSynchronize(ctx->stream));
  *out = d.release();
  return LMH_OK;
}

int lmh_dofmap_destroy(lmh_dofmap *dm) {
  if (!dm) return LMH_OK;
  (void)hipStreamSynchronize(dm->ctx->stream);
  (void)hipFree(dm->edofs);
  (void)hipFree(dm->cons);
  (void)hipFree(dm->inc_ptr);
  (void)hipFree(dm->inc);
  (void)release_constraints(dm->cr);
  delete dm;
  return LMH_OK;
}

}  // extern "C"
