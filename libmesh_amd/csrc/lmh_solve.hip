// lmh_solve.hip -- CSR SpMV, Jacobi-preconditioned CG (FP64) and the
// mixed-precision (FP32 inner / FP64 refined) solver.
//
// CG follows Eigen::internal::conjugate_gradient (ConjugateGradient.h:28-90)
// with DiagonalPreconditioner (BasicPreconditioners.h:66-76), the solver
// libMesh runs for CG + JACOBI_PRECOND (eigen_sparse_linear_solver.C:106-160).
// All scalars live on the device; each iteration is three kernels
//   spmv+dot(p,Ap) -> alpha | x,r update + |r|^2 + (r,z) -> beta | p = z + beta p
// with a deterministic "last block reduces the partials" finish, so the host
// only polls a converged flag every few iterations.
#include <algorithm>
#include <cmath>
#include <functional>

#include "lmh_internal.h"

namespace lmh {

// solver scalar slots in ctx->d_scal
enum {
  S_BB = 0, S_THR, S_ABSNEW, S_ABSOLD, S_PAP, S_ALPHA, S_BETA, S_RR, S_ITS, S_DONE, S_REASON,
  S_MAXRR = 16, S_REFRESH, S_RRTRUE, S_NREF, S_DELTA2, S_MODE, S_RZ
};



constexpr int BLK = 256;
constexpr int NXCD = 8;
constexpr int RB_CAP = 2048;  // stored entries per SpMV row block

// Blocks are dispatched round-robin over the XCDs; remap so each XCD owns one
// contiguous range of rows (its L2 then holds the x entries its rows gather).
__device__ inline int logical_block() {
  const int G = gridDim.x, b = blockIdx.x;
  if (G % NXCD) return b;
  return (b % NXCD) * (G / NXCD) + b / NXCD;
}

__device__ inline void block_range(int64_t n, int64_t &lo, int64_t &hi) {
  const int lb = logical_block();
  const int64_t per = (n + gridDim.x - 1) / gridDim.x;
  lo = (int64_t)lb * per;
  hi = lo + per < n ? lo + per : n;
}

template <typename T>
__device__ inline T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// block-wide sum of up to 3 values (deterministic); result valid in thread 0
template <int NV>
__device__ inline void block_sum(double (&v)[NV], double (*sh)[BLK / 64]) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double s = wave_sum(v[k]);
    if (lane == 0) sh[k][wv] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double s = 0;
      for (int w = 0; w < BLK / 64; ++w) s += sh[k][w];
      v[k] = s;
    }
  }
}

// Publish the block's NV partial sums to part[k*gridDim.x + blockIdx.x].  A
// one-block finalize kernel launched next on the same stream reduces them in
// block order (deterministic); the kernel boundary makes the partials visible
// without any device-scope fence inside the producer.
template <int NV>
__device__ inline void publish(double (&v)[NV], double *__restrict__ part) {
  __shared__ double sh[NV][BLK / 64];
  block_sum<NV>(v, sh);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) part[k * gridDim.x + blockIdx.x] = v[k];
  }
}

// In a one-block kernel: totals of G published partials; valid in thread 0.
template <int NV>
__device__ inline void reduce_partials(const double *__restrict__ part, int G, double (&t)[NV]) {
  __shared__ double sh[NV][BLK / 64];
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double s = 0;
    for (int b = threadIdx.x; b < G; b += BLK) s += part[k * G + b];
    t[k] = s;
  }
  block_sum<NV>(t, sh);
}

// Row-block ("CSR stream") SpMV core.  Rows are grouped into blocks holding at
// most RB_CAP stored entries; a workgroup streams a block's values and column
// indices with coalesced, independent loads (RB_CAP/BLK per thread), gathers x,
// writes the products to LDS, then 8 lanes per row sum each row's products in a
// fixed order.  A single row longer than RB_CAP is one block, reduced directly.
// epi(row, sum) consumes each row result in the calling thread.
template <typename VT, typename XT, typename AT, typename Epi>
__device__ inline void rowblock_spmv(const int32_t *__restrict__ rb, int64_t nrb, const int32_t *__restrict__ rp,
                                     const int32_t *__restrict__ ci, const VT *__restrict__ v,
                                     const XT *__restrict__ x, Epi &&epi) {
  __shared__ AT prod[RB_CAP];
  __shared__ double red[BLK / 64];
  constexpr int PER = RB_CAP / BLK;
  int64_t lo, hi;
  block_range(nrb, lo, hi);
  const int tid = threadIdx.x;
  for (int64_t bk = lo; bk < hi; ++bk) {
    const int32_t r0 = rb[bk], r1 = rb[bk + 1];
    const int32_t k0 = rp[r0], cnt = rp[r1] - k0;
    if (cnt > RB_CAP) {
      AT s = 0;
      for (int32_t t = tid; t < cnt; t += BLK)
        s += (AT)__builtin_nontemporal_load(&v[k0 + t]) * (AT)x[__builtin_nontemporal_load(&ci[k0 + t])];
      double w = wave_sum((double)s);
      if ((tid & 63) == 0) red[tid >> 6] = w;
      __syncthreads();
      if (tid == 0) {
        double t2 = 0;
        for (int k = 0; k < BLK / 64; ++k) t2 += red[k];
        epi(r0, (AT)t2);
      }
      __syncthreads();
      continue;
    }
    // branch-free loads (out-of-range slots re-read entry k0 and are dropped)
    // so all PER value/index loads, then all PER gathers, are in flight at once
    VT vv[PER];
    int32_t cc[PER];
    AT xx[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const int32_t t = tid + j * BLK;
      const int32_t k = k0 + (t < cnt ? t : 0);
      vv[j] = __builtin_nontemporal_load(&v[k]);
      cc[j] = __builtin_nontemporal_load(&ci[k]);
    }
#pragma unroll
    for (int j = 0; j < PER; ++j) xx[j] = (AT)x[cc[j]];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const int32_t t = tid + j * BLK;
      if (t < cnt) prod[t] = (AT)vv[j] * xx[j];
    }
    __syncthreads();
    const int g = tid >> 3, sub = tid & 7;
    for (int32_t r = r0 + g; r < r1; r += BLK / 8) {
      const int32_t a = rp[r] - k0, e = rp[r + 1] - k0;
      AT s = 0;
      for (int32_t t = a + sub; t < e; t += 8) s += prod[t];
      s += __shfl_xor(s, 4, 8);
      s += __shfl_xor(s, 2, 8);
      s += __shfl_xor(s, 1, 8);
      if (sub == 0) epi(r, s);
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// plain SpMV y = A x
// ---------------------------------------------------------------------------
template <typename VT, typename XT, typename YT, typename AT>
__global__ __launch_bounds__(BLK) void k_spmv(const int32_t *__restrict__ rb, int64_t nrb,
                                              const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                              const VT *__restrict__ v, const XT *__restrict__ x,
                                              YT *__restrict__ y) {
  rowblock_spmv<VT, XT, AT>(rb, nrb, rp, ci, v, x, [&](int32_t r, AT s) { y[r] = (YT)s; });
}


// r = b - A x (FP64); partials |r|^2, |b|^2
__global__ __launch_bounds__(BLK) void k_residual(const int32_t *__restrict__ rb, int64_t nrb,
                                                  const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                                  const double *__restrict__ v, const double *__restrict__ b,
                                                  const double *__restrict__ x, double *__restrict__ r,
                                                  double *__restrict__ part) {
  double acc[2] = {0, 0};
  rowblock_spmv<double, double, double>(rb, nrb, rp, ci, v, x, [&](int32_t row, double s) {
    const double bv = b[row], rv = bv - s;
    r[row] = rv;
    acc[0] += rv * rv;
    acc[1] += bv * bv;
  });
  publish<2>(acc, part);
}

__global__ __launch_bounds__(BLK) void k_fin_residual(const double *__restrict__ part, int G, double *__restrict__ S) {
  double t[2];
  reduce_partials<2>(part, G, t);
  if (threadIdx.x == 0) {
    S[S_RR] = t[0];
    S[S_BB] = t[1];
  }
}

// Jacobi inverse diagonal: 1/a_rr, or 1 when the diagonal is absent or zero
template <typename VT, typename OT>
__global__ void k_invdiag(int64_t n, const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                          const VT *__restrict__ v, OT *__restrict__ dinv) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    int32_t lo = rp[r], hi = rp[r + 1];
    const int32_t end = hi;
    while (lo < hi) {
      const int32_t mid = (lo + hi) >> 1;
      if (ci[mid] < r)
        lo = mid + 1;
      else
        hi = mid;
    }
    OT d = (OT)1;
    if (lo < end && ci[lo] == r && v[lo] != (VT)0) d = (OT)1 / (OT)v[lo];
    dinv[r] = d;
  }
}

// ---------------------------------------------------------------------------
// FP64 CG
// ---------------------------------------------------------------------------
// after k_residual + k_fin_residual: p = dinv r (x = p = 0 when b = 0); partial r.p
__global__ __launch_bounds__(BLK) void k_cg_init(int64_t n, const double *__restrict__ dinv,
                                                 const double *__restrict__ r, double *__restrict__ p,
                                                 double *__restrict__ x, double *__restrict__ part,
                                                 const double *__restrict__ S) {
  int64_t lo, hi;
  block_range(n, lo, hi);
  const bool zero_rhs = S[S_BB] == 0.0;
  double acc[1] = {0};
  for (int64_t i = lo + threadIdx.x; i < hi; i += BLK) {
    if (zero_rhs) {
      x[i] = 0.0;
      p[i] = 0.0;
    } else {
      const double z = dinv[i] * r[i];
      p[i] = z;
      acc[0] += r[i] * z;
    }
  }
  publish<1>(acc, part);
}

// the early exits of ConjugateGradient.h:44-60
__global__ __launch_bounds__(BLK) void k_fin_cg_init(const double *__restrict__ part, int G, double tol,
                                                     double *__restrict__ S) {
  double t[1];
  reduce_partials<1>(part, G, t);
  if (threadIdx.x == 0) {
    const double bb = S[S_BB];
    const double thr = fmax(tol * tol * bb, 2.2250738585072014e-308);
    S[S_THR] = thr;
    S[S_ABSNEW] = t[0];
    S[S_ITS] = 0;
    S[S_DONE] = 0;
    S[S_REASON] = 0;
    if (bb == 0.0) {
      S[S_RR] = 0;
      S[S_DONE] = 1;
      S[S_REASON] = LMH_CONVERGED_RTOL;
    } else if (S[S_RR] < thr) {
      S[S_DONE] = 1;
      S[S_REASON] = LMH_CONVERGED_RTOL;
    }
  }
}

// q = A p; partial p.q  (template VT/AT: FP64 for CG, FP32 for the mixed inner solver)
template <typename VT, typename AT>
__global__ __launch_bounds__(BLK) void k_spmv_dot(const int32_t *__restrict__ rb, int64_t nrb,
                                                  const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                                  const VT *__restrict__ v, const VT *__restrict__ p,
                                                  VT *__restrict__ q, double *__restrict__ part,
                                                  const double *__restrict__ S) {
  if (S[S_DONE] != 0.0) return;
  double acc[1] = {0};
  rowblock_spmv<VT, VT, AT>(rb, nrb, rp, ci, v, p, [&](int32_t r, AT s) {
    q[r] = (VT)s;
    acc[0] += (double)p[r] * (double)s;
  });
  publish<1>(acc, part);
}

// alpha = (r,z) / (p,Ap); slot holds (r,z): S_ABSNEW for CG, S_RZ for the mixed solver
__global__ __launch_bounds__(BLK) void k_fin_alpha(const double *__restrict__ part, int G, int rz_slot,
                                                   double *__restrict__ S) {
  if (S[S_DONE] != 0.0) return;
  double t[1];
  reduce_partials<1>(part, G, t);
  if (threadIdx.x == 0) {
    const double pap = t[0];
    S[S_PAP] = pap;
    const double alpha = S[rz_slot] / pap;
    S[S_ALPHA] = alpha;
    if (!(pap != 0.0) || !isfinite(alpha)) {
      S[S_DONE] = 1;
      S[S_REASON] = LMH_DIVERGED_BREAKDOWN;
    }
  }
}

__global__ __launch_bounds__(BLK) void k_cg_update(int64_t n, const double *__restrict__ dinv,
                                                   const double *__restrict__ p, const double *__restrict__ q,
                                                   double *__restrict__ x, double *__restrict__ r,
                                                   double *__restrict__ part, const double *__restrict__ S) {
  if (S[S_DONE] != 0.0) return;
  int64_t lo, hi;
  block_range(n, lo, hi);
  const double alpha = S[S_ALPHA];
  double acc[2] = {0, 0};
  for (int64_t i = lo + threadIdx.x; i < hi; i += BLK) {
    x[i] += alpha * p[i];
    const double rv = r[i] - alpha * q[i];
    r[i] = rv;
    acc[0] += rv * rv;
    acc[1] += rv * (dinv[i] * rv);
  }
  publish<2>(acc, part);
}

// residualNorm2 test, then absOld/absNew/beta and i++ (ConjugateGradient.h:76-87)
__global__ __launch_bounds__(BLK) void k_fin_cg_beta(const double *__restrict__ part, int G, double *__restrict__ S) {
  if (S[S_DONE] != 0.0) return;
  double t[2];
  reduce_partials<2>(part, G, t);
  if (threadIdx.x == 0) {
    S[S_RR] = t[0];
    if (!isfinite(t[0])) {
      S[S_DONE] = 1;
      S[S_REASON] = LMH_DIVERGED_NANORINF;
    } else if (t[0] < S[S_THR]) {
      S[S_DONE] = 1;
      S[S_REASON] = LMH_CONVERGED_RTOL;
    } else {
      const double absOld = S[S_ABSNEW];
      S[S_ABSOLD] = absOld;
      S[S_ABSNEW] = t[1];
      S[S_BETA] = t[1] / absOld;
      S[S_ITS] += 1;
    }
  }
}

__global__ __launch_bounds__(BLK) void k_cg_pupdate(int64_t n, const double *__restrict__ dinv,
                                                    const double *__restrict__ r, double *__restrict__ p,
                                                    const double *__restrict__ S) {
  if (S[S_DONE] != 0.0) return;
  int64_t lo, hi;
  block_range(n, lo, hi);
  const double beta = S[S_BETA];
  for (int64_t i = lo + threadIdx.x; i < hi; i += BLK) p[i] = dinv[i] * r[i] + beta * p[i];
}

// ---------------------------------------------------------------------------
// Mixed precision.  FP32 recurrence (x_s, r_s, p_s, q_s) on the FP32 operator;
// FP64 solution x and true residual r = b - A x on the FP64 operator.
// ---------------------------------------------------------------------------
// r64 = b - A64 x; r32 = (float) r64; z = dinv32 r32 (p32 = z on the initial
// call); partials |r|^2, |b|^2, (r32, z).  Used initially and on every refresh.
__global__ __launch_bounds__(BLK) void k_mx_residual(int init, const int32_t *__restrict__ rb, int64_t nrb,
                                                     const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                                     const double *__restrict__ v, const double *__restrict__ b,
                                                     const double *__restrict__ x, const float *__restrict__ dinv32,
                                                     float *__restrict__ r32, float *__restrict__ p32,
                                                     double *__restrict__ part, const double *__restrict__ S) {
  if (!init && (S[S_DONE] != 0.0 || S[S_REFRESH] == 0.0)) return;
  double acc[3] = {0, 0, 0};
  rowblock_spmv<double, double, double>(rb, nrb, rp, ci, v, x, [&](int32_t row, double s) {
    const double bv = b[row], rv = bv - s;
    const float rf = (float)rv;
    r32[row] = rf;
    const float z = dinv32[row] * rf;
    if (init) p32[row] = z;
    acc[0] += rv * rv;
    acc[1] += bv * bv;
    acc[2] += (double)rf * (double)z;
  });
  publish<3>(acc, part);
}

__global__ __launch_bounds__(BLK) void k_fin_mx_residual(const double *__restrict__ part, int G, int init,
                                                         double *__restrict__ S) {
  if (!init && (S[S_DONE] != 0.0 || S[S_REFRESH] == 0.0)) return;
  double t[3];
  reduce_partials<3>(part, G, t);
  if (threadIdx.x != 0) return;
  const double rr = t[0];
  S[S_RRTRUE] = rr;
  if (init) {
    const double tol2 = S[S_THR];  // rtol^2 on entry
    S[S_BB] = t[1];
    S[S_THR] = tol2 * t[1];
    S[S_ITS] = 0;
    S[S_NREF] = 0;
    S[S_DONE] = 0;
    S[S_REASON] = 0;
    S[S_REFRESH] = 0;
    S[S_RZ] = t[2];
    S[S_MAXRR] = rr;
    if (t[1] == 0.0 || rr <= S[S_THR]) {
      S[S_DONE] = 1;
      S[S_REASON] = LMH_CONVERGED_RTOL;
    }
    return;
  }
  S[S_NREF] += 1;
  if (rr <= S[S_THR]) {
    S[S_DONE] = 1;
    S[S_REASON] = LMH_CONVERGED_RTOL;
  } else if (!isfinite(rr)) {
    S[S_DONE] = 1;
    S[S_REASON] = LMH_DIVERGED_NANORINF;
  } else {
    // keep the search direction (mode 0, reliable update) or restart (mode 1)
    S[S_MAXRR] = rr;
    S[S_BETA] = (S[S_MODE] == 0.0) ? t[2] / S[S_RZ] : 0.0;
    S[S_RZ] = t[2];
  }
}

__global__ __launch_bounds__(BLK) void k_mx_update(int64_t n, const float *__restrict__ dinv32,
                                                   const float *__restrict__ p, const float *__restrict__ q,
                                                   float *__restrict__ xs, float *__restrict__ r,
                                                   double *__restrict__ part, const double *__restrict__ S) {
  if (S[S_DONE] != 0.0) return;
  int64_t lo, hi;
  block_range(n, lo, hi);
  const float alpha = (float)S[S_ALPHA];
  double acc[2] = {0, 0};
  for (int64_t i = lo + threadIdx.x; i < hi; i += BLK) {
    xs[i] += alpha * p[i];
    const float rv = r[i] - alpha * q[i];
    r[i] = rv;
    const float z = dinv32[i] * rv;
    acc[0] += (double)rv * (double)rv;
    acc[1] += (double)rv * (double)z;
  }
  publish<2>(acc, part);
}

// refresh when the FP32 residual fell below delta times the last true residual
// (or below the final threshold); otherwise the ordinary CG beta
__global__ __launch_bounds__(BLK) void k_fin_mx_update(const double *__restrict__ part, int G,
                                                       double *__restrict__ S) {
  if (S[S_DONE] != 0.0) return;
  double t[2];
  reduce_partials<2>(part, G, t);
  if (threadIdx.x != 0) return;
  S[S_ITS] += 1;
  const double rr = t[0];
  if (!isfinite(rr)) {
    S[S_DONE] = 1;
    S[S_REASON] = LMH_DIVERGED_NANORINF;
  } else if (rr < S[S_DELTA2] * S[S_MAXRR] || rr < S[S_THR]) {
    S[S_REFRESH] = 1;
  } else {
    S[S_REFRESH] = 0;
    S[S_BETA] = t[1] / S[S_RZ];
    S[S_RZ] = t[1];
  }
}

// x += xs; xs = 0 (on a refresh, or when forced)
__global__ __launch_bounds__(BLK) void k_mx_fold(int64_t n, int force, float *__restrict__ xs, double *__restrict__ x,
                                                 const double *__restrict__ S) {
  if (!force && (S[S_DONE] != 0.0 || S[S_REFRESH] == 0.0)) return;
  int64_t lo, hi;
  block_range(n, lo, hi);
  for (int64_t i = lo + threadIdx.x; i < hi; i += BLK) {
    x[i] += (double)xs[i];
    xs[i] = 0.0f;
  }
}

__global__ __launch_bounds__(BLK) void k_mx_pupdate(int64_t n, const float *__restrict__ dinv32,
                                                    const float *__restrict__ r, float *__restrict__ p,
                                                    const double *__restrict__ S) {
  if (S[S_DONE] != 0.0) return;
  int64_t lo, hi;
  block_range(n, lo, hi);
  const float beta = (float)S[S_BETA];
  for (int64_t i = lo + threadIdx.x; i < hi; i += BLK) p[i] = dinv32[i] * r[i] + beta * p[i];
}

__global__ void k_round32(int64_t n, const double *__restrict__ a, float *__restrict__ b) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}

__global__ __launch_bounds__(BLK) void k_norm2(int64_t n, const double *__restrict__ x, double *__restrict__ part) {
  int64_t lo, hi;
  block_range(n, lo, hi);
  double acc[1] = {0};
  for (int64_t i = lo + threadIdx.x; i < hi; i += BLK) acc[0] += x[i] * x[i];
  publish<1>(acc, part);
}

__global__ __launch_bounds__(BLK) void k_fin_sum(const double *__restrict__ part, int G, double *__restrict__ out) {
  double t[1];
  reduce_partials<1>(part, G, t);
  if (threadIdx.x == 0) *out = t[0];
}

static int vec_grid(lmh_context *ctx) {
  // the partial buffer holds 4 * n_part doubles; publish<3> uses 3 * grid
  const int g = ctx->cu_count * 8;
  return g < ctx->n_part ? g : ctx->n_part;
}

static int ensure_work(lmh_matrix *A, bool mixed) {
  if (!A->w64) LMH_HIP(hipMalloc((void **)&A->w64, sizeof(double) * 4 * A->rows));
  if (mixed && !A->w32) LMH_HIP(hipMalloc((void **)&A->w32, sizeof(float) * 5 * A->rows));
  return LMH_OK;
}

}  // namespace lmh

using namespace lmh;

extern "C" {

int lmh_spmv(lmh_matrix *A, const double *x_d, double *y_d) {
  LMH_CHECK(A && x_d && y_d, LMH_ERR_INVALID, "null argument");
  lmh_context *ctx = A->ctx;
  const int g = vec_grid(ctx);
  LMH_LAUNCH(ctx, "spmv_fp64", (k_spmv<double, double, double, double>), dim3(g),
                                          dim3(BLK), 0, A->rb, A->nrb, A->rp, A->ci, A->vals, x_d, y_d);
  return check_launch("k_spmv");
}

int lmh_spmv_fp32(lmh_matrix *A, const float *x_d, float *y_d) {
  LMH_CHECK(A && x_d && y_d, LMH_ERR_INVALID, "null argument");
  LMH_CHECK(A->have32 && A->vals32, LMH_ERR_INVALID, "matrix has no FP32 operator");
  lmh_context *ctx = A->ctx;
  const int g = vec_grid(ctx);
  LMH_LAUNCH(ctx, "spmv_fp32", (k_spmv<float, float, float, float>), dim3(g),
                                          dim3(BLK), 0, A->rb, A->nrb, A->rp, A->ci, A->vals32, x_d, y_d);
  return check_launch("k_spmv");
}

int lmh_residual(lmh_matrix *A, const double *b_d, const double *x_d, double *r_d) {
  LMH_CHECK(A && b_d && x_d && r_d, LMH_ERR_INVALID, "null argument");
  lmh_context *ctx = A->ctx;
  const int g = vec_grid(ctx);
  LMH_LAUNCH(ctx, "residual_fp64", k_residual, dim3(g), dim3(BLK), 0, A->rb, A->nrb, A->rp,
                                          A->ci, A->vals, b_d, x_d, r_d, ctx->d_part);
  LMH_LAUNCH(ctx, "fin", k_fin_residual, dim3(1), dim3(BLK), 0, ctx->d_part, g, ctx->d_scal + 32);
  return check_launch("k_residual");
}

double lmh_norm2(lmh_context *ctx, const double *x_d, int64_t n) {
  if (!ctx || !x_d) return NAN;
  const int g = vec_grid(ctx);
  hipLaunchKernelGGL(k_norm2, dim3(g), dim3(BLK), 0, ctx->stream, n, x_d, ctx->d_part);
  hipLaunchKernelGGL(k_fin_sum, dim3(1), dim3(BLK), 0, ctx->stream, ctx->d_part, g, ctx->d_scal + 48);
  double v = NAN;
  if (hipMemcpyAsync(&v, ctx->d_scal + 48, sizeof(double), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess)
    return NAN;
  if (hipStreamSynchronize(ctx->stream) != hipSuccess) return NAN;
  return std::sqrt(v);
}

int lmh_matrix_make_fp32(lmh_matrix *A) {
  LMH_CHECK(A, LMH_ERR_INVALID, "null matrix");
  lmh_context *ctx = A->ctx;
  if (!A->vals32) LMH_HIP(hipMalloc((void **)&A->vals32, sizeof(float) * A->nnz));
  LMH_LAUNCH(ctx, "round_fp32", k_round32, dim3(grid_for(A->nnz, 256, 16384)), dim3(256), 0, A->nnz, A->vals,
             A->vals32);
  A->have32 = true;
  return check_launch("k_round32");
}

// Host driver: launches iterations in batches and polls the device "done" flag
// of the previous batch (one batch of lag) through pinned memory.
static int run_batches(lmh_context *ctx, int64_t maxits, int batch, const std::function<void()> &one_iter) {
  double *flag = ctx->h_pinned;
  hipEvent_t ev[2] = {get_event(ctx), get_event(ctx)};
  int64_t launched = 0;
  int k = 0;
  bool pending = false;
  while (launched < maxits) {
    const int64_t nb = std::min<int64_t>(batch, maxits - launched);
    for (int64_t i = 0; i < nb; ++i) one_iter();
    launched += nb;
    LMH_HIP(hipMemcpyAsync(&flag[k], ctx->d_scal + S_DONE, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    LMH_HIP(hipEventRecord(ev[k], ctx->stream));
    if (pending) {
      LMH_HIP(hipEventSynchronize(ev[k ^ 1]));
      if (flag[k ^ 1] != 0.0) break;
    }
    pending = true;
    k ^= 1;
  }
  LMH_HIP(hipStreamSynchronize(ctx->stream));
  ctx->event_pool.push_back(ev[0]);
  ctx->event_pool.push_back(ev[1]);
  return check_launch("solver iteration");
}

int lmh_cg_solve(lmh_matrix *A, const double *b_d, double *x_d, double rtol, int64_t maxits, int64_t *iters,
                 double *rel_res, int *reason) {
  LMH_CHECK(A && b_d && x_d, LMH_ERR_INVALID, "null argument");
  LMH_CHECK(rtol >= 0 && maxits >= 0, LMH_ERR_INVALID, "bad tolerance or iteration limit");
  lmh_context *ctx = A->ctx;
  LMH_TRY(ensure_work(A, false));
  const int64_t n = A->rows;
  double *r = A->w64, *p = r + n, *q = p + n, *dinv = q + n;
  double *S = ctx->d_scal, *part = ctx->d_part;
  const int g = vec_grid(ctx);
  LMH_LAUNCH(ctx, "invdiag", (k_invdiag<double, double>), dim3(grid_for(n, 256, 8192)), dim3(256), 0, n, A->rp, A->ci,
             A->vals, dinv);
  LMH_LAUNCH(ctx, "residual_fp64", k_residual, dim3(g), dim3(BLK), 0, A->rb, A->nrb, A->rp, A->ci, A->vals,
                                b_d, x_d, r, part);
  LMH_LAUNCH(ctx, "fin", k_fin_residual, dim3(1), dim3(BLK), 0, part, g, S);
  LMH_LAUNCH(ctx, "cg_init", k_cg_init, dim3(g), dim3(BLK), 0, n, dinv, r, p, x_d, part, S);
  LMH_LAUNCH(ctx, "fin", k_fin_cg_init, dim3(1), dim3(BLK), 0, part, g, rtol, S);
  LMH_TRY(check_launch("cg setup"));
  auto iter = [&]() {
    LMH_LAUNCH(ctx, "cg_spmv_fp64", (k_spmv_dot<double, double>), dim3(g), dim3(BLK), 0, A->rb,
                                  A->nrb, A->rp, A->ci, A->vals, p, q, part, S);
    LMH_LAUNCH(ctx, "fin", k_fin_alpha, dim3(1), dim3(BLK), 0, part, g, (int)S_ABSNEW, S);
    LMH_LAUNCH(ctx, "cg_update_fp64", k_cg_update, dim3(g), dim3(BLK), 0, n, dinv, p, q, x_d, r, part, S);
    LMH_LAUNCH(ctx, "fin", k_fin_cg_beta, dim3(1), dim3(BLK), 0, part, g, S);
    LMH_LAUNCH(ctx, "cg_pupdate_fp64", k_cg_pupdate, dim3(g), dim3(BLK), 0, n, dinv, r, p, S);
  };
  LMH_TRY(run_batches(ctx, maxits, 4, iter));
  double h[16];
  LMH_HIP(hipMemcpy(h, S, sizeof(h), hipMemcpyDeviceToHost));
  const bool done = h[S_DONE] != 0.0;
  if (iters) *iters = done ? (int64_t)h[S_ITS] : maxits;
  if (rel_res) *rel_res = h[S_BB] > 0 ? std::sqrt(h[S_RR] / h[S_BB]) : 0.0;
  if (reason) *reason = done ? (int)h[S_REASON] : LMH_DIVERGED_ITS;
  return LMH_OK;
}

int lmh_mixed_solve(lmh_matrix *A, const double *b_d, double *x_d, double rtol, double inner_rtol, int mode,
                    int64_t maxits, int64_t *iters, double *rel_res, int *reason, double *stats_h) {
  LMH_CHECK(A && b_d && x_d, LMH_ERR_INVALID, "null argument");
  LMH_CHECK(A->have32 && A->vals32, LMH_ERR_INVALID, "mixed solve needs the FP32 operator (assemble FP32 or make_fp32)");
  LMH_CHECK(mode == 0 || mode == 1, LMH_ERR_INVALID, "mode must be 0 (reliable update) or 1 (restarted refinement)");
  LMH_CHECK(rtol > 0 && inner_rtol > 0 && inner_rtol < 1, LMH_ERR_INVALID, "bad tolerances");
  lmh_context *ctx = A->ctx;
  LMH_TRY(ensure_work(A, true));
  const int64_t n = A->rows;
  float *xs = A->w32, *r32 = xs + n, *p32 = r32 + n, *q32 = p32 + n, *d32 = q32 + n;
  double *S = ctx->d_scal, *part = ctx->d_part;
  const int g = vec_grid(ctx);
  double init[32] = {0};
  init[S_THR] = rtol * rtol;
  init[S_DELTA2] = inner_rtol * inner_rtol;
  init[S_MODE] = mode;
  LMH_HIP(hipMemcpyAsync(S, init, sizeof(init), hipMemcpyHostToDevice, ctx->stream));
  LMH_HIP(hipMemsetAsync(xs, 0, sizeof(float) * n, ctx->stream));
  LMH_LAUNCH(ctx, "invdiag", (k_invdiag<float, float>), dim3(grid_for(n, 256, 8192)), dim3(256), 0, n, A->rp, A->ci,
             A->vals32, d32);
  LMH_LAUNCH(ctx, "mx_residual_fp64", k_mx_residual, dim3(g), dim3(BLK), 0, 1, A->rb, A->nrb, A->rp, A->ci,
                                A->vals, b_d, x_d, d32, r32, p32, part, S);
  LMH_LAUNCH(ctx, "fin", k_fin_mx_residual, dim3(1), dim3(BLK), 0, part, g, 1, S);
  LMH_TRY(check_launch("mixed setup"));
  auto iter = [&]() {
    LMH_LAUNCH(ctx, "mx_spmv_fp32", (k_spmv_dot<float, float>), dim3(g), dim3(BLK), 0, A->rb, A->nrb, A->rp,
                                  A->ci, A->vals32, p32, q32, part, S);
    LMH_LAUNCH(ctx, "fin", k_fin_alpha, dim3(1), dim3(BLK), 0, part, g, (int)S_RZ, S);
    LMH_LAUNCH(ctx, "mx_update_fp32", k_mx_update, dim3(g), dim3(BLK), 0, n, d32, p32, q32, xs, r32, part, S);
    LMH_LAUNCH(ctx, "fin", k_fin_mx_update, dim3(1), dim3(BLK), 0, part, g, S);
    LMH_LAUNCH(ctx, "mx_fold", k_mx_fold, dim3(g), dim3(BLK), 0, n, 0, xs, x_d, S);
    LMH_LAUNCH(ctx, "mx_residual_fp64", k_mx_residual, dim3(g), dim3(BLK), 0, 0, A->rb, A->nrb, A->rp, A->ci,
                                  A->vals, b_d, x_d, d32, r32, p32, part, S);
    LMH_LAUNCH(ctx, "fin", k_fin_mx_residual, dim3(1), dim3(BLK), 0, part, g, 0, S);
    LMH_LAUNCH(ctx, "mx_pupdate_fp32", k_mx_pupdate, dim3(g), dim3(BLK), 0, n, d32, r32, p32, S);
  };
  LMH_TRY(run_batches(ctx, maxits, 4, iter));
  double h[32];
  LMH_HIP(hipMemcpy(h, S, sizeof(h), hipMemcpyDeviceToHost));
  const bool done = h[S_DONE] != 0.0;
  if (!done) {
    // iteration limit: fold the pending correction and report the true residual
    LMH_LAUNCH(ctx, "mx_fold", k_mx_fold, dim3(g), dim3(BLK), 0, n, 1, xs, x_d, S);
    LMH_LAUNCH(ctx, "residual_fp64", k_residual, dim3(g), dim3(BLK), 0, A->rb, A->nrb, A->rp, A->ci, A->vals,
                                  b_d, x_d, A->w64, part);
    LMH_LAUNCH(ctx, "fin", k_fin_residual, dim3(1), dim3(BLK), 0, part, g, ctx->d_scal + 32);
    LMH_TRY(check_launch("mixed finish"));
    LMH_HIP(hipMemcpy(&h[S_RRTRUE], ctx->d_scal + 32 + S_RR, sizeof(double), hipMemcpyDeviceToHost));
  }
  if (iters) *iters = (int64_t)h[S_ITS];
  if (rel_res) *rel_res = h[S_BB] > 0 ? std::sqrt(h[S_RRTRUE] / h[S_BB]) : 0.0;
  if (reason) *reason = done ? (int)h[S_REASON] : LMH_DIVERGED_ITS;
  if (stats_h) {
    stats_h[0] = h[S_ITS];
    stats_h[1] = h[S_NREF];
    stats_h[2] = h[S_BB] > 0 ? std::sqrt(h[S_RRTRUE] / h[S_BB]) : 0.0;
  }
  return LMH_OK;
}

}  // extern "C"
