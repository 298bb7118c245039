// lmh_vector.hip -- whole-vector operations of the device-resident
// NumericVector: elementwise updates and deterministic reductions.
//
// Elementwise: one grid-stride pass, HBM-bound (8-24 B per entry).
// Reductions: every block reduces a fixed slice in a fixed order (wavefront
// shuffles, then the block's waves), one block folds the block partials in
// index order: the result does not depend on scheduling.
#include <algorithm>
#include <cmath>

#include "lmh_internal.h"

namespace lmh {

constexpr int VBLK = 256;

template <int OP>
__global__ __launch_bounds__(VBLK) void k_vec_op(int64_t n, double a, const double *__restrict__ x,
                                                 const double *__restrict__ y, double *__restrict__ z) {
  for (int64_t i = blockIdx.x * (int64_t)VBLK + threadIdx.x; i < n; i += (int64_t)gridDim.x * VBLK) {
    double v;
    switch (OP) {
      case LMH_VEC_SET: v = a; break;
      case LMH_VEC_ADD_SCALAR: v = z[i] + a; break;
      case LMH_VEC_SCALE: v = z[i] * a; break;
      case LMH_VEC_AXPY: v = z[i] + a * x[i]; break;
      case LMH_VEC_MUL: v = z[i] * x[i]; break;
      case LMH_VEC_DIV: v = z[i] / x[i]; break;
      case LMH_VEC_RECIPROCAL: v = 1.0 / z[i]; break;
      case LMH_VEC_ABS: v = fabs(z[i]); break;
      case LMH_VEC_PMUL: v = x[i] * y[i]; break;
      case LMH_VEC_PDIV: v = x[i] / y[i]; break;
      default: v = x[i]; break;  // LMH_VEC_COPY
    }
    z[i] = v;
  }
}

template <int OP>
__device__ inline double vred_term(const double *__restrict__ x, const double *__restrict__ y, int64_t i) {
  switch (OP) {
    case LMH_VEC_SUM: return x[i];
    case LMH_VEC_L1: return fabs(x[i]);
    case LMH_VEC_L2: return x[i] * x[i];
    case LMH_VEC_LINF: return fabs(x[i]);
    case LMH_VEC_MIN:
    case LMH_VEC_MAX: return x[i];
    case LMH_VEC_DOT: return x[i] * y[i];
    case LMH_VEC_L1_DIFF: return fabs(x[i] - y[i]);
    default: {  // LMH_VEC_L2_DIFF
      const double d = x[i] - y[i];
      return d * d;
    }
  }
}

template <int OP>
__device__ inline double vred_identity() {
  return OP == LMH_VEC_MIN ? INFINITY : (OP == LMH_VEC_MAX ? -INFINITY : 0.0);
}

template <int OP>
__device__ inline double vred_join(double a, double b) {
  if (OP == LMH_VEC_LINF || OP == LMH_VEC_MAX) return fmax(a, b);
  if (OP == LMH_VEC_MIN) return fmin(a, b);
  return a + b;
}

template <int OP>
__device__ inline double vred_block(double v) {
  __shared__ double sh[VBLK / 64];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = vred_join<OP>(v, __shfl_xor(v, o, 64));
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) sh[wv] = v;
  __syncthreads();
  double t = sh[0];
#pragma unroll
  for (int k = 1; k < VBLK / 64; ++k) t = vred_join<OP>(t, sh[k]);
  return t;
}

// block b reduces the contiguous slice [b*chunk, min(n, (b+1)*chunk))
template <int OP>
__global__ __launch_bounds__(VBLK) void k_vec_reduce(int64_t n, int64_t chunk, const double *__restrict__ x,
                                                     const double *__restrict__ y, double *__restrict__ part) {
  const int64_t lo = blockIdx.x * chunk, hi = lo + chunk < n ? lo + chunk : n;
  double acc = vred_identity<OP>();
  for (int64_t i = lo + threadIdx.x; i < hi; i += VBLK) acc = vred_join<OP>(acc, vred_term<OP>(x, y, i));
  const double t = vred_block<OP>(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

template <int OP>
__global__ __launch_bounds__(VBLK) void k_vec_reduce_fin(int G, const double *__restrict__ part,
                                                         double *__restrict__ out) {
  // thread t folds partials t, t + VBLK, ... in order, then the block folds the threads
  double acc = vred_identity<OP>();
  for (int k = threadIdx.x; k < G; k += VBLK) acc = vred_join<OP>(acc, part[k]);
  const double t = vred_block<OP>(acc);
  if (threadIdx.x == 0) *out = t;
}

template <int OP>
static void launch_op(lmh_context *ctx, int64_t n, double a, const double *x, const double *y, double *z) {
  LMH_LAUNCH(ctx, "vector_op", k_vec_op<OP>, dim3(grid_for(n, VBLK, 8192)), dim3(VBLK), 0, n, a, x, y, z);
}

template <int OP>
static void launch_reduce(lmh_context *ctx, int64_t n, const double *x, const double *y, double *out_d) {
  const int64_t G = std::min<int64_t>(std::max<int64_t>(1, (n + VBLK * 4 - 1) / (VBLK * 4)), ctx->n_part);
  const int64_t chunk = (n + G - 1) / G > 0 ? (n + G - 1) / G : 1;
  LMH_LAUNCH(ctx, "vector_reduce", k_vec_reduce<OP>, dim3((unsigned)G), dim3(VBLK), 0, n, chunk, x, y,
                  ctx->d_part);
  LMH_LAUNCH(ctx, "vector_reduce", k_vec_reduce_fin<OP>, dim3(1), dim3(VBLK), 0, (int)G, ctx->d_part, out_d);
}

}  // namespace lmh

using namespace lmh;

extern "C" {

int lmh_vector_op(lmh_context *ctx, int op, int64_t n, double a, const double *x_d, const double *y_d,
                  double *z_d) {
  LMH_CHECK(ctx && z_d && n >= 0, LMH_ERR_INVALID, "lmh_vector_op: bad argument");
  const bool needs_x = op == LMH_VEC_AXPY || op == LMH_VEC_MUL || op == LMH_VEC_DIV || op == LMH_VEC_PMUL ||
                       op == LMH_VEC_PDIV || op == LMH_VEC_COPY;
  const bool needs_y = op == LMH_VEC_PMUL || op == LMH_VEC_PDIV;
  LMH_CHECK(op >= LMH_VEC_SET && op <= LMH_VEC_COPY, LMH_ERR_INVALID, "lmh_vector_op: unknown operation %d", op);
  LMH_CHECK((!needs_x || x_d) && (!needs_y || y_d), LMH_ERR_INVALID, "lmh_vector_op: missing operand");
  if (n == 0) return LMH_OK;
  switch (op) {
    case LMH_VEC_SET: launch_op<LMH_VEC_SET>(ctx, n, a, x_d, y_d, z_d); break;
    case LMH_VEC_ADD_SCALAR: launch_op<LMH_VEC_ADD_SCALAR>(ctx, n, a, x_d, y_d, z_d); break;
    case LMH_VEC_SCALE: launch_op<LMH_VEC_SCALE>(ctx, n, a, x_d, y_d, z_d); break;
    case LMH_VEC_AXPY: launch_op<LMH_VEC_AXPY>(ctx, n, a, x_d, y_d, z_d); break;
    case LMH_VEC_MUL: launch_op<LMH_VEC_MUL>(ctx, n, a, x_d, y_d, z_d); break;
    case LMH_VEC_DIV: launch_op<LMH_VEC_DIV>(ctx, n, a, x_d, y_d, z_d); break;
    case LMH_VEC_RECIPROCAL: launch_op<LMH_VEC_RECIPROCAL>(ctx, n, a, x_d, y_d, z_d); break;
    case LMH_VEC_ABS: launch_op<LMH_VEC_ABS>(ctx, n, a, x_d, y_d, z_d); break;
    case LMH_VEC_PMUL: launch_op<LMH_VEC_PMUL>(ctx, n, a, x_d, y_d, z_d); break;
    case LMH_VEC_PDIV: launch_op<LMH_VEC_PDIV>(ctx, n, a, x_d, y_d, z_d); break;
    default: launch_op<LMH_VEC_COPY>(ctx, n, a, x_d, y_d, z_d); break;
  }
  return check_launch("k_vec_op");
}

int lmh_vector_reduce(lmh_context *ctx, int op, int64_t n, const double *x_d, const double *y_d, double *out_h) {
  LMH_CHECK(ctx && out_h && n >= 0, LMH_ERR_INVALID, "lmh_vector_reduce: bad argument");
  LMH_CHECK(op >= LMH_VEC_SUM && op <= LMH_VEC_L2_DIFF, LMH_ERR_INVALID, "lmh_vector_reduce: unknown reduction %d",
            op);
  const bool needs_y = op == LMH_VEC_DOT || op == LMH_VEC_L1_DIFF || op == LMH_VEC_L2_DIFF;
  LMH_CHECK((n == 0 || x_d) && (n == 0 || !needs_y || y_d), LMH_ERR_INVALID, "lmh_vector_reduce: missing operand");
  if (n == 0) {
    *out_h = op == LMH_VEC_MIN ? INFINITY : (op == LMH_VEC_MAX ? -INFINITY : 0.0);
    return LMH_OK;
  }
  double *out_d = ctx->d_scal + 57;
  switch (op) {
    case LMH_VEC_SUM: launch_reduce<LMH_VEC_SUM>(ctx, n, x_d, y_d, out_d); break;
    case LMH_VEC_L1: launch_reduce<LMH_VEC_L1>(ctx, n, x_d, y_d, out_d); break;
    case LMH_VEC_L2: launch_reduce<LMH_VEC_L2>(ctx, n, x_d, y_d, out_d); break;
    case LMH_VEC_LINF: launch_reduce<LMH_VEC_LINF>(ctx, n, x_d, y_d, out_d); break;
    case LMH_VEC_MIN: launch_reduce<LMH_VEC_MIN>(ctx, n, x_d, y_d, out_d); break;
    case LMH_VEC_MAX: launch_reduce<LMH_VEC_MAX>(ctx, n, x_d, y_d, out_d); break;
    case LMH_VEC_DOT: launch_reduce<LMH_VEC_DOT>(ctx, n, x_d, y_d, out_d); break;
    case LMH_VEC_L1_DIFF: launch_reduce<LMH_VEC_L1_DIFF>(ctx, n, x_d, y_d, out_d); break;
    default: launch_reduce<LMH_VEC_L2_DIFF>(ctx, n, x_d, y_d, out_d); break;
  }
  LMH_TRY(check_launch("k_vec_reduce"));
  LMH_HIP(hipMemcpyAsync(ctx->h_pinned, out_d, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  LMH_HIP(hipStreamSynchronize(ctx->stream));
  const double v = ctx->h_pinned[0];
  *out_h = (op == LMH_VEC_L2 || op == LMH_VEC_L2_DIFF) ? std::sqrt(v) : v;
  return LMH_OK;
}

}  // extern "C"
