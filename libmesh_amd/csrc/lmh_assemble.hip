// lmh_assemble.hip -- batched FE::reinit + element contraction + element
// constraints + deterministic CSR scatter (the body of System::assemble for
// -div(kappa grad u) = f), plus the QoI derivative and adjoint kernels.
//
// Per element e and quadrature point q (fe_map.C:958-1090, h1_fe_transformation.C:64):
//   J_q[r][c] = sum_a X_a[c] dphi_a/dxi_r(q),  det_q = det J_q,  G_q = J_q^{-1}
//   JxW_q = det_q w_q,  dphi_a(q) = G_q^T dref_a(q)
// so Ke_ij = sum_q JxW_q kappa_e dphi_i(q).dphi_j(q)
//          = sum_q dref_i(q)^T M_q dref_j(q),   M_q = JxW_q kappa_e G_q G_q^T,
// i.e. Ke = Gref^T H_e with Gref[(q,r)][i] = dref_i(q)[r] constant over elements
// and H_e[(q,r)][j] = (M_q dref_j(q))[r].  Geometry (J, det, inverse) is FP64.
// The contraction runs on MFMA for HEX27 (v_mfma_f64_16x16x4f64 / f32), on
// wavefront FMA for HEX8, and on a generic per-element kernel otherwise.
#include <cmath>

#include "lmh_internal.h"

namespace lmh {

typedef double double4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

__device__ inline double forcing(int f_kind, double x, double y, double z) {
  if (f_kind == 0) return 1.0;
  const double pi = 3.14159265358979323846264338327950288;
  return 3.0 * pi * pi * sin(pi * x) * sin(pi * y) * sin(pi * z);
}

// Geometry at one quadrature point from LDS coordinates X[a*3+c] and the
// reference derivative table dref[(a*nq+q)*3+r].  Writes M (9, symmetric,
// scaled by JxW*kappa) and JxW (which is negative for the negative-weight tet
// rule), and returns det J (<= 0 signals a degenerate map, fe_map.C:467-505).
__device__ inline double qp_geometry(const double *X, const double *dref, int n, int nq, int q, double w,
                                     double kap, double *M, double &jxw_out) {
  double G[3][3];
  const double det = qp_map(X, dref, n, nq, q, G);
  jxw_out = 0.0;
  if (!(det > 0.0)) return -1.0;
  const double jxw = det * w;
  const double s = jxw * kap;
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int t = r; t < 3; ++t) {
      const double v = s * (G[r][0] * G[t][0] + G[r][1] * G[t][1] + G[r][2] * G[t][2]);
      M[r * 3 + t] = v;
      M[t * 3 + r] = v;
    }
  jxw_out = jxw;
  return det;
}

// ---------------------------------------------------------------------------
// Generic Poisson element kernel: one wave per element (any type, nq <= 64).
// ---------------------------------------------------------------------------
template <typename CT>
__global__ __launch_bounds__(64) void k_poisson_generic(const double *__restrict__ tab, int n, int nq,
                                                        const double *__restrict__ xyz, const int32_t *__restrict__ conn,
                                                        const int32_t *__restrict__ edofs,
                                                        const uint8_t *__restrict__ cons,
                                                        const double *__restrict__ kappa, int f_kind, int constrain,
                                                        int64_t ne, CT *__restrict__ Ke, double *__restrict__ Fe,
                                                        int *__restrict__ err) {
  extern __shared__ double sm[];
  const int lane = threadIdx.x;
  double *w = sm;
  double *phi = w + nq;
  double *dref = phi + n * nq;
  double *X = dref + n * nq * 3;
  double *M = X + 81;
  double *Jf = M + 9 * nq;
  for (int t = lane; t < nq + n * nq * 4; t += 64) sm[t] = tab[t];
  __syncthreads();
  for (int64_t e = blockIdx.x; e < ne; e += gridDim.x) {
    if (lane < n) {
      const int32_t nd = conn[e * n + lane];
      X[lane * 3] = xyz[3 * (int64_t)nd];
      X[lane * 3 + 1] = xyz[3 * (int64_t)nd + 1];
      X[lane * 3 + 2] = xyz[3 * (int64_t)nd + 2];
    }
    const bool ci = (lane < n) ? (cons[edofs[e * n + lane]] != 0) : false;
    const unsigned long long cmask = constrain ? __ballot(ci) : 0ull;
    const double kap = kappa ? kappa[e] : 1.0;
    __syncthreads();
    if (lane < nq) {
      double Mq[9];
      double jxw;
      if (qp_geometry(X, dref, n, nq, lane, w[lane], kap, Mq, jxw) <= 0.0) atomicOr(err, 1);
      for (int t = 0; t < 9; ++t) M[lane * 9 + t] = Mq[t];
      double f = 1.0;
      if (f_kind != 0) {
        double x = 0, y = 0, z = 0;
        for (int a = 0; a < n; ++a) {
          const double p = phi[a * nq + lane];
          x += X[a * 3] * p;
          y += X[a * 3 + 1] * p;
          z += X[a * 3 + 2] * p;
        }
        f = forcing(f_kind, x, y, z);
      }
      Jf[lane] = jxw * f;
    }
    __syncthreads();
    if (lane < n) {
      double row[27];
      for (int j = 0; j < n; ++j) row[j] = 0.0;
      for (int q = 0; q < nq; ++q) {
        const double *di = &dref[(lane * nq + q) * 3];
        const double *Mq = &M[q * 9];
        double v[3];
        for (int s = 0; s < 3; ++s) v[s] = di[0] * Mq[s] + di[1] * Mq[3 + s] + di[2] * Mq[6 + s];
        for (int j = 0; j < n; ++j) {
          const double *dj = &dref[(j * nq + q) * 3];
          row[j] += (CT)v[0] * (CT)dj[0] + (CT)v[1] * (CT)dj[1] + (CT)v[2] * (CT)dj[2];
        }
      }
      double fe = 0.0;
      for (int q = 0; q < nq; ++q) fe += Jf[q] * phi[lane * nq + q];
      const bool myc = (cmask >> lane) & 1ull;
      for (int j = 0; j < n; ++j) {
        double v = row[j];
        if (myc || ((cmask >> j) & 1ull)) v = (j == lane && myc) ? 1.0 : 0.0;
        Ke[(e * n + lane) * n + j] = (CT)v;
      }
      Fe[e * n + lane] = myc ? 0.0 : fe;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Q1 HEX8, 2x2x2 Gauss: wavefront-level FMA, 8 elements per wave, lane = 8*el + t
// (t = quadrature point in the geometry phase and matrix row in the contraction).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_poisson_hex8(const double *__restrict__ tab, const double *__restrict__ xyz,
                                                      const int32_t *__restrict__ conn,
                                                      const int32_t *__restrict__ edofs,
                                                      const uint8_t *__restrict__ cons,
                                                      const double *__restrict__ kappa, int f_kind, int constrain,
                                                      int64_t ne, double *__restrict__ Ke, double *__restrict__ Fe,
                                                      int *__restrict__ err) {
  __shared__ double s_tab[8 + 64 + 192];
  __shared__ double s_X[4][8][24];
  __shared__ double s_M[4][8][8 * 9];
  __shared__ double s_Jf[4][8][8];
  const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63, el = lane >> 3, t = lane & 7;
  for (int i = tid; i < 8 + 64 + 192; i += 256) s_tab[i] = tab[i];
  __syncthreads();
  const double *w = s_tab, *phi = s_tab + 8, *dref = s_tab + 72;
  const int64_t groups = (ne + 31) / 32;
  for (int64_t g = blockIdx.x; g < groups; g += gridDim.x) {
    const int64_t e = g * 32 + wv * 8 + el;
    const bool valid = e < ne;
    double *X = s_X[wv][el];
    bool ci = false;
    double kap = 1.0;
    if (valid) {
      const int32_t nd = conn[e * 8 + t];
      X[t * 3] = xyz[3 * (int64_t)nd];
      X[t * 3 + 1] = xyz[3 * (int64_t)nd + 1];
      X[t * 3 + 2] = xyz[3 * (int64_t)nd + 2];
      ci = cons[edofs[e * 8 + t]] != 0;
      if (kappa) kap = kappa[e];
    }
    const unsigned long long bal = __ballot(ci);
    const unsigned cm = constrain ? (unsigned)((bal >> (8 * el)) & 0xffull) : 0u;
    __syncthreads();
    if (valid) {
      double Mq[9];
      double jxw;
      if (qp_geometry(X, dref, 8, 8, t, w[t], kap, Mq, jxw) <= 0.0) atomicOr(err, 1);
      for (int k = 0; k < 9; ++k) s_M[wv][el][t * 9 + k] = Mq[k];
      double f = 1.0;
      if (f_kind != 0) {
        double x = 0, y = 0, z = 0;
        for (int a = 0; a < 8; ++a) {
          const double p = phi[a * 8 + t];
          x += X[a * 3] * p;
          y += X[a * 3 + 1] * p;
          z += X[a * 3 + 2] * p;
        }
        f = forcing(f_kind, x, y, z);
      }
      s_Jf[wv][el][t] = jxw * f;
    }
    __syncthreads();
    if (valid) {
      double row[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      const double *Mb = s_M[wv][el];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const double *di = &dref[(t * 8 + q) * 3];
        const double *Mq = &Mb[q * 9];
        double v[3];
#pragma unroll
        for (int s = 0; s < 3; ++s) v[s] = di[0] * Mq[s] + di[1] * Mq[3 + s] + di[2] * Mq[6 + s];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const double *dj = &dref[(j * 8 + q) * 3];
          row[j] += v[0] * dj[0] + v[1] * dj[1] + v[2] * dj[2];
        }
      }
      double fe = 0.0;
#pragma unroll
      for (int q = 0; q < 8; ++q) fe += s_Jf[wv][el][q] * phi[t * 8 + q];
      const bool myc = (cm >> t) & 1u;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        double v = row[j];
        if (myc || ((cm >> j) & 1u)) v = (j == t && myc) ? 1.0 : 0.0;
        Ke[(e * 8 + t) * 8 + j] = v;
      }
      Fe[e * 8 + t] = myc ? 0.0 : fe;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Q2 HEX27, 3x3x3 Gauss, MFMA contraction.  One element per wave, 4 waves per
// block, persistent.  Ke (27x27, padded 32x32) = Gref^T (32x84) * H (84x32):
// all four 16x16 tiles run on v_mfma_*_16x16x4 (21 k-steps each), so every
// lane stores its accumulators straight to Ke in row-major order (16
// consecutive entries per lane group: coalesced), with no LDS staging of Ke.
// MFMA operand layout (16x16x4): lane l supplies A[l%16][l/16] and
// B[l/16][l%16]; accumulator element r of lane l holds D[drow(l/16, r)][l%16]
// (see MfmaT::drow).  The constant Gref^T operand is read per k-step from the
// LDS derivative table; the shape values for Fe come from the FE table in
// global memory (27x27, cache resident).  LDS holds the derivative table and
// per wave the coordinates, the 27 M_q and JxW_q f_q: 29 KB per block, so the
// occupancy is set by registers, not LDS.  (Measured: per-wave instead of
// block barriers let the waves drift apart and made the FP64 kernel slower,
// 1.65 -> 2.22 ms at 64^3.)
// ---------------------------------------------------------------------------
template <typename CT>
struct MfmaT;
template <>
struct MfmaT<double> {
  typedef double4_t acc_t;
  __device__ static inline acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // f64 16x16x4: the four accumulator registers of lane group g hold rows g, g+4, g+8, g+12
  __device__ static inline int drow(int g, int r) { return g + 4 * r; }
};
template <>
struct MfmaT<float> {
  typedef float4_t acc_t;
  __device__ static inline acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  // f32 16x16x4: lane group g holds rows 4g .. 4g+3
  __device__ static inline int drow(int g, int r) { return 4 * g + r; }
};

constexpr int Q2N = 27, Q2Q = 27, Q2K = 81, Q2KS = 21;

// The forcing kind is a template argument: the f = 1 launch carries none of
// the FP64 sine's code or registers.
template <typename CT, int FK>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(FK ? 2 : 3))) void k_poisson_hex27_mfma(
    const double *__restrict__ tab, const double *__restrict__ xyz, const int32_t *__restrict__ conn,
    const int32_t *__restrict__ edofs, const uint8_t *__restrict__ cons, const double *__restrict__ kappa,
    int constrain, int64_t ne, CT *__restrict__ Ke, double *__restrict__ Fe, int *__restrict__ err) {
  typedef typename MfmaT<CT>::acc_t acc_t;
  __shared__ double s_w[Q2Q];
  __shared__ double s_dref[Q2N * Q2Q * 3];
  __shared__ double s_X[4][Q2N * 3];
  __shared__ double s_M[4][Q2Q * 9];
  __shared__ double s_Jf[4][Q2Q];
  const double *g_phi = tab + Q2Q;  // phi[a * nq + q]
  const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
  for (int i = tid; i < Q2Q; i += 256) s_w[i] = tab[i];
  for (int i = tid; i < Q2N * Q2Q * 3; i += 256) s_dref[i] = tab[Q2Q + Q2N * Q2Q + i];
  __syncthreads();

  double *X = s_X[wv];
  double *M = s_M[wv];
  double *Jf = s_Jf[wv];
  const int64_t groups = (ne + 3) / 4;
  for (int64_t g = blockIdx.x; g < groups; g += gridDim.x) {
    const int64_t e = g * 4 + wv;
    const bool valid = e < ne;
    bool ci = false;
    double kap = 1.0;
    if (valid) {
      if (lane < Q2N) {
        const int32_t nd = conn[e * Q2N + lane];
        X[lane * 3] = xyz[3 * (int64_t)nd];
        X[lane * 3 + 1] = xyz[3 * (int64_t)nd + 1];
        X[lane * 3 + 2] = xyz[3 * (int64_t)nd + 2];
        ci = cons[edofs[e * Q2N + lane]] != 0;
      }
      if (kappa) kap = kappa[e];
    }
    const unsigned long long cmask = constrain ? __ballot(ci) : 0ull;
    __syncthreads();
    if (valid && lane < Q2Q) {
      double Mq[9];
      double jxw;
      if (qp_geometry(X, s_dref, Q2N, Q2Q, lane, s_w[lane], kap, Mq, jxw) <= 0.0) atomicOr(err, 1);
#pragma unroll
      for (int t = 0; t < 9; ++t) M[lane * 9 + t] = Mq[t];
      double f = 1.0;
      if constexpr (FK != 0) {
        double x = 0, y = 0, z = 0;
        for (int a = 0; a < Q2N; ++a) {
          const double p = g_phi[a * Q2Q + lane];
          x += X[a * 3] * p;
          y += X[a * 3 + 1] * p;
          z += X[a * 3 + 2] * p;
        }
        f = forcing(FK, x, y, z);
      }
      Jf[lane] = jxw * f;
    }
    __syncthreads();
    if (valid) {
      acc_t c00 = {0, 0, 0, 0}, c01 = {0, 0, 0, 0}, c10 = {0, 0, 0, 0}, c11 = {0, 0, 0, 0};
      const int li = lane & 15, lk = lane >> 4;
      const int j0 = li, j1 = 16 + li;
#pragma unroll 1
      for (int ks = 0; ks < Q2KS; ++ks) {
        const int k = 4 * ks + lk;
        double b0 = 0.0, b1 = 0.0;
        CT a0 = 0, a1 = 0;
        if (k < Q2K) {
          const int q = k / 3, r = k - 3 * (k / 3);
          const double m0 = M[q * 9 + r * 3], m1 = M[q * 9 + r * 3 + 1], m2 = M[q * 9 + r * 3 + 2];
          const double *d0 = &s_dref[(j0 * Q2Q + q) * 3];
          b0 = m0 * d0[0] + m1 * d0[1] + m2 * d0[2];
          a0 = (CT)d0[r];
          if (j1 < Q2N) {
            const double *d1 = &s_dref[(j1 * Q2Q + q) * 3];
            b1 = m0 * d1[0] + m1 * d1[1] + m2 * d1[2];
            a1 = (CT)d1[r];
          }
        }
        // A[I][ks] = Gref^T[16I + l%16][4ks + l/16] (the same LDS reads as the B operand's)
        c00 = MfmaT<CT>::mma(a0, (CT)b0, c00);
        c01 = MfmaT<CT>::mma(a0, (CT)b1, c01);
        c10 = MfmaT<CT>::mma(a1, (CT)b0, c10);
        c11 = MfmaT<CT>::mma(a1, (CT)b1, c11);
      }
      CT *Kr = Ke + e * (Q2N * Q2N);
      const bool cj0 = (cmask >> j0) & 1ull, cj1 = j1 < Q2N && ((cmask >> j1) & 1ull);
      auto put = [&](int i, int j, bool cj, CT v) {
        const bool ri = (cmask >> i) & 1ull;
        if (ri || cj) v = (i == j && ri) ? (CT)1 : (CT)0;
        Kr[i * Q2N + j] = v;
      };
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i0 = MfmaT<CT>::drow(lk, r), i1 = 16 + i0;
        put(i0, j0, cj0, c00[r]);
        if (j1 < Q2N) put(i0, j1, cj1, c01[r]);
        if (i1 < Q2N) {
          put(i1, j0, cj0, c10[r]);
          if (j1 < Q2N) put(i1, j1, cj1, c11[r]);
        }
      }
      if (lane < Q2N) {
        double fe = 0.0;
        for (int q = 0; q < Q2Q; ++q) fe += Jf[q] * g_phi[lane * Q2Q + q];
        Fe[e * Q2N + lane] = ((cmask >> lane) & 1ull) ? 0.0 : fe;
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Scatter Ke/Fe -> CSR values / rhs.
// Gather form: one wave per row; the row's slots accumulate in LDS over the
// row's incidences in ascending (element, local index) order -- the same
// summation order as the element loop of the reference -- then the row is
// written once.  Coloured form: one launch per colour, every (element, i, j)
// adds into its slot; elements of one colour share no DoF.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(64) void k_scatter_gather(const int32_t *__restrict__ inc_ptr,
                                                       const int32_t *__restrict__ inc, int dpe, int64_t n_rows,
                                                       const int32_t *__restrict__ rp, const uint16_t *__restrict__ off,
                                                       const T *__restrict__ Ke, const double *__restrict__ Fe,
                                                       T *__restrict__ vals, double *__restrict__ rhs) {
  extern __shared__ unsigned char smraw[];
  T *acc = (T *)smraw;
  const int lane = threadIdx.x;
  int64_t rlo, rhi;
  block_range(n_rows, rlo, rhi);  // contiguous rows per workgroup, per XCD
  for (int64_t r = rlo; r < rhi; ++r) {
    const int32_t b = rp[r], len = rp[r + 1] - b;
    for (int t = lane; t < len; t += 64) acc[t] = (T)0;
    __syncthreads();
    const int32_t k0 = inc_ptr[r], k1 = inc_ptr[r + 1];
    double f = 0.0;
    for (int32_t k = k0; k < k1; ++k) {
      const int64_t ei = inc[k];  // e*dpe + i
      const int64_t base = ei * dpe;
      for (int j = lane; j < dpe; j += 64) acc[off[base + j]] += Ke[base + j];
      if (rhs) f += Fe[ei];
      __syncthreads();
    }
    for (int t = lane; t < len; t += 64) vals[b + t] = acc[t];
    if (rhs && lane == 0) rhs[r] = f;
    __syncthreads();
  }
}

// The same gather for element DoF counts up to 32: each half-wave owns a row
// (two rows per wave), and a row's incidences are taken in chunks of GC
// whose element matrix rows and slot offsets are all loaded before any is
// added, so a row waits on one round of loads instead of one per incidence.
// The adds still go in ascending incidence order (one wave executes its LDS
// operations in program order, so no barrier is needed between them): the
// sums are bit-identical to k_scatter_gather's.
template <typename T>
__global__ __launch_bounds__(64) void k_scatter_gather2(const int32_t *__restrict__ inc_ptr,
                                                        const int32_t *__restrict__ inc, int dpe, int64_t n_rows,
                                                        const int32_t *__restrict__ rp,
                                                        const uint16_t *__restrict__ off, const T *__restrict__ Ke,
                                                        const double *__restrict__ Fe, T *__restrict__ vals,
                                                        double *__restrict__ rhs, int stride) {
  constexpr int GC = 8;
  extern __shared__ unsigned char smraw[];
  const int lane = threadIdx.x, h = lane >> 5, hl = lane & 31;
  T *acc = (T *)smraw + h * stride;
  int64_t rlo, rhi;
  block_range(n_rows, rlo, rhi);
  for (int64_t rb = rlo; rb < rhi; rb += 2) {
    const int64_t r = rb + h;
    const bool live = r < rhi;
    const int32_t b = live ? rp[r] : 0, len = live ? rp[r + 1] - b : 0;
    for (int t = hl; t < len; t += 32) acc[t] = (T)0;
    const int32_t k0 = live ? inc_ptr[r] : 0, k1 = live ? inc_ptr[r + 1] : 0;
    double f = 0.0;
    for (int32_t kc = k0; kc < k1; kc += GC) {
      const int nc = k1 - kc < GC ? k1 - kc : GC;
      int64_t ei[GC];
      T v[GC];
      uint16_t o[GC];
#pragma unroll
      for (int c = 0; c < GC; ++c) ei[c] = inc[kc + (c < nc ? c : 0)];
#pragma unroll
      for (int c = 0; c < GC; ++c) {
        const int64_t at = ei[c] * dpe + (hl < dpe ? hl : 0);
        v[c] = Ke[at];
        o[c] = off[at];
      }
      double fv[GC];
      if (rhs)
#pragma unroll
        for (int c = 0; c < GC; ++c) fv[c] = Fe[ei[c]];
#pragma unroll
      for (int c = 0; c < GC; ++c)
        if (c < nc) {
          if (hl < dpe) acc[o[c]] += v[c];
          if (rhs) f += fv[c];
        }
    }
    for (int t = hl; t < len; t += 32) vals[b + t] = acc[t];
    if (rhs && live && hl == 0) rhs[r] = f;
  }
}

template <typename T>
__global__ void k_scatter_color(const int32_t *__restrict__ elems, int64_t n, int dpe,
                                const int32_t *__restrict__ edofs, const int32_t *__restrict__ rp,
                                const uint16_t *__restrict__ off, const T *__restrict__ Ke,
                                const double *__restrict__ Fe, T *__restrict__ vals, double *__restrict__ rhs) {
  const int64_t dd = (int64_t)dpe * dpe;
  const int64_t total = n * dd;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t k = t / dd;
    const int rem = (int)(t - k * dd);
    const int i = rem / dpe, j = rem - i * dpe;
    const int64_t e = elems[k];
    const int32_t r = edofs[e * dpe + i];
    const int64_t ij = e * dd + rem;
    vals[rp[r] + off[ij]] += Ke[ij];
    if (rhs && j == 0) rhs[r] += Fe[e * dpe + i];
  }
}

template <typename T>
__global__ void k_fill(T *__restrict__ p, int64_t n, T v) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) p[t] = v;
}

// generic add of a batch of element matrices (binary search for the slot),
// gather form over the batch's own incidence lists; accumulates onto vals
__global__ __launch_bounds__(64) void k_add_gather_bs(const int32_t *__restrict__ inc_ptr,
                                                      const int32_t *__restrict__ inc, int n, int64_t n_rows,
                                                      const int32_t *__restrict__ dofs, const int32_t *__restrict__ rp,
                                                      const int32_t *__restrict__ ci, const double *__restrict__ Ke,
                                                      double *__restrict__ vals, int *__restrict__ err) {
  extern __shared__ double acc[];
  const int lane = threadIdx.x;
  for (int64_t r = blockIdx.x; r < n_rows; r += gridDim.x) {
    const int32_t k0 = inc_ptr[r], k1 = inc_ptr[r + 1];
    if (k0 == k1) continue;
    const int32_t b = rp[r], len = rp[r + 1] - b;
    for (int t = lane; t < len; t += 64) acc[t] = vals[b + t];
    __syncthreads();
    for (int32_t k = k0; k < k1; ++k) {
      const int64_t ei = inc[k];
      const int64_t e = ei / n;
      for (int j = lane; j < n; j += 64) {
        const int32_t c = dofs[e * n + j];
        int32_t lo = b, hi = b + len;
        while (lo < hi) {
          const int32_t mid = (lo + hi) >> 1;
          if (ci[mid] < c)
            lo = mid + 1;
          else
            hi = mid;
        }
        if (lo < b + len && ci[lo] == c)
          acc[lo - b] += Ke[ei * n + j];
        else
          atomicOr(err, 1);
      }
      __syncthreads();
    }
    for (int t = lane; t < len; t += 64) vals[b + t] = acc[t];
    __syncthreads();
  }
}

__global__ void k_check_range(const int32_t *__restrict__ v, int64_t n, int64_t hi, int *__restrict__ err) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
    if (v[t] < 0 || v[t] >= hi) atomicOr(err, 2);
}

__global__ void k_add_vec_gather(const int32_t *__restrict__ inc_ptr, const int32_t *__restrict__ inc, int64_t n_rows,
                                 const double *__restrict__ Fe, double *__restrict__ v) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n_rows; r += (int64_t)gridDim.x * blockDim.x) {
    const int32_t k0 = inc_ptr[r], k1 = inc_ptr[r + 1];
    if (k0 == k1) continue;
    double s = v[r];
    for (int32_t k = k0; k < k1; ++k) s += Fe[inc[k]];
    v[r] = s;
  }
}

// Element residuals of the Poisson physics, R_e = K_e u_e - F_e from the
// unconstrained element matrices, with C^T applied for empty constraint rows
// (a constrained DoF's entry is 0); written over Fe.  One wave per element:
// the element matrix is staged in LDS with coalesced loads, lane i forms row i.
__global__ __launch_bounds__(256) void k_elem_residual(const double *__restrict__ Ke, double *__restrict__ Fe,
                                                       const int32_t *__restrict__ edofs,
                                                       const uint8_t *__restrict__ cons,
                                                       const double *__restrict__ u, int dpe, int64_t ne) {
  extern __shared__ double rsm[];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int dd = dpe * dpe;
  double *K = rsm + wv * (dd + dpe), *ue = K + dd;
  for (int64_t e = (int64_t)blockIdx.x * 4 + wv; e < ne; e += (int64_t)gridDim.x * 4) {
    const double *src = Ke + e * dd;
    for (int t = lane; t < dd; t += 64) K[t] = src[t];
    const int32_t di = lane < dpe ? edofs[e * dpe + lane] : 0;
    if (lane < dpe) ue[lane] = u[di];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (lane < dpe) {
      double s = 0.0;
      for (int j = 0; j < dpe; ++j) s += K[lane * dpe + j] * ue[j];
      const int64_t t = e * dpe + lane;
      Fe[t] = cons[di] ? 0.0 : s - Fe[t];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

// ---------------------------------------------------------------------------
// QoI derivative element vectors: Q_i = sum_{q: x_q in box} JxW_q phi_i(q)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_qoi_box(const double *__restrict__ tab, int n, int nq,
                                                const double *__restrict__ xyz, const int32_t *__restrict__ conn,
                                                int64_t ne, double lo, double hi, double *__restrict__ Q,
                                                int *__restrict__ err) {
  extern __shared__ double sm[];
  const int lane = threadIdx.x;
  double *w = sm, *phi = w + nq, *dref = phi + n * nq, *X = dref + n * nq * 3, *Jc = X + 81;
  for (int t = lane; t < nq + n * nq * 4; t += 64) sm[t] = tab[t];
  __syncthreads();
  for (int64_t e = blockIdx.x; e < ne; e += gridDim.x) {
    if (lane < n) {
      const int32_t nd = conn[e * n + lane];
      X[lane * 3] = xyz[3 * (int64_t)nd];
      X[lane * 3 + 1] = xyz[3 * (int64_t)nd + 1];
      X[lane * 3 + 2] = xyz[3 * (int64_t)nd + 2];
    }
    __syncthreads();
    if (lane < nq) {
      double Mq[9];
      double jxw;
      if (qp_geometry(X, dref, n, nq, lane, w[lane], 1.0, Mq, jxw) <= 0.0) atomicOr(err, 1);
      double x = 0, y = 0, z = 0;
      for (int a = 0; a < n; ++a) {
        const double p = phi[a * nq + lane];
        x += X[a * 3] * p;
        y += X[a * 3 + 1] * p;
        z += X[a * 3 + 2] * p;
      }
      const bool in = x >= lo && x <= hi && y >= lo && y <= hi && z >= lo && z <= hi;
      Jc[lane] = in ? jxw : 0.0;
    }
    __syncthreads();
    if (lane < n) {
      double s = 0.0;
      for (int q = 0; q < nq; ++q) s += Jc[q] * phi[lane * nq + q];
      Q[e * n + lane] = s;
    }
    __syncthreads();
  }
}

__global__ void k_zero_constrained(double *__restrict__ v, const uint8_t *__restrict__ cons, int64_t n) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x)
    if (cons[r]) v[r] = 0.0;
}

// ---------------------------------------------------------------------------
// Adjoint backward kernel: dJ/dkappa_e = -lambda_e^T (dKc_e/dkappa_e) u_e.
// Kc_e is linear in kappa_e on free rows/columns and kappa-independent on
// constrained ones, and u, lambda vanish on constrained DoFs, so
// dJ/dkappa_e = -(1/kappa_e) lambda_e^T Ke_e u_e = -sum_q JxW_q grad(lambda).grad(u)
// with the nodal values masked to the free DoFs.  One wave per element.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_adjoint_kappa(const double *__restrict__ tab, int n, int nq,
                                                      const double *__restrict__ xyz, const int32_t *__restrict__ conn,
                                                      const int32_t *__restrict__ edofs,
                                                      const uint8_t *__restrict__ cons, const double *__restrict__ u,
                                                      const double *__restrict__ lam, int64_t ne,
                                                      double *__restrict__ grad, int *__restrict__ err) {
  extern __shared__ double sm[];
  const int lane = threadIdx.x;
  double *w = sm, *phi = w + nq, *dref = phi + n * nq, *X = dref + n * nq * 3, *U = X + 81, *L = U + 27,
         *part = L + 27;
  for (int t = lane; t < nq + n * nq * 4; t += 64) sm[t] = tab[t];
  __syncthreads();
  for (int64_t e = blockIdx.x; e < ne; e += gridDim.x) {
    if (lane < n) {
      const int32_t nd = conn[e * n + lane];
      X[lane * 3] = xyz[3 * (int64_t)nd];
      X[lane * 3 + 1] = xyz[3 * (int64_t)nd + 1];
      X[lane * 3 + 2] = xyz[3 * (int64_t)nd + 2];
      const int32_t d = edofs[e * n + lane];
      const bool c = cons[d] != 0;
      U[lane] = c ? 0.0 : u[d];
      L[lane] = c ? 0.0 : lam[d];
    }
    __syncthreads();
    double s = 0.0;
    if (lane < nq) {
      double Mq[9];
      double jxw;
      if (qp_geometry(X, dref, n, nq, lane, w[lane], 1.0, Mq, jxw) <= 0.0) atomicOr(err, 1);
      double gu[3] = {0, 0, 0}, gl[3] = {0, 0, 0};
      for (int a = 0; a < n; ++a) {
        const double *d = &dref[(a * nq + lane) * 3];
        for (int r = 0; r < 3; ++r) {
          gu[r] += U[a] * d[r];
          gl[r] += L[a] * d[r];
        }
      }
      for (int r = 0; r < 3; ++r) s += gl[r] * (Mq[r * 3] * gu[0] + Mq[r * 3 + 1] * gu[1] + Mq[r * 3 + 2] * gu[2]);
    }
    part[lane] = s;
    __syncthreads();
    if (lane == 0) {
      double t = 0.0;
      for (int q = 0; q < nq; ++q) t += part[q];
      grad[e] = -t;
    }
    __syncthreads();
  }
}

static int read_flag(lmh_context *ctx, const char *what) {
  int flag = 0;
  LMH_HIP(hipMemcpyAsync(&flag, ctx->d_flags, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  LMH_HIP(hipStreamSynchronize(ctx->stream));
  if (flag) return set_error(LMH_ERR_DEGENERATE, "%s: Jacobian determinant <= 0 (degenerate map)", what);
  return LMH_OK;
}

// v <- C^T v for general constraint rows: constrain_element_vector per element
// (R^T F with constrained entries zeroed, dof_map_constraints.C:2981-3020)
// summed over elements.  Each constrained entry is distributed to the DoFs of
// its row, then zeroed; processed rows hold no chains.  Applied on the host in
// ascending DoF order (deterministic); a no-op without general rows, where the
// element kernels' mask already zeroes the constrained entries.
static int apply_constraint_transpose(lmh_context *ctx, const lmh_dofmap *dm, double *v_d) {
  const lmh_constraints &c = dm->cr;
  if (c.n_general == 0) return LMH_OK;
  std::vector<double> h((size_t)dm->n_dofs);
  LMH_HIP(hipMemcpyAsync(h.data(), v_d, sizeof(double) * h.size(), hipMemcpyDeviceToHost, ctx->stream));
  LMH_HIP(hipStreamSynchronize(ctx->stream));
  for (size_t r = 0; r < c.dof.size(); ++r) {
    const double q = h[c.dof[r]];
    for (int32_t k = c.ptr[r]; k < c.ptr[r + 1]; ++k) h[c.col[k]] += c.coef[k] * q;
    h[c.dof[r]] = 0.0;
  }
  LMH_HIP(hipMemcpyAsync(v_d, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, ctx->stream));
  LMH_HIP(hipStreamSynchronize(ctx->stream));
  return LMH_OK;
}

// Element matrices for a scalar Poisson problem into Ke (CT) and Fe (FP64).
template <typename CT>
static int element_poisson_impl(lmh_context *ctx, const lmh_mesh *m, const lmh_dofmap *dm, int npts,
                                const double *kappa, int f_kind, int constrain, CT *Ke, double *Fe) {
  LMH_CHECK(dm->nv == 1, LMH_ERR_UNSUPPORTED, "Poisson assembly needs one variable");
  const FETable *t = nullptr;
  LMH_TRY(fe_table(ctx, m->type, npts, &t));
  LMH_CHECK(t->nq <= 64 && t->n <= 27, LMH_ERR_UNSUPPORTED, "element too large for the generic kernel");
  const int64_t ne = m->n_elem;
  const int maxg = ctx->cu_count * 8;
  if (m->type == LMH_HEX27 && npts == 3) {
    // one resident wave of workgroups: 3 per CU at the f = 1 kernel's registers, 2 with the sine
    const int g = grid_for((ne + 3) / 4, 1, ctx->cu_count * (f_kind == 0 ? 3 : 2));
    const char *name = sizeof(CT) == 8 ? "element_hex27_mfma_fp64" : "element_hex27_mfma_fp32";
    if (f_kind == 0)
      LMH_LAUNCH(ctx, name, (k_poisson_hex27_mfma<CT, 0>), dim3(g), dim3(256), 0, t->d_data, m->xyz, m->conn,
                 dm->edofs, dm->cons, kappa, constrain, ne, Ke, Fe, ctx->d_flags);
    else
      LMH_LAUNCH(ctx, name, (k_poisson_hex27_mfma<CT, 1>), dim3(g), dim3(256), 0, t->d_data, m->xyz, m->conn,
                 dm->edofs, dm->cons, kappa, constrain, ne, Ke, Fe, ctx->d_flags);
    return check_launch("k_poisson_hex27_mfma");
  }
  if (m->type == LMH_HEX8 && npts == 2 && sizeof(CT) == 8) {
    const int g = grid_for((ne + 31) / 32, 1, ctx->cu_count * 4);
    LMH_LAUNCH(ctx, "element_hex8_fma", k_poisson_hex8, dim3(g), dim3(256), 0, t->d_data, m->xyz, m->conn, dm->edofs,
               dm->cons, kappa, f_kind, constrain, ne, (double *)Ke, Fe, ctx->d_flags);
    return check_launch("k_poisson_hex8");
  }
  const size_t shm = sizeof(double) * (t->nq + (size_t)t->n * t->nq * 4 + 81 + 9 * t->nq + t->nq);
  const int g = grid_for(ne, 1, maxg);
  LMH_LAUNCH(ctx, "element_generic", k_poisson_generic<CT>, dim3(g), dim3(64), shm, t->d_data, t->n, t->nq, m->xyz,
             m->conn, dm->edofs, dm->cons, kappa, f_kind, constrain, ne, Ke, Fe, ctx->d_flags);
  return check_launch("k_poisson_generic");
}

#ifndef LMH_SCATTER_PER_CU
#define LMH_SCATTER_PER_CU 256
#endif
template <typename T>
static int scatter_impl(lmh_matrix *A, const T *Ke, const double *Fe, T *vals, double *rhs, int scatter) {
  lmh_context *ctx = A->ctx;
  const lmh_dofmap *dm = A->dm;
  const int stride = (int)((A->max_row + 1) & ~1);
  if (scatter == LMH_SCATTER_GATHER && dm->dpe <= 32 && sizeof(T) * 2 * (size_t)stride <= 32768) {
    const int g = (int)std::min<int64_t>((A->rows + 1) / 2, (int64_t)ctx->cu_count * LMH_SCATTER_PER_CU);
    LMH_LAUNCH(ctx, sizeof(T) == 8 ? "scatter_gather_fp64" : "scatter_gather_fp32", k_scatter_gather2<T>, dim3(g),
               dim3(64), sizeof(T) * 2 * stride, dm->inc_ptr, dm->inc, dm->dpe, A->rows, A->rp, A->off, Ke, Fe, vals,
               rhs, stride);
    return check_launch("k_scatter_gather2");
  }
  if (scatter == LMH_SCATTER_GATHER) {
    const int g = (int)std::min<int64_t>(A->rows, (int64_t)ctx->cu_count * LMH_SCATTER_PER_CU);
    LMH_LAUNCH(ctx, sizeof(T) == 8 ? "scatter_gather_fp64" : "scatter_gather_fp32", k_scatter_gather<T>, dim3(g),
               dim3(64), sizeof(T) * A->max_row, dm->inc_ptr, dm->inc, dm->dpe, A->rows, A->rp, A->off, Ke, Fe, vals,
               rhs);
    return check_launch("k_scatter_gather");
  }
  LMH_CHECK(scatter == LMH_SCATTER_COLORED, LMH_ERR_INVALID, "unknown scatter mode %d", scatter);
  LMH_CHECK(A->ncolors > 0 || dm->mesh->n_elem == 0, LMH_ERR_UNSUPPORTED, "matrix was created without a colouring");
  LMH_LAUNCH(ctx, "fill", k_fill<T>, dim3(grid_for(A->nnz, 256, 8192)), dim3(256), 0, vals, A->nnz, (T)0);
  if (rhs) LMH_LAUNCH(ctx, "fill", k_fill<double>, dim3(grid_for(A->rows, 256, 8192)), dim3(256), 0, rhs, A->rows, 0.0);
  for (int c = 0; c < A->ncolors; ++c) {
    const int64_t n = A->color_ptr[c + 1] - A->color_ptr[c];
    const int64_t work = n * dm->dpe * dm->dpe;
    LMH_LAUNCH(ctx, sizeof(T) == 8 ? "scatter_colored_fp64" : "scatter_colored_fp32", k_scatter_color<T>,
               dim3(grid_for(work, 256, 65536)), dim3(256), 0, A->color_elems + A->color_ptr[c], n, dm->dpe,
               dm->edofs, A->rp, A->off, Ke, Fe, vals, rhs);
  }
  return check_launch("k_scatter_color");
}

}  // namespace lmh

using namespace lmh;

extern "C" {

int lmh_element_poisson(lmh_context *ctx, const lmh_mesh *m, const lmh_dofmap *dm, int npts, const double *kappa_d,
                        int f_kind, int constrain, double *Ke_d, double *Fe_d) {
  LMH_CHECK(ctx && m && dm && Ke_d && Fe_d, LMH_ERR_INVALID, "null argument");
  LMH_CHECK(!constrain || dm->cr.n_general == 0, LMH_ERR_UNSUPPORTED,
            "element matrices with general constraint rows grow past the element's DoFs: assemble the matrix");
  LMH_HIP(hipMemsetAsync(ctx->d_flags, 0, sizeof(int), ctx->stream));
  LMH_TRY(element_poisson_impl<double>(ctx, m, dm, npts, kappa_d, f_kind, constrain, Ke_d, Fe_d));
  return read_flag(ctx, "lmh_element_poisson");
}

int lmh_assemble_poisson(lmh_matrix *A, const lmh_mesh *m, const lmh_dofmap *dm, int npts, const double *kappa_d,
                         int f_kind, double *rhs_d, int precision, int scatter) {
  LMH_CHECK(A && m && dm, LMH_ERR_INVALID, "null argument");
  LMH_CHECK(A->dm == dm && dm->mesh == m, LMH_ERR_INVALID, "matrix, dofmap and mesh do not belong together");
  LMH_CHECK(A->vals, LMH_ERR_UNSUPPORTED, "matrix holds FP32 values only");
  LMH_CHECK(A->cons_version == dm->cr.version, LMH_ERR_INVALID,
            "the constraints changed after the matrix was created: create the matrix again");
  lmh_context *ctx = A->ctx;
  const int64_t ne = m->n_elem, dd = (int64_t)dm->dpe * dm->dpe;
  if (!A->Fe) LMH_HIP(hipMalloc((void **)&A->Fe, sizeof(double) * ne * dm->dpe));
  LMH_HIP(hipMemsetAsync(ctx->d_flags, 0, sizeof(int), ctx->stream));
  if (precision == LMH_FP64) {
    if (!A->Ke) LMH_HIP(hipMalloc((void **)&A->Ke, sizeof(double) * ne * dd));
    LMH_TRY(element_poisson_impl<double>(ctx, m, dm, npts, kappa_d, f_kind, 1, A->Ke, A->Fe));
    LMH_TRY(constrain_general_impl<double>(A, A->Ke, A->Fe));
    LMH_TRY(scatter_impl<double>(A, A->Ke, A->Fe, A->vals, rhs_d, scatter));
    LMH_TRY(scatter_extra_impl<double>(A, A->vals, rhs_d));
    A->have32 = false;
  } else {
    LMH_CHECK(precision == LMH_FP32, LMH_ERR_INVALID, "unknown precision %d", precision);
    if (!A->Ke32) LMH_HIP(hipMalloc((void **)&A->Ke32, sizeof(float) * ne * dd));
    if (!A->vals32) LMH_HIP(hipMalloc((void **)&A->vals32, sizeof(float) * (A->nnz + LMH_PAD)));
    LMH_TRY(element_poisson_impl<float>(ctx, m, dm, npts, kappa_d, f_kind, 1, A->Ke32, A->Fe));
    LMH_TRY(constrain_general_impl<float>(A, A->Ke32, A->Fe));
    LMH_TRY(scatter_impl<float>(A, A->Ke32, A->Fe, A->vals32, rhs_d, scatter));
    LMH_TRY(scatter_extra_impl<float>(A, A->vals32, rhs_d));
    A->have32 = true;
  }
  return read_flag(ctx, "lmh_assemble_poisson");
}

int lmh_matrix_get_values_fp32(const lmh_matrix *A, float *vals_h) {
  LMH_CHECK(A && vals_h, LMH_ERR_INVALID, "null argument");
  LMH_CHECK(A->have32 && A->vals32, LMH_ERR_INVALID, "matrix has no FP32 operator");
  LMH_HIP(hipMemcpyAsync(vals_h, A->vals32, sizeof(float) * A->nnz, hipMemcpyDeviceToHost, A->ctx->stream));
  LMH_HIP(hipStreamSynchronize(A->ctx->stream));
  return LMH_OK;
}

int lmh_matrix_add_element_matrices(lmh_matrix *A, int64_t n_elem, int n, const double *Ke_d, const int32_t *dofs_d) {
  LMH_CHECK(A && Ke_d && dofs_d && n > 0, LMH_ERR_INVALID, "bad argument");
  LMH_CHECK(A->vals, LMH_ERR_UNSUPPORTED, "matrix holds FP32 values only");
  if (n_elem == 0) return LMH_OK;
  lmh_context *ctx = A->ctx;
  LMH_HIP(hipMemsetAsync(ctx->d_flags, 0, sizeof(int), ctx->stream));
  LMH_LAUNCH(ctx, "check_range", k_check_range, dim3(grid_for(n_elem * n, 256, 8192)), dim3(256), 0, dofs_d,
             n_elem * n, A->rows, ctx->d_flags);
  {
    int bad = 0;
    LMH_HIP(hipMemcpyAsync(&bad, ctx->d_flags, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    LMH_HIP(hipStreamSynchronize(ctx->stream));
    LMH_CHECK(!bad, LMH_ERR_INVALID, "add_matrix: dof index out of range");
  }
  int32_t *ip = nullptr, *lst = nullptr;
  LMH_TRY(build_incidence(ctx, dofs_d, n_elem, n, A->rows, &ip, &lst));
  const int g = (int)std::min<int64_t>(A->rows, (int64_t)ctx->cu_count * 64);
  LMH_LAUNCH(ctx, "add_element_matrices", k_add_gather_bs, dim3(g), dim3(64), sizeof(double) * A->max_row, ip, lst, n,
             A->rows, dofs_d, A->rp, A->ci, Ke_d, A->vals, ctx->d_flags);
  LMH_TRY(check_launch("k_add_gather_bs"));
  int bad = 0;
  LMH_HIP(hipMemcpyAsync(&bad, ctx->d_flags, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  LMH_HIP(hipStreamSynchronize(ctx->stream));
  LMH_HIP(hipFree(ip));
  LMH_HIP(hipFree(lst));
  A->have32 = false;
  LMH_CHECK(!bad, LMH_ERR_INVALID, "add_matrix: entry outside the sparsity pattern");
  return LMH_OK;
}

int lmh_vector_add_element_vectors(lmh_context *ctx, double *v_d, int64_t n_elem, int n, const double *Fe_d,
                                   const int32_t *dofs_d) {
  LMH_CHECK(ctx && v_d && Fe_d && dofs_d && n > 0, LMH_ERR_INVALID, "bad argument");
  if (n_elem == 0) return LMH_OK;
  // rows = max dof + 1 is unknown here; use a host scan of the dof range
  std::vector<int32_t> d((size_t)n_elem * n);
  LMH_HIP(hipMemcpyAsync(d.data(), dofs_d, sizeof(int32_t) * d.size(), hipMemcpyDeviceToHost, ctx->stream));
  LMH_HIP(hipStreamSynchronize(ctx->stream));
  int64_t rows = 0;
  for (int32_t x : d) {
    LMH_CHECK(x >= 0, LMH_ERR_INVALID, "negative dof index");
    rows = std::max<int64_t>(rows, (int64_t)x + 1);
  }
  int32_t *ip = nullptr, *lst = nullptr;
  LMH_TRY(build_incidence(ctx, dofs_d, n_elem, n, rows, &ip, &lst));
  LMH_LAUNCH(ctx, "add_element_vectors", k_add_vec_gather, dim3(grid_for(rows, 256, 8192)), dim3(256), 0, ip, lst,
             rows, Fe_d, v_d);
  LMH_TRY(check_launch("k_add_vec_gather"));
  LMH_HIP(hipStreamSynchronize(ctx->stream));
  LMH_HIP(hipFree(ip));
  LMH_HIP(hipFree(lst));
  return LMH_OK;
}

int lmh_assemble_residual_poisson(lmh_matrix *A, const lmh_mesh *m, const lmh_dofmap *dm, int npts,
                                  const double *kappa_d, int f_kind, const double *u_d, double *R_d) {
  LMH_CHECK(A && m && dm && u_d && R_d, LMH_ERR_INVALID, "null argument");
  LMH_CHECK(A->dm == dm && dm->mesh == m, LMH_ERR_INVALID, "matrix, dofmap and mesh do not belong together");
  LMH_CHECK(dm->dpe <= 32, LMH_ERR_UNSUPPORTED, "residual assembly needs at most 32 DoFs per element");
  lmh_context *ctx = A->ctx;
  const int64_t ne = m->n_elem, dd = (int64_t)dm->dpe * dm->dpe;
  if (!A->Fe) LMH_HIP(hipMalloc((void **)&A->Fe, sizeof(double) * ne * dm->dpe));
  if (!A->Ke) LMH_HIP(hipMalloc((void **)&A->Ke, sizeof(double) * ne * dd));
  LMH_HIP(hipMemsetAsync(ctx->d_flags, 0, sizeof(int), ctx->stream));
  LMH_TRY(element_poisson_impl<double>(ctx, m, dm, npts, kappa_d, f_kind, 0, A->Ke, A->Fe));
  LMH_LAUNCH(ctx, "element_residual", k_elem_residual, dim3(grid_for((ne + 3) / 4, 1, ctx->cu_count * 8)), dim3(256),
             sizeof(double) * 4 * ((size_t)dm->dpe * dm->dpe + dm->dpe), (const double *)A->Ke, A->Fe, dm->edofs,
             dm->cons, u_d, dm->dpe, ne);
  LMH_LAUNCH(ctx, "fill", k_fill<double>, dim3(grid_for(dm->n_dofs, 256, 8192)), dim3(256), 0, R_d, dm->n_dofs, 0.0);
  LMH_LAUNCH(ctx, "add_element_vectors", k_add_vec_gather, dim3(grid_for(dm->n_dofs, 256, 8192)), dim3(256), 0,
             dm->inc_ptr, dm->inc, dm->n_dofs, (const double *)A->Fe, R_d);
  LMH_TRY(check_launch("residual"));
  LMH_TRY(read_flag(ctx, "lmh_assemble_residual_poisson"));
  // general rows: u holds the enforced values (C u + H), so the summed element
  // residuals get C^T, as constrain_element_residual does per element
  return apply_constraint_transpose(ctx, dm, R_d);
}

int lmh_qoi_box_derivative(const lmh_mesh *m, const lmh_dofmap *dm, int npts, double lo, double hi, double *dq_d) {
  LMH_CHECK(m && dm && dq_d, LMH_ERR_INVALID, "null argument");
  LMH_CHECK(dm->nv == 1, LMH_ERR_UNSUPPORTED, "QoI derivative needs one variable");
  lmh_context *ctx = m->ctx;
  const FETable *t = nullptr;
  LMH_TRY(fe_table(ctx, m->type, npts, &t));
  double *Q = nullptr;
  LMH_HIP(hipMallocAsync((void **)&Q, sizeof(double) * m->n_elem * t->n, ctx->stream));
  LMH_HIP(hipMemsetAsync(ctx->d_flags, 0, sizeof(int), ctx->stream));
  const size_t shm = sizeof(double) * (t->nq + (size_t)t->n * t->nq * 4 + 81 + t->nq);
  LMH_LAUNCH(ctx, "qoi_box", k_qoi_box, dim3(grid_for(m->n_elem, 1, ctx->cu_count * 8)), dim3(64), shm, t->d_data,
             t->n, t->nq, m->xyz, m->conn, m->n_elem, lo, hi, Q, ctx->d_flags);
  LMH_TRY(check_launch("k_qoi_box"));
  LMH_LAUNCH(ctx, "fill", k_fill<double>, dim3(grid_for(dm->n_dofs, 256, 8192)), dim3(256), 0, dq_d, dm->n_dofs, 0.0);
  LMH_LAUNCH(ctx, "add_element_vectors", k_add_vec_gather, dim3(grid_for(dm->n_dofs, 256, 8192)), dim3(256), 0,
             dm->inc_ptr, dm->inc, dm->n_dofs, (const double *)Q, dq_d);
  LMH_LAUNCH(ctx, "zero_constrained", k_zero_constrained, dim3(grid_for(dm->n_dofs, 256, 8192)), dim3(256), 0, dq_d,
             dm->cons, dm->n_dofs);
  LMH_TRY(check_launch("qoi"));
  LMH_HIP(hipFreeAsync(Q, ctx->stream));
  LMH_TRY(read_flag(ctx, "lmh_qoi_box_derivative"));
  // what assemble_qoi_derivative with apply_constraints = true gives
  return apply_constraint_transpose(ctx, dm, dq_d);
}

int lmh_adjoint_kappa_gradient(const lmh_mesh *m, const lmh_dofmap *dm, int npts, const double *kappa_d,
                               const double *u_d, const double *lambda_d, double *grad_d) {
  (void)kappa_d;  // dJ/dkappa_e does not depend on kappa_e for a per-element constant kappa
  LMH_CHECK(m && dm && u_d && lambda_d && grad_d, LMH_ERR_INVALID, "null argument");
  LMH_CHECK(dm->nv == 1, LMH_ERR_UNSUPPORTED, "adjoint gradient needs one variable");
  // With general constraint rows, u and lambda are the constraint-enforced
  // vectors (heterogeneous for u, homogeneous for lambda): -(C lambda)^T
  // (K_e / kappa_e)(C u + H) is then the element term over the element's own
  // DoFs, which the kernel forms; only empty-row DoFs are masked.
  lmh_context *ctx = m->ctx;
  const FETable *t = nullptr;
  LMH_TRY(fe_table(ctx, m->type, npts, &t));
  LMH_CHECK(t->nq <= 64 && t->n <= 27, LMH_ERR_UNSUPPORTED, "element too large");
  LMH_HIP(hipMemsetAsync(ctx->d_flags, 0, sizeof(int), ctx->stream));
  const size_t shm = sizeof(double) * (t->nq + (size_t)t->n * t->nq * 4 + 81 + 27 + 27 + 64);
  LMH_LAUNCH(ctx, "adjoint_kappa", k_adjoint_kappa, dim3(grid_for(m->n_elem, 1, ctx->cu_count * 8)), dim3(64), shm,
             t->d_data, t->n, t->nq, m->xyz, m->conn, dm->edofs, dm->cons, u_d, lambda_d, m->n_elem, grad_d,
             ctx->d_flags);
  LMH_TRY(check_launch("k_adjoint_kappa"));
  return read_flag(ctx, "lmh_adjoint_kappa_gradient");
}

// ---------------------------------------------------------------------------
// User physics (lmh_fem_physics.h): arrays for the template element loop and
// the element-order scatter of its results.
// ---------------------------------------------------------------------------
int lmh_fem_element_data(const lmh_mesh *m, const lmh_dofmap *dm, int npts, lmh_fem_data *out) {
  LMH_CHECK(m && dm && out, LMH_ERR_INVALID, "null argument");
  LMH_CHECK(dm->mesh == m, LMH_ERR_INVALID, "dofmap and mesh do not belong together");
  LMH_CHECK(dm->cr.n_general == 0, LMH_ERR_UNSUPPORTED,
            "user physics: the element loop applies homogeneous Dirichlet rows only, not general constraint rows");
  lmh_context *ctx = m->ctx;
  const FETable *t = nullptr;
  LMH_TRY(fe_table(ctx, m->type, npts, &t));
  out->elem_type = m->type;
  out->n_nodes = m->npe;
  out->n_vars = dm->nv;
  out->n_dofs = dm->dpe;
  out->n_qp = t->nq;
  out->n_elem = m->n_elem;
  out->n_dofs_total = dm->n_dofs;
  out->xyz = m->xyz;
  out->conn = m->conn;
  out->elem_dofs = dm->edofs;
  out->constrained = dm->cons;
  out->w = t->d_w();
  out->phi = t->d_phi();
  out->dref = t->d_dref();
  out->flag = ctx->d_flags + 15;  // slot of the user element loop
  out->stream = (void *)ctx->stream;
  LMH_HIP(hipMemsetAsync(out->flag, 0, sizeof(int), ctx->stream));
  return LMH_OK;
}

int lmh_fem_check_flag(const lmh_fem_data *d) {
  LMH_CHECK(d && d->flag, LMH_ERR_INVALID, "null argument");
  hipStream_t s = (hipStream_t)d->stream;
  int flag = 0;
  LMH_HIP(hipMemcpyAsync(&flag, d->flag, sizeof(int), hipMemcpyDeviceToHost, s));
  LMH_HIP(hipStreamSynchronize(s));
  if (flag) {
    LMH_HIP(hipMemsetAsync(d->flag, 0, sizeof(int), s));
    return set_error(LMH_ERR_DEGENERATE, "element loop: Jacobian determinant <= 0 (degenerate map)");
  }
  return LMH_OK;
}

int lmh_matrix_scatter_elements(lmh_matrix *A, const double *Ke_d, const double *Fe_d, double *R_d) {
  LMH_CHECK(A && A->dm && (Ke_d || Fe_d), LMH_ERR_INVALID, "null argument");
  LMH_CHECK(!R_d || Fe_d, LMH_ERR_INVALID, "a residual needs element vectors");
  const lmh_dofmap *dm = A->dm;
  LMH_CHECK(dm->cr.n_general == 0, LMH_ERR_UNSUPPORTED, "element scatter with general constraint rows");
  LMH_CHECK(A->cons_version == dm->cr.version, LMH_ERR_INVALID,
            "the constraints changed after the matrix was created: create the matrix again");
  lmh_context *ctx = A->ctx;
  if (Ke_d) {
    LMH_CHECK(A->vals, LMH_ERR_UNSUPPORTED, "matrix holds FP32 values only");
    LMH_TRY(scatter_impl<double>(A, Ke_d, Fe_d, A->vals, R_d, LMH_SCATTER_GATHER));
    A->have32 = false;
    return LMH_OK;
  }
  if (R_d) {
    // residual only: the DoF map's incidence lists hold (element, local) in
    // ascending order, so each entry sums in element order
    LMH_LAUNCH(ctx, "fill", k_fill<double>, dim3(grid_for(A->rows, 256, 8192)), dim3(256), 0, R_d, A->rows, 0.0);
    LMH_LAUNCH(ctx, "add_element_vectors", k_add_vec_gather, dim3(grid_for(A->rows, 256, 8192)), dim3(256), 0,
               dm->inc_ptr, dm->inc, A->rows, Fe_d, R_d);
  }
  return check_launch("lmh_matrix_scatter_elements");
}

}  // extern "C"
