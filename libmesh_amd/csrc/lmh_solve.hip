// lmh_solve.hip -- CSR SpMV, Jacobi-preconditioned CG (FP64) and the
// mixed-precision (FP32 inner / FP64 refined) solver.
//
// CG follows Eigen::internal::conjugate_gradient (ConjugateGradient.h:28-90)
// with DiagonalPreconditioner (BasicPreconditioners.h:66-76), the solver
// libMesh runs for CG + JACOBI_PRECOND (eigen_sparse_linear_solver.C:106-160).
// All scalars live on the device; each iteration is three kernels
//   spmv+dot(p,Ap) -> alpha | x,r update + |r|^2 + (r,z) -> beta | p = z + beta p
// where each kernel reduces the previous kernel's per-block partials itself
// (deterministic), so the host only polls a converged flag every few iterations.
#include <algorithm>
#include <climits>
#include <cmath>
#include <functional>
#include <mutex>
#include <type_traits>

#include "lmh_internal.h"

namespace lmh {

constexpr int BLK = 256;
constexpr int RB_CAP = LMH_RB_CAP;  // stored entries per SpMV row block

template <typename T>
__device__ inline T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// block-wide sum of up to 3 values (deterministic); result valid in thread 0
template <int NV, int NT = BLK>
__device__ inline void block_sum(double (&v)[NV], double (*sh)[NT / 64]) {
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
      for (int w = 0; w < NT / 64; ++w) s += sh[k][w];
      v[k] = s;
    }
  }
}

// Publish the block's NV partial sums to part[k*gridDim.x + blockIdx.x].  A
// one-block finalize kernel launched next on the same stream reduces them in
// block order (deterministic); the kernel boundary makes the partials visible
// without any device-scope fence inside the producer.
template <int NV, int NT = BLK>
__device__ inline void publish(double (&v)[NV], double *__restrict__ part) {
  __shared__ double sh[NV][NT / 64];
  block_sum<NV, NT>(v, sh);
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
// pre(row, out) loads the NPRE per-row values an epilogue needs (b[row],
// p[row], ...); epi(row, sum, pre_values) consumes each row result in the
// calling thread.
struct NoPre {
  __device__ inline void operator()(int32_t, double *) const {}
};

// Software-pipelined over the workgroup's row blocks: the value/column loads
// and the row pointers of block bk+1 are issued right after the x gathers and
// the per-row epilogue values of block bk, and nothing of block bk's own work
// loads from global memory after them, so they stay in flight while block
// bk's products are formed and its rows reduced.  The main loop issues that
// fetch unconditionally (the last block is peeled): behind a condition, the
// compiler's wait counts merge both paths and every later wait drains the
// fetch.  Row pointers and epilogue values of a block with fewer than BLK rows
// are staged in LDS; FP32 products are double-buffered so one barrier per
// block suffices.
template <typename VT, typename XT, typename AT, int NPRE, typename Pre, typename Epi>
__device__ inline void rowblock_spmv(const int32_t *__restrict__ rb, int64_t nrb, const int32_t *__restrict__ rp,
                                     const int32_t *__restrict__ ci, const VT *__restrict__ v,
                                     const XT *__restrict__ x, Pre &&pre, Epi &&epi) {
  constexpr int NBUF = sizeof(AT) == 4 ? 2 : 1;
  constexpr int NP = NPRE > 0 ? NPRE : 1;
  __shared__ AT prodbuf[NBUF][RB_CAP];
  __shared__ int32_t rowbuf[NBUF][BLK];
  __shared__ double prebuf[NBUF][NP][NPRE > 0 ? BLK : 1];
  __shared__ double red[BLK / 64];
  constexpr int PER = RB_CAP / BLK;
  int64_t lo, hi;
  block_range(nrb, lo, hi);
  const int tid = threadIdx.x;
  // state of the block whose loads are in flight
  int32_t r0 = 0, r1 = 0, k0 = 0, cnt = 0, rpn = 0;
  VT vv[PER];
  int32_t cr[PER];
  // branch-free loads (out-of-range slots re-read entry k0 and are dropped)
  auto fetch = [&](int64_t bk) {
    r0 = rb[bk];
    r1 = rb[bk + 1];
    k0 = rp[r0];
    cnt = rp[r1] - k0;
    rpn = rp[r0 + (tid <= r1 - r0 ? tid : 0)];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const int32_t t = tid + j * BLK;
      const int32_t k = k0 + (t < cnt ? t : 0);
      vv[j] = __builtin_nontemporal_load(&v[k]);
      cr[j] = __builtin_nontemporal_load(&ci[k]);
    }
  };
  int buf = 0;
  auto body = [&](int64_t bk, auto more) {
    const int32_t R0 = r0, R1 = r1, K0 = k0, CNT = cnt, RPN = rpn;
    if (CNT > RB_CAP) {  // one long row: reduce it directly
      AT s = 0;
      for (int32_t t = tid; t < CNT; t += BLK)
        s += (AT)__builtin_nontemporal_load(&v[K0 + t]) * (AT)x[__builtin_nontemporal_load(&ci[K0 + t])];
      double w = wave_sum((double)s);
      __syncthreads();
      if ((tid & 63) == 0) red[tid >> 6] = w;
      __syncthreads();
      if (tid == 0) {
        double t2 = 0;
        for (int k = 0; k < BLK / 64; ++k) t2 += red[k];
        double pv[NP];
        pre(R0, pv);
        epi(R0, (AT)t2, pv);
      }
      if constexpr (decltype(more)::value) fetch(bk + 1);
      return;
    }
    AT *prod = prodbuf[buf];
    int32_t *rows = rowbuf[buf];
    double(*pb)[NPRE > 0 ? BLK : 1] = prebuf[buf];
    const bool staged = R1 - R0 < BLK;
    AT xx[PER];
    VT cv[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      xx[j] = (AT)x[cr[j]];
      cv[j] = vv[j];
    }
    double pv[NP];
    if (NPRE > 0 && staged && R1 > R0) pre(R0 + (tid < R1 - R0 ? tid : 0), pv);
    if constexpr (decltype(more)::value) fetch(bk + 1);
    if (NBUF == 1) __syncthreads();  // the previous block's rows are reduced
    if (staged && tid <= R1 - R0) rows[tid] = RPN - K0;
    if (NPRE > 0 && staged && tid < R1 - R0)
#pragma unroll
      for (int k = 0; k < NP; ++k) pb[k][tid] = pv[k];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const int32_t t = tid + j * BLK;
      if (t < CNT) prod[t] = (AT)cv[j] * xx[j];
    }
    __syncthreads();
    const int g = tid >> 3, sub = tid & 7;
    if (staged) {
      for (int32_t r = R0 + g; r < R1; r += BLK / 8) {
        const int32_t a = rows[r - R0], e = rows[r - R0 + 1];
        AT s = 0;
        for (int32_t t = a + sub; t < e; t += 8) s += prod[t];
        s += __shfl_xor(s, 4, 8);
        s += __shfl_xor(s, 2, 8);
        s += __shfl_xor(s, 1, 8);
        if (sub == 0) {
          double p2[NP];
#pragma unroll
          for (int k = 0; k < NP; ++k) p2[k] = NPRE > 0 ? pb[k][r - R0] : 0.0;
          epi(r, s, p2);
        }
      }
    } else {
      for (int32_t r = R0 + g; r < R1; r += BLK / 8) {
        const int32_t a = rp[r] - K0, e = rp[r + 1] - K0;
        AT s = 0;
        for (int32_t t = a + sub; t < e; t += 8) s += prod[t];
        s += __shfl_xor(s, 4, 8);
        s += __shfl_xor(s, 2, 8);
        s += __shfl_xor(s, 1, 8);
        if (sub == 0) {
          double p2[NP];
          pre(r, p2);
          epi(r, s, p2);
        }
      }
    }
    buf ^= NBUF - 1;
  };
  if (lo < hi) {
    fetch(lo);
    for (int64_t bk = lo; bk + 1 < hi; ++bk) body(bk, std::true_type{});
    body(hi - 1, std::false_type{});
  }
}

// FP32 values with x windows.  A row block of ~2000 entries references only a
// few hundred distinct columns (about 300 for Q2 hexahedra, under 520 for every
// Lagrange type here), so the block's x values are gathered once, in sorted
// column order, into an LDS window, and each entry reads its x there through a
// 16-bit window index.  The sorted gather touches each x cache line about once
// per block instead of once per entry that uses it.  Values are read as 16-byte
// chunks and the window indices as 8-entry chunks, both from the block's start
// rounded down to 8 entries; the indices pass through LDS, so every product
// lands at its entry's slot and the rows are reduced exactly as in
// rowblock_spmv (same products, same summation order, bit-identical sums).
// Row blocks hold at most RB_CAP - 8 entries, so a block's chunks fit one
// 8-entry index chunk and two 4-entry value chunks per thread.
// T = double (the FP64 operator) runs the same scheme with 2-entry value
// chunks; products and sums are formed as in rowblock_spmv<double>, so the
// FP64 results are bit-identical to the int32-column kernel too.
typedef uint32_t vu4 __attribute__((ext_vector_type(4)));
constexpr int UCAP = 768;  // distinct columns of one row block (x window)
// 8-bit windows: blocks of at most W8_CAP - 8 entries and 256 distinct columns
// (a Q2 hexahedral operator's 1024-entry blocks hold about 200); an entry then
// costs one index byte instead of two
constexpr int W8_CAP = 1024, W8_UCAP = 256;
// 8-bit blocks run on 128-thread workgroups (8 entries per thread, as the
// 2048-entry blocks on 256 threads), twice as many of them
#ifndef LMH_W8_GRID
#define LMH_W8_GRID 2
#endif
constexpr int W8_NT = 128, W8_GRID = LMH_W8_GRID;
template <int IW>
using widx_t = typename std::conditional<IW == 1, uint8_t, uint16_t>::type;
// Pipelining.  Every block's loads are addressed from registers only, so no
// load waits on another in flight: a 32-byte descriptor per block (rows,
// entry range, window meta, built with the windows) is loaded three blocks
// ahead, the block's window offsets two blocks ahead, and its values, window
// indices, row pointers and x-window gathers one block ahead, while the
// current block runs its LDS phases.
// CAPW: the block capacity (RB_CAP with 16-bit indices, W8_CAP with 8-bit);
// IW: index bytes per entry; NT: threads per workgroup.
template <typename T, bool DOT, int CAPW = RB_CAP, int IW = 2, int NT = BLK>
__device__ inline double rowblock_spmv16(int64_t nrb, const int4 *__restrict__ desc, const int32_t *__restrict__ rp,
                                         const widx_t<IW> *__restrict__ widx, const uint16_t *__restrict__ woff,
                                         const T *__restrict__ v, const T *__restrict__ x, T *__restrict__ y) {
  constexpr int CH = 16 / sizeof(T);  // entries per 16-byte value chunk
  typedef T vch __attribute__((ext_vector_type(CH)));
  // the chunk's CH window indices (IW * CH bytes: 8 for FP32 16-bit, 4 for
  // FP64 16-bit and FP32 8-bit, 2 for FP64 8-bit)
  typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
  typedef typename std::conditional<
      IW * CH == 8, u32x2, typename std::conditional<IW * CH == 4, uint32_t, uint16_t>::type>::type ich;
  constexpr int UC = IW == 1 ? W8_UCAP : UCAP;
  // products at their chunk positions from the block's start rounded down to 8
  // entries (slots outside the block hold unread values)
  __shared__ __attribute__((aligned(16))) T prod[CAPW];
  __shared__ T xw[UC];
  // Row offsets and the x of the block's rows (for the (x, Ax) partials).
  // FP32: double-buffered, written before the first barrier.  FP64: one
  // buffer, written after it (the previous block's rows are reduced by then),
  // which keeps the LDS under 32 KB for 5 resident workgroups per CU.
  constexpr int NB = sizeof(T) == 8 ? 1 : 2;
  __shared__ int32_t rowbuf[NB][NT];
  __shared__ T xrow[NB][NT];
  constexpr int PER = CAPW / NT, UPER = UC / NT, NCH = PER / CH;
  static_assert(PER % CH == 0 && UC % NT == 0, "whole value chunks and window slots per thread");
  int64_t lo, hi;
  block_range(nrb, lo, hi);
  double acc = 0;
  if (lo >= hi) return acc;
  const int tid = threadIdx.x;
  // descriptor of block bk (clamped to the workgroup's last block): geometry
  // {r0, r1, k0, cnt} and window {b0, b1, s0 | nu << 16, w0}
  auto dsc = [&](int64_t bk, int4 &g, int4 &w) {
    const int64_t b = bk < hi ? bk : hi - 1;
    g = desc[2 * b];
    w = desc[2 * b + 1];
  };
  // branch-free loads: out-of-range slots re-read a valid address and are dropped
  auto load_woff = [&](const int4 &W, uint32_t(&uo)[UPER]) {
    const int32_t nu = W.z >> 16;
#pragma unroll
    for (int j = 0; j < UPER; ++j) {
      const int32_t u = tid + j * NT;
      uo[j] = __builtin_nontemporal_load(&woff[W.w + (u < nu ? u : 0)]);
    }
  };
  auto gather = [&](const int4 &G, const int4 &W, const uint32_t(&uo)[UPER], T(&xu)[UPER], T &xo) {
    const int32_t NU = W.z >> 16, S0 = W.z & 0xffff;
#pragma unroll
    for (int j = 0; j < UPER; ++j) {
      const int32_t u = tid + j * NT, o = (int32_t)uo[j];
      xu[j] = x[u < NU ? o + (o <= S0 ? W.x : W.y) : 0];
    }
    if (DOT) xo = x[G.x + (tid < G.y - G.x ? tid : 0)];
  };
  auto fetch = [&](const int4 &G, vch(&vv)[NCH], ich(&iv)[NCH], int32_t &rpn) {
    const int32_t r0 = G.x, r1 = G.y, k0 = G.z, cnt = G.w;
    rpn = rp[r0 + (tid <= r1 - r0 ? tid : 0)];
    const int32_t a0 = k0 & ~7, span = (k0 - a0) + cnt;
#pragma unroll
    for (int j = 0; j < NCH; ++j) {
      const int32_t c = CH * tid + CH * NT * j;  // chunks starting past the block re-read chunk 0
      const int32_t o = c < span ? c : 0;
      vv[j] = __builtin_nontemporal_load((const vch *)(v + a0 + o));
      iv[j] = __builtin_nontemporal_load((const ich *)(widx + a0 + o));
    }
  };
  // current block (c), the next (n), the one after (nn)
  int4 Gc, Wc, Gn, Wn, Gnn, Wnn;
  vch vvc[NCH], vvn[NCH];
  ich ivc[NCH], ivn[NCH];
  int32_t rpnc = 0, rpnn = 0;
  T xuc[UPER], xun[UPER], xoc = 0, xon = 0;
  uint32_t uon[UPER], uonn[UPER];
  dsc(lo, Gc, Wc);
  dsc(lo + 1, Gn, Wn);
  dsc(lo + 2, Gnn, Wnn);
  {
    uint32_t uoc[UPER];
    load_woff(Wc, uoc);
    gather(Gc, Wc, uoc, xuc, xoc);
  }
  fetch(Gc, vvc, ivc, rpnc);
  load_woff(Wn, uon);
  int buf = 0;
  // One block from registers: window, indices and row data to LDS, products,
  // rows reduced in a fixed order.
  auto compute = [&]() {
    int32_t *rows = rowbuf[buf];
    T *xr = xrow[buf];
    buf ^= NB - 1;
    const int32_t R0 = Gc.x, R1 = Gc.y, K0 = Gc.z, CNT = Gc.w, RPN = rpnc, lead = K0 & 7, A0 = K0 - lead;
    const int32_t NU = Wc.z >> 16;
    const bool staged = R1 - R0 < NT;
    if (NB == 2 && staged && tid <= R1 - R0) rows[tid] = RPN - A0;
#pragma unroll
    for (int j = 0; j < UPER; ++j)
      if (tid + j * NT < NU) xw[tid + j * NT] = xuc[j];
    if (NB == 2 && DOT) xr[tid] = xoc;
    __syncthreads();  // window and row data visible; the previous block's rows are reduced
    if (NB == 1 && staged && tid <= R1 - R0) rows[tid] = RPN - A0;
    if (NB == 1 && DOT) xr[tid] = xoc;
#pragma unroll
    for (int j = 0; j < NCH; ++j) {
      vch xs;
#pragma unroll
      for (int e = 0; e < CH; ++e) {
        const int32_t t = CH * tid + CH * NT * j + e - lead;
        uint32_t word;
        int sh;
        if constexpr (IW * CH == 8) {
          word = ivc[j][e >> 1];
          sh = 16 * (e & 1);
        } else {
          word = (uint32_t)ivc[j];
          sh = 8 * IW * e;
        }
        const int32_t wi = (int32_t)((word >> sh) & (IW == 1 ? 0xffu : 0xffffu));
        xs[e] = xw[t >= 0 && t < CNT ? wi : 0];
      }
      *(vch *)&prod[CH * tid + CH * NT * j] = vvc[j] * xs;
    }
    __syncthreads();
    const int g = tid >> 3, sub = tid & 7;
    if (staged) {
      for (int32_t r = R0 + g; r < R1; r += NT / 8) {
        const int32_t a = rows[r - R0], e = rows[r - R0 + 1];
        T s = 0;
        for (int32_t t = a + sub; t < e; t += 8) s += prod[t];
        s += __shfl_xor(s, 4, 8);
        s += __shfl_xor(s, 2, 8);
        s += __shfl_xor(s, 1, 8);
        if (sub == 0) {
          y[r] = s;
          if (DOT) acc += (double)xr[r - R0] * (double)s;
        }
      }
    } else {
      for (int32_t r = R0 + g; r < R1; r += NT / 8) {
        const int32_t a = rp[r] - A0, e = rp[r + 1] - A0;
        T s = 0;
        for (int32_t t = a + sub; t < e; t += 8) s += prod[t];
        s += __shfl_xor(s, 4, 8);
        s += __shfl_xor(s, 2, 8);
        s += __shfl_xor(s, 1, 8);
        if (sub == 0) {
          y[r] = s;
          if (DOT) acc += (double)x[r] * (double)s;
        }
      }
    }
  };
  for (int64_t bk = lo; bk + 1 < hi; ++bk) {
    // the next block's x window and row x (its offsets arrived with this
    // block's loads), then its values and indices, the offsets of the block
    // after it and the descriptor three ahead
    gather(Gn, Wn, uon, xun, xon);
    fetch(Gn, vvn, ivn, rpnn);
    load_woff(Wnn, uonn);
    int4 G3, W3;
    dsc(bk + 3, G3, W3);
    compute();
    Gc = Gn;
    Wc = Wn;
#pragma unroll
    for (int j = 0; j < NCH; ++j) {
      vvc[j] = vvn[j];
      ivc[j] = ivn[j];
    }
    rpnc = rpnn;
#pragma unroll
    for (int j = 0; j < UPER; ++j) {
      xuc[j] = xun[j];
      uon[j] = uonn[j];
    }
    xoc = xon;
    Gn = Gnn;
    Wn = Wnn;
    Gnn = G3;
    Wnn = W3;
  }
  compute();
  return acc;
}


// ---------------------------------------------------------------------------
// plain SpMV y = A x
// ---------------------------------------------------------------------------
template <typename VT, typename XT, typename YT, typename AT>
__global__ __launch_bounds__(BLK) void k_spmv(const int32_t *__restrict__ rb, int64_t nrb,
                                              const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                              const VT *__restrict__ v, const XT *__restrict__ x,
                                              YT *__restrict__ y) {
  rowblock_spmv<VT, XT, AT, 0>(rb, nrb, rp, ci, v, x, NoPre{}, [&](int32_t r, AT s, const double *) { y[r] = (YT)s; });
}


// r = b - A x (FP64); partials |r|^2, |b|^2
__global__ __launch_bounds__(BLK) void k_residual(const int32_t *__restrict__ rb, int64_t nrb,
                                                  const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                                  const double *__restrict__ v, const double *__restrict__ b,
                                                  const double *__restrict__ x, double *__restrict__ r,
                                                  double *__restrict__ part) {
  double acc[2] = {0, 0};
  rowblock_spmv<double, double, double, 1>(
      rb, nrb, rp, ci, v, x, [&](int32_t row, double *o) { o[0] = b[row]; },
      [&](int32_t row, double s, const double *pr) {
        const double bv = pr[0], rv = bv - s;
        r[row] = rv;
        acc[0] += rv * rv;
        acc[1] += bv * bv;
      });
  publish<2>(acc, part);
}

// out[0] = |r|^2, out[1] = |b|^2
__global__ __launch_bounds__(BLK) void k_fin_residual(const double *__restrict__ part, int G, double *__restrict__ out) {
  double t[2];
  reduce_partials<2>(part, G, t);
  if (threadIdx.x == 0) {
    out[0] = t[0];
    out[1] = t[1];
  }
}

// y += a x
__global__ void k_axpy(int64_t n, double a, const double *__restrict__ x, double *__restrict__ y) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] += a * x[i];
}

// Position of each row's diagonal entry (binary search of the sorted columns),
// -1 when the row stores none.  Depends on the pattern only.
__global__ void k_diag_index(int64_t n, const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                             int32_t *__restrict__ diag) {
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
    diag[r] = (lo < end && ci[lo] == r) ? lo : -1;
  }
}

// Jacobi inverse diagonal: 1/a_rr, or 1 when the diagonal is absent or zero
template <typename VT, typename OT>
__global__ void k_invdiag(int64_t n, const int32_t *__restrict__ diag, const VT *__restrict__ v, int jacobi,
                          OT *__restrict__ dinv) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    const int32_t k = diag[r];
    OT d = (OT)1;
    if (jacobi && k >= 0 && v[k] != (VT)0) d = (OT)1 / (OT)v[k];
    dinv[r] = d;
  }
}

// ---------------------------------------------------------------------------
// Solver state: two banks of scalars.  Iteration k reads bank k&1 and its last
// kernel writes bank (k+1)&1, so no block ever reads a scalar that another
// block of the same kernel writes.  Each kernel that needs a global sum
// reduces the previous kernel's per-block partials itself (every block, same
// order), so an iteration is three kernels: spmv | update | p-update.
// ---------------------------------------------------------------------------
enum { B_RZ = 0, B_RR, B_ITS, B_DONE, B_REASON, B_REFRESH, B_MAXRR, B_NREF, B_SIZE = 16 };
enum { G_THR = 32, G_BB, G_DELTA2, G_MODE, G_RRTRUE };

template <int NV>
__device__ inline void reduce_all(const double *__restrict__ part, int G, double (&t)[NV]) {
  __shared__ double sh[NV][BLK / 64];
  __shared__ double res[NV];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  // eight independent loads per thread in flight (a consumer reads up to
  // 8192 partials; one dependent load per step costs microseconds of latency)
  constexpr int U = 8;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    const double *pk = part + k * G;
    double a[U];
#pragma unroll
    for (int u = 0; u < U; ++u) a[u] = 0;
    int b = threadIdx.x;
    for (; b + (U - 1) * BLK < G; b += U * BLK) {
#pragma unroll
      for (int u = 0; u < U; ++u) a[u] += pk[b + u * BLK];
    }
#pragma unroll
    for (int u = 0; u < U - 1; ++u)
      if (b + u * BLK < G) a[u] += pk[b + u * BLK];
    double s = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
    s = wave_sum(s);
    if (lane == 0) sh[k][wv] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double s = 0;
      for (int w = 0; w < BLK / 64; ++w) s += sh[k][w];
      res[k] = s;
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) t[k] = res[k];
}

// q = A p; partials (p, q)
template <typename VT>
__global__ __launch_bounds__(BLK) void k_it_spmv(const int32_t *__restrict__ rb, int64_t nrb,
                                                 const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                                 const VT *__restrict__ v, const VT *__restrict__ p,
                                                 VT *__restrict__ q, double *__restrict__ part,
                                                 const double *__restrict__ S, int bank) {
  if (S[bank * B_SIZE + B_DONE] != 0.0) return;
  double acc[1] = {0};
  rowblock_spmv<VT, VT, VT, 1>(
      rb, nrb, rp, ci, v, p, [&](int32_t r, double *o) { o[0] = (double)p[r]; },
      [&](int32_t r, VT s, const double *pr) {
        q[r] = s;
        acc[0] += pr[0] * (double)s;
      });
  publish<1>(acc, part);
}

// q = A32 p on the x-window FP32 operator (IW-byte window indices on blocks
// of CAPW entries); partials (p, q).  Held to 64 VGPRs (8 waves per SIMD): the
// pipelined loop needs every resident wave.
template <int CAPW, int IW, int NT>
__global__ __launch_bounds__(NT, 2048 / NT) void k_it_spmv16(int64_t nrb, const int32_t *__restrict__ rp,
                                                          const void *__restrict__ widx,
                                                          const uint16_t *__restrict__ woff,
                                                          const int4 *__restrict__ desc, const float *__restrict__ v,
                                                          const float *__restrict__ p, float *__restrict__ q,
                                                          double *__restrict__ part, const double *__restrict__ S,
                                                          int bank) {
  if (S[bank * B_SIZE + B_DONE] != 0.0) return;
  double acc[1] = {
      rowblock_spmv16<float, true, CAPW, IW, NT>(nrb, desc, rp, (const widx_t<IW> *)widx, woff, v, p, q)};
  publish<1, NT>(acc, part);
}

// q = A p on the FP64 operator through the same x windows; partials (p, q).
// Bit-identical to k_it_spmv<double>.
template <int CAPW, int IW, int NT>
__global__ __launch_bounds__(NT) void k_it_spmv16d(int64_t nrb, const int32_t *__restrict__ rp,
                                                   const void *__restrict__ widx, const uint16_t *__restrict__ woff,
                                                   const int4 *__restrict__ desc, const double *__restrict__ v,
                                                   const double *__restrict__ p, double *__restrict__ q,
                                                   double *__restrict__ part, const double *__restrict__ S,
                                                   int bank) {
  if (S[bank * B_SIZE + B_DONE] != 0.0) return;
  double acc[1] = {
      rowblock_spmv16<double, true, CAPW, IW, NT>(nrb, desc, rp, (const widx_t<IW> *)widx, woff, v, p, q)};
  publish<1, NT>(acc, part);
}

template <int CAPW, int IW, int NT>
__global__ __launch_bounds__(NT, 2048 / NT) void k_spmv16(int64_t nrb, const int32_t *__restrict__ rp,
                                                       const void *__restrict__ widx,
                                                       const uint16_t *__restrict__ woff,
                                                       const int4 *__restrict__ desc, const float *__restrict__ v,
                                                       const float *__restrict__ x, float *__restrict__ y) {
  rowblock_spmv16<float, false, CAPW, IW, NT>(nrb, desc, rp, (const widx_t<IW> *)widx, woff, v, x, y);
}

// Build the x windows, one row block per workgroup: sort the block's
// (column, position) pairs in LDS; its distinct columns in ascending order form
// the window.  They are split at the largest gap into two clusters of spans s0,
// s1, which fit 16-bit offsets when s0 + s1 < 65535 (or the whole block spans
// < 65536).  COUNT: store each block's window size in wsz.  Otherwise wsz holds
// their exclusive prefix sums (the window starts): write the window offsets,
// each entry's window index and the block's descriptor (rows, entries, meta).  A block that does not fit
// (more than UCAP distinct columns, too wide a span, too many entries) raises
// *bad.  CAP: block capacity; IW: index bytes (the window limit is UCAP for
// 16-bit and W8_UCAP for 8-bit indices).  The counting pass reports every
// block's window size without judging it, so the host can split blocks whose
// window is too large.
template <bool COUNT, int CAP, int IW>
__global__ __launch_bounds__(BLK) void k_build_windows(const int32_t *__restrict__ rb, int64_t nrb,
                                                       const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                                       int32_t *__restrict__ wsz, widx_t<IW> *__restrict__ widx,
                                                       uint16_t *__restrict__ woff, int4 *__restrict__ meta,
                                                       int *__restrict__ bad) {
  __shared__ int64_t keys[CAP];
  __shared__ int32_t gv[BLK], gi[BLK];
  constexpr int PER = CAP / BLK;
  constexpr int UC = IW == 1 ? W8_UCAP : UCAP;
  static_assert(CAP <= 4096, "positions are kept in 12 key bits");
  const int tid = threadIdx.x;
  for (int64_t bk = blockIdx.x; bk < nrb; bk += gridDim.x) {
    const int32_t k0 = rp[rb[bk]], cnt = rp[rb[bk + 1]] - k0;
    if ((k0 & 7) + cnt > CAP) {  // the block from its start rounded down to 8 entries exceeds CAP
      if (tid == 0) atomicOr(bad, 1);
      continue;
    }
    for (int t = tid; t < CAP; t += BLK) keys[t] = t < cnt ? ((int64_t)ci[k0 + t] << 12) | t : INT64_MAX;
    __syncthreads();
    for (int k = 2; k <= CAP; k <<= 1)
      for (int j = k >> 1; j > 0; j >>= 1) {
        for (int i = tid; i < CAP; i += BLK) {
          const int pj = i ^ j;
          if (pj > i) {
            const int64_t a = keys[i], c = keys[pj];
            if ((a > c) == ((i & k) == 0)) {
              keys[i] = c;
              keys[pj] = a;
            }
          }
        }
        __syncthreads();
      }
    auto col = [&](int i) { return (int32_t)(keys[i] >> 12); };
    // thread tid owns the sorted entries PER*tid .. PER*tid + PER - 1
    int32_t nfl = 0, bestg = -1, besti = 0;
    for (int e = 0; e < PER; ++e) {
      const int i = PER * tid + e;
      if (i < cnt && (i == 0 || col(i) != col(i - 1))) ++nfl;
      if (i + 1 < cnt) {
        const int32_t g = col(i + 1) - col(i);
        if (g > bestg) {
          bestg = g;
          besti = i;
        }
      }
    }
    gv[tid] = bestg;
    gi[tid] = besti;
    __syncthreads();
    for (int w = BLK / 2; w > 0; w >>= 1) {
      if (tid < w && (gv[tid + w] > gv[tid] || (gv[tid + w] == gv[tid] && gi[tid + w] < gi[tid]))) {
        gv[tid] = gv[tid + w];
        gi[tid] = gi[tid + w];
      }
      __syncthreads();
    }
    const int32_t cmin = cnt > 0 ? col(0) : 0, cmax = cnt > 0 ? col(cnt - 1) : 0;
    int32_t b0 = cmin, b1 = cmin, s0 = 65535, split = INT_MAX;
    bool ok = true;
    if (cmax - cmin > 65535) {
      const int32_t glo = col(gi[0]), ghi = col(gi[0] + 1);
      s0 = glo - cmin;
      const int32_t s1 = cmax - ghi;
      ok = s0 + s1 + 1 <= 65535;
      b1 = ghi - (s0 + 1);
      split = glo;
    }
    __syncthreads();  // gi[0] read by every thread before gv is reused
    // window size and each thread's first window index: scan of the counts
    gv[tid] = nfl;
    __syncthreads();
    for (int d = 1; d < BLK; d <<= 1) {
      const int32_t add = tid >= d ? gv[tid - d] : 0;
      __syncthreads();
      gv[tid] += add;
      __syncthreads();
    }
    const int32_t nu = gv[BLK - 1], before = gv[tid] - nfl;
    if (!COUNT) ok = ok && nu <= UC;
    if (!ok) {
      if (tid == 0) atomicOr(bad, 1);
    } else if (COUNT) {
      if (tid == 0) wsz[bk] = nu;
    } else {
      const int32_t w0 = wsz[bk];
      int32_t u = before - 1;  // window index of the last distinct column before this thread's entries
      for (int e = 0; e < PER; ++e) {
        const int i = PER * tid + e;
        if (i < cnt) {
          const int32_t c = col(i);
          if (i == 0 || c != col(i - 1)) {
            ++u;
            woff[w0 + u] = (uint16_t)(c <= split ? c - b0 : c - b1);
          }
          widx[k0 + (int32_t)(keys[i] & 4095)] = (widx_t<IW>)u;
        }
      }
      if (tid == 0) {
        meta[2 * bk] = make_int4(rb[bk], rb[bk + 1], k0, cnt);
        meta[2 * bk + 1] = make_int4(b0, b1, s0 | (nu << 16), w0);
      }
    }
    __syncthreads();
  }
}

// alpha = (r,z)/(p,Ap); x += alpha p (x in XT, FP64 for both solvers);
// r -= alpha q; partials |r|^2 and (r, dinv r)
template <typename VT, typename XT>
__global__ __launch_bounds__(BLK) void k_it_update(int64_t n, const double *__restrict__ part_in, int G,
                                                   const VT *__restrict__ dinv, const VT *__restrict__ p,
                                                   const VT *__restrict__ q, XT *__restrict__ x,
                                                   VT *__restrict__ r, double *__restrict__ part,
                                                   double *__restrict__ S, int bank, int xvec) {
  double *Sb = S + bank * B_SIZE;
  if (Sb[B_DONE] != 0.0) return;
  double t[1];
  reduce_all<1>(part_in, G, t);
  const double alpha = Sb[B_RZ] / t[0];
  if (!(t[0] != 0.0) || !isfinite(alpha)) {  // every block sees the same totals
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      Sb[B_DONE] = 1;
      Sb[B_REASON] = LMH_DIVERGED_BREAKDOWN;
    }
    return;
  }
  const VT a = (VT)alpha;
  // 16-byte chunks of W entries (dinv, p, q, r start 16-byte aligned; x too
  // when xvec), then the workgroup's remaining entries one by one
  constexpr int W = 16 / sizeof(VT);
  typedef VT vt __attribute__((ext_vector_type(W)));
  typedef XT xt __attribute__((ext_vector_type(W)));
  int64_t lo, hi;
  block_range_aligned<W>(n, lo, hi);
  const int64_t hv = lo + ((hi - lo) / W) * W;
  double acc[2] = {0, 0};
  for (int64_t i = lo + W * threadIdx.x; i < hv; i += W * BLK) {
    const vt pv = *(const vt *)(p + i), qv = *(const vt *)(q + i), dv = *(const vt *)(dinv + i);
    vt rv = *(const vt *)(r + i);
    xt xv;
    if (xvec)
      xv = *(const xt *)(x + i);
    else
#pragma unroll
      for (int k = 0; k < W; ++k) xv[k] = x[i + k];
#pragma unroll
    for (int k = 0; k < W; ++k) {
      xv[k] += (XT)(alpha * (double)pv[k]);
      const VT rk = rv[k] - a * qv[k];
      rv[k] = rk;
      acc[0] += (double)rk * (double)rk;
      acc[1] += (double)rk * (double)(dv[k] * rk);
    }
    *(vt *)(r + i) = rv;
    if (xvec)
      *(xt *)(x + i) = xv;
    else
#pragma unroll
      for (int k = 0; k < W; ++k) x[i + k] = xv[k];
  }
  for (int64_t i = hv + threadIdx.x; i < hi; i += BLK) {
    x[i] += (XT)(alpha * (double)p[i]);
    const VT rv = r[i] - a * q[i];
    r[i] = rv;
    acc[0] += (double)rv * (double)rv;
    acc[1] += (double)rv * (double)(dinv[i] * rv);
  }
  publish<2>(acc, part);
}

// Convergence test / refresh request, then beta and p = dinv r + beta p
// (ConjugateGradient.h:76-87).  mixed: convergence is decided only on the FP64
// true residual; an FP32 residual below delta * (last true residual) or below
// the threshold raises the refresh request that the host serves.
template <typename VT>
__global__ __launch_bounds__(BLK) void k_it_pupdate(int64_t n, const double *__restrict__ part_in, int G,
                                                    const VT *__restrict__ dinv, const VT *__restrict__ r,
                                                    VT *__restrict__ p, double *__restrict__ S, int bank,
                                                    int mixed) {
  const double *Sb = S + bank * B_SIZE;
  double *Sn = S + (bank ^ 1) * B_SIZE;
  if (Sb[B_DONE] != 0.0) {
    if (blockIdx.x == 0 && threadIdx.x < B_SIZE) Sn[threadIdx.x] = Sb[threadIdx.x];
    return;
  }
  double t[2];
  reduce_all<2>(part_in, G, t);
  const double rr = t[0], rz = t[1];
  bool done = false;
  int reason = 0;
  double refresh = Sb[B_REFRESH];
  if (!isfinite(rr)) {
    done = true;
    reason = LMH_DIVERGED_NANORINF;
  } else if (!mixed && rr < S[G_THR]) {
    done = true;
    reason = LMH_CONVERGED_RTOL;
  } else if (mixed && (rr < S[G_DELTA2] * Sb[B_MAXRR] || rr < S[G_THR])) {
    refresh = 1;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    Sn[B_RR] = rr;
    Sn[B_DONE] = done ? 1.0 : 0.0;
    Sn[B_REASON] = reason;
    Sn[B_ITS] = Sb[B_ITS] + (done ? 0.0 : 1.0);
    Sn[B_RZ] = done ? Sb[B_RZ] : rz;
    Sn[B_REFRESH] = refresh;
    Sn[B_MAXRR] = Sb[B_MAXRR];
    Sn[B_NREF] = Sb[B_NREF];
  }
  if (done) return;
  const VT beta = (VT)(rz / Sb[B_RZ]);
  constexpr int W = 16 / sizeof(VT);
  typedef VT vt __attribute__((ext_vector_type(W)));
  int64_t lo, hi;
  block_range_aligned<W>(n, lo, hi);
  const int64_t hv = lo + ((hi - lo) / W) * W;
  for (int64_t i = lo + W * threadIdx.x; i < hv; i += W * BLK)
    *(vt *)(p + i) = *(const vt *)(dinv + i) * *(const vt *)(r + i) + beta * *(const vt *)(p + i);
  for (int64_t i = hv + threadIdx.x; i < hi; i += BLK) p[i] = dinv[i] * r[i] + beta * p[i];
}

// FP64 CG start: after k_residual (r = b - A x), p = dinv r; partials (r, p).
__global__ __launch_bounds__(BLK) void k_cg_init(int64_t n, const double *__restrict__ dinv,
                                                 const double *__restrict__ r, double *__restrict__ p,
                                                 double *__restrict__ part) {
  int64_t lo, hi;
  block_range(n, lo, hi);
  double acc[1] = {0};
  for (int64_t i = lo + threadIdx.x; i < hi; i += BLK) {
    const double z = dinv[i] * r[i];
    p[i] = z;
    acc[0] += r[i] * z;
  }
  publish<1>(acc, part);
}

// the early exits of ConjugateGradient.h:44-60; rr, bb from k_residual's partials
__global__ __launch_bounds__(BLK) void k_cg_init_fin(const double *__restrict__ part_res, int Gres,
                                                     const double *__restrict__ part_rz, int Grz, double tol,
                                                     double *__restrict__ S, double *__restrict__ x, int64_t n) {
  double t[2], u[1];
  reduce_all<2>(part_res, Gres, t);
  reduce_all<1>(part_rz, Grz, u);
  const double rr = t[0], bb = t[1];
  if (bb == 0.0)  // b = 0: x = 0 (this kernel runs with a full grid for that case)
    for (int64_t i = blockIdx.x * (int64_t)BLK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLK) x[i] = 0.0;
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const double thr = fmax(tol * tol * bb, 2.2250738585072014e-308);
  S[G_THR] = thr;
  S[G_BB] = bb;
  double *B0 = S;
  for (int k = 0; k < B_SIZE; ++k) B0[k] = 0.0;
  B0[B_RZ] = u[0];
  B0[B_RR] = rr;
  if (bb == 0.0) {
    B0[B_RR] = 0;
    B0[B_DONE] = 1;
    B0[B_REASON] = LMH_CONVERGED_RTOL;
  } else if (!isfinite(rr) || !isfinite(bb)) {  // NaN/Inf in b or x: stop before iterating
    B0[B_DONE] = 1;
    B0[B_REASON] = LMH_DIVERGED_NANORINF;
  } else if (rr < thr) {
    B0[B_DONE] = 1;
    B0[B_REASON] = LMH_CONVERGED_RTOL;
  }
}

// ---------------------------------------------------------------------------
// Mixed precision refresh: r64 = b - A64 x (FP64 operator); r32 = (float) r64;
// z = dinv32 r32 (p32 = z on the initial call); partials |r|^2, |b|^2, (r32, z).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BLK) void k_refresh(int init, const int32_t *__restrict__ rb, int64_t nrb,
                                                 const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                                 const double *__restrict__ v, const double *__restrict__ b,
                                                 const double *__restrict__ x, const float *__restrict__ dinv32,
                                                 float *__restrict__ r32, float *__restrict__ p32,
                                                 double *__restrict__ part, const double *__restrict__ S, int bank) {
  if (!init && S[bank * B_SIZE + B_DONE] != 0.0) return;
  double acc[3] = {0, 0, 0};
  rowblock_spmv<double, double, double, 2>(
      rb, nrb, rp, ci, v, x,
      [&](int32_t row, double *o) {
        o[0] = b[row];
        o[1] = (double)dinv32[row];
      },
      [&](int32_t row, double s, const double *pr) {
        const double bv = pr[0], rv = bv - s;
        const float rf = (float)rv;
        r32[row] = rf;
        const float z = (float)pr[1] * rf;
        if (init) p32[row] = z;
        acc[0] += rv * rv;
        acc[1] += bv * bv;
        acc[2] += (double)rf * (double)z;
      });
  publish<3>(acc, part);
}

// One block.  init: |b|^2, threshold and the starting bank; otherwise the
// FP64 convergence test on the true residual and the restart of the FP32
// recurrence's (r, z) for the next alpha.
__global__ __launch_bounds__(BLK) void k_refresh_fin(const double *__restrict__ part, int G, int init,
                                                     double *__restrict__ S, int bank) {
  double t[3];
  reduce_partials<3>(part, G, t);
  if (threadIdx.x != 0) return;
  double *Sb = S + bank * B_SIZE;
  if (!init && Sb[B_DONE] != 0.0) return;
  const double rr = t[0];
  S[G_RRTRUE] = rr;
  if (init) {
    const double tol2 = S[G_THR];  // rtol^2 on entry
    S[G_BB] = t[1];
    S[G_THR] = tol2 * t[1];
    for (int k = 0; k < B_SIZE; ++k) Sb[k] = 0.0;
    Sb[B_RZ] = t[2];
    Sb[B_RR] = rr;
    Sb[B_MAXRR] = rr;
    // non-finite first: an Inf in b makes the threshold Inf too, and rr <= Inf would pass
    if (!isfinite(rr) || !isfinite(t[1])) {
      Sb[B_DONE] = 1;
      Sb[B_REASON] = LMH_DIVERGED_NANORINF;
    } else if (t[1] == 0.0 || rr <= S[G_THR]) {
      Sb[B_DONE] = 1;
      Sb[B_REASON] = LMH_CONVERGED_RTOL;
    }
    return;
  }
  Sb[B_NREF] += 1;
  Sb[B_REFRESH] = 0;
  if (rr <= S[G_THR]) {
    Sb[B_DONE] = 1;
    Sb[B_REASON] = LMH_CONVERGED_RTOL;
  } else if (!isfinite(rr)) {
    Sb[B_DONE] = 1;
    Sb[B_REASON] = LMH_DIVERGED_NANORINF;
  } else {
    Sb[B_MAXRR] = rr;
    Sb[B_RZ] = t[2];
  }
}

// x += xc (FP64), xc = 0
__global__ __launch_bounds__(BLK) void k_fold(int64_t n, float *__restrict__ xc, double *__restrict__ x) {
  int64_t lo, hi;
  block_range(n, lo, hi);
  for (int64_t i = lo + threadIdx.x; i < hi; i += BLK) {
    x[i] += (double)xc[i];
    xc[i] = 0.f;
  }
}

// restarted refinement (mode 1): the correction solve starts over, p = dinv r
__global__ __launch_bounds__(BLK) void k_restart_p(int64_t n, const float *__restrict__ dinv32,
                                                   const float *__restrict__ r32, float *__restrict__ p32,
                                                   const double *__restrict__ S, int bank) {
  if (S[bank * B_SIZE + B_DONE] != 0.0) return;
  int64_t lo, hi;
  block_range(n, lo, hi);
  for (int64_t i = lo + threadIdx.x; i < hi; i += BLK) p32[i] = dinv32[i] * r32[i];
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
  const int g = ctx->cu_count * LMH_GRID_PER_CU;
  return g < ctx->n_part ? g : ctx->n_part;
}

// SpMV-type kernels (row-block streams) want more resident waves than the
// vector kernels: their grid is sized separately; every kernel that reduces
// partials is told the producer's grid.
static int spmv_grid(lmh_context *ctx) {
  // k_refresh publishes 3 values per workgroup into the second half (2 * n_part)
  const int g = ctx->cu_count * LMH_SPMV_GRID_PER_CU, cap = 2 * ctx->n_part / 3;
  return g < cap ? g : cap;
}

// x windows of the SpMV operators, built once per pattern on the row
// partition rb (host copy): window sizes (counting pass), blocks whose window
// exceeds the index width split by rows when `split` (at most 8 rounds; a
// single row over the limit does not fit), prefix sums on the host, then
// offsets, indices and descriptors.  *fits = false leaves A unchanged.
template <int CAP, int IW>
static int try_windows(lmh_matrix *A, std::vector<int32_t> rb, bool split, bool *fits) {
  constexpr int UC = IW == 1 ? W8_UCAP : UCAP;
  lmh_context *ctx = A->ctx;
  *fits = false;
  struct Tmp {
    void *p = nullptr;
    ~Tmp() { (void)hipFree(p); }
  } wsz, rbd;
  std::vector<int32_t> w0;
  int64_t nrb = 0;
  for (int round = 0;; ++round) {
    nrb = (int64_t)rb.size() - 1;
    (void)hipFree(rbd.p);
    (void)hipFree(wsz.p);
    rbd.p = wsz.p = nullptr;
    LMH_HIP(hipMalloc(&rbd.p, sizeof(int32_t) * rb.size()));
    LMH_HIP(hipMalloc(&wsz.p, sizeof(int32_t) * (nrb + 1)));
    LMH_HIP(hipMemcpyAsync(rbd.p, rb.data(), sizeof(int32_t) * rb.size(), hipMemcpyHostToDevice, ctx->stream));
    LMH_HIP(hipMemsetAsync(ctx->d_flags, 0, sizeof(int), ctx->stream));
    LMH_LAUNCH(ctx, "build_windows", (k_build_windows<true, CAP, IW>), dim3(grid_for(nrb, 1, 16384)), dim3(BLK), 0,
               (const int32_t *)rbd.p, nrb, A->rp, A->ci, (int32_t *)wsz.p, nullptr, nullptr, nullptr,
               ctx->d_flags);
    LMH_TRY(check_launch("k_build_windows"));
    w0.assign(nrb + 1, 0);
    int bad = 0;
    LMH_HIP(hipMemcpyAsync(w0.data(), wsz.p, sizeof(int32_t) * nrb, hipMemcpyDeviceToHost, ctx->stream));
    LMH_HIP(hipMemcpyAsync(&bad, ctx->d_flags, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    LMH_HIP(hipStreamSynchronize(ctx->stream));
    if (bad) return LMH_OK;
    bool over = false;
    for (int64_t b = 0; b < nrb; ++b) over |= w0[b] > UC;
    if (!over) break;
    if (!split || round == 8) return LMH_OK;
    std::vector<int32_t> nb{0};
    for (int64_t b = 0; b < nrb; ++b) {
      const int32_t r0 = rb[b], r1 = rb[b + 1];
      if (w0[b] > UC) {
        if (r1 - r0 < 2) return LMH_OK;  // one row alone exceeds the window
        const int32_t parts = std::min<int32_t>(r1 - r0, (w0[b] + (UC * 7 / 8) - 1) / (UC * 7 / 8));
        for (int32_t k = 1; k < parts; ++k) nb.push_back(r0 + (int32_t)((int64_t)(r1 - r0) * k / parts));
      }
      nb.push_back(r1);
    }
    rb.swap(nb);
  }
  int64_t total = 0;
  for (int64_t b = 0; b < nrb; ++b) {
    const int32_t s = w0[b];
    w0[b] = (int32_t)total;
    total += s;
  }
  w0[nrb] = (int32_t)total;  // at most nnz < 2^31
  LMH_HIP(hipMemcpyAsync(wsz.p, w0.data(), sizeof(int32_t) * (nrb + 1), hipMemcpyHostToDevice, ctx->stream));
  widx_t<IW> *widx = nullptr;
  uint16_t *woff = nullptr;
  int4 *meta = nullptr;
  LMH_HIP(hipMalloc((void **)&widx, sizeof(widx_t<IW>) * (A->nnz + LMH_PAD)));
  LMH_HIP(hipMalloc((void **)&woff, sizeof(uint16_t) * (total + LMH_PAD)));
  LMH_HIP(hipMalloc((void **)&meta, sizeof(int4) * 2 * (nrb ? nrb : 1)));
  LMH_HIP(hipMemsetAsync(ctx->d_flags, 0, sizeof(int), ctx->stream));
  LMH_LAUNCH(ctx, "build_windows", (k_build_windows<false, CAP, IW>), dim3(grid_for(nrb, 1, 16384)), dim3(BLK), 0,
             (const int32_t *)rbd.p, nrb, A->rp, A->ci, (int32_t *)wsz.p, widx, woff, meta, ctx->d_flags);
  LMH_TRY(check_launch("k_build_windows"));
  int bad = 0;
  LMH_HIP(hipMemcpyAsync(&bad, ctx->d_flags, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  LMH_HIP(hipStreamSynchronize(ctx->stream));
  if (bad) {  // not expected: the counting pass accepted every block
    LMH_HIP(hipFree(widx));
    LMH_HIP(hipFree(woff));
    LMH_HIP(hipFree(meta));
    return LMH_OK;
  }
  if (IW == 1)
    A->widx8 = (uint8_t *)widx;
  else
    A->widx16 = (uint16_t *)widx;
  A->woff16 = woff;
  A->rbmeta = meta;
  A->rbw = (int32_t *)rbd.p;
  rbd.p = nullptr;
  A->nrbw = nrb;
  A->wbits = 8 * IW;
  *fits = true;
  return LMH_OK;
}

// 8-bit windows on blocks of up to W8_CAP - 8 entries (split further where a
// window exceeds 256 columns), else 16-bit windows on the SpMV row blocks,
// else none (the int32-column kernels).
static int ensure_windows(lmh_matrix *A) {
  if (A->idx16_state != 0 || !A->fp32_index) return LMH_OK;
  lmh_context *ctx = A->ctx;
  std::vector<int32_t> rp(A->rows + 1);
  LMH_HIP(hipMemcpyAsync(rp.data(), A->rp, sizeof(int32_t) * (A->rows + 1), hipMemcpyDeviceToHost, ctx->stream));
  LMH_HIP(hipStreamSynchronize(ctx->stream));
  std::vector<int32_t> rb{0};
  for (int64_t r = 0; r < A->rows; ++r)
    if (r > rb.back() && rp[r + 1] - rp[rb.back()] > W8_CAP - 8) rb.push_back((int32_t)r);
  rb.push_back((int32_t)A->rows);
  bool fits = false;
  LMH_TRY((try_windows<W8_CAP, 1>(A, rb, true, &fits)));
  if (!fits) {
    std::vector<int32_t> rb16(A->nrb + 1);
    LMH_HIP(hipMemcpy(rb16.data(), A->rb, sizeof(int32_t) * (A->nrb + 1), hipMemcpyDeviceToHost));
    LMH_TRY((try_windows<RB_CAP, 2>(A, rb16, false, &fits)));
  }
  A->idx16_state = fits ? 1 : -1;
  return LMH_OK;
}

// at matrix creation: every solver entry point then finds the windows built,
// the allocation-free one included
int matrix_build_windows(lmh_matrix *A) { return ensure_windows(A); }

// stride of the solvers' work vectors: a multiple of 64 entries, so every
// vector starts 256-byte aligned (the vector kernels read 16-byte chunks)
static int64_t work_stride(int64_t n) { return (n + 63) & ~(int64_t)63; }

static int ensure_diag(lmh_matrix *A) {
  if (A->diag) return LMH_OK;
  LMH_HIP(hipMalloc((void **)&A->diag, sizeof(int32_t) * std::max<int64_t>(1, A->rows)));
  LMH_LAUNCH(A->ctx, "diag_index", k_diag_index, dim3(grid_for(A->rows, 256, 8192)), dim3(256), 0, A->rows, A->rp,
             A->ci, A->diag);
  return check_launch("k_diag_index");
}

static bool use_windows(const lmh_matrix *A) { return A->fp32_index && A->idx16_state == 1; }

static int ensure_work(lmh_matrix *A, bool mixed) {
  if (!A->w64) LMH_HIP(hipMalloc((void **)&A->w64, sizeof(double) * 4 * work_stride(A->rows)));
  if (mixed && !A->w32) LMH_HIP(hipMalloc((void **)&A->w32, sizeof(float) * 5 * work_stride(A->rows)));
  return LMH_OK;
}

}  // namespace lmh

using namespace lmh;

extern "C" {

int lmh_spmv(lmh_matrix *A, const double *x_d, double *y_d) {
  LMH_CHECK(A && x_d && y_d, LMH_ERR_INVALID, "null argument");
  LMH_CHECK(A->vals, LMH_ERR_UNSUPPORTED, "matrix holds FP32 values only");
  lmh_context *ctx = A->ctx;
  const int g = spmv_grid(ctx);
  LMH_LAUNCH(ctx, "spmv_fp64", (k_spmv<double, double, double, double>), dim3(g),
                                          dim3(BLK), 0, A->rb, A->nrb, A->rp, A->ci, A->vals, x_d, y_d);
  return check_launch("k_spmv");
}

int lmh_spmv_fp32(lmh_matrix *A, const float *x_d, float *y_d) {
  LMH_CHECK(A && x_d && y_d, LMH_ERR_INVALID, "null argument");
  LMH_CHECK(A->have32 && A->vals32, LMH_ERR_INVALID, "matrix has no FP32 operator");
  lmh_context *ctx = A->ctx;
  const int g = spmv_grid(ctx);
  LMH_TRY(ensure_windows(A));
  if (use_windows(A)) {
    if (A->wbits == 8)
      LMH_LAUNCH(ctx, "spmv_fp32", (k_spmv16<W8_CAP, 1, W8_NT>), dim3(W8_GRID * g), dim3(W8_NT), 0, A->nrbw,
                 A->rp, A->widx8, A->woff16, A->rbmeta, A->vals32, x_d, y_d);
    else
      LMH_LAUNCH(ctx, "spmv_fp32", (k_spmv16<RB_CAP, 2, BLK>), dim3(g), dim3(BLK), 0, A->nrbw, A->rp, A->widx16,
                 A->woff16, A->rbmeta, A->vals32, x_d, y_d);
    return check_launch("k_spmv16");
  }
  LMH_LAUNCH(ctx, "spmv_fp32", (k_spmv<float, float, float, float>), dim3(g),
                                          dim3(BLK), 0, A->rb, A->nrb, A->rp, A->ci, A->vals32, x_d, y_d);
  return check_launch("k_spmv");
}

int lmh_residual(lmh_matrix *A, const double *b_d, const double *x_d, double *r_d) {
  LMH_CHECK(A && b_d && x_d && r_d, LMH_ERR_INVALID, "null argument");
  LMH_CHECK(A->vals, LMH_ERR_UNSUPPORTED, "matrix holds FP32 values only");
  lmh_context *ctx = A->ctx;
  const int g = spmv_grid(ctx);
  LMH_LAUNCH(ctx, "residual_fp64", k_residual, dim3(g), dim3(BLK), 0, A->rb, A->nrb, A->rp,
                                          A->ci, A->vals, b_d, x_d, r_d, ctx->d_part);
  LMH_LAUNCH(ctx, "fin", k_fin_residual, dim3(1), dim3(BLK), 0, ctx->d_part, g, ctx->d_scal + 48);
  return check_launch("k_residual");
}

double lmh_norm2(lmh_context *ctx, const double *x_d, int64_t n) {
  if (!ctx || !x_d) return NAN;
  const int g = vec_grid(ctx);
  hipLaunchKernelGGL(k_norm2, dim3(g), dim3(BLK), 0, ctx->stream, n, x_d, ctx->d_part);
  hipLaunchKernelGGL(k_fin_sum, dim3(1), dim3(BLK), 0, ctx->stream, ctx->d_part, g, ctx->d_scal + 56);
  double v = NAN;
  if (hipMemcpyAsync(&v, ctx->d_scal + 56, sizeof(double), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess)
    return NAN;
  if (hipStreamSynchronize(ctx->stream) != hipSuccess) return NAN;
  return std::sqrt(v);
}

int lmh_axpy(lmh_context *ctx, int64_t n, double a, const double *x_d, double *y_d) {
  LMH_CHECK(ctx && x_d && y_d && n >= 0, LMH_ERR_INVALID, "bad argument");
  if (n == 0) return LMH_OK;
  LMH_LAUNCH(ctx, "axpy", k_axpy, dim3(grid_for(n, 256, 8192)), dim3(256), 0, n, a, x_d, y_d);
  return check_launch("k_axpy");
}

int lmh_matrix_make_fp32(lmh_matrix *A) {
  LMH_CHECK(A, LMH_ERR_INVALID, "null matrix");
  LMH_CHECK(A->vals, LMH_ERR_UNSUPPORTED, "matrix holds FP32 values only");
  lmh_context *ctx = A->ctx;
  if (!A->vals32) LMH_HIP(hipMalloc((void **)&A->vals32, sizeof(float) * (A->nnz + LMH_PAD)));
  LMH_LAUNCH(ctx, "round_fp32", k_round32, dim3(grid_for(A->nnz, 256, 16384)), dim3(256), 0, A->nnz, A->vals,
             A->vals32);
  A->have32 = true;
  return check_launch("k_round32");
}

// ---------------------------------------------------------------------------
// BiCGSTAB with the Jacobi preconditioner, FP64 (Eigen BiCGSTAB.h:28-112 as
// EigenSparseLinearSolver::solve runs it for BICGSTAB + JACOBI_PRECOND).  One
// iteration is five kernels: p/y update (reduces |r|^2 and (r0, r), tests
// convergence) | v = A y + (r0, v) | s, z | t = A z + (t, t), (t, s) | x, r
// update + |r|^2, (r0, r).  rho, alpha, w live in the banks next to the CG
// fields; each is written by block 0 of one kernel and read by later kernels.
// ---------------------------------------------------------------------------
enum { B_RHO = 8, B_ALPHA, B_W };
enum { G_R0SQ = 40 };

// r0 = r, p = v = 0; partials (r.r, r0.r)
__global__ __launch_bounds__(BLK) void k_bcg_init(int64_t n, const double *__restrict__ r, double *__restrict__ r0,
                                                  double *__restrict__ p, double *__restrict__ v,
                                                  double *__restrict__ part) {
  int64_t lo, hi;
  block_range(n, lo, hi);
  double acc[2] = {0, 0};
  for (int64_t i = lo + threadIdx.x; i < hi; i += BLK) {
    const double ri = r[i];
    r0[i] = ri;
    p[i] = 0.0;
    v[i] = 0.0;
    acc[0] += ri * ri;
    acc[1] += ri * ri;
  }
  publish<2>(acc, part);
}

// |b|^2 == 0: x = 0 and done; tol2 = tol^2 |b|^2, |r0|^2; rho = alpha = w = 1
__global__ __launch_bounds__(BLK) void k_bcg_init_fin(const double *__restrict__ part_res, int Gres, double tol,
                                                      double *__restrict__ S, double *__restrict__ x, int64_t n) {
  double t[2];
  reduce_all<2>(part_res, Gres, t);
  const double rr = t[0], bb = t[1];
  if (bb == 0.0)
    for (int64_t i = blockIdx.x * (int64_t)BLK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLK) x[i] = 0.0;
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  S[G_THR] = tol * tol * bb;
  S[G_BB] = bb;
  S[G_R0SQ] = rr;
  double *B0 = S;
  for (int k = 0; k < B_SIZE; ++k) B0[k] = 0.0;
  B0[B_RHO] = B0[B_ALPHA] = B0[B_W] = 1.0;
  B0[B_RR] = rr;
  if (bb == 0.0) {
    B0[B_RR] = 0.0;
    B0[B_DONE] = 1;
    B0[B_REASON] = LMH_CONVERGED_RTOL;
  } else if (!isfinite(rr) || !isfinite(bb)) {
    B0[B_DONE] = 1;
    B0[B_REASON] = LMH_DIVERGED_NANORINF;
  }
}

// top of the loop: rr = |r|^2, rho = (r0, r); stop when rr <= tol2; the
// restart of Eigen's loop (|rho| < eps^2 |r0|^2) is reported as a breakdown;
// else beta = (rho / rho_old)(alpha / w), p = r + beta (p - w v), y = dinv p
__global__ __launch_bounds__(BLK) void k_bcg_p(int64_t n, const double *__restrict__ part_in, int G,
                                               const double *__restrict__ dinv, const double *__restrict__ r,
                                               const double *__restrict__ v, double *__restrict__ p,
                                               double *__restrict__ y, double *__restrict__ S, int bank) {
  const double *Sb = S + bank * B_SIZE;
  double *Sn = S + (bank ^ 1) * B_SIZE;
  if (Sb[B_DONE] != 0.0) {
    if (blockIdx.x == 0 && threadIdx.x < B_SIZE) Sn[threadIdx.x] = Sb[threadIdx.x];
    return;
  }
  double t[2];
  reduce_all<2>(part_in, G, t);
  const double rr = t[0], rho = t[1];
  const double eps = 2.220446049250313e-16;
  bool done = true;
  int reason = 0;
  if (!isfinite(rr))
    reason = LMH_DIVERGED_NANORINF;
  else if (rr <= S[G_THR])
    reason = LMH_CONVERGED_RTOL;
  else if (fabs(rho) < eps * eps * S[G_R0SQ])
    reason = LMH_DIVERGED_BREAKDOWN;
  else
    done = false;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    for (int k = 0; k < B_SIZE; ++k) Sn[k] = Sb[k];
    Sn[B_RR] = rr;
    Sn[B_DONE] = done ? 1.0 : 0.0;
    Sn[B_REASON] = reason;
    Sn[B_RHO] = done ? Sb[B_RHO] : rho;
  }
  if (done) return;
  const double beta = (rho / Sb[B_RHO]) * (Sb[B_ALPHA] / Sb[B_W]), w = Sb[B_W];
  int64_t lo, hi;
  block_range(n, lo, hi);
  for (int64_t i = lo + threadIdx.x; i < hi; i += BLK) {
    const double pi = r[i] + beta * (p[i] - w * v[i]);
    p[i] = pi;
    y[i] = dinv[i] * pi;
  }
}

// v = A y; partials (r0, v)
__global__ __launch_bounds__(BLK) void k_bcg_spmv_v(const int32_t *__restrict__ rb, int64_t nrb,
                                                    const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                                    const double *__restrict__ vals, const double *__restrict__ y,
                                                    const double *__restrict__ r0, double *__restrict__ v,
                                                    double *__restrict__ part, const double *__restrict__ S,
                                                    int nbank) {
  if (S[nbank * B_SIZE + B_DONE] != 0.0) return;
  double acc[1] = {0};
  rowblock_spmv<double, double, double, 1>(
      rb, nrb, rp, ci, vals, y, [&](int32_t row, double *o) { o[0] = r0[row]; },
      [&](int32_t row, double sv, const double *pr) {
        v[row] = sv;
        acc[0] += pr[0] * sv;
      });
  publish<1>(acc, part);
}

// alpha = rho / (r0, v); s = r - alpha v; z = dinv s
__global__ __launch_bounds__(BLK) void k_bcg_s(int64_t n, const double *__restrict__ part_in, int G,
                                               const double *__restrict__ dinv, const double *__restrict__ r,
                                               const double *__restrict__ v, double *__restrict__ s,
                                               double *__restrict__ z, double *__restrict__ S, int nbank) {
  double *Sn = S + nbank * B_SIZE;
  if (Sn[B_DONE] != 0.0) return;
  double t[1];
  reduce_all<1>(part_in, G, t);
  const double alpha = Sn[B_RHO] / t[0];
  if (blockIdx.x == 0 && threadIdx.x == 0) Sn[B_ALPHA] = alpha;
  int64_t lo, hi;
  block_range(n, lo, hi);
  for (int64_t i = lo + threadIdx.x; i < hi; i += BLK) {
    const double si = r[i] - alpha * v[i];
    s[i] = si;
    z[i] = dinv[i] * si;
  }
}

// t = A z; partials (t, t), (t, s)
__global__ __launch_bounds__(BLK) void k_bcg_spmv_t(const int32_t *__restrict__ rb, int64_t nrb,
                                                    const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                                    const double *__restrict__ vals, const double *__restrict__ z,
                                                    const double *__restrict__ s, double *__restrict__ t,
                                                    double *__restrict__ part, const double *__restrict__ S,
                                                    int nbank) {
  if (S[nbank * B_SIZE + B_DONE] != 0.0) return;
  double acc[2] = {0, 0};
  rowblock_spmv<double, double, double, 1>(
      rb, nrb, rp, ci, vals, z, [&](int32_t row, double *o) { o[0] = s[row]; },
      [&](int32_t row, double tv, const double *pr) {
        t[row] = tv;
        acc[0] += tv * tv;
        acc[1] += tv * pr[0];
      });
  publish<2>(acc, part);
}

// w = (t, s) / (t, t) (0 when (t, t) = 0); x += alpha y + w z; r = s - w t;
// partials |r|^2, (r0, r)
__global__ __launch_bounds__(BLK) void k_bcg_update(int64_t n, const double *__restrict__ part_in, int G,
                                                    const double *__restrict__ y, const double *__restrict__ z,
                                                    const double *__restrict__ s, const double *__restrict__ t,
                                                    const double *__restrict__ r0, double *__restrict__ x,
                                                    double *__restrict__ r, double *__restrict__ part,
                                                    double *__restrict__ S, int nbank) {
  double *Sn = S + nbank * B_SIZE;
  if (Sn[B_DONE] != 0.0) return;
  double tt[2];
  reduce_all<2>(part_in, G, tt);
  const double w = tt[0] > 0.0 ? tt[1] / tt[0] : 0.0, alpha = Sn[B_ALPHA];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    Sn[B_W] = w;
    Sn[B_ITS] = Sn[B_ITS] + 1.0;
  }
  int64_t lo, hi;
  block_range(n, lo, hi);
  double acc[2] = {0, 0};
  for (int64_t i = lo + threadIdx.x; i < hi; i += BLK) {
    x[i] += alpha * y[i] + w * z[i];
    const double ri = s[i] - w * t[i];
    r[i] = ri;
    acc[0] += ri * ri;
    acc[1] += r0[i] * ri;
  }
  publish<2>(acc, part);
}

// Poll slots in pinned memory: the two banks are copied after every batch.
struct Poll {
  hipEvent_t ev = nullptr;
  int64_t end = 0;  // iterations launched when the copy was enqueued
};

}  // extern "C"

namespace lmh {

// Work vectors of the allocation-free CG: r, p, q, 1/a_ii (4 x LMH_CG_NOALLOC_MAX_ROWS
// doubles), then the diagonal slots (LMH_CG_NOALLOC_MAX_ROWS int32).  Each GPU
// holds its own copy until the process ends; host code reaches it only through
// the address hipGetSymbolAddress returns for that device (cg_scratch_setup).
static __device__ __align__(256) double scratch[4 * (size_t)LMH_CG_NOALLOC_MAX_ROWS + LMH_CG_NOALLOC_MAX_ROWS / 2];

static std::mutex g_scratch_mutex;
static void *g_scratch_ptr[LMH_MAX_DEVICES] = {};

int cg_scratch_setup(lmh_context *ctx) {
  LMH_CHECK(ctx->device >= 0 && ctx->device < LMH_MAX_DEVICES, LMH_ERR_UNSUPPORTED, "device %d beyond %d",
            ctx->device, LMH_MAX_DEVICES);
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[ctx->device]) {
    // the caller made ctx->device current
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return set_error(LMH_ERR_HIP, "hipGetSymbolAddress of the CG scratch failed on device %d: %s", ctx->device,
                       hipGetErrorString(e));
    g_scratch_ptr[ctx->device] = ptr;
  }
  ctx->cg_scratch = g_scratch_ptr[ctx->device];
  return LMH_OK;
}

// Jacobi/identity CG iterations on A (FP64).  r, p, q, dinv: n doubles each;
// diag: A's diagonal slots.  NA (no allocation): launches without the
// profiling events and uses the x windows only when they are already built.
template <bool NA>
static int cg_impl(lmh_matrix *A, const double *b_d, double *x_d, double rtol, int64_t maxits, double *r, double *p,
                   double *q, double *dinv, const int32_t *diag, int64_t *iters, double *rel_res, int *reason) {
#define CG_LAUNCH_NT(name, kernel, grid, nt, ...)                                            \
  do {                                                                                      \
    if (NA)                                                                                 \
      hipLaunchKernelGGL(kernel, grid, dim3(nt), 0, ctx->stream, __VA_ARGS__);              \
    else                                                                                    \
      LMH_LAUNCH(ctx, name, kernel, grid, dim3(nt), 0, __VA_ARGS__);                        \
  } while (0)
#define CG_LAUNCH(name, kernel, grid, ...) CG_LAUNCH_NT(name, kernel, grid, BLK, __VA_ARGS__)
  lmh_context *ctx = A->ctx;
  if (!NA) LMH_TRY(ensure_windows(A));
  const bool c16 = use_windows(A);
  const int64_t n = A->rows;
  double *S = ctx->d_scal;
  double *part1 = ctx->d_part, *part2 = ctx->d_part + 2 * ctx->n_part;  // 1 and 2 values per block
  const int g = vec_grid(ctx), gs = spmv_grid(ctx);
  const int gsp = c16 && A->wbits == 8 ? W8_GRID * gs : gs;  // the iteration SpMV's grid
  const int xvec = ((uintptr_t)x_d & 15) == 0;               // the caller's x takes 16-byte chunks
  CG_LAUNCH("invdiag", (k_invdiag<double, double>), dim3(grid_for(n, BLK, 8192)), n, diag, A->vals,
            A->precond == LMH_PRECOND_JACOBI, dinv);
  CG_LAUNCH("residual_fp64", k_residual, dim3(gs), A->rb, A->nrb, A->rp, A->ci, A->vals, b_d, x_d, r, part2);
  CG_LAUNCH("cg_init", k_cg_init, dim3(g), n, dinv, r, p, part1);
  CG_LAUNCH("cg_init", k_cg_init_fin, dim3(g), part2, gs, part1, g, rtol, S, x_d, n);
  LMH_TRY(check_launch("cg setup"));
  // The scalar banks are copied to pinned memory after every batch of
  // iterations; the host reads a batch's copy while the next batch runs.
  const int batch = 4;
  int64_t end[2] = {0, 0};
  double *snap = ctx->h_pinned;  // 2 x (2 banks)
  int64_t K = 0;
  int slot = 0;
  bool pending = false, done = false;
  while (K < maxits && !done) {
    const int64_t nb = std::min<int64_t>(batch, maxits - K);
    for (int64_t i = 0; i < nb; ++i, ++K) {
      const int bank = (int)(K & 1);
      if (c16 && A->wbits == 8)
        CG_LAUNCH_NT("cg_spmv_fp64", (k_it_spmv16d<W8_CAP, 1, W8_NT>), dim3(gsp), W8_NT, A->nrbw, A->rp, A->widx8,
                     A->woff16, A->rbmeta, A->vals, p, q, part1, S, bank);
      else if (c16)
        CG_LAUNCH("cg_spmv_fp64", (k_it_spmv16d<RB_CAP, 2, BLK>), dim3(gs), A->nrbw, A->rp, A->widx16, A->woff16,
                  A->rbmeta, A->vals, p, q, part1, S, bank);
      else
        CG_LAUNCH("cg_spmv_fp64", k_it_spmv<double>, dim3(gs), A->rb, A->nrb, A->rp, A->ci, A->vals, p, q, part1, S,
                  bank);
      CG_LAUNCH("cg_update_fp64", (k_it_update<double, double>), dim3(g), n, part1, gsp, dinv, p, q, x_d, r, part2, S,
                bank, xvec);
      CG_LAUNCH("cg_pupdate_fp64", k_it_pupdate<double>, dim3(g), n, part2, g, dinv, r, p, S, bank, 0);
    }
    LMH_HIP(hipMemcpyAsync(snap + 32 * slot, S, sizeof(double) * 2 * B_SIZE, hipMemcpyDeviceToHost, ctx->stream));
    LMH_HIP(hipEventRecord(ctx->ev_poll[slot], ctx->stream));
    end[slot] = K;
    if (pending) {
      const int o = slot ^ 1;
      LMH_HIP(hipEventSynchronize(ctx->ev_poll[o]));
      done = snap[32 * o + (end[o] & 1) * B_SIZE + B_DONE] != 0.0;
    }
    pending = true;
    slot ^= 1;
  }
#undef CG_LAUNCH
  LMH_TRY(check_launch("cg iterations"));
  // both banks and |b|^2 (G_BB follows the banks) into pinned memory
  static_assert(G_BB < 64, "scalar slots beyond the pinned copy");
  LMH_HIP(hipMemcpyAsync(snap, S, sizeof(double) * (G_BB + 1), hipMemcpyDeviceToHost, ctx->stream));
  LMH_HIP(hipStreamSynchronize(ctx->stream));
  const double *Bk = snap + (K & 1) * B_SIZE, gbb = snap[G_BB];
  const bool conv = Bk[B_DONE] != 0.0;
  if (iters) *iters = conv ? (int64_t)Bk[B_ITS] : maxits;
  if (rel_res) *rel_res = gbb > 0 ? std::sqrt(Bk[B_RR] / gbb) : 0.0;
  if (reason) *reason = conv ? (int)Bk[B_REASON] : LMH_DIVERGED_ITS;
  return LMH_OK;
}

}  // namespace lmh

extern "C" {

int lmh_cg_solve(lmh_matrix *A, const double *b_d, double *x_d, double rtol, int64_t maxits, int64_t *iters,
                 double *rel_res, int *reason) {
  LMH_CHECK(A && b_d && x_d, LMH_ERR_INVALID, "null argument");
  LMH_CHECK(rtol >= 0 && maxits >= 0, LMH_ERR_INVALID, "bad tolerance or iteration limit");
  LMH_CHECK(A->rows == A->cols, LMH_ERR_INVALID, "CG needs a square matrix");
  LMH_CHECK(A->vals, LMH_ERR_UNSUPPORTED, "matrix holds FP32 values only");
  LMH_TRY(ensure_work(A, false));
  LMH_TRY(ensure_diag(A));
  const int64_t n = A->rows;
  double *r = A->w64;
  const int64_t ns = work_stride(n);
  return cg_impl<false>(A, b_d, x_d, rtol, maxits, r, r + ns, r + 2 * ns, r + 3 * ns, A->diag, iters, rel_res,
                        reason);
}

int lmh_cg_solve_noalloc(lmh_matrix *A, const double *b_d, double *x_d, double rtol, int64_t maxits,
                         int64_t *iters, double *rel_res, int *reason) {
  LMH_CHECK(A && b_d && x_d, LMH_ERR_INVALID, "null argument");
  LMH_CHECK(rtol >= 0 && maxits >= 0, LMH_ERR_INVALID, "bad tolerance or iteration limit");
  LMH_CHECK(A->rows == A->cols, LMH_ERR_INVALID, "CG needs a square matrix");
  LMH_CHECK(A->vals, LMH_ERR_UNSUPPORTED, "matrix holds FP32 values only");
  LMH_CHECK(A->rows <= LMH_CG_NOALLOC_MAX_ROWS, LMH_ERR_INVALID,
            "lmh_cg_solve_noalloc: %lld rows exceed its scratch limit of %d rows (use lmh_cg_solve)",
            (long long)A->rows, (int)LMH_CG_NOALLOC_MAX_ROWS);
  lmh_context *ctx = A->ctx;
  LMH_CHECK(ctx->cg_scratch, LMH_ERR_HIP, "CG scratch address not set up for device %d", ctx->device);
  const int64_t n = A->rows;
  const size_t cap = LMH_CG_NOALLOC_MAX_ROWS;
  double *w = static_cast<double *>(ctx->cg_scratch);
  int32_t *diag = reinterpret_cast<int32_t *>(w + 4 * cap);
  hipLaunchKernelGGL(k_diag_index, dim3(grid_for(n, 256, 8192)), dim3(256), 0, ctx->stream, n, A->rp, A->ci, diag);
  LMH_TRY(check_launch("k_diag_index"));
  return cg_impl<true>(A, b_d, x_d, rtol, maxits, w, w + cap, w + 2 * cap, w + 3 * cap, diag, iters, rel_res,
                       reason);
}

}  // extern "C"

namespace lmh {

// Jacobi/identity BiCGSTAB on A's pattern with the values `vals` (A's own, or
// the mirrored values of a structurally symmetric A, i.e. A transposed).
static int bicgstab_impl(lmh_matrix *A, const double *vals, const double *b_d, double *x_d, double rtol,
                         int64_t maxits, int64_t *iters, double *rel_res, int *reason) {
  lmh_context *ctx = A->ctx;
  const int64_t n = A->rows;
  if (!A->wbcg) LMH_HIP(hipMalloc((void **)&A->wbcg, sizeof(double) * 9 * std::max<int64_t>(1, n)));
  double *r = A->wbcg, *r0 = r + n, *p = r0 + n, *v = p + n, *y = v + n, *z = y + n, *sv = z + n, *tv = sv + n,
         *dinv = tv + n;
  double *S = ctx->d_scal;
  double *part1 = ctx->d_part, *part2 = ctx->d_part + 2 * ctx->n_part;
  const int g = vec_grid(ctx), gs = spmv_grid(ctx);
  LMH_TRY(ensure_diag(A));
  LMH_LAUNCH(ctx, "invdiag", (k_invdiag<double, double>), dim3(grid_for(n, 256, 8192)), dim3(256), 0, n, A->diag,
             vals, A->precond == LMH_PRECOND_JACOBI, dinv);
  LMH_LAUNCH(ctx, "residual_fp64", k_residual, dim3(gs), dim3(BLK), 0, A->rb, A->nrb, A->rp, A->ci, vals, b_d,
             x_d, r, part2);
  LMH_LAUNCH(ctx, "bcg_init", k_bcg_init, dim3(g), dim3(BLK), 0, n, (const double *)r, r0, p, v, part1);
  LMH_LAUNCH(ctx, "bcg_init", k_bcg_init_fin, dim3(g), dim3(BLK), 0, part2, gs, rtol, S, x_d, n);
  LMH_TRY(check_launch("bicgstab setup"));
  const int batch = 4;
  Poll poll[2];
  double *snap = ctx->h_pinned;
  int64_t K = 0;
  int slot = 0;
  bool pending = false, done = false;
  auto iteration = [&](int bank) {
    const int nb = bank ^ 1;
    LMH_LAUNCH(ctx, "bcg_p", k_bcg_p, dim3(g), dim3(BLK), 0, n, part1, g, dinv, r, v, p, y, S, bank);
    LMH_LAUNCH(ctx, "bcg_spmv_fp64", k_bcg_spmv_v, dim3(gs), dim3(BLK), 0, A->rb, A->nrb, A->rp, A->ci, vals, y,
               r0, v, part1, S, nb);
    LMH_LAUNCH(ctx, "bcg_s", k_bcg_s, dim3(g), dim3(BLK), 0, n, part1, gs, dinv, r, v, sv, z, S, nb);
    LMH_LAUNCH(ctx, "bcg_spmv_fp64", k_bcg_spmv_t, dim3(gs), dim3(BLK), 0, A->rb, A->nrb, A->rp, A->ci, vals, z,
               sv, tv, part2, S, nb);
    LMH_LAUNCH(ctx, "bcg_update", k_bcg_update, dim3(g), dim3(BLK), 0, n, part2, gs, y, z, sv, tv, r0, x_d, r, part1,
               S, nb);
  };
  while (K < maxits && !done) {
    const int64_t nb = std::min<int64_t>(batch, maxits - K);
    for (int64_t i = 0; i < nb; ++i, ++K) iteration((int)(K & 1));
    LMH_HIP(hipMemcpyAsync(snap + 32 * slot, S, sizeof(double) * 2 * B_SIZE, hipMemcpyDeviceToHost, ctx->stream));
    if (!poll[slot].ev) poll[slot].ev = get_event(ctx);
    LMH_HIP(hipEventRecord(poll[slot].ev, ctx->stream));
    poll[slot].end = K;
    if (pending) {
      const int o = slot ^ 1;
      LMH_HIP(hipEventSynchronize(poll[o].ev));
      done = snap[32 * o + (poll[o].end & 1) * B_SIZE + B_DONE] != 0.0;
    }
    pending = true;
    slot ^= 1;
  }
  // the loop test after the last iteration (Eigen: while (|r|^2 > tol2 && i < maxIters))
  LMH_LAUNCH(ctx, "bcg_p", k_bcg_p, dim3(g), dim3(BLK), 0, n, part1, g, dinv, r, v, p, y, S, (int)(K & 1));
  LMH_HIP(hipStreamSynchronize(ctx->stream));
  LMH_TRY(check_launch("bicgstab iterations"));
  for (auto &pl : poll)
    if (pl.ev) ctx->event_pool.push_back(pl.ev);
  double h[2 * B_SIZE], gbb = 0;
  LMH_HIP(hipMemcpy(h, S, sizeof(h), hipMemcpyDeviceToHost));
  LMH_HIP(hipMemcpy(&gbb, S + G_BB, sizeof(double), hipMemcpyDeviceToHost));
  const double *Bk = h + ((K + 1) & 1) * B_SIZE;
  const bool conv = Bk[B_DONE] != 0.0;
  if (iters) *iters = (int64_t)Bk[B_ITS];
  if (rel_res) *rel_res = gbb > 0 ? std::sqrt(Bk[B_RR] / gbb) : 0.0;
  if (reason) *reason = conv ? (int)Bk[B_REASON] : LMH_DIVERGED_ITS;
  return LMH_OK;
}

// Position of each entry's mirror: tpos[k] = position of (c, r) for entry
// k = (r, c), found by binary search in row c; *bad is raised when a mirror is
// missing (the pattern is not structurally symmetric).
__global__ __launch_bounds__(256) void k_transpose_pos(int64_t n, const int32_t *__restrict__ rp,
                                                       const int32_t *__restrict__ ci, int32_t *__restrict__ tpos,
                                                       int *__restrict__ bad) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x)
    for (int32_t k = rp[r]; k < rp[r + 1]; ++k) {
      const int32_t c = ci[k];
      int32_t lo = rp[c], hi = rp[c + 1];
      while (lo < hi) {
        const int32_t mid = lo + (hi - lo) / 2;
        if (ci[mid] < r)
          lo = mid + 1;
        else
          hi = mid;
      }
      const bool found = lo < rp[c + 1] && ci[lo] == r;
      tpos[k] = found ? lo : k;
      if (!found) atomicOr(bad, 1);
    }
}

__global__ __launch_bounds__(256) void k_gather_mirror(int64_t nnz, const int32_t *__restrict__ tpos,
                                                       const double *__restrict__ v, double *__restrict__ vt) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < nnz; k += (int64_t)gridDim.x * blockDim.x)
    vt[k] = v[tpos[k]];
}

}  // namespace lmh

extern "C" {

int lmh_bicgstab_solve(lmh_matrix *A, const double *b_d, double *x_d, double rtol, int64_t maxits, int64_t *iters,
                       double *rel_res, int *reason) {
  LMH_CHECK(A && b_d && x_d, LMH_ERR_INVALID, "null argument");
  LMH_CHECK(rtol >= 0 && maxits >= 0, LMH_ERR_INVALID, "bad tolerance or iteration limit");
  LMH_CHECK(A->rows == A->cols, LMH_ERR_INVALID, "BiCGSTAB needs a square matrix");
  LMH_CHECK(A->vals, LMH_ERR_UNSUPPORTED, "matrix holds FP32 values only");
  return bicgstab_impl(A, A->vals, b_d, x_d, rtol, maxits, iters, rel_res, reason);
}

int lmh_bicgstab_solve_transpose(lmh_matrix *A, const double *b_d, double *x_d, double rtol, int64_t maxits,
                                 int64_t *iters, double *rel_res, int *reason) {
  LMH_CHECK(A && b_d && x_d, LMH_ERR_INVALID, "null argument");
  LMH_CHECK(rtol >= 0 && maxits >= 0, LMH_ERR_INVALID, "bad tolerance or iteration limit");
  LMH_CHECK(A->rows == A->cols, LMH_ERR_INVALID, "BiCGSTAB needs a square matrix");
  LMH_CHECK(A->vals, LMH_ERR_UNSUPPORTED, "matrix holds FP32 values only");
  lmh_context *ctx = A->ctx;
  if (A->tpos_state == 0) {
    LMH_HIP(hipMalloc((void **)&A->tpos, sizeof(int32_t) * std::max<int64_t>(1, A->nnz)));
    LMH_HIP(hipMemsetAsync(ctx->d_flags, 0, sizeof(int), ctx->stream));
    LMH_LAUNCH(ctx, "transpose_pos", k_transpose_pos, dim3(grid_for(A->rows, 256, 8192)), dim3(256), 0, A->rows,
               A->rp, A->ci, A->tpos, ctx->d_flags);
    LMH_TRY(check_launch("k_transpose_pos"));
    int bad = 0;
    LMH_HIP(hipMemcpyAsync(&bad, ctx->d_flags, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    LMH_HIP(hipStreamSynchronize(ctx->stream));
    A->tpos_state = bad ? -1 : 1;
  }
  LMH_CHECK(A->tpos_state == 1, LMH_ERR_UNSUPPORTED,
            "transpose solve needs a structurally symmetric pattern (every (i, j) with its (j, i))");
  if (!A->vals_t) LMH_HIP(hipMalloc((void **)&A->vals_t, sizeof(double) * (A->nnz + LMH_PAD)));
  LMH_LAUNCH(ctx, "gather_mirror", k_gather_mirror, dim3(grid_for(A->nnz, 256, 16384)), dim3(256), 0, A->nnz,
             A->tpos, A->vals, A->vals_t);
  LMH_TRY(check_launch("k_gather_mirror"));
  return bicgstab_impl(A, A->vals_t, b_d, x_d, rtol, maxits, iters, rel_res, reason);
}

int lmh_mixed_solve(lmh_matrix *A, const double *b_d, double *x_d, double rtol, double inner_rtol, int mode,
                    int64_t maxits, int64_t *iters, double *rel_res, int *reason, double *stats_h) {
  LMH_CHECK(A && b_d && x_d, LMH_ERR_INVALID, "null argument");
  LMH_CHECK(A->vals, LMH_ERR_UNSUPPORTED, "matrix holds FP32 values only");
  LMH_CHECK(A->have32 && A->vals32, LMH_ERR_INVALID, "mixed solve needs the FP32 operator (assemble FP32 or make_fp32)");
  LMH_CHECK(mode == 0 || mode == 1, LMH_ERR_INVALID, "mode must be 0 (reliable update) or 1 (restarted refinement)");
  LMH_CHECK(A->rows == A->cols, LMH_ERR_INVALID, "the solver needs a square matrix");
  LMH_CHECK(rtol > 0 && inner_rtol > 0 && inner_rtol < 1, LMH_ERR_INVALID, "bad tolerances");
  lmh_context *ctx = A->ctx;
  LMH_TRY(ensure_work(A, true));
  LMH_TRY(ensure_windows(A));
  const bool c16 = use_windows(A);
  const int64_t n = A->rows;
  // xc: FP32 solution correction accumulated since the last refresh, folded
  // into the FP64 x before each refresh
  const int64_t ns = work_stride(n);
  float *r32 = A->w32, *p32 = r32 + ns, *q32 = p32 + ns, *d32 = q32 + ns, *xc = d32 + ns;
  double *S = ctx->d_scal;
  double *part1 = ctx->d_part, *part2 = ctx->d_part + 2 * ctx->n_part;
  const int g = vec_grid(ctx), gs = spmv_grid(ctx);
  const int gsp = c16 && A->wbits == 8 ? W8_GRID * gs : gs;  // the iteration SpMV's grid
  double init[40] = {0};
  init[G_THR] = rtol * rtol;
  init[G_DELTA2] = inner_rtol * inner_rtol;
  init[G_MODE] = mode;
  LMH_HIP(hipMemcpyAsync(S, init, sizeof(init), hipMemcpyHostToDevice, ctx->stream));
  LMH_TRY(ensure_diag(A));
  LMH_LAUNCH(ctx, "invdiag", (k_invdiag<float, float>), dim3(grid_for(n, 256, 8192)), dim3(256), 0, n, A->diag,
             A->vals32, A->precond == LMH_PRECOND_JACOBI, d32);
  LMH_HIP(hipMemsetAsync(xc, 0, sizeof(float) * n, ctx->stream));
  LMH_LAUNCH(ctx, "mx_refresh_fp64", k_refresh, dim3(gs), dim3(BLK), 0, 1, A->rb, A->nrb, A->rp, A->ci, A->vals,
             b_d, x_d, d32, r32, p32, part2, S, 0);
  LMH_LAUNCH(ctx, "fin", k_refresh_fin, dim3(1), dim3(BLK), 0, part2, gs, 1, S, 0);
  LMH_TRY(check_launch("mixed setup"));
  auto refresh = [&](int bank) {
    LMH_LAUNCH(ctx, "mx_fold", k_fold, dim3(g), dim3(BLK), 0, n, xc, x_d);
    LMH_LAUNCH(ctx, "mx_refresh_fp64", k_refresh, dim3(gs), dim3(BLK), 0, 0, A->rb, A->nrb, A->rp, A->ci, A->vals,
               b_d, x_d, d32, r32, p32, part2, S, bank);
    LMH_LAUNCH(ctx, "fin", k_refresh_fin, dim3(1), dim3(BLK), 0, part2, gs, 0, S, bank);
    if (mode == 1) LMH_LAUNCH(ctx, "mx_restart", k_restart_p, dim3(g), dim3(BLK), 0, n, d32, r32, p32, S, bank);
  };
  const int batch = 4;
  Poll poll[2];
  double *snap = ctx->h_pinned;
  int64_t K = 0, last_refresh = 0;
  int slot = 0;
  bool pending = false, done = false;
  while (K < maxits && !done) {
    const int64_t nb = std::min<int64_t>(batch, maxits - K);
    for (int64_t i = 0; i < nb; ++i, ++K) {
      const int bank = (int)(K & 1);
      if (c16 && A->wbits == 8)
        LMH_LAUNCH(ctx, "mx_spmv_fp32", (k_it_spmv16<W8_CAP, 1, W8_NT>), dim3(gsp), dim3(W8_NT), 0, A->nrbw, A->rp,
                   A->widx8, A->woff16, A->rbmeta, A->vals32, p32, q32, part1, S, bank);
      else if (c16)
        LMH_LAUNCH(ctx, "mx_spmv_fp32", (k_it_spmv16<RB_CAP, 2, BLK>), dim3(gs), dim3(BLK), 0, A->nrbw, A->rp,
                   A->widx16, A->woff16, A->rbmeta, A->vals32, p32, q32, part1, S, bank);
      else
        LMH_LAUNCH(ctx, "mx_spmv_fp32", k_it_spmv<float>, dim3(gs), dim3(BLK), 0, A->rb, A->nrb, A->rp, A->ci,
                   A->vals32, p32, q32, part1, S, bank);
      LMH_LAUNCH(ctx, "mx_update_fp32", (k_it_update<float, float>), dim3(g), dim3(BLK), 0, n, part1, gsp, d32, p32,
                 q32, xc, r32, part2, S, bank, 1);
      LMH_LAUNCH(ctx, "mx_pupdate_fp32", k_it_pupdate<float>, dim3(g), dim3(BLK), 0, n, part2, g, d32, r32, p32, S,
                 bank, 1);
    }
    LMH_HIP(hipMemcpyAsync(snap + 32 * slot, S, sizeof(double) * 2 * B_SIZE, hipMemcpyDeviceToHost, ctx->stream));
    if (!poll[slot].ev) poll[slot].ev = get_event(ctx);
    LMH_HIP(hipEventRecord(poll[slot].ev, ctx->stream));
    poll[slot].end = K;
    if (pending) {
      const int o = slot ^ 1;
      LMH_HIP(hipEventSynchronize(poll[o].ev));
      const double *Bo = snap + 32 * o + (poll[o].end & 1) * B_SIZE;
      done = Bo[B_DONE] != 0.0;
      // serve a refresh request raised after the last refresh was enqueued
      if (!done && Bo[B_REFRESH] != 0.0 && poll[o].end > last_refresh) {
        refresh((int)(K & 1));
        last_refresh = K;
      }
    }
    pending = true;
    slot ^= 1;
  }
  LMH_LAUNCH(ctx, "mx_fold", k_fold, dim3(g), dim3(BLK), 0, n, xc, x_d);
  LMH_HIP(hipStreamSynchronize(ctx->stream));
  LMH_TRY(check_launch("mixed iterations"));
  for (auto &pl : poll)
    if (pl.ev) ctx->event_pool.push_back(pl.ev);
  double h[2 * B_SIZE], gl[8];
  LMH_HIP(hipMemcpy(h, S, sizeof(h), hipMemcpyDeviceToHost));
  const double *Bk = h + (K & 1) * B_SIZE;
  bool conv = Bk[B_DONE] != 0.0;
  if (!conv) {
    // iteration limit: report the FP64 true residual of the current x
    LMH_LAUNCH(ctx, "residual_fp64", k_residual, dim3(gs), dim3(BLK), 0, A->rb, A->nrb, A->rp, A->ci, A->vals, b_d,
               x_d, A->w64, part2);
    LMH_LAUNCH(ctx, "fin", k_fin_residual, dim3(1), dim3(BLK), 0, part2, gs, S + G_RRTRUE);
    LMH_TRY(check_launch("mixed finish"));
  }
  LMH_HIP(hipMemcpy(gl, S + G_THR, sizeof(gl), hipMemcpyDeviceToHost));
  const double bb = gl[G_BB - G_THR], rrt = gl[G_RRTRUE - G_THR];
  const double res = bb > 0 ? std::sqrt(rrt / bb) : 0.0;
  if (iters) *iters = (int64_t)Bk[B_ITS];
  if (rel_res) *rel_res = res;
  if (reason) *reason = conv ? (int)Bk[B_REASON] : LMH_DIVERGED_ITS;
  if (stats_h) {
    stats_h[0] = Bk[B_ITS];
    stats_h[1] = Bk[B_NREF];
    stats_h[2] = res;
  }
  return LMH_OK;
}

int lmh_matrix_set_fp32_compression(lmh_matrix *A, int on) {
  LMH_CHECK(A, LMH_ERR_INVALID, "null matrix");
  A->fp32_index = on ? 1 : 0;
  return LMH_OK;
}

int lmh_matrix_set_preconditioner(lmh_matrix *A, int type) {
  LMH_CHECK(A, LMH_ERR_INVALID, "null matrix");
  LMH_CHECK(type == LMH_PRECOND_JACOBI || type == LMH_PRECOND_IDENTITY, LMH_ERR_UNSUPPORTED,
            "preconditioner %d: only JACOBI_PRECOND and IDENTITY_PRECOND run on the device", type);
  A->precond = type;
  return LMH_OK;
}

int lmh_matrix_fp32_compression(lmh_matrix *A) {
  if (!A) return -1;
  if (ensure_windows(A) != LMH_OK) return -1;
  return use_windows(A) ? 1 : 0;
}

int lmh_matrix_window_bits(lmh_matrix *A) {
  if (!A) return -1;
  if (ensure_windows(A) != LMH_OK) return -1;
  return use_windows(A) ? A->wbits : 0;
}

}  // extern "C"
