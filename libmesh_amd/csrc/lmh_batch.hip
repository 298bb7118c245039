// lmh_batch.hip -- FP32 assembly of many independent linear-elasticity problems
// (uneven sizes, mixed HEX8/HEX27/TET4/TET10) in one hipGraph launch: a
// persistent element kernel walks a tile list that covers every problem, then
// a persistent gather-scatter kernel writes every problem's CSR values and
// right-hand side.  Tile shapes (lanes per element, elements per tile) are
// chosen per element type by timing candidates on the actual batch.
//
// Element operator (systems_of_equations_ex6.C:127-150, 220-236):
//   Ke[(i,a),(k,b)] = sum_q JxW_q sum_jl C_ijkl dphi_a(q)_j dphi_b(q)_l,
//   C_ijkl = l1 d_ij d_kl + l2 (d_ik d_jl + d_il d_jk),
// which with T_ab[i][k] = sum_q JxW_q dphi_a(q)_i dphi_b(q)_k is
//   Ke[(i,a),(k,b)] = l1 T_ab[i][k] + l2 T_ab[k][i] + d_ik l2 tr T_ab,
// and Fe[(i,a)] = sum_q JxW_q f_i phi_a(q) with f = (0, 0, fz).  Element DoFs are
// variable-major as DofMap::dof_indices returns them (all nodes of u, then v,
// then w).  Geometry (J, det, inverse) is FP64; physical gradients are rounded
// to FP32 and the contraction accumulates in FP32.  Fe stays FP64.
#include <algorithm>
#include <cmath>
#include <vector>

#include "lmh_internal.h"

namespace lmh {

struct BProb {
  const double *xyz;
  const int32_t *conn;
  const int32_t *edofs;
  const uint8_t *cons;
  const double *fe;  // w[nq] | phi[n*nq] | dref[n*nq*3]
  float *Ke;         // n_elem * dpe * dpe
  double *Fe;        // n_elem * dpe
  int64_t n_elem;
  int n, nq;
};

struct BTile {
  int32_t p, e0, e1, lpe;
};

struct BScat {
  const int32_t *inc_ptr, *inc, *rp;
  const uint16_t *off;
  const float *Ke;
  const double *Fe;
  float *vals;
  double *rhs;
  int64_t rows;
  int dpe, w;  // element DoFs, lanes per row (16, 32 or 64)
  int stride;  // LDS accumulator row of each sub-group: the problem's longest row, rounded up to 4
  int pad;
};

struct BRTile {
  int32_t p, r0, r1, pad;
};

constexpr int kBatchBlock = 256;
#ifndef LMH_BATCH_SCAT_PER_CU
#define LMH_BATCH_SCAT_PER_CU 32
#endif
constexpr int kBatchLds = 32 * 1024;  // bytes of LDS per element-kernel block
constexpr int kRowsPerTile = 128;

// type slots in the order HEX8, HEX27, TET4, TET10
inline int type_slot(int type) {
  switch (type) {
    case 10: return 0;
    case 12: return 1;
    case 8: return 2;
    case 9: return 3;
  }
  return -1;
}

// LDS bytes of one element slot: float4 gradients [nq][n] (coordinates overlay
// them before the gradients are formed), J^{-1} [nq][9] and JxW [nq] in FP64,
// body-force vector [n] in FP64, constraint flags of the 3n element DoFs.
__host__ __device__ inline int slot_bytes(int n, int nq) {
  int gp = nq * n * 16;
  if (gp < n * 24) gp = n * 24;
  const int s = gp + nq * 80 + n * 8 + ((3 * n + 15) & ~15);
  return (s + 15) & ~15;
}

__global__ void k_batch_clear(int *flag) {
  if (threadIdx.x == 0 && blockIdx.x == 0) flag[0] = 0;
}

__global__ __launch_bounds__(kBatchBlock) void k_batch_elasticity(const BProb *__restrict__ probs,
                                                                  const BTile *__restrict__ tiles, int ntiles,
                                                                  float lam1, float lam2, double fz,
                                                                  int *__restrict__ err) {
  extern __shared__ __align__(16) unsigned char sm[];
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const BTile tl = tiles[t];
    const BProb P = probs[tl.p];
    const int L = tl.lpe, epp = kBatchBlock / L;
    const int sub = threadIdx.x / L, l = threadIdx.x - sub * L;
    const int n = P.n, nq = P.nq, dpe = 3 * n;
    int gpb = nq * n * 16;
    if (gpb < n * 24) gpb = n * 24;
    unsigned char *S = sm + sub * slot_bytes(n, nq);
    float4 *Gp = (float4 *)S;
    double *X = (double *)S;
    double *Gq = (double *)(S + gpb);
    double *Jw = Gq + nq * 9;
    double *Fz = Jw + nq;
    uint8_t *cf = (uint8_t *)(Fz + n);  // constraint flag per element DoF
    const double *w = P.fe, *phi = P.fe + nq, *dref = P.fe + nq + (size_t)n * nq;
    for (int64_t eb = tl.e0; eb < tl.e1; eb += epp) {
      const int64_t e = eb + sub;
      const bool act = e < tl.e1;
      if (act)
        for (int k = l; k < n * 3; k += L) {
          const int a = k / 3, c = k - a * 3;
          X[k] = P.xyz[(int64_t)P.conn[e * n + a] * 3 + c];
          cf[k] = P.cons[P.edofs[e * dpe + k]];
        }
      __syncthreads();
      if (act)
        for (int q = l; q < nq; q += L) {
          double G[3][3];
          const double det = qp_map(X, dref, n, nq, q, G);
          if (!(det > 0.0)) {
            atomicOr(err, 1);
            for (int r = 0; r < 3; ++r)
              for (int c = 0; c < 3; ++c) G[r][c] = 0.0;
          }
          for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) Gq[q * 9 + r * 3 + c] = G[r][c];
          Jw[q] = det * w[q];
        }
      __syncthreads();
      if (act) {
        // dphi_a(q)_c = sum_r dref_a(q)_r G[r][c]  (X is dead from here on)
        for (int k = l; k < nq * n; k += L) {
          const int q = k / n, a = k - q * n;
          const double *d = &dref[(a * nq + q) * 3];
          const double *G = &Gq[q * 9];
          float4 g;
          g.x = (float)(d[0] * G[0] + d[1] * G[3] + d[2] * G[6]);
          g.y = (float)(d[0] * G[1] + d[1] * G[4] + d[2] * G[7]);
          g.z = (float)(d[0] * G[2] + d[1] * G[5] + d[2] * G[8]);
          g.w = 0.f;
          Gp[q * n + a] = g;
        }
        for (int a = l; a < n; a += L) {
          double s = 0.0;
          for (int q = 0; q < nq; ++q) s += Jw[q] * (fz * phi[a * nq + q]);
          Fz[a] = s;
        }
      }
      __syncthreads();
      if (act) {
        float *K = P.Ke + e * (int64_t)dpe * dpe;
        const int npair = n * (n + 1) / 2;
        for (int k = l; k < npair; k += L) {
          int a = 0, rem = k;
          while (rem >= n - a) {
            rem -= n - a;
            ++a;
          }
          const int b = a + rem;
          float T[9];
#pragma unroll
          for (int m = 0; m < 9; ++m) T[m] = 0.f;
          for (int q = 0; q < nq; ++q) {
            const float wq = (float)Jw[q];
            const float4 ga = Gp[q * n + a], gb = Gp[q * n + b];
            const float ax = wq * ga.x, ay = wq * ga.y, az = wq * ga.z;
            T[0] += ax * gb.x;
            T[1] += ax * gb.y;
            T[2] += ax * gb.z;
            T[3] += ay * gb.x;
            T[4] += ay * gb.y;
            T[5] += ay * gb.z;
            T[6] += az * gb.x;
            T[7] += az * gb.y;
            T[8] += az * gb.z;
          }
          const float tr = T[0] + T[4] + T[8];
#pragma unroll
          for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
              float v = lam1 * T[i * 3 + c] + lam2 * T[c * 3 + i];
              if (i == c) v += lam2 * tr;
              const int r = i * n + a, s = c * n + b;
              // element constraint (dof_map_constraints.C:2518-2616): zero row
              // and column of a constrained DoF, unit diagonal
              if (cf[r] | cf[s]) v = (r == s) ? 1.f : 0.f;
              K[(int64_t)r * dpe + s] = v;
              if (a != b) K[(int64_t)s * dpe + r] = v;
            }
        }
        for (int k = l; k < dpe; k += L) {
          const int i = k / n, a = k - i * n;
          double f = (i == 2) ? Fz[a] : 0.0;
          if (cf[k]) f = 0.0;
          P.Fe[e * dpe + k] = f;
        }
      }
      __syncthreads();
    }
  }
}

// Row-gather scatter over all problems (SparseMatrix::add_matrix /
// NumericVector::add_vector for every element): each CSR row is summed by a
// sub-group of W lanes (W >= the element's DoF count up to 64; 64/W rows per
// wave) in an LDS accumulator, contributions added in element order, the row
// written once.  A row's incidences are taken in chunks: the chunk's element
// ids are loaded at once, then all of its (offset, value) pairs (kScR per lane),
// then they are added one incidence after the other from registers.
constexpr int kScR = 8;      // contributions per lane per chunk
constexpr int kScChunk = 64; // incidences per chunk at most

#ifndef LMH_BATCH_SCAT_WAVES
#define LMH_BATCH_SCAT_WAVES 6
#endif
__global__ __launch_bounds__(64, LMH_BATCH_SCAT_WAVES) void k_batch_scatter(const BScat *__restrict__ probs,
                                                      const BRTile *__restrict__ tiles, int ntiles) {
  extern __shared__ __align__(16) unsigned char scsm[];
  const int lane = threadIdx.x;
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const BRTile tl = tiles[t];
    const BScat P = probs[tl.p];
    const int dpe = P.dpe, W = P.w, G = 64 / W, stride = P.stride;
    const int sg = lane / W, l = lane - sg * W;
    // per sub-group: acc[stride] floats | element ids [kScChunk] | Fe [kScChunk]
    unsigned char *base = scsm + (size_t)sg * (stride * 4 + kScChunk * 12);
    float *acc = (float *)base;
    double *febuf = (double *)(base + stride * 4);
    int32_t *incbuf = (int32_t *)(base + stride * 4 + kScChunk * 8);
    const int ch = min(min((kScR * W) / dpe, W), kScChunk);  // incidences per chunk
    for (int32_t rb = tl.r0; rb < tl.r1; rb += G) {
      const int32_t r = rb + sg;
      const bool act = r < tl.r1;
      int32_t b = 0, len = 0, k0 = 0, cnt = 0;
      if (act) {
        b = P.rp[r];
        len = P.rp[r + 1] - b;
        k0 = P.inc_ptr[r];
        cnt = P.inc_ptr[r + 1] - k0;
      }
      for (int k = l; k < len; k += W) acc[k] = 0.f;
      int mx = cnt;
      for (int o = W; o < 64; o <<= 1) mx = max(mx, __shfl_xor(mx, o, 64));
      double f = 0.0;
      for (int c0 = 0; c0 < mx; c0 += ch) {
        const int nch = min(ch, cnt - c0);  // incidences of this chunk (<= 0: none)
        __builtin_amdgcn_wave_barrier();
        if (l < nch) {
          const int32_t ei = P.inc[k0 + c0 + l];
          incbuf[l] = ei;
          febuf[l] = P.Fe[ei];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        uint16_t oq[kScR];
        float vq[kScR];
#pragma unroll
        for (int q = 0; q < kScR; ++q) {
          const int tq = q * W + l;
          const int kk = tq / dpe;
          oq[q] = 0;
          vq[q] = 0.f;
          if (kk < nch) {
            const int64_t at = (int64_t)incbuf[kk] * dpe + (tq - kk * dpe);
            oq[q] = P.off[at];
            vq[q] = P.Ke[at];
          }
        }
#pragma unroll
        for (int q = 0; q < kScR; ++q) {
          const int tq = q * W + l;
          const int kk = tq / dpe;
          const int lo = (q * W) / dpe, hi = (q * W + W - 1) / dpe;
          for (int k = lo; k <= hi; ++k) {
            if (k == kk && kk < nch) acc[oq[q]] += vq[q];
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
          }
        }
        if (l == 0)
          for (int k = 0; k < nch; ++k) f += febuf[k];
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      for (int k = l; k < len; k += W) P.vals[b + k] = acc[k];
      if (act && l == 0) P.rhs[r] = f;
      __builtin_amdgcn_wave_barrier();
    }
  }
}

}  // namespace lmh

using namespace lmh;

struct lmh_batch {
  lmh_context *ctx = nullptr;
  int np = 0;
  std::vector<const lmh_dofmap *> dms;
  std::vector<lmh_matrix *> mats;
  std::vector<double *> rhs;
  std::vector<BProb> hprob;
  std::vector<int> slot;  // type slot per problem
  BProb *dprob = nullptr;
  BScat *dscat = nullptr;
  BRTile *drt = nullptr;
  int nrt = 0;
  BTile *dtile = nullptr;
  int ntile = 0, tile_cap = 0;
  int lpe[4] = {16, 128, 16, 32};   // lanes per element
  int ppt[4] = {4, 4, 4, 4};        // passes (groups of concurrent elements) per tile
  int max_row = 0;
  size_t scat_lds = 0;
  int grid_elem = 0, grid_scat = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  double gE = 0, gnu = 0, gfz = 0;
};

namespace lmh {

// LDS of the scatter: 64/W sub-groups, each an accumulator row (stride floats,
// a multiple of 4) plus a chunk's element ids and Fe values; the launch takes
// the largest need over the problems
static int scatter_stride(int max_row) { return (max_row + 3) & ~3; }
static size_t scatter_lds(const BScat &s) { return (size_t)(64 / s.w) * ((size_t)s.stride * 4 + kScChunk * 12); }

static bool config_fits(int n, int nq, int L) {
  return L >= 16 && L <= kBatchBlock && (kBatchBlock / L) * slot_bytes(n, nq) <= kBatchLds;
}

// tiles of the problems whose type slot is in `mask`, heavy element types first
static void make_tiles(const lmh_batch *b, unsigned mask, const int *lpe, const int *ppt, std::vector<BTile> &out) {
  out.clear();
  const int order[4] = {1, 3, 0, 2};  // HEX27, TET10, HEX8, TET4
  for (int oi = 0; oi < 4; ++oi) {
    const int s = order[oi];
    if (!(mask >> s & 1u)) continue;
    const int ept = (kBatchBlock / lpe[s]) * ppt[s];
    for (int p = 0; p < b->np; ++p) {
      if (b->slot[p] != s) continue;
      const int64_t ne = b->hprob[p].n_elem;
      for (int64_t e = 0; e < ne; e += ept)
        out.push_back(BTile{p, (int32_t)e, (int32_t)std::min<int64_t>(ne, e + ept), lpe[s]});
    }
  }
}

static int upload_tiles(lmh_batch *b, const std::vector<BTile> &t) {
  if ((int)t.size() > b->tile_cap) {
    if (b->dtile) LMH_HIP(hipFree(b->dtile));
    b->tile_cap = (int)t.size();
    LMH_HIP(hipMalloc((void **)&b->dtile, sizeof(BTile) * std::max(1, b->tile_cap)));
  }
  if (!t.empty())
    LMH_HIP(hipMemcpyAsync(b->dtile, t.data(), sizeof(BTile) * t.size(), hipMemcpyHostToDevice, b->ctx->stream));
  b->ntile = (int)t.size();
  return LMH_OK;
}

static void lame(double E, double nu, float &l1, float &l2) {
  l1 = (float)((E * nu) / ((1. + nu) * (1. - 2. * nu)));
  l2 = (float)(E / (2. * (1. + nu)));
}

static int build_graph(lmh_batch *b, double E, double nu, double fz) {
  lmh_context *ctx = b->ctx;
  if (b->exec) LMH_HIP(hipGraphExecDestroy(b->exec));
  if (b->graph) LMH_HIP(hipGraphDestroy(b->graph));
  b->exec = nullptr;
  b->graph = nullptr;
  float l1, l2;
  lame(E, nu, l1, l2);
  LMH_HIP(hipStreamSynchronize(ctx->stream));
  LMH_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  hipLaunchKernelGGL(k_batch_clear, dim3(1), dim3(64), 0, ctx->stream, ctx->d_flags);
  if (b->ntile > 0)
    hipLaunchKernelGGL(k_batch_elasticity, dim3(b->grid_elem), dim3(kBatchBlock), kBatchLds, ctx->stream,
                       (const BProb *)b->dprob, (const BTile *)b->dtile, b->ntile, l1, l2, fz, ctx->d_flags);
  if (b->nrt > 0)
    hipLaunchKernelGGL(k_batch_scatter, dim3(b->grid_scat), dim3(64), b->scat_lds, ctx->stream,
                       (const BScat *)b->dscat, (const BRTile *)b->drt, b->nrt);
  const hipError_t ce = hipStreamEndCapture(ctx->stream, &b->graph);
  LMH_HIP(ce);
  LMH_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  b->gE = E;
  b->gnu = nu;
  b->gfz = fz;
  return LMH_OK;
}

}  // namespace lmh

extern "C" {

int lmh_batch_destroy(lmh_batch *b);

static int batch_setup(lmh_batch *b, lmh_context *ctx, int n_prob, const lmh_dofmap *const *dms) {
  LMH_HIP(hipSetDevice(ctx->device));
  for (int p = 0; p < n_prob; ++p) {
    const lmh_dofmap *dm = dms[p];
    LMH_CHECK(dm && dm->ctx == ctx, LMH_ERR_INVALID, "problem %d: null DoF map or other context", p);
    LMH_CHECK(dm->nv == 3, LMH_ERR_INVALID, "problem %d: elasticity needs 3 variables per node", p);
    LMH_CHECK(dm->cr.n_general == 0, LMH_ERR_UNSUPPORTED, "problem %d: general constraint rows", p);
    const int s = type_slot(dm->mesh->type);
    LMH_CHECK(s >= 0, LMH_ERR_UNSUPPORTED, "problem %d: element type %d", p, dm->mesh->type);
    LMH_CHECK(dm->mesh->n_elem * (int64_t)dm->dpe < INT32_MAX, LMH_ERR_UNSUPPORTED, "problem %d too large", p);
    // QGauss at FEType::default_quadrature_order (2p+1): HEX8 2^3, HEX27 3^3, TET4 order 3, TET10 order 5
    const int arg = (s == 0) ? 2 : (s == 1) ? 3 : (s == 2) ? 3 : 5;
    const FETable *t = nullptr;
    LMH_TRY(fe_table(ctx, dm->mesh->type, arg, &t));
    lmh_matrix *A = nullptr;
    LMH_TRY(matrix_create_impl(ctx, dm, false, false, &A));
    b->mats.push_back(A);
    b->dms.push_back(dm);
    b->slot.push_back(s);
    b->max_row = std::max(b->max_row, A->max_row);
    const int64_t ne = dm->mesh->n_elem;
    const int dpe = dm->dpe;
    LMH_HIP(hipMalloc((void **)&A->Ke32, sizeof(float) * std::max<int64_t>(1, ne * dpe * dpe)));
    LMH_HIP(hipMalloc((void **)&A->Fe, sizeof(double) * std::max<int64_t>(1, ne * dpe)));
    LMH_HIP(hipMalloc((void **)&A->vals32, sizeof(float) * (A->nnz + LMH_PAD)));
    LMH_HIP(hipMemsetAsync(A->vals32, 0, sizeof(float) * (A->nnz + LMH_PAD), ctx->stream));
    double *r = nullptr;
    LMH_HIP(hipMalloc((void **)&r, sizeof(double) * std::max<int64_t>(1, dm->n_dofs)));
    LMH_HIP(hipMemsetAsync(r, 0, sizeof(double) * std::max<int64_t>(1, dm->n_dofs), ctx->stream));
    b->rhs.push_back(r);
    BProb bp;
    bp.xyz = dm->mesh->xyz;
    bp.conn = dm->mesh->conn;
    bp.edofs = dm->edofs;
    bp.cons = dm->cons;
    bp.fe = t->d_data;
    bp.Ke = A->Ke32;
    bp.Fe = A->Fe;
    bp.n_elem = ne;
    bp.n = t->n;
    bp.nq = t->nq;
    LMH_CHECK(config_fits(bp.n, bp.nq, kBatchBlock), LMH_ERR_UNSUPPORTED, "element type %d does not fit in LDS",
              dm->mesh->type);
    b->hprob.push_back(bp);
  }
  // default tile shapes: the smallest lanes-per-element that fits the LDS budget
  for (int p = 0; p < n_prob; ++p) {
    const int s = b->slot[p];
    int L = b->lpe[s];
    while (!config_fits(b->hprob[p].n, b->hprob[p].nq, L)) L *= 2;
    b->lpe[s] = L;
  }
  LMH_HIP(hipMalloc((void **)&b->dprob, sizeof(BProb) * n_prob));
  LMH_HIP(hipMemcpy(b->dprob, b->hprob.data(), sizeof(BProb) * n_prob, hipMemcpyHostToDevice));
  std::vector<BScat> sc(n_prob);
  std::vector<BRTile> rt;
  for (int p = 0; p < n_prob; ++p) {
    lmh_matrix *A = b->mats[p];
    const int dpe = b->dms[p]->dpe;
    sc[p] = BScat{b->dms[p]->inc_ptr, b->dms[p]->inc, A->rp, A->off, A->Ke32, A->Fe, A->vals32, b->rhs[p],
                  A->rows, dpe, dpe <= 16 ? 16 : dpe <= 32 ? 32 : 64, scatter_stride(A->max_row), 0};
    b->scat_lds = std::max(b->scat_lds, scatter_lds(sc[p]));
    for (int64_t r = 0; r < A->rows; r += kRowsPerTile)
      rt.push_back(BRTile{p, (int32_t)r, (int32_t)std::min<int64_t>(A->rows, r + kRowsPerTile), 0});
  }
  // interleave row tiles of different problems so each part of the grid sees a mix
  std::vector<BRTile> rt2;
  rt2.reserve(rt.size());
  {
    std::vector<size_t> start(n_prob + 1, 0);
    for (size_t k = 0; k < rt.size(); ++k) start[rt[k].p + 1]++;
    for (int p = 0; p < n_prob; ++p) start[p + 1] += start[p];
    std::vector<size_t> pos(start.begin(), start.end() - 1);
    bool any = true;
    while (any) {
      any = false;
      for (int p = 0; p < n_prob; ++p)
        if (pos[p] < start[p + 1]) {
          rt2.push_back(rt[pos[p]++]);
          any = true;
        }
    }
  }
  b->nrt = (int)rt2.size();
  LMH_HIP(hipMalloc((void **)&b->dscat, sizeof(BScat) * n_prob));
  LMH_HIP(hipMemcpy(b->dscat, sc.data(), sizeof(BScat) * n_prob, hipMemcpyHostToDevice));
  LMH_HIP(hipMalloc((void **)&b->drt, sizeof(BRTile) * std::max(1, b->nrt)));
  if (b->nrt) LMH_HIP(hipMemcpy(b->drt, rt2.data(), sizeof(BRTile) * b->nrt, hipMemcpyHostToDevice));
  int occ = 0;
  LMH_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_batch_elasticity, kBatchBlock, kBatchLds));
  b->grid_elem = ctx->cu_count * std::max(1, occ);
  // one-wave workgroups, as many per CU as can be resident at once (registers
  // and LDS): the tiles are dealt round-robin, so a second round of
  // workgroups would only wait (measured: 24 resident per CU 71 ms, 32 launched 84 ms)
  {
    int socc = 0;
    LMH_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&socc, k_batch_scatter, 64, b->scat_lds));
    b->grid_scat = ctx->cu_count * std::max(1, std::min(socc, (int)LMH_BATCH_SCAT_PER_CU));
  }
  std::vector<BTile> tiles;
  make_tiles(b, 0xFu, b->lpe, b->ppt, tiles);
  LMH_TRY(upload_tiles(b, tiles));
  LMH_HIP(hipStreamSynchronize(ctx->stream));
  return LMH_OK;
}

int lmh_batch_create(lmh_context *ctx, int n_prob, const lmh_dofmap *const *dms, lmh_batch **out) {
  LMH_CHECK(ctx && dms && out && n_prob > 0, LMH_ERR_INVALID, "bad argument");
  lmh_batch *b = new lmh_batch;
  b->ctx = ctx;
  b->np = n_prob;
  const int rc = batch_setup(b, ctx, n_prob, dms);
  if (rc != LMH_OK) {
    b->np = (int)b->mats.size();
    lmh_batch_destroy(b);
    return rc;
  }
  *out = b;
  return LMH_OK;
}

int lmh_batch_destroy(lmh_batch *b) {
  if (!b) return LMH_OK;
  (void)hipStreamSynchronize(b->ctx->stream);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  for (lmh_matrix *A : b->mats) lmh_matrix_destroy(A);
  for (double *r : b->rhs) (void)hipFree(r);
  (void)hipFree(b->dprob);
  (void)hipFree(b->dscat);
  (void)hipFree(b->drt);
  (void)hipFree(b->dtile);
  delete b;
  return LMH_OK;
}

int lmh_batch_size(const lmh_batch *b) { return b ? b->np : -1; }

int lmh_batch_matrix(const lmh_batch *b, int p, lmh_matrix **A) {
  LMH_CHECK(b && A && p >= 0 && p < b->np, LMH_ERR_INVALID, "bad argument");
  *A = b->mats[p];
  return LMH_OK;
}

int lmh_batch_rhs(const lmh_batch *b, int p, double **rhs_d) {
  LMH_CHECK(b && rhs_d && p >= 0 && p < b->np, LMH_ERR_INVALID, "bad argument");
  *rhs_d = b->rhs[p];
  return LMH_OK;
}

int lmh_batch_set_tiles(lmh_batch *b, const int *lpe, const int *passes) {
  LMH_CHECK(b && lpe && passes, LMH_ERR_INVALID, "null argument");
  for (int s = 0; s < 4; ++s) {
    LMH_CHECK(lpe[s] == 16 || lpe[s] == 32 || lpe[s] == 64 || lpe[s] == 128 || lpe[s] == 256, LMH_ERR_INVALID,
              "lanes per element must be 16, 32, 64, 128 or 256");
    LMH_CHECK(passes[s] >= 1 && passes[s] <= 1024, LMH_ERR_INVALID, "passes per tile out of range");
  }
  for (int p = 0; p < b->np; ++p) {
    const int s = b->slot[p];
    LMH_CHECK(config_fits(b->hprob[p].n, b->hprob[p].nq, lpe[s]), LMH_ERR_INVALID,
              "%d lanes per element exceed the LDS budget for element type %d", lpe[s], b->dms[p]->mesh->type);
  }
  for (int s = 0; s < 4; ++s) {
    b->lpe[s] = lpe[s];
    b->ppt[s] = passes[s];
  }
  std::vector<BTile> tiles;
  make_tiles(b, 0xFu, b->lpe, b->ppt, tiles);
  LMH_TRY(upload_tiles(b, tiles));
  if (b->exec) LMH_TRY(build_graph(b, b->gE, b->gnu, b->gfz));
  LMH_HIP(hipStreamSynchronize(b->ctx->stream));
  return LMH_OK;
}

int lmh_batch_get_tiles(const lmh_batch *b, int *lpe, int *passes) {
  LMH_CHECK(b && lpe && passes, LMH_ERR_INVALID, "null argument");
  for (int s = 0; s < 4; ++s) {
    lpe[s] = b->lpe[s];
    passes[s] = b->ppt[s];
  }
  return LMH_OK;
}

int lmh_batch_autotune(lmh_batch *b, double E, double nu, double fz, int reps, double *best_ms) {
  LMH_CHECK(b && reps >= 1, LMH_ERR_INVALID, "bad argument");
  lmh_context *ctx = b->ctx;
  float l1, l2;
  lame(E, nu, l1, l2);
  const int cand_l[5] = {16, 32, 64, 128, 256};
  const int cand_p[3] = {1, 4, 16};
  hipEvent_t e0, e1;
  LMH_HIP(hipEventCreate(&e0));
  LMH_HIP(hipEventCreate(&e1));
  int lpe[4], ppt[4];
  for (int s = 0; s < 4; ++s) {
    lpe[s] = b->lpe[s];
    ppt[s] = b->ppt[s];
  }
  for (int s = 0; s < 4; ++s) {
    if (best_ms) best_ms[s] = -1.0;
    int pn = -1, pnq = -1;
    for (int p = 0; p < b->np; ++p)
      if (b->slot[p] == s) {
        pn = b->hprob[p].n;
        pnq = b->hprob[p].nq;
        break;
      }
    if (pn < 0) continue;
    double best = 1e300;
    for (int li = 0; li < 5; ++li) {
      if (!config_fits(pn, pnq, cand_l[li])) continue;
      for (int pi = 0; pi < 3; ++pi) {
        int tl[4], tp[4];
        for (int k = 0; k < 4; ++k) {
          tl[k] = lpe[k];
          tp[k] = ppt[k];
        }
        tl[s] = cand_l[li];
        tp[s] = cand_p[pi];
        std::vector<BTile> tiles;
        make_tiles(b, 1u << s, tl, tp, tiles);
        LMH_TRY(upload_tiles(b, tiles));
        float ms_min = 1e30f;
        for (int r = 0; r <= reps; ++r) {
          LMH_HIP(hipEventRecord(e0, ctx->stream));
          hipLaunchKernelGGL(k_batch_elasticity, dim3(b->grid_elem), dim3(kBatchBlock), kBatchLds, ctx->stream,
                             (const BProb *)b->dprob, (const BTile *)b->dtile, b->ntile, l1, l2, fz, ctx->d_flags);
          LMH_HIP(hipEventRecord(e1, ctx->stream));
          LMH_HIP(hipEventSynchronize(e1));
          float ms = 0;
          LMH_HIP(hipEventElapsedTime(&ms, e0, e1));
          if (r > 0) ms_min = std::min(ms_min, ms);  // first run warms caches
        }
        if (ms_min < best) {
          best = ms_min;
          lpe[s] = cand_l[li];
          ppt[s] = cand_p[pi];
        }
      }
    }
    if (best_ms) best_ms[s] = best;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  LMH_TRY(check_launch("k_batch_elasticity"));
  for (int s = 0; s < 4; ++s) {
    b->lpe[s] = lpe[s];
    b->ppt[s] = ppt[s];
  }
  std::vector<BTile> tiles;
  make_tiles(b, 0xFu, b->lpe, b->ppt, tiles);
  LMH_TRY(upload_tiles(b, tiles));
  LMH_TRY(build_graph(b, E, nu, fz));
  LMH_HIP(hipStreamSynchronize(ctx->stream));
  return LMH_OK;
}

int lmh_batch_assemble_elasticity(lmh_batch *b, double E, double nu, double fz) {
  LMH_CHECK(b, LMH_ERR_INVALID, "null batch");
  LMH_CHECK(E > 0 && nu > -1.0 && nu < 0.5, LMH_ERR_INVALID, "need E > 0 and -1 < nu < 1/2");
  lmh_context *ctx = b->ctx;
  if (!b->exec || E != b->gE || nu != b->gnu || fz != b->gfz) LMH_TRY(build_graph(b, E, nu, fz));
  {
    ProfScope ps(ctx, "batch_graph");
    LMH_HIP(hipGraphLaunch(b->exec, ctx->stream));
  }
  int bad = 0;
  LMH_HIP(hipMemcpyAsync(&bad, ctx->d_flags, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  LMH_HIP(hipStreamSynchronize(ctx->stream));
  LMH_CHECK(!bad, LMH_ERR_DEGENERATE, "degenerate element map (det J <= 0) in the batch");
  for (lmh_matrix *A : b->mats) A->have32 = true;
  return LMH_OK;
}

}  // extern "C"
