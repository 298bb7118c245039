// Static condensation of element-interior DoFs on the device.
//
// libMesh's StaticCondensation (include/numerics/static_condensation.h:64,
// src/numerics/static_condensation.C) condenses the DoFs on an element's
// internal nodes (static_condensation_dof_map.C:187-199: elem.is_internal);
// for each element it factors Acc, forms Auu - Auc Acc^-1 Acu
// (static_condensation.C:205-211) and sums these into the reduced matrix,
// condenses the right-hand side (forward_elimination, :383-417) and recovers
// the condensed values afterwards (backwards_substitution, :419-456).
//
// Here the same operator is formed algebraically from the assembled CSR
// matrix.  A condensed DoF c lives on one element only, so every assembled
// entry K_ic, K_cj equals that element's entry and the assembled condensed
// block is diagonal when each element condenses one DoF per coupled block
// (a scalar variable on HEX27).  Summing the element Schur complements is then
//     S_ij = K_ij - sum_{c in row i, condensed} K_ic K_cj / K_cc,
// computed row by row (one thread owns one reduced row: no atomics, the
// result does not depend on scheduling).  The reduced pattern is A's pattern
// without the condensed rows and columns, numbered in the full order.
#include <algorithm>
#include <memory>
#include <vector>

#include "lmh_internal.h"

struct lmh_condensation {
  lmh_context *ctx = nullptr;  // A's context: destroy needs no access to A
  lmh_matrix *A = nullptr;     // full operator (not owned)
  lmh_matrix *R = nullptr;  // reduced operator (owned)
  int64_t n_full = 0, n_cond = 0, n_red = 0;
  int32_t *ridx = nullptr;   // n_full: reduced index, or -1 - q for the q-th condensed DoF
  int32_t *rfull = nullptr;  // n_red: full index of each reduced DoF
  int32_t *cdof = nullptr;   // n_cond: condensed DoFs, ascending
  int32_t *dpos = nullptr;   // n_cond: position of K_cc in A's row c
  double *g = nullptr, *y = nullptr;  // reduced right-hand side and solution
  bool pattern_checked = false;       // k_sc_values found no fill-in (then k_sc_values_wave is used)
  // Position map (pattern only, built after the fill check): for reduced row r
  // and each structurally condensed column c of its full row i (ascending),
  // the full-row-i position of every entry of row c, -1 for condensed or
  // absent columns.  pmoff[r] .. pmoff[r + 1] are row r's entries.
  int64_t *pmoff = nullptr;
  int16_t *pmap = nullptr;
};

namespace lmh {

enum { SC_OFFDIAG = 1, SC_NODIAG = 2, SC_ZEROPIVOT = 4, SC_FILL = 8 };

// Per condensed DoF: the diagonal position, and a check that row c couples to
// no other condensed DoF (the condensed block must be diagonal).
__global__ __launch_bounds__(256) void k_sc_check(int64_t n_cond, const int32_t *__restrict__ cdof,
                                                  const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                                  const double *__restrict__ v, const int32_t *__restrict__ ridx,
                                                  int32_t *__restrict__ dpos, int *__restrict__ flags) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n_cond; q += (int64_t)gridDim.x * blockDim.x) {
    const int32_t c = cdof[q];
    int32_t d = -1;
    int f = 0;
    for (int32_t k = rp[c]; k < rp[c + 1]; ++k) {
      const int32_t j = ci[k];
      if (j == c)
        d = k;
      else if (ridx[j] < 0 && v[k] != 0.0)
        f |= SC_OFFDIAG;
    }
    if (d < 0)
      f |= SC_NODIAG;
    else if (v[d] == 0.0)
      f |= SC_ZEROPIVOT;
    dpos[q] = d < 0 ? rp[c] : d;
    if (f) atomicOr(flags, f);
  }
}

// Reduced values: copy row i of A without its condensed columns, then for
// each condensed column c of row i subtract (K_ic / K_cc) * row c, merged
// into the (sorted) reduced row.  Runs once per condensation: it also checks
// that every uncondensed column j of every condensed row c of row i lies in
// reduced row i, structurally (whatever the values), so that later refreshes
// (k_sc_values_wave, which has no such check) can never drop an entry that
// becomes nonzero after reassembly.
__global__ __launch_bounds__(256) void k_sc_values(int64_t n_red, const int32_t *__restrict__ rfull,
                                                   const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                                   const double *__restrict__ v, const int32_t *__restrict__ ridx,
                                                   const int32_t *__restrict__ dpos, const int32_t *__restrict__ srp,
                                                   const int32_t *__restrict__ sci, double *__restrict__ sv,
                                                   int *__restrict__ flags) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n_red; r += (int64_t)gridDim.x * blockDim.x) {
    const int32_t i = rfull[r];
    const int32_t s0 = srp[r], s1 = srp[r + 1];
    int32_t p = s0;
    for (int32_t k = rp[i]; k < rp[i + 1]; ++k)
      if (ridx[ci[k]] >= 0) sv[p++] = v[k];
    for (int32_t k = rp[i]; k < rp[i + 1]; ++k) {
      const int32_t c = ci[k], qc = ridx[c];
      if (qc >= 0) continue;
      const double f = v[k] / v[dpos[-1 - qc]];
      const bool sub = v[k] != 0.0;
      p = s0;
      for (int32_t m = rp[c]; m < rp[c + 1]; ++m) {
        const int32_t rj = ridx[ci[m]];
        if (rj < 0) continue;
        while (p < s1 && sci[p] < rj) ++p;
        if (p == s1 || sci[p] != rj) {
          atomicOr(flags, SC_FILL);
          continue;
        }
        if (sub) sv[p] -= f * v[m];
      }
    }
  }
}

// The same values with one wavefront per reduced row (used once the pattern
// has passed k_sc_values' fill check): lane t owns entry (i, j) of the row
// and subtracts f_c K_cj for the condensed columns c of row i in ascending
// order (wave-uniform loop), K_cj found by binary search in row c.  The
// kept entries are compacted to the reduced row with a ballot prefix count.
__global__ __launch_bounds__(256) void k_sc_values_wave(int64_t n_red, const int32_t *__restrict__ rfull,
                                                        const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                                        const double *__restrict__ v, const int32_t *__restrict__ ridx,
                                                        const int32_t *__restrict__ dpos,
                                                        const int32_t *__restrict__ srp, double *__restrict__ sv) {
  const int lane = threadIdx.x & 63;
  const uint64_t below = (1ull << lane) - 1;
  const int64_t nw = (int64_t)gridDim.x * (blockDim.x >> 6);
  for (int64_t r = blockIdx.x * (int64_t)(blockDim.x >> 6) + (threadIdx.x >> 6); r < n_red; r += nw) {
    const int32_t i = rfull[r], k0 = rp[i], k1 = rp[i + 1];
    int32_t out = srp[r];
    for (int32_t kb = k0; kb < k1; kb += 64) {
      const int32_t k = kb + lane;
      const int32_t j = k < k1 ? ci[k] : 0;
      const bool keep = k < k1 && ridx[j] >= 0;
      double acc = keep ? v[k] : 0.0;
      for (int32_t cb = k0; cb < k1; cb += 64) {
        const int32_t kc = cb + lane;
        const int32_t cc = kc < k1 ? ci[kc] : 0;
        const int32_t qc = kc < k1 ? ridx[cc] : 0;
        uint64_t m = __ballot(kc < k1 && qc < 0 && v[kc] != 0.0);
        while (m) {
          const int b = __builtin_ctzll(m);
          m &= m - 1;
          const int32_t c = __shfl(cc, b), q = -1 - __shfl(qc, b);
          const double f = v[cb + b] / v[dpos[q]];
          if (keep) {
            int32_t lo = rp[c], hi = rp[c + 1];
            while (lo < hi) {
              const int32_t mid = (lo + hi) >> 1;
              if (ci[mid] < j)
                lo = mid + 1;
              else
                hi = mid;
            }
            if (lo < rp[c + 1] && ci[lo] == j) acc -= f * v[lo];
          }
        }
      }
      const uint64_t km = __ballot(keep);
      if (keep) sv[out + __popcll(km & below)] = acc;
      out += __popcll(km);
    }
  }
}

// The same values with reduced row i held in LDS: its columns and running
// values are staged once, then for each condensed column c of row i
// (ascending, wave-uniform) the lanes walk row c (the condensed DoF's
// element, 27 entries for Q2) and find each column in the staged row by a
// binary search in LDS instead of in global memory.  Every entry receives the
// same subtractions in the same order as in k_sc_values_wave, so the values
// are bit-identical.  LDS: max_row int32 + max_row doubles per wavefront.
__device__ inline void sc_wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ __launch_bounds__(256) void k_sc_values_lds(int64_t n_red, const int32_t *__restrict__ rfull,
                                                       const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                                       const double *__restrict__ v, const int32_t *__restrict__ ridx,
                                                       const int32_t *__restrict__ dpos,
                                                       const int32_t *__restrict__ srp, double *__restrict__ sv,
                                                       int max_row) {
  extern __shared__ double sc_sm[];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double *acc = sc_sm + (size_t)w * max_row;
  int32_t *cols = (int32_t *)(sc_sm + (size_t)(blockDim.x >> 6) * max_row) + (size_t)w * max_row;
  const uint64_t below = (1ull << lane) - 1;
  const int64_t nw = (int64_t)gridDim.x * (blockDim.x >> 6);
  for (int64_t r = blockIdx.x * (int64_t)(blockDim.x >> 6) + w; r < n_red; r += nw) {
    const int32_t i = rfull[r], k0 = rp[i], k1 = rp[i + 1], len = k1 - k0;
    for (int32_t t = lane; t < len; t += 64) {
      cols[t] = ci[k0 + t];
      acc[t] = v[k0 + t];
    }
    sc_wave_lds_sync();
    for (int32_t cb = k0; cb < k1; cb += 64) {
      const int32_t kc = cb + lane;
      const int32_t cc = kc < k1 ? ci[kc] : 0;
      const int32_t qc = kc < k1 ? ridx[cc] : 0;
      uint64_t m = __ballot(kc < k1 && qc < 0 && v[kc] != 0.0);
      while (m) {
        const int b = __builtin_ctzll(m);
        m &= m - 1;
        const int32_t c = __shfl(cc, b), q = -1 - __shfl(qc, b);
        const double f = v[cb + b] / v[dpos[q]];
        for (int32_t mm = rp[c] + lane; mm < rp[c + 1]; mm += 64) {
          const int32_t j = ci[mm];
          if (ridx[j] < 0) continue;
          int32_t lo = 0, hi = len;
          while (lo < hi) {
            const int32_t mid = (lo + hi) >> 1;
            if (cols[mid] < j)
              lo = mid + 1;
            else
              hi = mid;
          }
          if (lo < len && cols[lo] == j) acc[lo] -= f * v[mm];
        }
        sc_wave_lds_sync();  // this condensed column's updates land before the next one's
      }
    }
    int32_t out = srp[r];
    for (int32_t t0 = 0; t0 < len; t0 += 64) {
      const int32_t t = t0 + lane;
      const bool keep = t < len && ridx[cols[t]] >= 0;
      const uint64_t km = __ballot(keep);
      if (keep) sv[out + __popcll(km & below)] = acc[t];
      out += __popcll(km);
    }
    sc_wave_lds_sync();  // the staged row is read before the next row overwrites it
  }
}

// g_r = b_i - sum_{c in row i, condensed} K_ic b_c / K_cc; one wavefront per
// reduced row, lanes over the row's entries (coalesced), then a fixed-order
// shuffle reduction.
__global__ __launch_bounds__(256) void k_sc_forward(int64_t n_red, const int32_t *__restrict__ rfull,
                                                    const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                                    const double *__restrict__ v, const int32_t *__restrict__ ridx,
                                                    const int32_t *__restrict__ dpos, const double *__restrict__ b,
                                                    double *__restrict__ g) {
  const int lane = threadIdx.x & 63;
  const int64_t nw = (int64_t)gridDim.x * (blockDim.x >> 6);
  for (int64_t r = blockIdx.x * (int64_t)(blockDim.x >> 6) + (threadIdx.x >> 6); r < n_red; r += nw) {
    const int32_t i = rfull[r], k1 = rp[i + 1];
    double s = 0.0;
    for (int32_t k = rp[i] + lane; k < k1; k += 64) {
      const int32_t c = ci[k], qc = ridx[c];
      if (qc < 0) s += v[k] * (b[c] / v[dpos[-1 - qc]]);
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off);
    if (lane == 0) g[r] = b[i] - s;
  }
}

// y_r = x_{rfull[r]} (TO_RED) or x_{rfull[r]} = y_r
template <bool TO_RED>
__global__ __launch_bounds__(256) void k_sc_map(int64_t n_red, const int32_t *__restrict__ rfull, double *__restrict__ x,
                                                double *__restrict__ y) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n_red; r += (int64_t)gridDim.x * blockDim.x) {
    if (TO_RED)
      y[r] = x[rfull[r]];
    else
      x[rfull[r]] = y[r];
  }
}

// x_c = (b_c - sum_{j uncondensed} K_cj x_j) / K_cc, after the uncondensed x_j
// are in place.  Columns of other condensed DoFs are skipped: k_sc_check
// admits them only as exact zeros, and those x_j are being written by other
// threads of this launch (their old contents may even be NaN).
__global__ __launch_bounds__(256) void k_sc_backward(int64_t n_cond, const int32_t *__restrict__ cdof,
                                                     const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                                     const double *__restrict__ v, const int32_t *__restrict__ ridx,
                                                     const int32_t *__restrict__ dpos, const double *__restrict__ b,
                                                     double *__restrict__ x) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n_cond; q += (int64_t)gridDim.x * blockDim.x) {
    const int32_t c = cdof[q], d = dpos[q];
    double s = b[c];
    for (int32_t k = rp[c]; k < rp[c + 1]; ++k) {
      const int32_t j = ci[k];
      if (ridx[j] >= 0) s -= v[k] * x[j];
    }
    x[c] = s / v[d];
  }
}

// Inclusive prefix sum over the wavefront's lanes.
__device__ inline int32_t sc_wave_scan(int32_t x) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int32_t y = __shfl_up(x, d);
    if (lane >= d) x += y;
  }
  return x;
}

// Position-map sizes: per reduced row, the entries of all rows c of its
// structurally condensed columns.  One wavefront per reduced row.
__global__ __launch_bounds__(256) void k_sc_map_count(int64_t n_red, const int32_t *__restrict__ rfull,
                                                      const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                                      const int32_t *__restrict__ ridx, int64_t *__restrict__ cnt) {
  const int lane = threadIdx.x & 63;
  const int64_t nw = (int64_t)gridDim.x * (blockDim.x >> 6);
  for (int64_t r = blockIdx.x * (int64_t)(blockDim.x >> 6) + (threadIdx.x >> 6); r < n_red; r += nw) {
    const int32_t i = rfull[r], k1 = rp[i + 1];
    int64_t tot = 0;
    for (int32_t cb = rp[i]; cb < k1; cb += 64) {
      const int32_t kc = cb + lane;
      const int32_t cc = kc < k1 ? ci[kc] : 0;
      const int32_t lenc = kc < k1 && ridx[cc] < 0 ? rp[cc + 1] - rp[cc] : 0;
      tot += __shfl(sc_wave_scan(lenc), 63);
    }
    if (lane == 0) cnt[r] = tot;
  }
}

// Position-map entries (same walk as k_sc_values_lds, positions instead of
// updates).  LDS: max_row int32 per wavefront.
__global__ __launch_bounds__(256) void k_sc_map_fill(int64_t n_red, const int32_t *__restrict__ rfull,
                                                     const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                                     const int32_t *__restrict__ ridx, const int64_t *__restrict__ pmoff,
                                                     int16_t *__restrict__ pmap, int max_row) {
  extern __shared__ int32_t sc_cols[];
  const int lane = threadIdx.x & 63;
  int32_t *cols = sc_cols + (size_t)(threadIdx.x >> 6) * max_row;
  const int64_t nw = (int64_t)gridDim.x * (blockDim.x >> 6);
  for (int64_t r = blockIdx.x * (int64_t)(blockDim.x >> 6) + (threadIdx.x >> 6); r < n_red; r += nw) {
    const int32_t i = rfull[r], k0 = rp[i], k1 = rp[i + 1], len = k1 - k0;
    for (int32_t t = lane; t < len; t += 64) cols[t] = ci[k0 + t];
    sc_wave_lds_sync();
    int64_t base = pmoff[r];
    for (int32_t cb = k0; cb < k1; cb += 64) {
      const int32_t kc = cb + lane;
      const int32_t cc = kc < k1 ? ci[kc] : 0;
      const bool isc = kc < k1 && ridx[cc] < 0;
      const int32_t lenc = isc ? rp[cc + 1] - rp[cc] : 0;
      const int32_t inc = sc_wave_scan(lenc), moff = inc - lenc;
      uint64_t m = __ballot(isc);
      while (m) {
        const int b = __builtin_ctzll(m);
        m &= m - 1;
        const int32_t c = __shfl(cc, b), mo = __shfl(moff, b);
        for (int32_t mm = rp[c] + lane; mm < rp[c + 1]; mm += 64) {
          const int32_t j = ci[mm];
          int32_t pos = -1;
          if (ridx[j] >= 0) {
            int32_t lo = 0, hi = len;
            while (lo < hi) {
              const int32_t mid = (lo + hi) >> 1;
              if (cols[mid] < j)
                lo = mid + 1;
              else
                hi = mid;
            }
            if (lo < len && cols[lo] == j) pos = lo;
          }
          pmap[base + mo + (mm - rp[c])] = (int16_t)pos;
        }
      }
      base += __shfl(inc, 63);
    }
    sc_wave_lds_sync();  // cols are read before the next row overwrites them
  }
}

// Inclusive prefix sum over the 16 lanes of a lane group.
__device__ inline int32_t sc_group_scan(int32_t x) {
  const int sl = threadIdx.x & 15;
#pragma unroll
  for (int d = 1; d < 16; d <<= 1) {
    const int32_t y = __shfl_up(x, d, 16);
    if (sl >= d) x += y;
  }
  return x;
}

// Reduced values through the position map: the same subtractions in the same
// order as k_sc_values_lds (condensed columns of row i ascending, those with a
// zero K_ic skipped, f_c = K_ic / K_cc), so the values are bit-identical.
// Each row c is walked with coalesced loads of its values and map entries, no
// searches.  A wavefront runs four reduced rows, one per 16-lane group, so
// four rows' chains of dependent loads are in flight at once.  LDS: max_row
// doubles per row (16 per workgroup).
__global__ __launch_bounds__(256) void k_sc_values_map(int64_t n_red, const int32_t *__restrict__ rfull,
                                                       const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                                       const double *__restrict__ v, const int32_t *__restrict__ ridx,
                                                       const int32_t *__restrict__ dpos,
                                                       const int64_t *__restrict__ pmoff,
                                                       const int16_t *__restrict__ pmap,
                                                       const int32_t *__restrict__ srp, double *__restrict__ sv,
                                                       int max_row) {
  extern __shared__ double sc_sm[];
  const int lane = threadIdx.x & 63, sl = lane & 15, gb = lane & 48;
  const int slot = (threadIdx.x >> 6) * 4 + (lane >> 4);
  double *acc = sc_sm + (size_t)slot * max_row;
  const uint32_t below = (1u << sl) - 1;
  auto gballot = [&](bool p) { return (uint32_t)(__ballot(p) >> gb) & 0xffffu; };
  const int64_t ng = (int64_t)gridDim.x * (blockDim.x >> 4);
  for (int64_t r = blockIdx.x * (int64_t)(blockDim.x >> 4) + slot; r < n_red; r += ng) {
    const int32_t i = rfull[r], k0 = rp[i], k1 = rp[i + 1], len = k1 - k0;
    for (int32_t t = sl; t < len; t += 16) acc[t] = v[k0 + t];
    sc_wave_lds_sync();
    int64_t base = pmoff[r];
    for (int32_t cb = k0; cb < k1; cb += 16) {
      const int32_t kc = cb + sl;
      const int32_t cc = kc < k1 ? ci[kc] : 0;
      const int32_t qc = kc < k1 ? ridx[cc] : 0;
      const bool isc = kc < k1 && qc < 0;
      const int32_t c0 = isc ? rp[cc] : 0, lenc = isc ? rp[cc + 1] - c0 : 0;
      const double vic = isc ? v[kc] : 0.0;
      const double fl = vic != 0.0 ? vic / v[dpos[-1 - qc]] : 0.0;
      const int32_t inc = sc_group_scan(lenc), moff = inc - lenc;
      uint32_t m = gballot(vic != 0.0);
      while (m) {
        const int b = __builtin_ctz(m);
        m &= m - 1;
        const double f = __shfl(fl, b, 16);
        const int32_t cs = __shfl(c0, b, 16), L = __shfl(lenc, b, 16);
        const int16_t *mp = pmap + base + __shfl(moff, b, 16);
        for (int32_t t = sl; t < L; t += 16) {
          const int32_t pos = mp[t];
          if (pos >= 0) acc[pos] -= f * v[cs + t];
        }
        sc_wave_lds_sync();  // this condensed column's updates land before the next one's
      }
      base += __shfl(inc, 15, 16);
    }
    int32_t out = srp[r];
    for (int32_t t0 = 0; t0 < len; t0 += 16) {
      const int32_t t = t0 + sl;
      const bool keep = t < len && ridx[ci[k0 + t]] >= 0;
      const uint32_t km = gballot(keep);
      if (keep) sv[out + __popc(km & below)] = acc[t];
      out += __popc(km);
    }
    sc_wave_lds_sync();  // the staged row is read before the next row overwrites it
  }
}

// Position map, built once the pattern has passed the fill check; skipped
// (the refresh keeps k_sc_values_lds) when rows are too long for int16
// positions or the map would exceed 4 GiB.
static int sc_build_map(lmh_condensation *sc) {
  lmh_matrix *A = sc->A;
  lmh_context *ctx = A->ctx;
  // the refresh stages max_row doubles per reduced row (16 per workgroup)
  if (!sc->n_cond || A->max_row >= 32768 || (size_t)A->max_row * 8 * 16 > 48 * 1024) return LMH_OK;
  const int64_t nr = sc->n_red;
  if (hipMalloc((void **)&sc->pmoff, sizeof(int64_t) * (nr + 1)) != hipSuccess) {
    (void)hipGetLastError();
    sc->pmoff = nullptr;
    return LMH_OK;  // no room: the refresh keeps the search kernel
  }
  LMH_LAUNCH(ctx, "sc_map_count", k_sc_map_count, dim3(grid_for(nr, 4, 65536)), dim3(256), 0, nr, sc->rfull, A->rp,
             A->ci, sc->ridx, sc->pmoff);
  std::vector<int64_t> h(nr + 1);
  LMH_HIP(hipMemcpyAsync(h.data(), sc->pmoff, sizeof(int64_t) * nr, hipMemcpyDeviceToHost, ctx->stream));
  LMH_HIP(hipStreamSynchronize(ctx->stream));
  int64_t tot = 0;
  for (int64_t r = 0; r < nr; ++r) {
    const int64_t c = h[r];
    h[r] = tot;
    tot += c;
  }
  h[nr] = tot;
  if (tot * (int64_t)sizeof(int16_t) > (int64_t(4) << 30) ||
      hipMalloc((void **)&sc->pmap, sizeof(int16_t) * std::max<int64_t>(1, tot)) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipFree(sc->pmoff);
    sc->pmoff = nullptr;
    sc->pmap = nullptr;
    return LMH_OK;
  }
  LMH_HIP(hipMemcpyAsync(sc->pmoff, h.data(), sizeof(int64_t) * (nr + 1), hipMemcpyHostToDevice, ctx->stream));
  LMH_LAUNCH(ctx, "sc_map_fill", k_sc_map_fill, dim3(grid_for(nr, 4, 65536)), dim3(256), (size_t)A->max_row * 4 * 4,
             nr, sc->rfull, A->rp, A->ci, sc->ridx, sc->pmoff, sc->pmap, A->max_row);
  LMH_TRY(check_launch("k_sc_map_fill"));
  LMH_HIP(hipStreamSynchronize(ctx->stream));  // h is read by the copy above
  return LMH_OK;
}

static int sc_setup(lmh_condensation *sc) {
  lmh_matrix *A = sc->A, *R = sc->R;
  lmh_context *ctx = A->ctx;
  LMH_HIP(hipSetDevice(ctx->device));
  LMH_HIP(hipMemsetAsync(ctx->d_flags, 0, sizeof(int), ctx->stream));
  if (sc->n_cond) {
    LMH_LAUNCH(ctx, "sc_check", k_sc_check, dim3(grid_for(sc->n_cond, 256, 8192)), dim3(256), 0, sc->n_cond, sc->cdof,
               A->rp, A->ci, A->vals, sc->ridx, sc->dpos, ctx->d_flags);
    LMH_TRY(check_launch("k_sc_check"));
  }
  if (!sc->pattern_checked) {
    LMH_LAUNCH(ctx, "sc_values_check", k_sc_values, dim3(grid_for(sc->n_red, 256, 16384)), dim3(256), 0, sc->n_red,
               sc->rfull, A->rp, A->ci, A->vals, sc->ridx, sc->dpos, R->rp, R->ci, R->vals, ctx->d_flags);
    LMH_TRY(check_launch("k_sc_values"));
  } else if (sc->pmap) {
    LMH_LAUNCH(ctx, "sc_values", k_sc_values_map, dim3(grid_for(sc->n_red, 16, 65536)), dim3(256),
               (size_t)A->max_row * 8 * 16, sc->n_red, sc->rfull, A->rp, A->ci, A->vals, sc->ridx, sc->dpos, sc->pmoff,
               sc->pmap, R->rp, R->vals, A->max_row);
    LMH_TRY(check_launch("k_sc_values_map"));
  } else if ((size_t)A->max_row * 12 * 4 <= 48 * 1024) {
    LMH_LAUNCH(ctx, "sc_values", k_sc_values_lds, dim3(grid_for(sc->n_red, 4, 65536)), dim3(256),
               (size_t)A->max_row * 12 * 4, sc->n_red, sc->rfull, A->rp, A->ci, A->vals, sc->ridx, sc->dpos, R->rp,
               R->vals, A->max_row);
    LMH_TRY(check_launch("k_sc_values_lds"));
  } else {
    LMH_LAUNCH(ctx, "sc_values", k_sc_values_wave, dim3(grid_for(sc->n_red, 4, 65536)), dim3(256), 0, sc->n_red,
               sc->rfull, A->rp, A->ci, A->vals, sc->ridx, sc->dpos, R->rp, R->vals);
    LMH_TRY(check_launch("k_sc_values_wave"));
  }
  int flags = 0;
  LMH_HIP(hipMemcpyAsync(&flags, ctx->d_flags, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  LMH_HIP(hipStreamSynchronize(ctx->stream));
  R->have32 = false;
  LMH_CHECK(!(flags & SC_OFFDIAG), LMH_ERR_UNSUPPORTED,
            "condensed DoFs couple to each other: only a diagonal condensed block is supported");
  LMH_CHECK(!(flags & SC_NODIAG), LMH_ERR_INVALID, "a condensed row has no diagonal entry");
  LMH_CHECK(!(flags & SC_ZEROPIVOT), LMH_ERR_DEGENERATE, "zero diagonal on a condensed DoF");
  LMH_CHECK(!(flags & SC_FILL), LMH_ERR_UNSUPPORTED,
            "condensation needs an entry outside the matrix pattern (the condensed DoF is not element-interior)");
  if (!sc->pattern_checked) {
    sc->pattern_checked = true;
    LMH_TRY(sc_build_map(sc));
  }
  return LMH_OK;
}

}  // namespace lmh

using namespace lmh;

extern "C" {

int lmh_condensation_create(lmh_matrix *A, int64_t n_cond, const int32_t *cond_dofs_h, lmh_condensation **out) {
  LMH_CHECK(A && out && n_cond >= 0 && (n_cond == 0 || cond_dofs_h), LMH_ERR_INVALID, "bad argument");
  LMH_CHECK(A->rows == A->cols, LMH_ERR_INVALID, "condensation needs a square matrix");
  LMH_CHECK(A->vals, LMH_ERR_UNSUPPORTED, "matrix holds FP32 values only");
  const int64_t n = A->rows;
  LMH_CHECK(n_cond < n, LMH_ERR_INVALID, "every DoF condensed");
  std::vector<int32_t> cd(cond_dofs_h, cond_dofs_h + n_cond);
  std::sort(cd.begin(), cd.end());
  for (int64_t q = 0; q < n_cond; ++q) {
    LMH_CHECK(cd[q] >= 0 && cd[q] < n, LMH_ERR_INVALID, "condensed DoF %d out of range", cd[q]);
    LMH_CHECK(q == 0 || cd[q] != cd[q - 1], LMH_ERR_INVALID, "condensed DoF %d listed twice", cd[q]);
  }
  lmh_context *ctx = A->ctx;
  LMH_HIP(hipSetDevice(ctx->device));
  std::vector<int32_t> rp(n + 1), ci(A->nnz);
  LMH_TRY(lmh_matrix_get_pattern(A, rp.data(), ci.data()));
  std::vector<int32_t> ridx(n), rfull;
  rfull.reserve(n - n_cond);
  for (int64_t i = 0, q = 0; i < n; ++i) {
    if (q < n_cond && cd[q] == i) {
      ridx[i] = (int32_t)(-1 - q);
      ++q;
    } else {
      ridx[i] = (int32_t)rfull.size();
      rfull.push_back((int32_t)i);
    }
  }
  const int64_t nr = (int64_t)rfull.size();
  std::vector<int32_t> srp(nr + 1, 0), sci;
  sci.reserve(A->nnz);
  for (int64_t r = 0; r < nr; ++r) {
    const int32_t i = rfull[r];
    for (int32_t k = rp[i]; k < rp[i + 1]; ++k)
      if (ridx[ci[k]] >= 0) sci.push_back(ridx[ci[k]]);
    srp[r + 1] = (int32_t)sci.size();
  }
  std::unique_ptr<lmh_condensation> sc(new lmh_condensation);
  sc->ctx = ctx;
  sc->A = A;
  sc->n_full = n;
  sc->n_cond = n_cond;
  sc->n_red = nr;
  const int32_t dummy = 0;
  LMH_TRY(lmh_matrix_create_csr2(ctx, nr, nr, srp.data(), sci.empty() ? &dummy : sci.data(), &sc->R));
  sc->R->precond = A->precond;
  auto fail = [&](int rc) {
    lmh_condensation_destroy(sc.release());
    return rc;
  };
  if (hipMalloc((void **)&sc->ridx, sizeof(int32_t) * n) != hipSuccess ||
      hipMalloc((void **)&sc->rfull, sizeof(int32_t) * nr) != hipSuccess ||
      hipMalloc((void **)&sc->cdof, sizeof(int32_t) * std::max<int64_t>(1, n_cond)) != hipSuccess ||
      hipMalloc((void **)&sc->dpos, sizeof(int32_t) * std::max<int64_t>(1, n_cond)) != hipSuccess ||
      hipMalloc((void **)&sc->g, sizeof(double) * nr) != hipSuccess ||
      hipMalloc((void **)&sc->y, sizeof(double) * nr) != hipSuccess)
    return fail(set_error(LMH_ERR_ALLOC, "hipMalloc failed for the condensation"));
  if (hipMemcpy(sc->ridx, ridx.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(sc->rfull, rfull.data(), sizeof(int32_t) * nr, hipMemcpyHostToDevice) != hipSuccess ||
      (n_cond && hipMemcpy(sc->cdof, cd.data(), sizeof(int32_t) * n_cond, hipMemcpyHostToDevice) != hipSuccess))
    return fail(set_error(LMH_ERR_HIP, "hipMemcpy failed for the condensation"));
  const int rc = sc_setup(sc.get());
  if (rc != LMH_OK) return fail(rc);
  *out = sc.release();
  return LMH_OK;
}

int lmh_condensation_setup(lmh_condensation *sc) {
  LMH_CHECK(sc, LMH_ERR_INVALID, "null condensation");
  return sc_setup(sc);
}

lmh_matrix *lmh_condensation_reduced_matrix(lmh_condensation *sc) { return sc ? sc->R : nullptr; }
int64_t lmh_condensation_n_condensed(const lmh_condensation *sc) { return sc ? sc->n_cond : -1; }

int lmh_condensation_reduced_dofs(const lmh_condensation *sc, int32_t *full_index_h) {
  LMH_CHECK(sc && full_index_h, LMH_ERR_INVALID, "null argument");
  LMH_HIP(hipMemcpy(full_index_h, sc->rfull, sizeof(int32_t) * sc->n_red, hipMemcpyDeviceToHost));
  return LMH_OK;
}

int lmh_condensation_forward(lmh_condensation *sc, const double *full_rhs_d, double *reduced_rhs_d) {
  LMH_CHECK(sc && full_rhs_d && reduced_rhs_d, LMH_ERR_INVALID, "null argument");
  lmh_matrix *A = sc->A;
  lmh_context *ctx = A->ctx;
  LMH_LAUNCH(ctx, "sc_forward", k_sc_forward, dim3(grid_for(sc->n_red, 4, 65536)), dim3(256), 0, sc->n_red,
             sc->rfull, A->rp, A->ci, A->vals, sc->ridx, sc->dpos, full_rhs_d, reduced_rhs_d);
  return check_launch("k_sc_forward");
}

int lmh_condensation_backward(lmh_condensation *sc, const double *full_rhs_d, const double *reduced_sol_d,
                              double *full_sol_d) {
  LMH_CHECK(sc && full_rhs_d && reduced_sol_d && full_sol_d, LMH_ERR_INVALID, "null argument");
  lmh_matrix *A = sc->A;
  lmh_context *ctx = A->ctx;
  LMH_LAUNCH(ctx, "sc_map", k_sc_map<false>, dim3(grid_for(sc->n_red, 256, 16384)), dim3(256), 0, sc->n_red,
             sc->rfull, full_sol_d, const_cast<double *>(reduced_sol_d));
  LMH_TRY(check_launch("k_sc_map"));
  if (!sc->n_cond) return LMH_OK;
  LMH_LAUNCH(ctx, "sc_backward", k_sc_backward, dim3(grid_for(sc->n_cond, 256, 8192)), dim3(256), 0, sc->n_cond,
             sc->cdof, A->rp, A->ci, A->vals, sc->ridx, sc->dpos, full_rhs_d, full_sol_d);
  return check_launch("k_sc_backward");
}

int lmh_condensation_solve(lmh_condensation *sc, const double *b_d, double *x_d, double rtol, int64_t maxits,
                           int64_t *iters, double *rel_res, int *reason) {
  LMH_CHECK(sc && b_d && x_d, LMH_ERR_INVALID, "null argument");
  lmh_context *ctx = sc->A->ctx;
  LMH_TRY(lmh_condensation_forward(sc, b_d, sc->g));
  LMH_LAUNCH(ctx, "sc_map", k_sc_map<true>, dim3(grid_for(sc->n_red, 256, 16384)), dim3(256), 0, sc->n_red, sc->rfull,
             x_d, sc->y);
  LMH_TRY(check_launch("k_sc_map"));
  LMH_TRY(lmh_cg_solve(sc->R, sc->g, sc->y, rtol, maxits, iters, rel_res, reason));
  return lmh_condensation_backward(sc, b_d, sc->y, x_d);
}

int lmh_condensation_destroy(lmh_condensation *sc) {
  if (!sc) return LMH_OK;
  (void)hipSetDevice(sc->ctx->device);
  (void)hipStreamSynchronize(sc->ctx->stream);
  lmh_matrix_destroy(sc->R);
  (void)hipFree(sc->ridx);
  (void)hipFree(sc->rfull);
  (void)hipFree(sc->cdof);
  (void)hipFree(sc->dpos);
  (void)hipFree(sc->g);
  (void)hipFree(sc->y);
  (void)hipFree(sc->pmoff);
  (void)hipFree(sc->pmap);
  delete sc;
  return LMH_OK;
}

int lmh_dofmap_condensed_dofs(const lmh_dofmap *dm, int64_t *n_h, int32_t *dofs_h) {
  LMH_CHECK(dm && n_h, LMH_ERR_INVALID, "null argument");
  const lmh_mesh *m = dm->mesh;
  const int internal = m->type == LMH_HEX27 ? 26 : -1;  // the only internal node among the supported types
  const int64_t n = internal < 0 ? 0 : m->n_elem * dm->nv;
  if (dofs_h && n) {
    std::vector<int32_t> ed((size_t)m->n_elem * dm->dpe);
    LMH_HIP(hipMemcpy(ed.data(), dm->edofs, sizeof(int32_t) * ed.size(), hipMemcpyDeviceToHost));
    for (int64_t e = 0, q = 0; e < m->n_elem; ++e)
      for (int v = 0; v < dm->nv; ++v) dofs_h[q++] = ed[(size_t)e * dm->dpe + (size_t)v * m->npe + internal];
  }
  *n_h = n;
  return LMH_OK;
}

}  // extern "C"
