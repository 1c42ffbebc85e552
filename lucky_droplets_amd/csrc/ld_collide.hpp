// Droplet-droplet collision pairs (cell list) and coalescence events.
//
// No counterpart in the reference (its droplets only grow by collecting the
// continuum density, particles.py:828).  Specification, shared with the CPU
// brute-force checker bit for bit:
//   d = x_j - x_i per axis on wrapped positions, d -= L * rint(d / L),
//   d2 = (dx*dx + dy*dy) + dz*dz, pair iff d2 < (r_i + r_j)^2, i < j,
//   IEEE binary64 without fused multiply-add; pairs sorted by (i, j).
// Coalescence visits the sorted pairs once, merging (i, j) only if neither
// droplet merged earlier in the call: survivor i takes m = m_i + m_j, the
// mass-weighted velocity and x_i + (m_j/m) d; droplet j is left with zero
// mass and velocity.
#pragma once
#include "ld_particles.hpp"

namespace ld {

// Cell of a wrapped position on an nc^3 cell grid of side `cell`.
__device__ __forceinline__ int collCell(double x, double cell, int nc) {
  int c = (int)(x / cell);
  return c < 0 ? 0 : (c >= nc ? nc - 1 : c);
}

static __global__ void __launch_bounds__(256)
k_coll_count(const double* __restrict__ rows, int np, double L, double cell, int nc, int* __restrict__ key,
             int* __restrict__ rank, int* __restrict__ hist) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const double* r = rows + (size_t)p * 7;
  const int cx = collCell(wrapL(r[0], L), cell, nc), cy = collCell(wrapL(r[1], L), cell, nc),
            cz = collCell(wrapL(r[2], L), cell, nc);
  const int k = (cx * nc + cy) * nc + cz;
  key[p] = k;
  rank[p] = atomicAdd(&hist[k], 1);
}

static __global__ void __launch_bounds__(256)
k_coll_scatter(int np, const int* __restrict__ key, const int* __restrict__ rank, const int* __restrict__ start,
               int* __restrict__ order) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < np) order[start[key[p]] + rank[p]] = p;
}

// Minimum-image separation and the pair test, no contraction into FMA.
__device__ __forceinline__ bool collideTest(const double* __restrict__ ri, const double* __restrict__ rj, double L,
                                            double radi, double radj) {
#pragma clang fp contract(off)
  double d[3];
  for (int a = 0; a < 3; ++a) {
    double t = wrapL(rj[a], L) - wrapL(ri[a], L);
    t = t - L * rint(t / L);
    d[a] = t;
  }
  const double d2 = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2];
  const double s = radi + radj;
  return d2 < s * s;
}

// One thread per droplet i: every droplet j > i in the 27 neighbouring cells
// (nc >= 3) or in the single cell (nc == 1).  Pairs are appended as
// (i << 32 | j) through an atomic counter; the caller sorts them.
static __global__ void __launch_bounds__(256)
k_coll_pairs(const double* __restrict__ rows, const double* __restrict__ rad, int np, double L, double cell, int nc,
             const int* __restrict__ start, const int* __restrict__ order, const int* __restrict__ key,
             unsigned long long* __restrict__ out, unsigned long long max_pairs,
             unsigned long long* __restrict__ count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const double* xi = rows + (size_t)i * 7;
  const double ri = rad[i];
  const int ki = key[i];
  const int cx = ki / (nc * nc), cy = (ki / nc) % nc, cz = ki % nc;
  const int span = nc >= 3 ? 1 : 0;
  const int ncells = nc * nc * nc;
  for (int ox = -span; ox <= span; ++ox)
    for (int oy = -span; oy <= span; ++oy)
      for (int oz = -span; oz <= span; ++oz) {
        const int c = ((((cx + ox) % nc + nc) % nc) * nc + (((cy + oy) % nc + nc) % nc)) * nc + (((cz + oz) % nc + nc) % nc);
        const int b = start[c], e = (c + 1 < ncells) ? start[c + 1] : np;
        for (int t = b; t < e; ++t) {
          const int j = order[t];
          if (j <= i) continue;
          if (collideTest(xi, rows + (size_t)j * 7, L, ri, rad[j])) {
            const unsigned long long slot = atomicAdd(count, 1ull);
            if (slot < max_pairs) out[slot] = ((unsigned long long)i << 32) | (unsigned long long)(unsigned)j;
          }
        }
      }
}

// Pair statistics within a cutoff (no reference counterpart; the input of the
// radial distribution function and the kinematic collision kernel).  For every
// pair i < j with d2 < rmax^2 (minimum image, the pair-test formula above):
//   dist = sqrt(d2), bin k = min((int)(dist / dr), nbins - 1),
//   w_r = ((dv_x d_x + dv_y d_y) + dv_z d_z) / dist  (dv = v_j - v_i; 0 if dist == 0),
//   counts[k] += 1, wsum[k] += |w_r|, wneg[k] += max(-w_r, 0).
// Per-block LDS histograms, then one global atomic per bin and block; counts
// are exact, the FP64 sums depend on the summation order.
static __global__ void __launch_bounds__(256)
k_pair_stats(const double* __restrict__ rows, int np, double L, double cell, int nc, const int* __restrict__ start,
             const int* __restrict__ order, const int* __restrict__ key, double rmax, double dr, int nbins,
             unsigned long long* __restrict__ counts, double* __restrict__ wsum, double* __restrict__ wneg) {
#pragma clang fp contract(off)
  __shared__ unsigned long long hc[512];
  __shared__ double hw[512], hn[512];
  for (int k = threadIdx.x; k < nbins; k += blockDim.x) { hc[k] = 0ull; hw[k] = 0.0; hn[k] = 0.0; }
  __syncthreads();
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < np) {
    const double* xi = rows + (size_t)i * 7;
    double pi[3];
    for (int a = 0; a < 3; ++a) pi[a] = wrapL(xi[a], L);
    const double r2max = rmax * rmax;
    const int ki = key[i];
    const int cx = ki / (nc * nc), cy = (ki / nc) % nc, cz = ki % nc;
    const int span = nc >= 3 ? 1 : 0;
    const int ncells = nc * nc * nc;
    for (int ox = -span; ox <= span; ++ox)
      for (int oy = -span; oy <= span; ++oy)
        for (int oz = -span; oz <= span; ++oz) {
          const int c = ((((cx + ox) % nc + nc) % nc) * nc + (((cy + oy) % nc + nc) % nc)) * nc + (((cz + oz) % nc + nc) % nc);
          const int b = start[c], e = (c + 1 < ncells) ? start[c + 1] : np;
          for (int t = b; t < e; ++t) {
            const int j = order[t];
            if (j <= i) continue;
            const double* xj = rows + (size_t)j * 7;
            double d[3];
            for (int a = 0; a < 3; ++a) {
              double u = wrapL(xj[a], L) - pi[a];
              u = u - L * rint(u / L);
              d[a] = u;
            }
            const double d2 = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2];
            if (!(d2 < r2max)) continue;
            const double dist = sqrt(d2);
            int k = (int)(dist / dr);
            k = k < nbins ? k : nbins - 1;
            double wr = 0.0;
            if (dist > 0.0) {
              const double dv0 = xj[3] - xi[3], dv1 = xj[4] - xi[4], dv2 = xj[5] - xi[5];
              wr = ((dv0 * d[0] + dv1 * d[1]) + dv2 * d[2]) / dist;
            }
            atomicAdd(&hc[k], 1ull);
            atomicAdd(&hw[k], fabs(wr));
            if (wr < 0.0) atomicAdd(&hn[k], -wr);
          }
        }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < nbins; k += blockDim.x) {
    if (hc[k]) {
      atomicAdd(&counts[k], hc[k]);
      unsafeAtomicAdd(&wsum[k], hw[k]);
      unsafeAtomicAdd(&wneg[k], hn[k]);
    }
  }
}

// One compare-exchange step of a bitonic sort over n = 2^m keys.
static __global__ void __launch_bounds__(256)
k_bitonic(unsigned long long* __restrict__ a, int n, int j, int k) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int ixj = i ^ j;
  if (ixj > i) {
    const bool up = (i & k) == 0;
    const unsigned long long x = a[i], y = a[ixj];
    if ((x > y) == up) { a[i] = y; a[ixj] = x; }
  }
}

static __global__ void k_pairs_out(const unsigned long long* __restrict__ keys, int n, int64_t* __restrict__ pairs) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  pairs[2 * t] = (int64_t)(keys[t] >> 32);
  pairs[2 * t + 1] = (int64_t)(keys[t] & 0xffffffffull);
}

// Sequential greedy coalescence over the sorted pairs (one thread; the pair
// list is short).  used[] must be zero on entry.
static __global__ void k_coalesce(double* rows, const int64_t* __restrict__ pairs, int npairs, double L,
                                  int* __restrict__ used, int64_t* __restrict__ events, int* __restrict__ nevents) {
#pragma clang fp contract(off)
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  int ne = 0;
  for (int t = 0; t < npairs; ++t) {
    const int64_t i = pairs[2 * t], j = pairs[2 * t + 1];
    if (used[i] || used[j]) continue;
    double* a = rows + (size_t)i * 7;
    double* b = rows + (size_t)j * 7;
    const double mi = a[6], mj = b[6];
    const double m = mi + mj;
    const double w = mj / m;
    for (int c = 0; c < 3; ++c) {
      const double xi = wrapL(a[c], L), xj = wrapL(b[c], L);
      double d = xj - xi;
      d = d - L * rint(d / L);
      a[c] = wrapL(xi + w * d, L);
      a[3 + c] = (mi * a[3 + c] + mj * b[3 + c]) / m;
      b[3 + c] = 0.0;
    }
    a[6] = m;
    b[6] = 0.0;
    used[i] = 1;
    used[j] = 1;
    events[2 * ne] = i;
    events[2 * ne + 1] = j;
    ++ne;
  }
  *nevents = ne;
}

}  // namespace ld
