// Diagnostics of the reference's save() and plot.py, on the device:
//   shell spectrum  e3d_to_e1d(w * sum_c |f_c|^2 * normalize)   (forced-dns-sm.py:232-233, 423, 431; plot.py:37-38)
//   energy flux     e3d_to_e1d(Re(conj(uk) . ku) * dealias * normalize), cumulated from high k
//                                                                   (forced-dns-sm.py:425-428)
//   field stats     sum, min, max, sum of squares of a real field  (forced-dns-sm.py:467-471)
// Shell binning follows np.histogram with bins [0, 0.5), [0.5, 1.5), ...,
// [N/2 - 0.5, N/2 + 0.5]: bin s holds |k| in [s - 1/2, s + 1/2) (s >= 1) and
// |k| < 1/2 for s = 0; modes with |k| > N/2 + 1/2 are not counted.
// Reductions are deterministic: per-block partial sums, then a fixed-order
// final sum.
#pragma once
#include "ld_kernels.hpp"

namespace ld {

// bin of integer |k|^2 (or -1 outside the last bin): round(|k|) is the bin
// for |k| < N/2 + 1/2; |k| = N/2 + 1/2 exactly cannot occur for integer k2.
template <int N> __device__ __forceinline__ int shellBin(int k2) {
  const int b = kround(k2);
  return b <= N / 2 ? b : -1;
}

// mode: 0 -> w * sum_c |f_c|^2 * normalize (nf fields); 1 -> Re(conj(a).b) * mask * normalize (3 fields each);
// 2 -> w * |k|^2 * sum_c |f_c|^2 * normalize (nf fields: the dissipation spectrum for w = nu)
template <typename T, int N>
__global__ void __launch_bounds__(256)
k_shell_partial(const C<T>* __restrict__ a, const C<T>* __restrict__ b, int nf, size_t fs, int mode, double w,
                double norm0, double* __restrict__ part) {
  constexpr int Nf = N / 2 + 1;
  // one set of bins per wavefront (256 threads = 4), summed in a fixed order at
  // the end: additions into one bin come only from the lanes of one
  // instruction, whose order the hardware fixes, so the sums are reproducible
  __shared__ double bins[4][Nf];
  for (int i = threadIdx.x; i < 4 * Nf; i += blockDim.x) bins[i / Nf][i % Nf] = 0.0;
  __syncthreads();
  double* wbins = bins[threadIdx.x >> 6];
  const size_t per = (fs + gridDim.x - 1) / gridDim.x;
  const size_t i0 = blockIdx.x * per, i1 = (i0 + per < fs) ? i0 + per : fs;
  for (size_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
    const int iz = (int)(i % Nf);
    const size_t rem = i / Nf;
    const int kx = kwave((int)(rem / N), N), ky = kwave((int)(rem % N), N);
    const int k2 = kx * kx + ky * ky + iz * iz;
    const int s = shellBin<N>(k2);
    if (s < 0) continue;
    const double nrm = (iz == 0 || iz == N / 2) ? norm0 : 2.0 * norm0;
    double v = 0.0;
    if (mode != 1) {
      for (int c = 0; c < nf; ++c) {
        const C<T> z = a[c * fs + i];
        v += (double)z.x * (double)z.x + (double)z.y * (double)z.y;
      }
      v *= (mode == 2 ? w * (double)k2 : w) * nrm;
    } else {
      const int ki = kround(k2);
      if (3 * ki > N) continue;   // dealias mask
      for (int c = 0; c < 3; ++c) {
        const C<T> x = a[c * fs + i], y = b[c * fs + i];
        v += (double)x.x * (double)y.x + (double)x.y * (double)y.y;
      }
      v *= nrm;
    }
    atomicAdd(&wbins[s], v);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < Nf; i += blockDim.x)
    part[(size_t)blockIdx.x * Nf + i] = ((bins[0][i] + bins[1][i]) + bins[2][i]) + bins[3][i];
}

// out[s] = sum over blocks (fixed order); cumulate != 0 -> cumulative sum from the highest shell down
template <int N>
__global__ void k_shell_final(const double* __restrict__ part, int nblk, int cumulate, double* __restrict__ out) {
  constexpr int Nf = N / 2 + 1;
  __shared__ double tot[Nf];
  for (int s = threadIdx.x; s < Nf; s += blockDim.x) {
    double acc = 0.0;
    for (int b = 0; b < nblk; ++b) acc += part[(size_t)b * Nf + s];
    tot[s] = acc;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    if (cumulate) {
      double run = 0.0;
      for (int s = Nf - 1; s >= 0; --s) { run += tot[s]; out[s] = run; }
    } else {
      for (int s = 0; s < Nf; ++s) out[s] = tot[s];
    }
  }
}

// Spectral divergence i (kx u_x + ky u_y + kz u_z) of a vector field, with the
// derivative of each axis's Nyquist mode taken as zero: the reference's
// diff_x / diff_y / diff_z (forced-dns-sm.py:216-229) differentiate with a
// 1-D rfft along the axis, multiply by i k (k = 0..N/2) and irfft, which drops
// the imaginary Nyquist coefficient that i (N/2) times a real one makes.
template <typename T, int N>
__global__ void __launch_bounds__(256)
k_div_spec(const C<T>* __restrict__ u, size_t fs, C<T>* __restrict__ out) {
  constexpr int Nf = N / 2 + 1;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < fs; i += (size_t)gridDim.x * blockDim.x) {
    const int iz = (int)(i % Nf);
    const size_t rem = i / Nf;
    const int ix = (int)(rem / N), iy = (int)(rem % N);
    const T kx = ix == N / 2 ? (T)0 : (T)kwave(ix, N);
    const T ky = iy == N / 2 ? (T)0 : (T)kwave(iy, N);
    const T kz = iz == N / 2 ? (T)0 : (T)iz;
    const C<T> a = u[i], b = u[fs + i], c = u[2 * fs + i];
    const T re = kx * a.x + ky * b.x + kz * c.x, im = kx * a.y + ky * b.y + kz * c.y;
    out[i] = mkc<T>(-im, re);      // i (k . u)
  }
}

// per block: sum, min, max, sum of squares
template <typename T>
__global__ void __launch_bounds__(256)
k_stats_partial(const T* __restrict__ x, size_t n, double* __restrict__ part) {
  __shared__ double sh[256];
  const size_t per = (n + gridDim.x - 1) / gridDim.x;
  const size_t i0 = blockIdx.x * per, i1 = (i0 + per < n) ? i0 + per : n;
  double s = 0.0, q = 0.0, mn = INFINITY, mx = -INFINITY;
  for (size_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
    const double v = (double)x[i];
    s += v; q += v * v; mn = fmin(mn, v); mx = fmax(mx, v);
  }
  const double a = blockSum<256>(s, sh);
  __syncthreads();
  const double b = blockSum<256>(q, sh);
  __syncthreads();
  const double c = -blockMax<256>(-mn, sh);
  __syncthreads();
  const double d = blockMax<256>(mx, sh);
  if (threadIdx.x == 0) { part[4 * blockIdx.x] = a; part[4 * blockIdx.x + 1] = c; part[4 * blockIdx.x + 2] = d; part[4 * blockIdx.x + 3] = b; }
}

static __global__ void k_stats_final(const double* __restrict__ part, int nblk, double* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s = 0.0, q = 0.0, mn = INFINITY, mx = -INFINITY;
  for (int b = 0; b < nblk; ++b) {
    s += part[4 * b]; mn = fmin(mn, part[4 * b + 1]); mx = fmax(mx, part[4 * b + 2]); q += part[4 * b + 3];
  }
  out[0] = s; out[1] = mn; out[2] = mx; out[3] = q;
}

}  // namespace ld
