// Kernels of the learned subgrid form (no counterpart in the reference): moving
// fields between a coarse and a fine grid, the CNN's input and target layouts,
// the reconstruction of a fine-grid velocity from a coarse one plus the
// network's subgrid part, and droplets advanced in a given velocity field.
#pragma once
#include "ld_particles.hpp"

namespace ld {

// Spectral resampling of nf half-spectrum fields from an Ns^3 grid to this
// N^3 grid: modes with |kx|, |ky|, kz < min(Ns, N)/2 are copied, scaled by
// (N/Ns)^3 (forward transforms are unnormalised), every other mode is zero
// (the Nyquist modes of the smaller grid included, so the result stays the
// transform of a real field).  Truncation when Ns > N, zero padding when Ns < N.
template <typename T, int N>
__global__ void __launch_bounds__(256)
k_resample(const C<T>* __restrict__ src, int Ns, C<T>* __restrict__ dst, int nf, T scale) {
  constexpr int Nf = N / 2 + 1;
  const size_t fs = (size_t)N * N * Nf, fss = (size_t)Ns * Ns * (Ns / 2 + 1);
  const int kmax = (Ns < N ? Ns : N) / 2;      // |k| < kmax is kept
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < (size_t)nf * fs; i += (size_t)gridDim.x * blockDim.x) {
    const int f = (int)(i / fs);
    const size_t e = i % fs;
    const int kz = (int)(e % Nf);
    const int ky = kwave((int)((e / Nf) % N), N), kx = kwave((int)(e / Nf / N), N);
    C<T> v = mkc<T>((T)0, (T)0);
    if (kz < kmax && ky < kmax && ky > -kmax && kx < kmax && kx > -kmax) {
      const int sx = kx < 0 ? kx + Ns : kx, sy = ky < 0 ? ky + Ns : ky;
      const C<T> a = src[f * fss + ((size_t)sx * Ns + sy) * (Ns / 2 + 1) + kz];
      v = mkc<T>(a.x * scale, a.y * scale);
    }
    dst[i] = v;
  }
}

// Velocity-gradient spectra: out[3 i + j] = s * i k_j u_i for the three
// components u_i of a half-spectrum velocity (the derivative of each axis's
// Nyquist mode taken as zero).
template <typename T, int N>
__global__ void __launch_bounds__(256)
k_grad_spec(const C<T>* __restrict__ u, C<T>* __restrict__ out, T s) {
  constexpr int Nf = N / 2 + 1;
  constexpr size_t fs = (size_t)N * N * Nf;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < fs; e += (size_t)gridDim.x * blockDim.x) {
    const int iz = (int)(e % Nf), iy = (int)((e / Nf) % N), ix = (int)(e / Nf / N);
    const T k[3] = {ix == N / 2 ? (T)0 : (T)kwave(ix, N), iy == N / 2 ? (T)0 : (T)kwave(iy, N),
                    iz == N / 2 ? (T)0 : (T)iz};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const C<T> a = u[i * fs + e];
#pragma unroll
      for (int j = 0; j < 3; ++j) out[(3 * i + j) * fs + e] = mkc<T>(-a.y * k[j] * s, a.x * k[j] * s);
    }
  }
}

// Planar fields f[c][node] (nc fields of N^3) -> rows out[node][c] * s (FP32).
template <typename T, int N>
__global__ void __launch_bounds__(256)
k_fields_to_rows(const T* __restrict__ f, int nc, float* __restrict__ out, float s) {
  constexpr size_t R = (size_t)N * N * N;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < R * nc; i += (size_t)gridDim.x * blockDim.x) {
    const size_t node = i / nc;
    const int c = (int)(i % nc);
    out[i] = (float)f[c * R + node] * s;
  }
}

// Subgrid channels.  With ratio r the fine grid (this N) has r^3 nodes per
// coarse cell; fine node (X, Y, Z) is coarse node (X/r, Y/r, Z/r) with sub-node
// sub = ((X%r) r + Y%r) r + Z%r, and channel c r^3 + sub of the coarse node's
// row (3 r^3 channels) holds velocity component c there.
//   compose:  out[c][X] = base[c][X] + s * net[node][c r^3 + sub]
//   residual: net[node][c r^3 + sub] = (fine[c][X] - base[c][X]) * s
template <typename T, int N>
__global__ void __launch_bounds__(256)
k_subgrid(const float* __restrict__ net_in, float* __restrict__ net_out, int r, const T* __restrict__ base,
          const T* __restrict__ fine, T* __restrict__ out, float s) {
  constexpr size_t R = (size_t)N * N * N;
  const int Nc = N / r, r3 = r * r * r;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < 3 * R; i += (size_t)gridDim.x * blockDim.x) {
    const int c = (int)(i / R);
    const size_t e = i % R;
    const int X = (int)(e / ((size_t)N * N)), Y = (int)((e / N) % N), Z = (int)(e % N);
    const size_t node = ((size_t)(X / r) * Nc + Y / r) * Nc + Z / r;
    const int sub = ((X % r) * r + Y % r) * r + Z % r;
    const size_t ch = node * (size_t)(3 * r3) + (size_t)c * r3 + sub;
    if (net_out)
      net_out[ch] = (float)(fine[i] - base[i]) * s;
    else
      out[i] = base[i] + (T)(s * net_in[ch]);
  }
}

// Droplets advanced nsteps classical RK4 steps of size h in a frozen velocity
// field u (3 physical fields on this grid): x' = v, v' = (I_u(x) - v) / (St
// tau_eta) - g e_z with I_u the cosine-kernel gather (particles.py:435-511) and
// St = (m / factor)^(2/3) (particles.py:823); the mass does not change.  Rows
// (Np, 7) float64 in place, positions unwrapped.  One thread per droplet.
template <typename T, int N>
__global__ void __launch_bounds__(256)
k_frozen_rk4(double* __restrict__ rows, int np, const T* __restrict__ u, double h, int nsteps, DropPhys P) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  constexpr size_t R = (size_t)N * N * N;
  double* row = rows + (size_t)p * 7;
  double x[3], v[3];
  for (int k = 0; k < 3; ++k) { x[k] = row[k]; v[k] = row[3 + k]; }
  const double st = pow(row[6] / P.factor, 2.0 / 3.0);
  const double rden = 1.0 / (st * P.tau_eta);
  const double a[4] = {0.0, 0.5, 0.5, 1.0}, w[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
  for (int it = 0; it < nsteps; ++it) {
    double kx[3] = {0, 0, 0}, kv[3] = {0, 0, 0}, sx[3], sv[3];
    for (int k = 0; k < 3; ++k) { sx[k] = x[k]; sv[k] = v[k]; }
    for (int s = 0; s < 4; ++s) {
      double xs[3], vs[3], xw[3];
      for (int k = 0; k < 3; ++k) {
        xs[k] = x[k] + a[s] * h * kx[k];
        vs[k] = v[k] + a[s] * h * kv[k];
        xw[k] = wrapL(xs[k], P.L);
      }
      Stencil<N> sten;
      sten.set(xw, P.dx, P.pi_2dx);
      double I[3];
      gatherN<T, N, 3>(sten, u, R, I);
      for (int k = 0; k < 3; ++k) {
        kx[k] = vs[k];
        kv[k] = (I[k] - vs[k]) * rden - (k == 2 ? P.g : 0.0);
        sx[k] += w[s] * h * kx[k];
        sv[k] += w[s] * h * kv[k];
      }
    }
    for (int k = 0; k < 3; ++k) { x[k] = sx[k]; v[k] = sv[k]; }
  }
  for (int k = 0; k < 3; ++k) { row[k] = x[k]; row[3 + k] = v[k]; }
}

}  // namespace ld
