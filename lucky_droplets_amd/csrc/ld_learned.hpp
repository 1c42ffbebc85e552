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

// The velocity u (3 physical fields) as the droplet gather's u records
// {u_x, u_y, u_z, 0} (recU: z-quads of four 16-byte records, so the four z
// nodes of a stencil column span one or two 128-byte lines).
template <typename T, int N>
__global__ void __launch_bounds__(256)
k_u_records(const T* __restrict__ u, V4<T>* __restrict__ g) {
  constexpr size_t R = (size_t)N * N * N;
  for (size_t o = blockIdx.x * (size_t)blockDim.x + threadIdx.x; o < R; o += (size_t)gridDim.x * blockDim.x) {
    V4<T> q;
    q.x = u[o]; q.y = u[R + o]; q.z = u[2 * R + o]; q.w = (T)0;
    g[recU((uint32_t)o)] = q;
  }
}

// Droplets advanced nsteps classical RK4 steps of size h in a frozen velocity
// field (its records g from k_u_records): x' = v, v' = (I_u(x) - v) / (St
// tau_eta) - g e_z with I_u the cosine-kernel gather (particles.py:435-511) and
// St = (m / factor)^(2/3) (particles.py:823); the mass does not change.  Rows
// (Np, 7) float64 in place, positions unwrapped.  Four lanes per droplet, lane
// c owning z node c of the 4x4x4 stencil, as in k_particles_stage (the same
// wrap, stencil origin and weights: sines and cosines and the gather sums in
// the field precision, positions and the RK update in FP64).  Blocks map
// XCD-aware: each XCD takes one contiguous range of rows, so rows sorted by
// position share its L2.
template <typename T, int N>
__global__ void __launch_bounds__(256)
k_frozen_rk4(double* __restrict__ rows, int np, const V4<T>* __restrict__ g, double h, int nsteps, DropPhys P) {
  const int nb = gridDim.x, per = nb / 8;
  const int lb = (blockIdx.x % 8) * per + blockIdx.x / 8;
  const int gt = lb * blockDim.x + threadIdx.x;
  const int p = gt >> 2, c = gt & 3;
  const bool live = p < np;
  double* row = rows + (size_t)(live ? p : 0) * 7;
  using GT = std::conditional_t<sizeof(T) == 4, float, double>;
  double x[3], v[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) { x[k] = row[k]; v[k] = row[3 + k]; }
  const double st = pow(row[6] / P.factor, 2.0 / 3.0);
  const double rden = 1.0 / (st * P.tau_eta);
  const double a_[4] = {0.0, 0.5, 0.5, 1.0}, w_[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
  const GT q = (GT)0.25;
  for (int it = 0; it < nsteps; ++it) {
    double kx[3] = {0, 0, 0}, kv[3] = {0, 0, 0}, sx[3], sv[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) { sx[k] = x[k]; sv[k] = v[k]; }
#pragma unroll 1
    for (int s = 0; s < 4; ++s) {
      double xs[3], vs[3], xw[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        xs[k] = x[k] + a_[s] * h * kx[k];
        vs[k] = v[k] + a_[s] * h * kv[k];
        xw[k] = wrapFast(xs[k], P.L, P.invL);
      }
      const int i0 = originOf(xw[0], P.inv_dx), j0 = originOf(xw[1], P.inv_dx), k0 = originOf(xw[2], P.inv_dx);
      GT xsn, xcs, ysn, ycs, zsn, zcs;
      sincosGT((xw[0] - (double)i0 * P.dx) * P.pi_2dx, xsn, xcs);
      sincosGT((xw[1] - (double)j0 * P.dx) * P.pi_2dx, ysn, ycs);
      sincosGT((xw[2] - (double)k0 * P.dx) * P.pi_2dx, zsn, zcs);
      const GT wy[4] = {q * ((GT)1 + ycs), q * ((GT)1 + ysn), q * ((GT)1 - ycs), q * ((GT)1 - ysn)};
      const GT wz = q * ((GT)1 + ((c & 2) ? (GT)-1 : (GT)1) * ((c & 1) ? zsn : zcs));
      const uint32_t iz = (uint32_t)wrapIdx<N>(k0 + c);
      GT Ig[3] = {(GT)0, (GT)0, (GT)0};
      if (live) {
        V4<T> U[16];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          const uint32_t plane = (uint32_t)wrapIdx<N>(i0 + a) * N;
#pragma unroll
          for (int b = 0; b < 4; ++b)
            U[a * 4 + b] = ldRec<T, N>(g, recU((plane + (uint32_t)wrapIdx<N>(j0 + b)) * N + iz));
        }
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          const GT wx = q * ((GT)1 + ((a & 2) ? (GT)-1 : (GT)1) * ((a & 1) ? xsn : xcs));
#pragma unroll
          for (int b = 0; b < 4; ++b) {
            const GT w = wy[b] * wx;
            Ig[0] += (GT)U[a * 4 + b].x * w;
            Ig[1] += (GT)U[a * 4 + b].y * w;
            Ig[2] += (GT)U[a * 4 + b].z * w;
          }
        }
      }
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        GT t = Ig[k] * wz;
        t += __shfl_xor(t, 1, 64);
        t += __shfl_xor(t, 2, 64);
        kx[k] = vs[k];
        kv[k] = ((double)t - vs[k]) * rden - (k == 2 ? P.g : 0.0);
        sx[k] += w_[s] * h * kx[k];
        sv[k] += w_[s] * h * kv[k];
      }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) { x[k] = sx[k]; v[k] = sv[k]; }
  }
  if (!live || c != 0) return;
#pragma unroll
  for (int k = 0; k < 3; ++k) { row[k] = x[k]; row[3 + k] = v[k]; }
}

}  // namespace ld
