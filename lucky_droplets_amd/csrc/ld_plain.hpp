// Elementwise kernels of the "plain" step: the reference's dataflow with one
// kernel per array statement and every 3-D transform separate
// (forced-dns-sm.py:287-362, 582-620).  It is the unfused baseline the fused
// engine is compared against; the arithmetic is the same.
#pragma once
#include "ld_kernels.hpp"

namespace ld {

template <int N> __device__ __forceinline__ void specIndex(size_t i, int& ix, int& iy, int& iz) {
  constexpr int Nf = N / 2 + 1;
  iz = (int)(i % Nf);
  const size_t rem = i / Nf;
  iy = (int)(rem % N);
  ix = (int)(rem / N);
}

#define LD_SPEC_LOOP(N_)                                                          \
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;                 \
  if (i >= fs) return;                                                            \
  int ix, iy, iz;                                                                 \
  specIndex<N_>(i, ix, iy, iz);                                                   \
  const Mode<T> md = mode<T, N_>(ix, iy, iz);

// out_c = i (k x uk)_c   (forced-dns-sm.py:295-297, the argument of irfft)
template <typename T, int N>
__global__ void k_pl_curl(const C<T>* __restrict__ uk, C<T>* __restrict__ out, size_t fs) {
  LD_SPEC_LOOP(N)
  const C<T> u0 = uk[i], u1 = uk[fs + i], u2 = uk[2 * fs + i];
  out[i] = cmuli<T>(mkc<T>(md.ky * u2.x - md.kz * u1.x, md.ky * u2.y - md.kz * u1.y));
  out[fs + i] = cmuli<T>(mkc<T>(md.kz * u0.x - md.kx * u2.x, md.kz * u0.y - md.kx * u2.y));
  out[2 * fs + i] = cmuli<T>(mkc<T>(md.kx * u1.x - md.ky * u0.x, md.kx * u1.y - md.ky * u0.y));
}

// r = omg x u  (:299-301)
template <typename T>
__global__ void k_pl_cross(const T* __restrict__ u, const T* __restrict__ w, T* __restrict__ r, size_t R) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= R) return;
  const T ux = u[i], uy = u[R + i], uz = u[2 * R + i], wx = w[i], wy = w[R + i], wz = w[2 * R + i];
  r[i] = wz * uy - wy * uz;
  r[R + i] = wx * uz - wz * ux;
  r[2 * R + i] = wy * ux - wx * uy;
}

// rk = (rk + fk) * mask   (:304-306)
template <typename T, int N>
__global__ void k_pl_add_mask(C<T>* __restrict__ rk, const C<T>* __restrict__ fk, size_t fs) {
  LD_SPEC_LOOP(N)
  for (int c = 0; c < 3; ++c) rk[c * fs + i] = cscale<T>(cadd<T>(rk[c * fs + i], fk[c * fs + i]), md.mask);
}

// pk = i invlap (k . rk); ku = rk - i k pk - nu k^2 uk; vhat = (uk - tps(-i k pk - nu k^2 uk)) mask
// (:309, :314-316, :319-321)
template <typename T, int N>
__global__ void k_pl_pressure(const C<T>* __restrict__ rk, const C<T>* __restrict__ uk, C<T>* __restrict__ ku,
                              C<T>* __restrict__ vh, size_t fs, T nu_eff, T tps) {
  LD_SPEC_LOOP(N)
  const C<T> r0 = rk[i], r1 = rk[fs + i], r2 = rk[2 * fs + i];
  const C<T> D = mkc<T>(md.kx * r0.x + md.ky * r1.x + md.kz * r2.x, md.kx * r0.y + md.ky * r1.y + md.kz * r2.y);
  const C<T> pk = cmuli<T>(cscale<T>(D, md.invlap));
  const T kk[3] = {md.kx, md.ky, md.kz};
  const C<T> rr[3] = {r0, r1, r2};
  for (int c = 0; c < 3; ++c) {
    const C<T> u = uk[c * fs + i];
    const C<T> ikp = cmuli<T>(cscale<T>(pk, kk[c]));
    const C<T> vs = cscale<T>(u, nu_eff * md.k2);
    ku[c * fs + i] = csub<T>(csub<T>(rr[c], ikp), vs);
    vh[c * fs + i] = cscale<T>(csub<T>(u, cscale<T>(csub<T>(cscale<T>(ikp, (T)-1), vs), tps)), md.mask);
  }
}

// x -= a (v_z -= tps g, :321)
template <typename T>
__global__ void k_pl_shift(T* __restrict__ x, size_t n, T a) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) x[i] -= a;
}

// vn = n v  (:326)
template <typename T>
__global__ void k_pl_nv(const T* __restrict__ n, const T* __restrict__ v, T* __restrict__ vn, size_t R) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= R) return;
  const T nn = n[i];
  vn[i] = nn * v[i]; vn[R + i] = nn * v[R + i]; vn[2 * R + i] = nn * v[2 * R + i];
}

// x *= mask for nf fields (the `* dealias` of :325, :327-329)
template <typename T, int N>
__global__ void k_pl_mask(C<T>* __restrict__ x, size_t fs, int nf) {
  LD_SPEC_LOOP(N)
  for (int c = 0; c < nf; ++c) x[c * fs + i] = cscale<T>(x[c * fs + i], md.mask);
}

// out = (-i k . vnk - fck) * mask  (the argument of irfft at :330)
template <typename T, int N>
__global__ void k_pl_div(const C<T>* __restrict__ vnk, const C<T>* __restrict__ fck, C<T>* __restrict__ out, size_t fs) {
  LD_SPEC_LOOP(N)
  const C<T> a = vnk[i], b = vnk[fs + i], c = vnk[2 * fs + i];
  const C<T> E = mkc<T>(md.kx * a.x + md.ky * b.x + md.kz * c.x, md.kx * a.y + md.ky * b.y + md.kz * c.y);
  out[i] = cscale<T>(csub<T>(mkc<T>(E.y, -E.x), fck[i]), md.mask);
}

// y = x + s z (complex or real arrays of n elements)
template <typename X, typename S>
__global__ void k_pl_axpy(const X* __restrict__ x, const X* __restrict__ z, X* __restrict__ y, size_t n, S s) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  if constexpr (sizeof(X) == sizeof(S)) y[i] = x[i] + s * z[i];
  else { X a = x[i], b = z[i]; a.x += s * b.x; a.y += s * b.y; y[i] = a; }
}

// pk = invlap (k . uk); uk += k pk  (:607-610), with the per-block max of Re(uk) (:618)
template <typename T, int N>
__global__ void __launch_bounds__(256)
k_pl_project(C<T>* __restrict__ uk, size_t fs, double* __restrict__ maxpart) {
  __shared__ double sh[256];
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  double mx = -INFINITY;
  if (i < fs) {
    int ix, iy, iz;
    specIndex<N>(i, ix, iy, iz);
    const Mode<T> md = mode<T, N>(ix, iy, iz);
    const C<T> a = uk[i], b = uk[fs + i], c = uk[2 * fs + i];
    const C<T> d = mkc<T>(md.kx * a.x + md.ky * b.x + md.kz * c.x, md.kx * a.y + md.ky * b.y + md.kz * c.y);
    const C<T> pk = cscale<T>(d, md.invlap);
    const C<T> a2 = cadd<T>(a, cscale<T>(pk, md.kx)), b2 = cadd<T>(b, cscale<T>(pk, md.ky)),
               c2 = cadd<T>(c, cscale<T>(pk, md.kz));
    uk[i] = a2; uk[fs + i] = b2; uk[2 * fs + i] = c2;
    mx = fmax((double)a2.x, fmax((double)b2.x, (double)c2.x));
  }
  const double bm = blockMax<256>(mx, sh);
  if (threadIdx.x == 0) maxpart[blockIdx.x] = bm;
}

static __global__ void k_max_final_n(const double* __restrict__ part, int nblk, double thresh, int* flag,
                                     double* lastmax) {
  __shared__ double sh[256];
  double m = -INFINITY;
  for (int i = threadIdx.x; i < nblk; i += 256) m = fmax(m, part[i]);
  const double b = blockMax<256>(m, sh);
  if (threadIdx.x == 0) { *lastmax = b; if (b > thresh) *flag = 1; }
}

}  // namespace ld
