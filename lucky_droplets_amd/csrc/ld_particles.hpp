// Droplet kernels: cosine-kernel gather/deposit (particles.py:6-35, 435-577),
// droplet RHS (particles.py:817-837) and its RK4 stage update
// (forced-dns-sm.py:337-362 for the droplet rows).
//
// Engine state is structure-of-arrays: positions in FP64 (pos[c*cap + p],
// unwrapped, as the reference accumulates them), velocity and mass in PT
// (vm[c*cap + p], c = vx, vy, vz, m/M0).  One thread per droplet.
#pragma once
#include "ld_kernels.hpp"
#include <type_traits>

namespace ld {

// The 4-point cosine weights of one axis: nodes i-1..i+2 with i = floor(x/dx),
// w = (1 + cos((x - X_j) pi / (2 dx))) / 4 (particles.py:12, 16, 456).  X_j is
// taken unwrapped (j*dx), which equals the reference's wrapped X[j mod N]
// through the period 4*dx of the kernel (N % 4 == 0).
template <int N> __device__ __forceinline__ int wrapIdx(int j) {
  int jj = j % N;
  return jj < 0 ? jj + N : jj;
}
// The four weights of one axis from one sincos: with
// theta = (x - i0 dx) pi / (2 dx), i0 = floor(x/dx) - 1, node i0 + a has the
// argument theta - a pi/2, so its weight is (1 + {cos, sin, -cos, -sin}(theta))/4.
// The four weights then sum to 1 in exact arithmetic.  Returns i0.
__device__ __forceinline__ int cosW4(double x, double dx, double pi_2dx, double w[4]) {
  const int i0 = (int)floor(x / dx) - 1;
  double sn, cs;
  sincos((x - (double)i0 * dx) * pi_2dx, &sn, &cs);
  w[0] = 0.25 * (1.0 + cs);
  w[1] = 0.25 * (1.0 + sn);
  w[2] = 0.25 * (1.0 - cs);
  w[3] = 0.25 * (1.0 - sn);
  return i0;
}
template <int N>
__device__ __forceinline__ void cosWeights(double x, double dx, double pi_2dx, int idx[4], double w[4]) {
  const int i0 = cosW4(x, dx, pi_2dx, w);
#pragma unroll
  for (int a = 0; a < 4; ++a) idx[a] = wrapIdx<N>(i0 + a);
}

// sin and cos of an FP64 angle in precision G (FP32: of the angle rounded to FP32).
template <typename G> __device__ __forceinline__ void sincosGT(double th, G& s, G& c) {
  if constexpr (sizeof(G) == 4) {
    sincosf((float)th, &s, &c);
  } else {
    sincos(th, &s, &c);
  }
}

// cosW4 with the sines and cosines taken in precision G (the angle in FP64).
template <typename G>
__device__ __forceinline__ int cosW4G(double x, double dx, double pi_2dx, G w[4]) {
  const int i0 = (int)floor(x / dx) - 1;
  G sn, cs;
  sincosGT<G>((x - (double)i0 * dx) * pi_2dx, sn, cs);
  const G q = (G)0.25;
  w[0] = q * ((G)1 + cs);
  w[1] = q * ((G)1 + sn);
  w[2] = q * ((G)1 - cs);
  w[3] = q * ((G)1 - sn);
  return i0;
}

// Periodic wrap of one coordinate, x mod L with numpy's sign convention
// (particles.py:269 `x[:, :d] %= L`).
__device__ __forceinline__ double wrapL(double x, double L) {
  double r = fmod(x, L);
  if (r < 0.0) r += L;
  // a non-finite position must not become an out-of-range stencil index
  if (!(r >= 0.0 && r <= L)) r = 0.0;
  return r;
}

// The step's droplet kernels (stage gather, sort bins, brick and fallback
// deposits) wrap and bin positions with multiplications by 1/L and 1/dx
// instead of fmod and divisions; all of them use these same three functions,
// so their cells agree.  The wrapped position can differ from fmod's in the
// last bit, and a stencil origin can differ by one cell when x/dx is within
// an ulp of an integer; the cosine kernel gives the node at distance 2 dx
// zero weight, so both stencils give the same interpolant.
__device__ __forceinline__ double wrapFast(double x, double L, double invL) {
  double r = fma(-L, floor(x * invL), x);
  r = r < 0.0 ? r + L : r;
  r = r >= L ? r - L : r;
  if (!(r >= 0.0 && r < L)) r = 0.0;   // a non-finite position must not become an out-of-range index
  return r;
}
// cell of a wrapped position, clamped to [0, N)
template <int N> __device__ __forceinline__ int cellOf(double xw, double inv_dx) {
  const int c = (int)(xw * inv_dx);
  return c < 0 ? 0 : (c > N - 1 ? N - 1 : c);
}
// stencil origin i0 = floor(x / dx) - 1 (nodes i0 .. i0 + 3)
__device__ __forceinline__ int originOf(double x, double inv_dx) { return (int)floor(x * inv_dx) - 1; }

// cosW4G with the origin from originOf.
template <typename G>
__device__ __forceinline__ int cosW4F(double x, double dx, double inv_dx, double pi_2dx, G w[4]) {
  const int i0 = originOf(x, inv_dx);
  G sn, cs;
  sincosGT<G>((x - (double)i0 * dx) * pi_2dx, sn, cs);
  const G q = (G)0.25;
  w[0] = q * ((G)1 + cs);
  w[1] = q * ((G)1 + sn);
  w[2] = q * ((G)1 - cs);
  w[3] = q * ((G)1 - sn);
  return i0;
}

// Geometry of one droplet's stencil.
template <int N> struct Stencil {
  int i0;
  int iy[4], iz[4];
  double wx[4], wy[4], wz[4];
  __device__ __forceinline__ void set(const double xw[3], double dx, double pi_2dx) {
    i0 = cosW4(xw[0], dx, pi_2dx, wx);
    cosWeights<N>(xw[1], dx, pi_2dx, iy, wy);
    cosWeights<N>(xw[2], dx, pi_2dx, iz, wz);
  }
  __device__ __forceinline__ int xa(int a) const { return wrapIdx<N>(i0 + a); }
  __device__ __forceinline__ double wxa(int a) const { return wx[a]; }
};

// Gather of NC fields f[c*R + idx] (particles.py:6-20, 454-509): for each x
// node a, sum over the (y, z) nodes, then weight by w_x(a).
template <typename T, int N, int NC>
__device__ __forceinline__ void gatherN(const Stencil<N>& s, const T* __restrict__ f, size_t R, double out[NC]) {
#pragma unroll
  for (int c = 0; c < NC; ++c) out[c] = 0.0;
#pragma unroll 1
  for (int a = 0; a < 4; ++a) {
    double t[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) t[c] = 0.0;
    const size_t plane = (size_t)s.xa(a) * N;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const size_t row = (plane + s.iy[b]) * N;
#pragma unroll
      for (int cc = 0; cc < 4; ++cc) {
        const double wyz = s.wy[b] * s.wz[cc];
        const size_t o = row + s.iz[cc];
#pragma unroll
        for (int c = 0; c < NC; ++c) t[c] += (double)f[c * R + o] * wyz;
      }
    }
    const double wx = s.wxa(a);
#pragma unroll
    for (int c = 0; c < NC; ++c) out[c] += t[c] * wx;
  }
}

// Gather of the stage density with the lazy clip (nStage) applied per node.
template <typename T, int N>
__device__ __forceinline__ double gatherDensity(const Stencil<N>& s, const T* __restrict__ n, const double* __restrict__ clip) {
  double out = 0.0;
  const bool doclip = clip[2] != 0.0;
#pragma unroll 1
  for (int a = 0; a < 4; ++a) {
    double t = 0.0;
    const size_t plane = (size_t)s.xa(a) * N;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const size_t row = (plane + s.iy[b]) * N;
#pragma unroll
      for (int cc = 0; cc < 4; ++cc) {
        double v = (double)n[row + s.iz[cc]];
        if (doclip && v < 0.0) v = 0.0;
        t += v * (s.wy[b] * s.wz[cc]);
      }
    }
    out += t * s.wxa(a);
  }
  if (doclip) out = (out * clip[0]) / clip[1];
  return out;
}

// Deposit W into c with the same kernel, divided by the cell volume
// (particles.py:23-35, 529-575).  Atomic adds (the reference's prange loop
// races on the same updates).
template <typename T, int N>
__device__ __forceinline__ void deposit(const Stencil<N>& s, double W, double inv_vol, T* c) {
#pragma unroll 1
  for (int a = 0; a < 4; ++a) {
    const double wa = W * s.wxa(a);
    const size_t plane = (size_t)s.xa(a) * N;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const double wab = wa * s.wy[b];
      const size_t row = (plane + s.iy[b]) * N;
#pragma unroll
      for (int cc = 0; cc < 4; ++cc) {
        unsafeAtomicAdd(&c[row + s.iz[cc]], (T)(wab * s.wz[cc] * inv_vol));
      }
    }
  }
}

struct DropPhys {
  double L, dx, pi_2dx, inv_vol;
  double factor, growth, st_s, tau_eta, g;
  double invL, inv_dx, inv_factor;
};

// Droplet RHS at one stage coordinate (particles.py:823-835).  xw: wrapped
// position, vp: velocity, m: mass/M0.  I: the 7 interpolated components
// [u, v, n].  Returns rhs[7] and W (= mdot / M0).
__device__ __forceinline__ void dropletRhs(const DropPhys& P, const double vp[3], double m, const double I[7],
                                           double rhs[7], double& W) {
  const double st = pow(m / P.factor, 2.0 / 3.0);                          // :823
  const double d0 = vp[0] - I[3], d1 = vp[1] - I[4], d2 = vp[2] - I[5];
  const double nrm = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
  const double q = 1.0 + sqrt(P.st_s / st);
  W = P.growth * (q * q) * I[6] * nrm;                                      // :828
  rhs[0] = vp[0]; rhs[1] = vp[1]; rhs[2] = vp[2];                           // :831
  const double den = st * P.tau_eta;
  rhs[3] = (I[0] - vp[0]) / den;                                            // :832
  rhs[4] = (I[1] - vp[1]) / den;
  rhs[5] = (I[2] - vp[2]) / den - P.g;                                      // :834
  rhs[6] = W;                                                               // :835
}

// The same RHS in single precision for the FP32 form (velocities and masses
// are FP32 there); St from exp2(2/3 log2(m / factor)), divisions by one
// reciprocal.
__device__ __forceinline__ void dropletRhsF(const DropPhys& P, const float vp[3], float m, const float I[7],
                                            float rhs[7], float& W) {
  const float st = exp2f((2.0f / 3.0f) * log2f(m * (float)P.inv_factor));   // :823
  const float d0 = vp[0] - I[3], d1 = vp[1] - I[4], d2 = vp[2] - I[5];
  const float nrm = sqrtf(d0 * d0 + d1 * d1 + d2 * d2);
  const float q = 1.0f + sqrtf((float)P.st_s / st);
  W = (float)P.growth * (q * q) * I[6] * nrm;                                // :828
  rhs[0] = vp[0]; rhs[1] = vp[1]; rhs[2] = vp[2];                           // :831
  const float rden = 1.0f / (st * (float)P.tau_eta);
  rhs[3] = (I[0] - vp[0]) * rden;                                           // :832
  rhs[4] = (I[1] - vp[1]) * rden;
  rhs[5] = (I[2] - vp[2]) * rden - (float)P.g;                              // :834
  rhs[6] = W;                                                               // :835
}

// ------------------------------------------------------------ spatial sort
// Droplets are reordered by spatial bin once per step (counting sort), so
// that lanes of a wavefront gather from and deposit into the same cache
// lines.  Bins are cubes of BS = max(2, N/128) grid cells; bins are numbered
// tile-major (tiles of TS^3 bins = TC^3 cells, the deposit bricks), so the
// droplets of one tile are contiguous; tiles run along z in columns (long
// contiguous z-runs of the fields), and the columns follow a Morton (Z-order)
// curve in (x, y), so the columns a run of gather blocks covers are mostly
// neighbours and share their halo nodes in one L2.
template <int N> struct Bins {
  static constexpr int BS = (N / 128 > 2) ? N / 128 : 2;
  static constexpr int B = N / BS;
  static constexpr int NB = B * B * B;
  static constexpr int TC = (N >= 32) ? 16 : 8;     // tile side in cells
  static constexpr int TS = TC / BS;                 // tile side in bins
  static constexpr int NTt = B / TS;                 // tiles per side
  static_assert(TC % BS == 0 && B % TS == 0, "bin tiles");
  static_assert((NTt & (NTt - 1)) == 0 && NTt <= 256, "tiles per side: a power of two");
  // bits of v spread to every second position (v < 256)
  static __device__ __forceinline__ int spread2(int v) {
    int r = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) r |= ((v >> i) & 1) << (2 * i);
    return r;
  }
  // key of bin (bx, by, bz): tile columns along z, the columns in Morton order
  // of (tile x, tile y); x-major inside the tile
  static __device__ __forceinline__ int key(int bx, int by, int bz) {
    const int tile = ((spread2(bx / TS) << 1) | spread2(by / TS)) * NTt + bz / TS;
    return tile * (TS * TS * TS) + ((bx % TS) * TS + by % TS) * TS + bz % TS;
  }
};

// ------------------------------------------------------------ fused RK stage
template <typename T, typename PT> struct PStage {
  int np, cap, stage;          // stage 0..3
  double* x0; PT* v0;          // RK base (unwrapped); final stage writes the new state here
  double* sx; PT* sv;          // sump accumulator
  PT* kx; PT* kv;              // stage rhs (read: previous stage, write: this stage); dx/dt is
                               // the stage velocity, which PT holds exactly
  const V4<T>* gu; const V4<T>* gv;   // gather records {u, 0} (recU) and {v, n_stage} (recV); one array
  double4* dep;                // out: wrapped stage position and W per slot (brick deposit input)
  int* flist; int* fcount;     // out: slots whose stage bin moved > 1 bin since the sort
  int all_fallback;            // 1: slots are not bin-sorted, every slot takes the fallback deposit
  unsigned long long* wmax;    // out: largest W > 0 of the launch (bits of a positive double); zeroed before
  double ah, wh;               // a_s * h, w_s * h
  DropPhys P;
  size_t rs;                   // realization stride in bytes (batched launches)
  __device__ PStage shifted() const {
    PStage a = *this;
    a.x0 = rsel1(x0, rs); a.v0 = rsel1(v0, rs); a.sx = rsel1(sx, rs); a.sv = rsel1(sv, rs);
    a.kx = rsel1(kx, rs); a.kv = rsel1(kv, rs); a.gu = rsel1(gu, rs); a.gv = rsel1(gv, rs);
    a.dep = rsel1(dep, rs); a.flist = rsel1(flist, rs); a.fcount = rsel1(fcount, rs); a.wmax = rsel1(wmax, rs);
    return a;
  }
};

// One RK4 stage of every droplet: stage coordinate X0 + a h k_prev
// (forced-dns-sm.py:347, 352, 357), wrap (particles.py:269), gather of
// [u, v, n] (:826), growth rate and deposit (:828-829), RHS (:831-835) and
// sump accumulation (forced-dns-sm.py:342-358).
//
// Four lanes per droplet: lane c owns z-node c of the 4x4x4 stencil, so one
// wavefront load or atomic touches 16 droplets x 4 consecutive z cells (one
// 16-byte segment per droplet).  Partial sums are combined across the four
// lanes with xor shuffles.
// BAT: batched launch (blockIdx.z = realization); the single-realization
// instantiation skips the pointer offsets and their registers.
template <typename T, typename PT, int N, bool BAT>
__global__ void __launch_bounds__(256)
k_particles_stage(PStage<T, PT> A0) {
  const PStage<T, PT> A = BAT ? A0.shifted() : A0;
  // XCD-aware mapping: the 8 XCDs take consecutive blocks round-robin; give
  // each XCD one contiguous range of (spatially sorted) droplets so its L2
  // sees one region of the fields.
  const int nb = gridDim.x, per = nb / 8;
  const int lb = (blockIdx.x % 8) * per + blockIdx.x / 8;
  const int gt = lb * blockDim.x + threadIdx.x;
  const int p = gt >> 2, c = gt & 3;
  const bool live = p < A.np;
  const int pp = live ? p : 0;
  constexpr size_t R = (size_t)N * N * N;
  const int cap = A.cap;
  // Loads are grouped into two round trips (stage position; the gather with
  // velocity, mass and the sump inputs), with uniform pointer selects or
  // branches (stage 0 takes k = 0 and the sump inputs from the RK base).
  const bool s0 = A.stage == 0;
  const double ah = s0 ? 0.0 : A.ah;
  double x0v[3], xs[3], kxv[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int k = 0; k < 3; ++k) x0v[k] = A.x0[k * cap + pp];
  if (!s0) {
#pragma unroll
    for (int k = 0; k < 3; ++k) kxv[k] = (double)A.kx[k * cap + pp];
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) xs[k] = x0v[k] + ah * kxv[k];
  const double L = A.P.L, dx = A.P.dx, pi_2dx = A.P.pi_2dx, invL = A.P.invL, inv_dx = A.P.inv_dx;
  const double xw = wrapFast(xs[0], L, invL), yw = wrapFast(xs[1], L, invL), zw = wrapFast(xs[2], L, invL);
  // one sincos per axis (cosW4); node i0 + a has weight (1 +- {cos, sin})/4.
  // The angles and the sincos are FP64; the weights and the gather sums are
  // kept in the field precision GT (FP32 fields: FP32 sums of 64 terms).
  using GT = std::conditional_t<sizeof(T) == 4, float, double>;
  // deposit input for the brick pass (k_brick_deposit), written before the
  // gather so the wrapped position holds no registers across it: the wrapped
  // stage position now, W after the RHS.  Slots whose stage bin left the
  // +-1-bin window around their sorted bin go to the fallback list.
  bool far = false;
  if (live && c == 0) {
    far = A.all_fallback != 0;
    const double xb[3] = {xw, yw, zw};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      // bins exactly as k_bin_count forms them
      const int cs = cellOf<N>(xb[k], inv_dx), c0 = cellOf<N>(wrapFast(x0v[k], L, invL), inv_dx);
      const int d = ((cs / Bins<N>::BS - c0 / Bins<N>::BS) % Bins<N>::B + Bins<N>::B) % Bins<N>::B;
      far = far || (d > 1 && d < Bins<N>::B - 1);
    }
    double* dp = reinterpret_cast<double*>(A.dep + p);
    dp[0] = xw; dp[1] = yw; dp[2] = zw;
    if (far) A.flist[atomicAdd(A.fcount, 1)] = p;
  }
  const int i0 = originOf(xw, inv_dx), j0 = originOf(yw, inv_dx), k0 = originOf(zw, inv_dx);
  // the angles are formed in FP64; their sines and cosines are taken in the
  // weight precision GT (the FP32 form rounds the angle in [pi/2, pi) to FP32)
  GT xsn, xcs, ysn, ycs, zsn, zcs;
  sincosGT((xw - (double)i0 * dx) * pi_2dx, xsn, xcs);
  sincosGT((yw - (double)j0 * dx) * pi_2dx, ysn, ycs);
  sincosGT((zw - (double)k0 * dx) * pi_2dx, zsn, zcs);
  const GT q = (GT)0.25;
  const GT wy[4] = {q * ((GT)1 + ycs), q * ((GT)1 + ysn), q * ((GT)1 - ycs), q * ((GT)1 - ysn)};
  int iy[4];
#pragma unroll
  for (int b = 0; b < 4; ++b) iy[b] = wrapIdx<N>(j0 + b);
  const GT wz = q * ((GT)1 + ((c & 2) ? (GT)-1 : (GT)1) * ((c & 1) ? zsn : zcs));
  const GT gxs = xsn, gxc = xcs;
  const int iz = wrapIdx<N>(k0 + c);
  // stage velocity and mass (v0 + a h kv) and the sump inputs (x0, v0 at
  // stage 0, else the accumulators), issued before the gather so that their
  // round trip overlaps its loads (measured 0.320 -> 0.306 ms per launch at
  // 256^3 with 10^6 droplets, though occupancy drops from 5 to 4 waves/SIMD)
  const PT* kvp = s0 ? A.v0 : A.kv;
  const double* sxp = s0 ? A.x0 : A.sx;
  const PT* svp = s0 ? A.v0 : A.sv;
  PT v0v[4], kvv[4], prv[4];
  double prx[3];
#pragma unroll
  for (int k = 0; k < 4; ++k) { v0v[k] = A.v0[k * cap + pp]; kvv[k] = kvp[k * cap + pp]; prv[k] = svp[k * cap + pp]; }
#pragma unroll
  for (int k = 0; k < 3; ++k) prx[k] = sxp[k * cap + pp];
  // gather: this lane's z node, all (x, y) nodes; two 16-byte records per node
  GT Ig[7];
#pragma unroll
  for (int k = 0; k < 7; ++k) Ig[k] = (GT)0;
  // FP32 fields: all 32 record loads are written ahead of the sums, so the
  // compiler keeps about a dozen in flight per lane (at 5 waves per SIMD);
  // this measured faster than a plane-by-plane loop at 8 waves per SIMD.
  // FP64 records are twice the size; they keep the loop.
  if (sizeof(T) == 4 && live) {
    V4<T> U[16], V[16];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const uint32_t plane = (uint32_t)wrapIdx<N>(i0 + a) * N;
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const uint32_t o = (plane + (uint32_t)iy[b]) * N + (uint32_t)iz;
        U[a * 4 + b] = ldRec<T, N>(A.gu, recU(o));
        V[a * 4 + b] = ldRec<T, N>(A.gv, recV(o));
      }
    }
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const GT wx = (GT)0.25 * ((GT)1 + ((a & 2) ? (GT)-1 : (GT)1) * ((a & 1) ? gxs : gxc));
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const GT w = wy[b] * wx;
        const V4<T>& u = U[a * 4 + b];
        const V4<T>& v = V[a * 4 + b];
        Ig[0] += (GT)u.x * w;
        Ig[1] += (GT)u.y * w;
        Ig[2] += (GT)u.z * w;
        Ig[3] += (GT)v.x * w;
        Ig[4] += (GT)v.y * w;
        Ig[5] += (GT)v.z * w;
        Ig[6] += (GT)v.w * w;
      }
    }
  } else if (live) {
#pragma unroll 1
    for (int a = 0; a < 4; ++a) {
      const GT wx = (GT)0.25 * ((GT)1 + ((a & 2) ? (GT)-1 : (GT)1) * ((a & 1) ? gxs : gxc));
      const uint32_t plane = (uint32_t)wrapIdx<N>(i0 + a) * N;
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const uint32_t o = (plane + (uint32_t)iy[b]) * N + (uint32_t)iz;   // < N^3 <= 2^27
        const GT w = wy[b] * wx;
        const V4<T> U = ldRec<T, N>(A.gu, recU(o));
        const V4<T> V = ldRec<T, N>(A.gv, recV(o));
        Ig[0] += (GT)U.x * w;
        Ig[1] += (GT)U.y * w;
        Ig[2] += (GT)U.z * w;
        Ig[3] += (GT)V.x * w;
        Ig[4] += (GT)V.y * w;
        Ig[5] += (GT)V.z * w;
        Ig[6] += (GT)V.w * w;
      }
    }
  }
  double I[7];
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    GT v = Ig[k] * wz;
    v += __shfl_xor(v, 1, 64);
    v += __shfl_xor(v, 2, 64);
    I[k] = (double)v;
  }
  double rhs[7], W = 0.0;
  if constexpr (sizeof(PT) == 4) {
    float vpf[3], If[7], rf[7], Wf = 0.0f;
#pragma unroll
    for (int k = 0; k < 3; ++k) vpf[k] = v0v[k] + (PT)ah * kvv[k];
    const float mf = v0v[3] + (PT)ah * kvv[3];
#pragma unroll
    for (int k = 0; k < 7; ++k) If[k] = (float)I[k];
    if (live) dropletRhsF(A.P, vpf, mf, If, rf, Wf);
#pragma unroll
    for (int k = 0; k < 7; ++k) rhs[k] = (double)rf[k];
    W = (double)Wf;
  } else {
    double vp[3], m;
#pragma unroll
    for (int k = 0; k < 3; ++k) vp[k] = (double)(PT)(v0v[k] + (PT)ah * kvv[k]);
    m = (double)(PT)(v0v[3] + (PT)ah * kvv[3]);
    if (live) dropletRhs(A.P, vp, m, I, rhs, W);
  }
  // largest growth rate of the launch: the bound of the deposits' fixed-point
  // scale.  A NaN or negative W (not a growth) makes it infinite, so the
  // deposit writes NaN everywhere, as the reference's would spread it.
  {
    __shared__ double wm[4];
    double v = !live ? 0.0 : (W > 0.0 ? W : (W == 0.0 ? 0.0 : INFINITY));
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
      const double b = fmax(fmax(wm[0], wm[1]), fmax(wm[2], wm[3]));
      if (b > 0.0) atomicMax(A.wmax, (unsigned long long)__double_as_longlong(b));
    }
  }
  if (!live || c != 0) return;
  // a fallback slot stores -W so the brick pass (which takes W > 0) skips it
  reinterpret_cast<double*>(A.dep + p)[3] = far ? -W : W;
  // sump (forced-dns-sm.py:337, 342-358); the last stage writes the new state
  double* dxp = A.stage == 3 ? A.x0 : A.sx;
  PT* dvp = A.stage == 3 ? A.v0 : A.sv;
#pragma unroll
  for (int k = 0; k < 3; ++k) dxp[k * cap + p] = prx[k] + A.wh * rhs[k];
#pragma unroll
  for (int k = 0; k < 4; ++k) dvp[k * cap + p] = prv[k] + (PT)A.wh * (PT)rhs[3 + k];
  if (A.stage < 3) {
#pragma unroll
    for (int k = 0; k < 3; ++k) A.kx[k * cap + p] = (PT)rhs[k];
#pragma unroll
    for (int k = 0; k < 4; ++k) A.kv[k * cap + p] = (PT)rhs[3 + k];
  }
}

// ------------------------------------------------------------ brick deposit
// Deposit of W with the cosine kernel / dx^3 (particles.py:23-35, 529-575),
// organised by output: one workgroup per brick of S^3 grid nodes accumulates
// in a 64-bit fixed-point LDS tile and writes each node of the brick exactly
// once.  Integer sums do not depend on the order of the additions, so the
// deposit is bitwise reproducible, and ds_add_u64 is the fastest LDS atomic
// wide enough (gfx950: u64 15, f64 27, f32 196 cycles per wavefront).  The
// scale 2^e is the largest power of two with  total * Wmax / dx^3 / 8 * 2^e
// < 2^62, 2^52 for FP32 fields  (total = candidates of the brick, Wmax =
// largest W of the launch, 1/8 = largest product of three weights), so no sum
// can overflow; one unit is ~2^-62 (2^-52) of that bound.
//
// The droplets that can reach the brick are those whose stage cell lies in
// [c0 - 2, c0 + S] per axis; their sorted bins (bins of the position at the
// last sort) lie in that cell range's bins widened by one bin of motion
// (slots that moved further are on the fallback list and skipped here).  For
// each (bin x, bin y) of that window the z-run of bins is one contiguous range
// of sorted slots per bin tile it crosses, so the candidates form at most
// NZS * NBX^2 ranges.  Per round, CH candidates are tested for overlap with the
// brick; those that overlap are compacted into an LDS list, and each listed
// droplet is then deposited by one lane.
template <int N> struct Brick {
  static constexpr int S = Bins<N>::TC;              // nodes per side (one bin tile)
  static constexpr int NB = N / S;                   // bricks per side
  // threads: 512 measured against 256 and 1024 (deposit 0.143 / 0.130 / 0.335 ms per
  // stage at 256^3 with 10^6 droplets; 64^3 ensemble step 10.99 / 10.67 / 10.68 ms)
  static constexpr int NT = 512;
  // bins spanned by S + 3 stage cells from a bin-aligned start, +-1 bin of motion
  static constexpr int W0 = (S + 2 + Bins<N>::BS - 1) / Bins<N>::BS + 1 + 2;
  static constexpr int NBX = W0 < Bins<N>::B ? W0 : Bins<N>::B;
  // z-runs of NBX bins split at tile boundaries into at most NZS key ranges
  static constexpr int NZS = (NBX + Bins<N>::TS - 1) / Bins<N>::TS + 1;
  static constexpr int NR = NZS * NBX * NBX;         // slot ranges
  static constexpr int CH = 2 * NT;                  // candidates tested per round (LDS fits 4 blocks per CU)
  static_assert(N % S == 0 && (NB * NB * NB) % 8 == 0, "brick grid");
};

template <typename T, int N>
__global__ void __launch_bounds__(Brick<N>::NT)
k_brick_deposit(const double4* __restrict__ dep0, const int* __restrict__ start0, int np,
                const unsigned long long* __restrict__ wmax0, T* __restrict__ fc0, DropPhys P, size_t rs) {
  const double4* __restrict__ dep = rsel1(dep0, rs);
  const int* __restrict__ start = rsel1(start0, rs);
  const unsigned long long* __restrict__ wmax = rsel1(wmax0, rs);
  T* __restrict__ fc = rsel1(fc0, rs);
  using Bk = Brick<N>;
  constexpr int S = Bk::S, NBr = Bk::NB, NBX = Bk::NBX, NR = Bk::NR, NT = Bk::NT, NZS = Bk::NZS, CH = Bk::CH;
  constexpr int TS = Bins<N>::TS;
  constexpr int BS = Bins<N>::BS, B = Bins<N>::B;
  constexpr int NBIN = B * B * B;
  __shared__ unsigned long long tile[S * S * S];
  __shared__ int rb[NR], rn[NR], rl[NR + 1];
  __shared__ int list[CH];
  __shared__ int cnt;
  const int tid = threadIdx.x, lane = tid & 63;
  // XCD-aware: each of the 8 XCDs takes a contiguous run of bricks
  const int nb = gridDim.x, per = nb / 8;
  const int br = (blockIdx.x % 8) * per + blockIdx.x / 8;
  const int c0x = (br / (NBr * NBr)) * S, c0y = ((br / NBr) % NBr) * S, c0z = (br % NBr) * S;
  for (int i = tid; i < S * S * S; i += NT) tile[i] = 0ull;
  if (tid == 0) cnt = 0;
  // slot ranges of the window's z-runs
  const int lo = (NBX == B) ? 0 : 1;
  const int blx = ((c0x - 2 + N) % N) / BS - lo, bly = ((c0y - 2 + N) % N) / BS - lo;
  const int z0 = (NBX == B) ? 0 : (((c0z - 2 + N) % N) / BS - lo + B) % B;
  for (int r = tid; r < NBX * NBX; r += NT) {
    const int bx = (blx + r / NBX + B) % B, by = (bly + r % NBX + B) % B;
    int bz = z0, rem = NBX;
#pragma unroll
    for (int sgm = 0; sgm < NZS; ++sgm) {
      int b0 = 0, len = 0;
      if (rem > 0) {
        const int l = bz % TS;
        const int nz = (TS - l < rem) ? TS - l : rem;   // bins of this run inside one tile
        const int k0 = Bins<N>::key(bx, by, bz);        // keys k0 .. k0 + nz - 1 are consecutive
        b0 = start[k0];
        len = ((k0 + nz < NBIN) ? start[k0 + nz] : np) - b0;
        bz = (bz + nz) % B;
        rem -= nz;
      }
      rb[NZS * r + sgm] = b0; rn[NZS * r + sgm] = len;
    }
  }
  __syncthreads();
  // exclusive prefix of the range lengths (wavefront 0)
  if (tid < 64) {
    constexpr int PER = (NR + 63) / 64;
    int loc[PER];
    int sum = 0;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int idx = tid * PER + k;
      loc[k] = sum;
      sum += (idx < NR) ? rn[idx] : 0;
    }
    int inc = sum;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int t = __shfl_up(inc, off, 64);
      if (tid >= off) inc += t;
    }
    const int ex = inc - sum;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int idx = tid * PER + k;
      if (idx < NR) rl[idx] = ex + loc[k];
    }
    if (tid == 63) rl[NR] = inc;
  }
  __syncthreads();
  const int total = rl[NR];
  const double dx = P.dx, pi_2dx = P.pi_2dx, inv_dx = P.inv_dx;
  const double bound = (double)total * __longlong_as_double((long long)*wmax) * P.inv_vol * 0.125;
  const bool finite = isfinite(bound);
  // FP32 fields: sums below 2^52, so a value converts to its integer by one
  // addition of 2^52 (round to nearest even, as llrint); FP64 fields keep the
  // 2^62 range and llrint
  constexpr bool F32 = sizeof(T) == 4;
  constexpr int EMAX = F32 ? 51 : 61;
  const int e = (bound > 0.0 && finite) ? EMAX - ilogb(bound) : 0;
  const double scale = ldexp(1.0, e), unit = ldexp(1.0, -e);
  for (int base = 0; base < total && finite; base += CH) {
    // test and compact: growing candidates whose stencil meets the brick
#pragma unroll 1
    for (int k = 0; k < CH / NT; ++k) {
      const int f = base + k * NT + tid;
      bool take = false;
      int t = 0;
      if (f < total) {
        int a = 0, b = NR;                                  // rl[a] <= f < rl[b]
        while (b - a > 1) {
          const int mid = (a + b) >> 1;
          if (rl[mid] <= f) a = mid; else b = mid;
        }
        t = rb[a] + (f - rl[a]);
        const double4 d = dep[t];
        if (d.w > 0.0) {
          const int rx = ((originOf(d.x, inv_dx) - c0x) % N + N) % N;
          const int ry = ((originOf(d.y, inv_dx) - c0y) % N + N) % N;
          const int rz = ((originOf(d.z, inv_dx) - c0z) % N + N) % N;
          take = (rx < S || rx > N - 4) && (ry < S || ry > N - 4) && (rz < S || rz > N - 4);
        }
      }
      const unsigned long long m = __ballot(take);
      const int nw = __popcll(m);
      int wbase = 0;
      if (lane == 0 && nw) wbase = atomicAdd(&cnt, nw);
      wbase = __shfl(wbase, 0, 64);
      if (take)
        list[wbase + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u))] = t;
    }
    __syncthreads();
    const int n = cnt;
    for (int i = tid; i < n; i += NT) {
      const double4 d = dep[list[i]];
      // stencil weights in the field precision (FP32 fields: FP32 sines and cosines)
      using GT = std::conditional_t<sizeof(T) == 4, float, double>;
      GT wx[4], wy[4], wz[4];
      const int i0 = cosW4F<GT>(d.x, dx, inv_dx, pi_2dx, wx), j0 = cosW4F<GT>(d.y, dx, inv_dx, pi_2dx, wy),
                k0 = cosW4F<GT>(d.z, dx, inv_dx, pi_2dx, wz);
      const int rx = ((i0 - c0x) % N + N) % N, ry = ((j0 - c0y) % N + N) % N, rz = ((k0 - c0z) % N + N) % N;
      // scale is a power of two, so W * scale is exact and the products below
      // round as (W wx wy wz) scale would
      const double W = d.w * P.inv_vol * scale;
      const double wzd[4] = {(double)wz[0], (double)wz[1], (double)wz[2], (double)wz[3]};
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        const int nx = (rx + a) % N;
        if (nx >= S) continue;
        const double wa = W * (double)wx[a];
#pragma unroll
        for (int bb = 0; bb < 4; ++bb) {
          const int ny = (ry + bb) % N;
          if (ny >= S) continue;
          const double wab = wa * (double)wy[bb];
#pragma unroll
          for (int cc = 0; cc < 4; ++cc) {
            const int nz = (rz + cc) % N;
            if (nz >= S) continue;
            const double v = wab * wzd[cc];
            unsigned long long q;
            if constexpr (F32)
              q = (unsigned long long)(__double_as_longlong(v + 0x1p52) - __double_as_longlong(0x1p52));
            else
              q = (unsigned long long)llrint(v);
            atomicAdd(&tile[(nx * S + ny) * S + nz], q);
          }
        }
      }
    }
    __syncthreads();
    if (tid == 0) cnt = 0;
    __syncthreads();
  }
  for (int i = tid; i < S * S * S; i += NT) {
    const int nx = i / (S * S), ny = (i / S) % S, nz = i % S;
    fc[((size_t)(c0x + nx) * N + (c0y + ny)) * N + (c0z + nz)] = finite ? (T)((double)tile[i] * unit) : (T)NAN;
  }
}

// Fallback deposit of the slots that moved more than one bin since the sort
// (or of every slot when the slots are not sorted); runs after
// k_brick_deposit.  Deterministic like the bricks: k_fallback_accum adds each
// slot's 64 node values as 64-bit fixed-point integers into the scratch field
// fx (zero between steps; integer sums do not depend on the order of the
// additions), with the scale 2^e of the largest value any node can reach
// (slots x Wmax / dx^3 / 8); k_fallback_fold then moves each touched node's
// sum into fc and clears it (the first slot to exchange a node takes its whole
// sum; the others find zero).
struct FallbackScale {
  double scale, unit;
  bool finite;
  __device__ FallbackScale(int nf, unsigned long long wmax_bits, double inv_vol) {
    const double bound = (double)nf * __longlong_as_double((long long)wmax_bits) * inv_vol * 0.125;
    finite = isfinite(bound);
    const int e = (bound > 0.0 && finite) ? 61 - ilogb(bound) : 0;
    scale = ldexp(1.0, e);
    unit = ldexp(1.0, -e);
  }
};

template <typename T, int N>
__global__ void __launch_bounds__(256)
k_fallback_accum(const double4* __restrict__ dep0, const int* __restrict__ flist0, const int* __restrict__ fcount0,
                 const unsigned long long* __restrict__ wmax0, unsigned long long* fx0, DropPhys P, size_t rs) {
  const double4* __restrict__ dep = rsel1(dep0, rs);
  const int* __restrict__ flist = rsel1(flist0, rs);
  unsigned long long* fx = rsel1(fx0, rs);
  const int nf = *rsel1(fcount0, rs);
  if (nf == 0) return;
  const FallbackScale sc(nf, *rsel1(wmax0, rs), P.inv_vol);
  if (!sc.finite) return;
  using GT = std::conditional_t<sizeof(T) == 4, float, double>;
  for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < nf; f += gridDim.x * blockDim.x) {
    const double4 d = dep[flist[f]];
    const double w = -d.w;                       // fallback slots carry -W
    if (!(w > 0.0)) continue;
    GT wx[4], wy[4], wz[4];
    const int i0 = cosW4F<GT>(d.x, P.dx, P.inv_dx, P.pi_2dx, wx), j0 = cosW4F<GT>(d.y, P.dx, P.inv_dx, P.pi_2dx, wy),
              k0 = cosW4F<GT>(d.z, P.dx, P.inv_dx, P.pi_2dx, wz);
    const double W = w * P.inv_vol * sc.scale;
#pragma unroll 1
    for (int a = 0; a < 4; ++a) {
      const double wa = W * (double)wx[a];
      const size_t plane = (size_t)wrapIdx<N>(i0 + a) * N;
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const double wab = wa * (double)wy[b];
        const size_t row = (plane + wrapIdx<N>(j0 + b)) * N;
#pragma unroll
        for (int c = 0; c < 4; ++c)
          atomicAdd(&fx[row + wrapIdx<N>(k0 + c)], (unsigned long long)llrint(wab * (double)wz[c]));
      }
    }
  }
}

template <typename T, int N>
__global__ void __launch_bounds__(256)
k_fallback_fold(const double4* __restrict__ dep0, const int* __restrict__ flist0, const int* __restrict__ fcount0,
                const unsigned long long* __restrict__ wmax0, unsigned long long* fx0, T* fc0, DropPhys P, size_t rs) {
  const double4* __restrict__ dep = rsel1(dep0, rs);
  const int* __restrict__ flist = rsel1(flist0, rs);
  unsigned long long* fx = rsel1(fx0, rs);
  T* fc = rsel1(fc0, rs);
  const int nf = *rsel1(fcount0, rs);
  if (nf == 0) return;
  const FallbackScale sc(nf, *rsel1(wmax0, rs), P.inv_vol);
  for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < nf; f += gridDim.x * blockDim.x) {
    const double4 d = dep[flist[f]];
    const int i0 = originOf(d.x, P.inv_dx), j0 = originOf(d.y, P.inv_dx), k0 = originOf(d.z, P.inv_dx);
    // a NaN or negative growth somewhere: NaN at every node the slot reaches
    const bool poison = !sc.finite;
#pragma unroll 1
    for (int a = 0; a < 4; ++a) {
      const size_t plane = (size_t)wrapIdx<N>(i0 + a) * N;
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const size_t row = (plane + wrapIdx<N>(j0 + b)) * N;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const size_t o = row + wrapIdx<N>(k0 + c);
          if (poison) { fc[o] = (T)NAN; continue; }
          const unsigned long long v = atomicExch(&fx[o], 0ull);
          if (v) fc[o] = fc[o] + (T)((double)v * sc.unit);
        }
      }
    }
  }
}

// ------------------------------------------------------------ API kernels
// interp_cosine (particles.py:435-511) of NC fields at rows of coord (Np, 7).
template <typename T, int N>
__global__ void __launch_bounds__(256)
k_interp(const double* __restrict__ coord, int np, const T* __restrict__ f, int ncomp, double* __restrict__ out,
         DropPhys P) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  constexpr size_t R = (size_t)N * N * N;
  double xw[3];
  for (int c = 0; c < 3; ++c) xw[c] = wrapL(coord[(size_t)p * 7 + c], P.L);
  Stencil<N> s;
  s.set(xw, P.dx, P.pi_2dx);
  for (int c = 0; c < ncomp; ++c) {
    double v;
    gatherN<T, N, 1>(s, f + (size_t)c * R, R, &v);
    out[(size_t)p * ncomp + c] = v;
  }
}

// exterp_cosine_scalar (particles.py:514-577): deposit w[p] into c (c must be zeroed).
template <typename T, int N>
__global__ void __launch_bounds__(256)
k_exterp(const double* __restrict__ coord, int np, const double* __restrict__ w, T* c, DropPhys P) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  double xw[3];
  for (int k = 0; k < 3; ++k) xw[k] = wrapL(coord[(size_t)p * 7 + k], P.L);
  Stencil<N> s;
  s.set(xw, P.dx, P.pi_2dx);
  deposit<T, N>(s, w[p], P.inv_vol, c);
}

// pRHS (particles.py:817-837) on rows of a stage coordinate (Np, 7): writes
// rhs (Np, 7), interpmat (Np, 7), exterpmat (Np) and deposits into fc (zeroed).
// n is used as given (clip = nullptr) or through the lazy clip.
template <typename T, int N>
__global__ void __launch_bounds__(256)
k_prhs(const double* __restrict__ coord, int np, const T* __restrict__ u, const T* __restrict__ us,
       const T* __restrict__ n, const double* __restrict__ clip, T* fc, double* __restrict__ rhs,
       double* __restrict__ interp, double* __restrict__ exterp, DropPhys P) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  constexpr size_t R = (size_t)N * N * N;
  const double* row = coord + (size_t)p * 7;
  double xw[3], vp[3];
  for (int c = 0; c < 3; ++c) { xw[c] = wrapL(row[c], P.L); vp[c] = row[3 + c]; }
  Stencil<N> s;
  s.set(xw, P.dx, P.pi_2dx);
  double I[7];
  gatherN<T, N, 3>(s, u, R, I);
  gatherN<T, N, 3>(s, us, R, I + 3);
  I[6] = gatherDensity<T, N>(s, n, clip);
  double r[7], W;
  dropletRhs(P, vp, row[6], I, r, W);
  deposit<T, N>(s, W, P.inv_vol, fc);
  for (int c = 0; c < 7; ++c) { rhs[(size_t)p * 7 + c] = r[c]; interp[(size_t)p * 7 + c] = I[c]; }
  exterp[p] = W;
}

// AoS (Np, 7) rows <-> engine SoA state.
template <typename PT>
__global__ void k_rows_to_soa(const double* __restrict__ rows, int np, int cap, double* pos, PT* vm) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  for (int c = 0; c < 3; ++c) pos[c * cap + p] = rows[(size_t)p * 7 + c];
  for (int c = 0; c < 4; ++c) vm[c * cap + p] = (PT)rows[(size_t)p * 7 + 3 + c];
}
// rid[p]: the input row of engine slot p (slots are permuted by the spatial sort).
template <typename PT>
__global__ void k_soa_to_rows(const double* __restrict__ pos, const PT* __restrict__ vm, const int64_t* __restrict__ rid,
                              int np, int cap, double* rows) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const size_t r = (size_t)rid[p];
  for (int c = 0; c < 3; ++c) rows[r * 7 + c] = pos[c * cap + p];
  for (int c = 0; c < 4; ++c) rows[r * 7 + 3 + c] = (double)vm[c * cap + p];
}

// Stage coordinate rows X + a*kp (forced-dns-sm.py:347) for the API path.
static __global__ void k_axpy_rows(const double* __restrict__ x, const double* __restrict__ k, double a, int n, double* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = x[i] + a * k[i];
}

template <int N>
__global__ void __launch_bounds__(256)
k_bin_count(const double* __restrict__ pos0, int np, int cap, double L, double invL, double inv_dx, int* __restrict__ key0,
            int* __restrict__ rank0, int* __restrict__ hist0, size_t rs) {
  const double* __restrict__ pos = rsel1(pos0, rs);
  int* __restrict__ key = rsel1(key0, rs);
  int* __restrict__ rank = rsel1(rank0, rs);
  int* __restrict__ hist = rsel1(hist0, rs);
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const bool valid = p < np;
  int k = -1 - lane;                        // invalid lanes: keys no other lane has
  if (valid) {
    int b[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) b[c] = cellOf<N>(wrapFast(pos[c * cap + p], L, invL), inv_dx) / Bins<N>::BS;
    k = Bins<N>::key(b[0], b[1], b[2]);
    key[p] = k;
  }
  // The slots are in last step's bin order, so lanes with one key mostly form
  // runs: the first lane of each run of equal keys adds the run's length to
  // the bin count and the run takes consecutive ranks from the old count
  // (one atomic per run instead of one per droplet).
  const int kprev = __shfl_up(k, 1, 64);     // every lane takes part in the shuffle
  const bool head = lane == 0 || kprev != k;
  const unsigned long long hm = __ballot(head);
  const unsigned long long upto = (2ull << lane) - 1ull;         // bits 0 .. lane
  const int start = 63 - __clzll(hm & upto);
  const unsigned long long after = hm & ~upto;
  const int next = after ? __ffsll((long long)after) - 1 : 64;
  int base = 0;
  if (head && valid) base = atomicAdd(&hist[k], next - lane);
  base = __shfl(base, start, 64);
  if (valid) rank[p] = base + (lane - start);
}

// Exclusive scan of 4096 ints per block (1024 threads x 4); block totals to bsum.
__device__ __forceinline__ int blockExclusiveScan1024(int v, int* sh, int& total) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  if (lane == 63) sh[wid] = x;
  __syncthreads();
  if (wid == 0) {
    int w = lane < 16 ? sh[lane] : 0;
#pragma unroll
    for (int d = 1; d < 16; d <<= 1) {
      const int y = __shfl_up(w, d, 64);
      if (lane >= d) w += y;
    }
    if (lane < 16) sh[16 + lane] = w;   // inclusive wave prefix
  }
  __syncthreads();
  const int wpre = wid > 0 ? sh[16 + wid - 1] : 0;
  total = sh[16 + 15];
  __syncthreads();
  return wpre + x - v;
}

static __global__ void __launch_bounds__(1024)
k_scan_local(int* __restrict__ data0, int n, int* __restrict__ bsum0, size_t rs) {
  int* __restrict__ data = rsel1(data0, rs);
  int* __restrict__ bsum = rsel1(bsum0, rs);
  __shared__ int sh[32];
  const int base = blockIdx.x * 4096 + threadIdx.x * 4;
  int v[4], s = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) { v[i] = (base + i < n) ? data[base + i] : 0; s += v[i]; }
  int total;
  int pre = blockExclusiveScan1024(s, sh, total);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    if (base + i < n) data[base + i] = pre;
    pre += v[i];
  }
  if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

static __global__ void __launch_bounds__(1024)
k_scan_bsums(int* __restrict__ bsum0, int nb, size_t rs) {
  int* __restrict__ bsum = rsel1(bsum0, rs);
  __shared__ int sh[32];
  const int base = threadIdx.x * 4;
  int v[4], s = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) { v[i] = (base + i < nb) ? bsum[base + i] : 0; s += v[i]; }
  int total;
  int pre = blockExclusiveScan1024(s, sh, total);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    if (base + i < nb) bsum[base + i] = pre;
    pre += v[i];
  }
}

static __global__ void __launch_bounds__(256)
k_scan_add(int* __restrict__ data0, int n, const int* __restrict__ bsum0, size_t rs) {
  int* __restrict__ data = rsel1(data0, rs);
  const int* __restrict__ bsum = rsel1(bsum0, rs);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) data[i] += bsum[i / 4096];
}

// Scatter the droplet state to its sorted slot.
template <typename PT>
__global__ void __launch_bounds__(256)
k_bin_scatter(int np, int cap, const int* __restrict__ key0, const int* __restrict__ rank0,
              const int* __restrict__ start0, const double* __restrict__ pos0, const PT* __restrict__ vm0,
              const int64_t* __restrict__ id0, double* __restrict__ pos20, PT* __restrict__ vm20,
              int64_t* __restrict__ id20, size_t rs) {
  const int* __restrict__ key = rsel1(key0, rs);
  const int* __restrict__ rank = rsel1(rank0, rs);
  const int* __restrict__ start = rsel1(start0, rs);
  const double* __restrict__ pos = rsel1(pos0, rs);
  const PT* __restrict__ vm = rsel1(vm0, rs);
  const int64_t* __restrict__ id = rsel1(id0, rs);
  double* __restrict__ pos2 = rsel1(pos20, rs);
  PT* __restrict__ vm2 = rsel1(vm20, rs);
  int64_t* __restrict__ id2 = rsel1(id20, rs);
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const int d = start[key[p]] + rank[p];
#pragma unroll
  for (int c = 0; c < 3; ++c) pos2[c * cap + d] = pos[c * cap + p];
#pragma unroll
  for (int c = 0; c < 4; ++c) vm2[c * cap + d] = vm[c * cap + p];
  id2[d] = id[p];
}

template <typename X>
__global__ void k_copy(const X* __restrict__ a, X* __restrict__ b, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) b[i] = a[i];
}
static __global__ void k_iota64(int64_t* __restrict__ id, int np) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < np) id[p] = p;
}

}  // namespace ld
