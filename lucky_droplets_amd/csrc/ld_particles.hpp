// Droplet kernels: cosine-kernel gather/deposit (particles.py:6-35, 435-577),
// droplet RHS (particles.py:817-837) and its RK4 stage update
// (forced-dns-sm.py:337-362 for the droplet rows).
//
// Engine state is structure-of-arrays: positions in FP64 (pos[c*cap + p],
// unwrapped, as the reference accumulates them), velocity and mass in PT
// (vm[c*cap + p], c = vx, vy, vz, m/M0).  One thread per droplet.
#pragma once
#include "ld_kernels.hpp"

namespace ld {

// The 4-point cosine weights of one axis: nodes i-1..i+2 with i = floor(x/dx),
// w = (1 + cos((x - X_j) pi / (2 dx))) / 4 (particles.py:12, 16, 456).  X_j is
// taken unwrapped (j*dx), which equals the reference's wrapped X[j mod N]
// through the period 4*dx of the kernel (N % 4 == 0).
__device__ __forceinline__ double cosW(double x, int j, double dx, double pi_2dx) {
  return 0.25 * (1.0 + cos((x - (double)j * dx) * pi_2dx));
}
template <int N> __device__ __forceinline__ int wrapIdx(int j) {
  int jj = j % N;
  return jj < 0 ? jj + N : jj;
}
template <int N>
__device__ __forceinline__ void cosWeights(double x, double dx, double pi_2dx, int idx[4], double w[4]) {
  const int i = (int)floor(x / dx);
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    w[a] = cosW(x, i - 1 + a, dx, pi_2dx);
    idx[a] = wrapIdx<N>(i - 1 + a);
  }
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

// Geometry of one droplet's stencil.  The x weights are evaluated inside the
// x loop (xa(), wxa()) so the loop need not be unrolled.
template <int N> struct Stencil {
  double x, dx, pi_2dx;
  int i0;
  int iy[4], iz[4];
  double wy[4], wz[4];
  __device__ __forceinline__ void set(const double xw[3], double dx_, double pi_2dx_) {
    x = xw[0]; dx = dx_; pi_2dx = pi_2dx_;
    i0 = (int)floor(x / dx) - 1;
    cosWeights<N>(xw[1], dx, pi_2dx, iy, wy);
    cosWeights<N>(xw[2], dx, pi_2dx, iz, wz);
  }
  __device__ __forceinline__ int xa(int a) const { return wrapIdx<N>(i0 + a); }
  __device__ __forceinline__ double wxa(int a) const { return cosW(x, i0 + a, dx, pi_2dx); }
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

// ------------------------------------------------------------ spatial sort
// Droplets are reordered by spatial bin once per step (counting sort), so
// that lanes of a wavefront gather from and deposit into the same cache
// lines.  Bins are cubes of BS = max(2, N/128) grid cells.
template <int N> struct Bins {
  static constexpr int BS = (N / 128 > 2) ? N / 128 : 2;
  static constexpr int B = N / BS;
  static constexpr int NB = B * B * B;
};

// ------------------------------------------------------------ fused RK stage
template <typename T, typename PT> struct PStage {
  int np, cap, stage;          // stage 0..3
  double* x0; PT* v0;          // RK base (unwrapped); final stage writes the new state here
  double* sx; PT* sv;          // sump accumulator
  double* kx; PT* kv;          // stage rhs (read: previous stage, write: this stage)
  const V4<T>* gu; const V4<T>* gv;   // {u, 0} and {v, n_stage} per node
  double4* dep;                // out: wrapped stage position and W per slot (brick deposit input)
  int* flist; int* fcount;     // out: slots whose stage bin moved > 1 bin since the sort
  int all_fallback;            // 1: slots are not bin-sorted, every slot takes the fallback deposit
  double ah, wh;               // a_s * h, w_s * h
  DropPhys P;
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
template <typename T, typename PT, int N>
__global__ void __launch_bounds__(256)
k_particles_stage(PStage<T, PT> A) {
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
  double xs[3], vp[3], m;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double x = A.x0[k * cap + pp];
    double vv = (double)A.v0[k * cap + pp];
    if (A.stage > 0) {
      x = x + A.ah * A.kx[k * cap + pp];
      vv = (double)(PT)(A.v0[k * cap + pp] + (PT)A.ah * A.kv[k * cap + pp]);
    }
    xs[k] = x; vp[k] = vv;
  }
  m = (double)A.v0[3 * cap + pp];
  if (A.stage > 0) m = (double)(PT)(A.v0[3 * cap + pp] + (PT)A.ah * A.kv[3 * cap + pp]);
  const double L = A.P.L, dx = A.P.dx, pi_2dx = A.P.pi_2dx;
  const double xw = wrapL(xs[0], L), yw = wrapL(xs[1], L), zw = wrapL(xs[2], L);
  const int i0 = (int)floor(xw / dx) - 1, j0 = (int)floor(yw / dx) - 1, k0 = (int)floor(zw / dx) - 1;
  double wy[4];
  int iy[4];
#pragma unroll
  for (int b = 0; b < 4; ++b) { wy[b] = cosW(yw, j0 + b, dx, pi_2dx); iy[b] = wrapIdx<N>(j0 + b); }
  const double wz = cosW(zw, k0 + c, dx, pi_2dx);
  const int iz = wrapIdx<N>(k0 + c);
  // gather: this lane's z node, all (x, y) nodes; two 16-byte records per node
  double I[7];
#pragma unroll
  for (int k = 0; k < 7; ++k) I[k] = 0.0;
  if (live) {
#pragma unroll 1
    for (int a = 0; a < 4; ++a) {
      const double wx = cosW(xw, i0 + a, dx, pi_2dx);
      const size_t plane = (size_t)wrapIdx<N>(i0 + a) * N;
      double t[7];
#pragma unroll
      for (int k = 0; k < 7; ++k) t[k] = 0.0;
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const size_t o = (plane + iy[b]) * N + iz;
        const double w = wy[b];
        const V4<T> U = A.gu[o];
        const V4<T> V = A.gv[o];
        t[0] += (double)U.x * w;
        t[1] += (double)U.y * w;
        t[2] += (double)U.z * w;
        t[3] += (double)V.x * w;
        t[4] += (double)V.y * w;
        t[5] += (double)V.z * w;
        t[6] += (double)V.w * w;
      }
#pragma unroll
      for (int k = 0; k < 7; ++k) I[k] += t[k] * wx;
    }
  }
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    double v = I[k] * wz;
    v += __shfl_xor(v, 1, 64);
    v += __shfl_xor(v, 2, 64);
    I[k] = v;
  }
  if (!live) return;
  double rhs[7], W;
  dropletRhs(A.P, vp, m, I, rhs, W);
  if (c != 0) return;
  // deposit input for the brick pass (k_brick_deposit); slots whose stage
  // bin left the +-1-bin window around their sorted bin go to the fallback list
  {
    bool far = A.all_fallback != 0;
    const double xb[3] = {xw, yw, zw};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      // bins exactly as k_bin_count forms them
      int cs = (int)(xb[k] / dx), c0 = (int)(wrapL(A.x0[k * cap + p], L) / dx);
      cs = cs < 0 ? 0 : (cs > N - 1 ? N - 1 : cs);
      c0 = c0 < 0 ? 0 : (c0 > N - 1 ? N - 1 : c0);
      const int d = ((cs / Bins<N>::BS - c0 / Bins<N>::BS) % Bins<N>::B + Bins<N>::B) % Bins<N>::B;
      far = far || (d > 1 && d < Bins<N>::B - 1);
    }
    // a fallback slot stores -(W + 1) so the brick pass skips it
    A.dep[p] = make_double4(xw, yw, zw, far ? -(W + 1.0) : W);
    if (far) A.flist[atomicAdd(A.fcount, 1)] = p;
  }
  // sump (forced-dns-sm.py:337, 342-358); the last stage writes the new state
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double prev = A.stage == 0 ? A.x0[k * cap + p] : A.sx[k * cap + p];
    const double nv = prev + A.wh * rhs[k];
    if (A.stage == 3) A.x0[k * cap + p] = nv; else A.sx[k * cap + p] = nv;
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const PT prev = A.stage == 0 ? A.v0[k * cap + p] : A.sv[k * cap + p];
    const PT nv = prev + (PT)A.wh * (PT)rhs[3 + k];
    if (A.stage == 3) A.v0[k * cap + p] = nv; else A.sv[k * cap + p] = nv;
  }
  if (A.stage < 3) {
#pragma unroll
    for (int k = 0; k < 3; ++k) A.kx[k * cap + p] = rhs[k];
#pragma unroll
    for (int k = 0; k < 4; ++k) A.kv[k * cap + p] = (PT)rhs[3 + k];
  }
}

// ------------------------------------------------------------ brick deposit
// Deposit of W with the cosine kernel / dx^3 (particles.py:23-35, 529-575),
// organised by output: one workgroup per brick of 8^3 grid nodes scans the
// sorted-bin ranges that can reach the brick (stencil -1..+2 cells, plus one
// bin of motion since the sort), accumulates in LDS and writes each node of
// the brick exactly once.  Slots on the fallback list are skipped here.
template <int N> struct Brick {
  static constexpr int S = 8;                       // nodes per side
  static constexpr int NB = N / S;                  // bricks per side
};

template <typename T, int N>
__global__ void __launch_bounds__(256)
k_brick_deposit(const double4* __restrict__ dep, const int* __restrict__ start, int np, T* __restrict__ fc,
                DropPhys P) {
  constexpr int S = Brick<N>::S, NBr = Brick<N>::NB;
  constexpr int BS = Bins<N>::BS, B = Bins<N>::B;
  __shared__ float tile_f[(sizeof(T) == 4) ? S * S * S : 1];
  __shared__ double tile_d[(sizeof(T) == 8) ? S * S * S : 1];
  T* tile = (sizeof(T) == 4) ? (T*)tile_f : (T*)tile_d;
  const int br = blockIdx.x;
  const int c0x = (br / (NBr * NBr)) * S, c0y = ((br / NBr) % NBr) * S, c0z = (br % NBr) * S;
  for (int i = threadIdx.x; i < S * S * S; i += blockDim.x) tile[i] = (T)0;
  __syncthreads();
  // bins whose droplets can reach the brick: stage cells [c0-2, c0+S] -> bins, +-1 bin of motion
  const int blx = (c0x - 2 + N) % N / BS, bly = (c0y - 2 + N) % N / BS, blz = (c0z - 2 + N) % N / BS;
  // stage cells c0-2 .. c0+S span at most (S + 2 + BS - 1)/BS + 1 bins; one more on each side for motion
  int nbx = (S + 2 + BS - 1) / BS + 1 + 2;
  if (nbx > B) nbx = B;
  const int lo_shift = (nbx == B) ? 0 : 1;          // start one bin below the stencil's lowest bin
  const int nruns = nbx * nbx;
  const double dx = P.dx, pi_2dx = P.pi_2dx;
  for (int r = threadIdx.x >> 2; r < nruns; r += blockDim.x >> 2) {
    const int bx = (blx - lo_shift + r / nbx + B) % B, by = (bly - lo_shift + r % nbx + B) % B;
    for (int q = 0; q < nbx; ++q) {
      const int bz = (blz - lo_shift + q + B) % B;
      const int bin = (bx * B + by) * B + bz;
      const int b = start[bin], e = (bin + 1 < B * B * B) ? start[bin + 1] : np;
      for (int t = b + (threadIdx.x & 3); t < e; t += 4) {
        const double4 d = dep[t];
        if (!(d.w > 0.0)) continue;             // no growth, or a fallback slot
        const int i0 = (int)floor(d.x / dx) - 1, j0 = (int)floor(d.y / dx) - 1, k0 = (int)floor(d.z / dx) - 1;
        // brick-relative first node of the stencil (periodic); reject without overlap
        const int rx = ((i0 - c0x) % N + N) % N, ry = ((j0 - c0y) % N + N) % N, rz = ((k0 - c0z) % N + N) % N;
        const bool ox = rx < S || rx > N - 4, oy = ry < S || ry > N - 4, oz = rz < S || rz > N - 4;
        if (!(ox && oy && oz)) continue;
        const double W = d.w * P.inv_vol;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          const int nx = (rx + a) % N;
          if (nx >= S) continue;
          const double wa = W * cosW(d.x, i0 + a, dx, pi_2dx);
#pragma unroll
          for (int bb = 0; bb < 4; ++bb) {
            const int ny = (ry + bb) % N;
            if (ny >= S) continue;
            const double wab = wa * cosW(d.y, j0 + bb, dx, pi_2dx);
#pragma unroll
            for (int cc = 0; cc < 4; ++cc) {
              const int nz = (rz + cc) % N;
              if (nz >= S) continue;
              atomicAdd(&tile[(nx * S + ny) * S + nz], (T)(wab * cosW(d.z, k0 + cc, dx, pi_2dx)));
            }
          }
        }
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < S * S * S; i += blockDim.x) {
    const int nx = i / (S * S), ny = (i / S) % S, nz = i % S;
    fc[((size_t)(c0x + nx) * N + (c0y + ny)) * N + (c0z + nz)] = tile[i];
  }
}

// Fallback deposit (global atomics) of the slots that moved more than one
// bin since the sort; runs after k_brick_deposit.
template <typename T, int N>
__global__ void __launch_bounds__(256)
k_fallback_deposit(const double4* __restrict__ dep, const int* __restrict__ flist, const int* __restrict__ fcount,
                   T* fc, DropPhys P) {
  const int nf = *fcount;
  for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < nf; f += gridDim.x * blockDim.x) {
    const double4 d = dep[flist[f]];
    const double xw[3] = {d.x, d.y, d.z};
    Stencil<N> s;
    s.set(xw, P.dx, P.pi_2dx);
    deposit<T, N>(s, -d.w - 1.0, P.inv_vol, fc);
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
k_bin_count(const double* __restrict__ pos, int np, int cap, double L, double dx, int* __restrict__ key,
            int* __restrict__ rank, int* __restrict__ hist) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  int b[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    int ci = (int)(wrapL(pos[c * cap + p], L) / dx);
    ci = ci < 0 ? 0 : (ci > N - 1 ? N - 1 : ci);
    b[c] = ci / Bins<N>::BS;
  }
  const int k = (b[0] * Bins<N>::B + b[1]) * Bins<N>::B + b[2];
  key[p] = k;
  rank[p] = atomicAdd(&hist[k], 1);
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
k_scan_local(int* __restrict__ data, int n, int* __restrict__ bsum) {
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
k_scan_bsums(int* __restrict__ bsum, int nb) {
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
k_scan_add(int* __restrict__ data, int n, const int* __restrict__ bsum) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) data[i] += bsum[i / 4096];
}

// Scatter the droplet state to its sorted slot.
template <typename PT>
__global__ void __launch_bounds__(256)
k_bin_scatter(int np, int cap, const int* __restrict__ key, const int* __restrict__ rank,
              const int* __restrict__ start, const double* __restrict__ pos, const PT* __restrict__ vm,
              const int64_t* __restrict__ id, double* __restrict__ pos2, PT* __restrict__ vm2,
              int64_t* __restrict__ id2) {
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
