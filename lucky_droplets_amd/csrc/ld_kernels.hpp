// Kernels of the forced-DNS + droplet RK4 step (templated on precision T and
// grid size N).  Layouts follow the reference at one MPI rank:
//   physical field  [x][y][z]   (z contiguous), N^3 reals
//   spectral field  [kx][ky][kz] (kz contiguous), N*N*Nf complex, Nf = N/2+1
// A 3-D r2c transform is three line passes: z (real <-> complex, contiguous
// lines), y (stride Nf) and x (stride N*Nf); the inverse runs x, y, z.
// Pointwise work of the step is fused into the first/last pass that touches
// the data (see each kernel's comment for the reference lines it covers).
#pragma once
#include "ld_fft.hpp"

namespace ld {

// ===================================================================== geometry
template <int N> struct Geo {
  static constexpr int Nf = N / 2 + 1;
  static constexpr int NNf = N * Nf;           // complex elements per plane (x fixed)
  static constexpr size_t C = (size_t)N * N * Nf;   // complex elements per field
  static constexpr size_t R = (size_t)N * N * N;    // reals per field
};

// Node-interleaved 4-vector of the physical fields the droplets gather
// ({u_x, u_y, u_z, 0} and {v_x, v_y, v_z, n_stage}).
template <typename T> struct alignas(4 * sizeof(T)) V4 { T x, y, z, w; };

// Batched launches: blockIdx.z selects the realization of an ensemble, and
// every per-realization buffer lies rs bytes after the previous realization's
// copy (rs = 0 with one realization).  Null pointers stay null.
// (byte arithmetic on char pointers, not integers, so the compiler still
// sees global-memory pointers and emits global_* rather than flat_* accesses)
template <class X> __device__ __forceinline__ X* rsel1(X* p, size_t rs) {
  using Byte = std::conditional_t<std::is_const<X>::value, const char, char>;
  return reinterpret_cast<X*>(reinterpret_cast<Byte*>(p) + (size_t)blockIdx.z * rs);
}
// the same for an optional pointer: null stays null
template <class X> __device__ __forceinline__ X* rsel(X* p, size_t rs) {
  return p ? rsel1(p, rs) : p;
}

// Loads/stores at a 32-bit element index from a uniform base pointer: the
// byte offset stays 32-bit so the address is SGPR base + VGPR offset.
template <typename X>
__device__ __forceinline__ X ld32(const X* base, uint32_t idx) {
  return *reinterpret_cast<const X*>(reinterpret_cast<const char*>(base) + (size_t)(idx * (uint32_t)sizeof(X)));
}
template <typename X>
__device__ __forceinline__ void st32(X* base, uint32_t idx, X v) {
  *reinterpret_cast<X*>(reinterpret_cast<char*>(base) + (size_t)(idx * (uint32_t)sizeof(X))) = v;
}

// Strided passes: AXIS 0 = x (lines q in [0, N*Nf), element stride N*Nf),
// AXIS 1 = y (lines q = x*Nf + kz, element stride Nf).
template <int N, int AXIS> __device__ __forceinline__ int lineBase(int q) {
  constexpr int Nf = N / 2 + 1;
  if constexpr (AXIS == 0) return q;
  else return (q / Nf) * (N * Nf) + (q % Nf);
}
template <int N, int AXIS> struct JS { static constexpr int v = AXIS == 0 ? N * (N / 2 + 1) : (N / 2 + 1); };

// Spectral-space operator values at one mode (forced-dns-sm.py:125-136).
template <typename T> struct Mode {
  T kx, ky, kz;   // wavenumbers
  T k2;           // |k|^2 (exact integer)
  T mask;         // dealias: kint <= N/3
  T invlap;       // mask / lap  (0 at k = 0)
  int kint;       // clip(round(|k|), N/2)
};
template <typename T, int N>
__device__ __forceinline__ Mode<T> mode(int ix, int iy, int iz) {
  Mode<T> m;
  const int kx = kwave(ix, N), ky = kwave(iy, N), kz = iz;
  const int k2 = kx * kx + ky * ky + kz * kz;
  int ki = kround(k2);
  if (ki > N / 2) ki = N / 2;
  m.kx = (T)kx; m.ky = (T)ky; m.kz = (T)kz; m.k2 = (T)k2; m.kint = ki;
  m.mask = (3 * ki <= N) ? (T)1 : (T)0;
  m.invlap = (k2 == 0) ? (T)0 : m.mask / (T)(-k2);
  return m;
}

// ================================================================ block reductions
// Deterministic sum over the block (fixed tree order); result valid in thread 0.
template <int NT>
__device__ __forceinline__ double blockSum(double v, double* sh) {
  const int tid = threadIdx.x;
  sh[tid] = v;
  __syncthreads();
#pragma unroll
  for (int s = NT / 2; s > 0; s >>= 1) {
    if (tid < s) sh[tid] += sh[tid + s];
    __syncthreads();
  }
  return sh[0];
}
template <int NT>
__device__ __forceinline__ double blockMax(double v, double* sh) {
  const int tid = threadIdx.x;
  sh[tid] = v;
  __syncthreads();
#pragma unroll
  for (int s = NT / 2; s > 0; s >>= 1) {
    if (tid < s) sh[tid] = fmax(sh[tid], sh[tid + s]);
    __syncthreads();
  }
  return sh[0];
}

// ======================================================== plain strided pass
// One line pass along x or y for `gridDim.y` fields (field stride fstride).
// DIR < 0: forward (input natural order, output natural order);
// DIR > 0: inverse times `scale`.  In place is allowed (each block owns its lines).
// y passes run over the lines (x, kz) with kz < lw, numbered x * lw + kz,
// from qoff to qend (lw = N/2 + 1: every line).  Lines with kz >= lw are
// neither read nor written: callers pass lw < N/2 + 1 only in place (in ==
// out) and only where those lines are zero (their transform is zero, which
// they already hold) or are never read afterwards (modes the dealiasing mask
// removes).  With r2 >= 0, only the spectral entries with ky^2 + kz^2 <= r2
// are moved: a forward pass writes only those (the next pass reads no other),
// an inverse pass reads only those (the others are zero).
template <typename T, int N, int M, int AXIS, int DIR, int Q = 0>
__global__ void __launch_bounds__((M * Plan<N, Q>::N1))
k_pass(const C<T>* in, C<T>* out, size_t fstride, const C<T>* __restrict__ tw, T scale, int qoff, size_t rs,
       int lw, int qend, int r2) {
  constexpr int N1 = Plan<N, Q>::N1, N2 = Plan<N, Q>::N2, js = JS<N, AXIS>::v;
  __shared__ C<T> lds[M * (N + 1)];
  const int m = threadIdx.x % M, n1 = threadIdx.x / M;
  // XCD-aware: consecutive line groups go to one XCD, so row segments that
  // straddle cache lines (y rows start every Nf elements) meet in its L2
  const int nbx = gridDim.x;
  const int bx = (nbx % 8 == 0) ? (blockIdx.x % 8) * (nbx / 8) + blockIdx.x / 8 : blockIdx.x;
  const int qc = qoff + bx * M + m;
  int q = qc, kz2 = 0;
  bool live = true;
  if constexpr (AXIS == 1) {
    const int kz = qc % lw;
    if (lw != N / 2 + 1) q = (qc / lw) * (N / 2 + 1) + kz;
    kz2 = kz * kz;
    live = qc < qend;
  }
  // spectral entry at y index iy is moved
  auto keep = [&](int iy) {
    const int ky = kwave(iy, N);
    return AXIS != 1 || r2 < 0 || ky * ky + kz2 <= r2;
  };
  const C<T>* src = rsel1(in, rs) + (size_t)blockIdx.y * fstride;
  C<T>* dst = rsel1(out, rs) + (size_t)blockIdx.y * fstride;
  const uint32_t lb = (uint32_t)lineBase<N, AXIS>(q);
  C<T>* L = lds + m * (N + 1);
  C<T> a[N2];
  if constexpr (DIR < 0) {
    if (live) {
#pragma unroll
      for (int n2 = 0; n2 < N2; ++n2) a[n2] = ld32(src, lb + (uint32_t)(posIn<N, Q>(n1, n2) * js));
    } else {
#pragma unroll
      for (int n2 = 0; n2 < N2; ++n2) a[n2] = mkc<T>((T)0, (T)0);
    }
    lineFwd<T, N, Q>(a, L, n1, true, tw);
    if (live) {
#pragma unroll
      for (int r = 0; r < N2; ++r)
        if (keep(posOut<N, Q>(n1, r))) st32(dst, lb + (uint32_t)(posOut<N, Q>(n1, r) * js), a[r]);
    }
  } else {
    if (live) {
#pragma unroll
      for (int r = 0; r < N2; ++r)
        a[r] = keep(posOut<N, Q>(n1, r)) ? ld32(src, lb + (uint32_t)(posOut<N, Q>(n1, r) * js)) : mkc<T>((T)0, (T)0);
    } else {
#pragma unroll
      for (int r = 0; r < N2; ++r) a[r] = mkc<T>((T)0, (T)0);
    }
    lineInv<T, N, Q, true>(a, L, n1, true, tw);
    if (live) {
#pragma unroll
      for (int n2 = 0; n2 < N2; ++n2) st32(dst, lb + (uint32_t)(posIn<N, Q>(n1, n2) * js), cscale<T>(a[n2], scale));
    }
  }
}

// ============================================================= z-pass machinery
// A block owns P consecutive (x,y) lines of several fields.  Real lines are
// paired two-for-one into complex lines: lines (f, 2c) and (f, 2c+1) share one
// complex FFT.  Each pair lives in an LDS "region" of 2*Nf complex elements,
// which holds the pair's spectral lines, the FFT exchange and the two real
// lines in turn.
template <typename T, int N, int P>
struct ZTile {
  static constexpr int Nf = N / 2 + 1;
  static constexpr int REG = 2 * Nf;
  C<T>* lds;
  __device__ __forceinline__ C<T>* region(int r) const { return lds + r * REG; }
  __device__ __forceinline__ T& re(int f, int p, int j) const {
    T* Rr = (T*)region(f * (P / 2) + p / 2);
    return Rr[(p & 1) * N + j];
  }
  __device__ __forceinline__ C<T>& spec(int f, int p, int k) const {
    return region(f * (P / 2) + p / 2)[(p & 1) * Nf + k];
  }
  // four consecutive real entries j..j+3 (j % 4 == 0; regions are 32-byte aligned)
  __device__ __forceinline__ V4<T>& re4(int f, int p, int j) const {
    return *reinterpret_cast<V4<T>*>(&re(f, p, j));
  }
};

// Inverse z transform (c2r, unnormalised) of FIN fields already loaded as
// spectral lines into the tile; leaves real lines in the tile.  Imaginary
// parts of the kz = 0 and kz = N/2 modes are ignored, as a c2r transform does.
// w: this thread's twiddles (n1 = threadIdx.x % N1).
template <typename T, int N, int P, int FIN>
__device__ __forceinline__ void zInverse(const ZTile<T, N, P>& tile, const Twiddles<T, N>& w) {
  constexpr int N1 = Plan<N>::N1, N2 = Plan<N>::N2, Nf = N / 2 + 1;
  const int r = threadIdx.x / N1, n1 = threadIdx.x % N1;
  const bool active = r < FIN * P / 2;
  C<T> a[N2];
  C<T>* reg = tile.region(active ? r : 0);
  if (active) {
    const C<T>* Xa = reg;        // even line
    const C<T>* Xb = reg + Nf;   // odd line
#pragma unroll
    for (int rr = 0; rr < N2; ++rr) {
      const int j = posOut<N>(n1, rr);
      C<T> A, B;
      if (j <= N / 2) {
        A = Xa[j]; B = Xb[j];
        if (j == 0 || j == N / 2) { A.y = (T)0; B.y = (T)0; }
      } else {
        A = cconj<T>(Xa[N - j]); B = cconj<T>(Xb[N - j]);
      }
      a[rr] = mkc<T>(A.x - B.y, A.y + B.x);   // A + i B
    }
  }
  lineInvT<T, N, 0>(a, reg, n1, active, w);
  __syncthreads();
  if (active) {
    T* Rr = (T*)reg;
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2) {
      const int j = posIn<N>(n1, n2);
      Rr[j] = a[n2].x;
      Rr[N + j] = a[n2].y;
    }
  }
  __syncthreads();
}

// Forward z transform (r2c, unnormalised) of FOUT real fields in the tile;
// leaves each pair's complex spectrum Z (natural order, N entries) in its
// region.  Split with zSplit().  w as for zInverse.
template <typename T, int N, int P, int FOUT>
__device__ __forceinline__ void zForward(const ZTile<T, N, P>& tile, const Twiddles<T, N>& w) {
  constexpr int N1 = Plan<N>::N1, N2 = Plan<N>::N2;
  const int r = threadIdx.x / N1, n1 = threadIdx.x % N1;
  const bool active = r < FOUT * P / 2;
  C<T> a[N2];
  C<T>* reg = tile.region(active ? r : 0);
  if (active) {
    const T* Rr = (const T*)reg;
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2) {
      const int j = posIn<N>(n1, n2);
      a[n2] = mkc<T>(Rr[j], Rr[N + j]);
    }
  }
  lineFwdT<T, N, 0>(a, reg, n1, active, w);
  __syncthreads();
  if (active) {
#pragma unroll
    for (int rr = 0; rr < N2; ++rr) reg[posOut<N>(n1, rr)] = a[rr];
  }
  __syncthreads();
}

// Spectrum of real line p (of field f) at kz = k from the pair's complex FFT.
template <typename T, int N, int P>
__device__ __forceinline__ C<T> zSplit(const ZTile<T, N, P>& tile, int f, int p, int k) {
  const C<T>* Z = tile.region(f * (P / 2) + p / 2);
  const C<T> zk = Z[k];
  const C<T> zm = cconj<T>(Z[(N - k) & (N - 1)]);
  if ((p & 1) == 0) return mkc<T>((T)0.5 * (zk.x + zm.x), (T)0.5 * (zk.y + zm.y));
  // -i/2 (zk - zm)
  return mkc<T>((T)0.5 * (zk.y - zm.y), (T)-0.5 * (zk.x - zm.x));
}

template <int FIN, int FOUT> struct ZMax { static constexpr int v = FIN > FOUT ? FIN : FOUT; };
template <int N, int P, int FIN, int FOUT> struct ZThreads {
  static constexpr int v = ZMax<FIN, FOUT>::v * P / 2 * Plan<N>::N1;
};

// Two consecutive spectral entries, moved as one 16-byte (FP32) access.
template <typename T> struct alignas(2 * sizeof(C<T>)) C2 { C<T> a, b; };

// Generic z-pass kernel: [inverse of FIN spectral fields] -> Op::physical on
// the P real lines of every field -> [forward of FOUT fields].  Spectral
// field f of the Op is read from Op::specIn(f) and written to Op::specOut(f)
// ([x][y][kz] layout).  The P lines of a block are one contiguous run of
// P*Nf entries per field, kept in LDS in the same order (the regions of
// consecutive line pairs are adjacent), so they move as pairs of entries;
// line0 * Nf is even (P and the line offset are even), so every pair is
// aligned.  Input entries with kz > kcin are zero by the caller's contract
// (dealiasing truncation: no earlier pass writes them): they are set in LDS,
// not loaded.  Output pairs whose two entries both have kz > kcout are not
// written (no later pass reads them).
template <typename T, int N, int P, int FIN, int FOUT, class Op>
__global__ void __launch_bounds__((ZThreads<N, P, FIN, FOUT>::v))
k_z(Op op0, const C<T>* __restrict__ tw, int lineoff, size_t rs, int kcin, int kcout) {
  static_assert(P % 2 == 0, "pairs of lines");
  const Op op = op0.shifted(rs);
  constexpr int Nf = N / 2 + 1;
  constexpr int NREG = ZMax<FIN, FOUT>::v * P / 2;
  constexpr int NE = P * Nf;        // entries per field of a block
  __shared__ C2<T> lds2[NREG * Nf];
  C<T>* lds = &lds2[0].a;
  ZTile<T, N, P> tile{lds};
  const int line0 = lineoff + blockIdx.x * P;
  const size_t g0 = (size_t)line0 * Nf;
  const int nt = blockDim.x;
  // the thread's line-FFT twiddles (the same n1 in both directions), loaded
  // first so that their latency overlaps the tile loads
  const Twiddles<T, N> w((int)threadIdx.x % Plan<N>::N1, tw);
  if constexpr (FIN > 0) {
    // pair e = f (NE / 2) + v of the block's input goes to lds2[e].  Every load
    // of a thread is issued before the first LDS store: the loads are
    // unconditional (a branch around one would wait for it), a pair that is
    // not needed (both kz > kcin, or past the end) reads the block's first
    // pair of the field instead, which other lanes load anyway.
    constexpr int NT = ZThreads<N, P, FIN, FOUT>::v, NPR = FIN * (NE / 2), IT = (NPR + NT - 1) / NT;
    C2<T> x[IT];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int e0 = (int)threadIdx.x + it * NT;
      const int e = (NPR % NT == 0 || e0 < NPR) ? e0 : 0;
      const int f = e / (NE / 2), v = e % (NE / 2);
      const bool need = (2 * v) % Nf <= kcin || (2 * v + 1) % Nf <= kcin;
      x[it] = reinterpret_cast<const C2<T>*>(op.specIn(f) + g0)[need ? v : 0];
    }
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int e = (int)threadIdx.x + it * NT;
      if (NPR % NT == 0 || e < NPR) {
        const int v = e % (NE / 2);
        if ((2 * v) % Nf > kcin) x[it].a = mkc<T>((T)0, (T)0);
        if ((2 * v + 1) % Nf > kcin) x[it].b = mkc<T>((T)0, (T)0);
        lds2[e] = x[it];
      }
    }
    __syncthreads();
    zInverse<T, N, P, FIN>(tile, w);
  }
  op.physical(tile, line0);
  __syncthreads();
  if constexpr (FOUT > 0) {
    zForward<T, N, P, FOUT>(tile, w);
    for (int e = threadIdx.x; e < FOUT * (NE / 2); e += nt) {
      const int f = e / (NE / 2), q = 2 * (e % (NE / 2));
      const int za = q % Nf, zb = (q + 1) % Nf;
      if (za > kcout && zb > kcout) continue;
      C2<T> v;
      v.a = zSplit<T, N, P>(tile, f, q / Nf, za);
      v.b = zSplit<T, N, P>(tile, f, (q + 1) / Nf, zb);
      reinterpret_cast<C2<T>*>(op.specOut(f) + g0)[q / 2] = v;
    }
  }
}

// ------------------------------------------------------------- z-pass ops
// The physical-space work of an op visits its P lines (x F fields) with the
// block's compile-time thread count NT: item e of thread t is t + it * NT,
// it < IT, so the loop is unrolled and all of a thread's global loads are
// issued before any is used.  An item is a quad of 4 consecutive z entries
// (planar arrays: 16-byte FP32 accesses) or one entry (node-interleaved
// records, which are 16 bytes already).
template <int NT, int NQ> struct Quads {
  static constexpr int IT = (NQ + NT - 1) / NT;
  static __device__ __forceinline__ int e(int it) { return (int)threadIdx.x + it * NT; }
  static __device__ __forceinline__ bool ok(int it) { return NQ % NT == 0 || e(it) < NQ; }
};
template <typename T> __device__ __forceinline__ V4<T> scale4(const V4<T>& a, T s) {
  V4<T> r; r.x = a.x * s; r.y = a.y * s; r.z = a.z * s; r.w = a.w * s; return r;
}
// Plain c2r along z of F fields, times `scale` (last pass of irfft_mpi,
// forced-dns-sm.py:212).
template <typename T, int N, int P, int F>
struct OpZInv {
  const C<T>* in; T* out; size_t fin, fout; T scale;
  __device__ OpZInv shifted(size_t rs) const { OpZInv o = *this; o.in = rsel1(in, rs); o.out = rsel1(out, rs); return o; }
  __device__ const C<T>* specIn(int f) const { return in + f * fin; }
  __device__ C<T>* specOut(int) const { return nullptr; }
  __device__ void physical(const ZTile<T, N, P>& t, int line0) const {
    using Q = Quads<ZThreads<N, P, F, 0>::v, F * P * N / 4>;
#pragma unroll
    for (int it = 0; it < Q::IT; ++it) {
      if (!Q::ok(it)) break;
      const int e4 = 4 * Q::e(it), f = e4 / (P * N), rem = e4 % (P * N), p = rem / N, j = rem % N;
      *reinterpret_cast<V4<T>*>(out + f * fout + (size_t)(line0 + p) * N + j) = scale4(t.re4(f, p, j), scale);
    }
  }
};
// Plain r2c along z of F fields (first pass of rfft_mpi, forced-dns-sm.py:202).
template <typename T, int N, int P, int F>
struct OpZFwd {
  const T* in; C<T>* out; size_t fin, fout;
  __device__ OpZFwd shifted(size_t rs) const { OpZFwd o = *this; o.in = rsel1(in, rs); o.out = rsel1(out, rs); return o; }
  __device__ const C<T>* specIn(int) const { return nullptr; }
  __device__ C<T>* specOut(int f) const { return out + f * fout; }
  __device__ void physical(const ZTile<T, N, P>& t, int line0) const {
    using Q = Quads<ZThreads<N, P, 0, F>::v, F * P * N / 4>;
    // one loop (global loads straight into the tile): with a separate register
    // array the compiler kept it in scratch memory
#pragma unroll
    for (int it = 0; it < Q::IT; ++it) {
      if (Q::ok(it)) {
        const int e4 = 4 * Q::e(it), f = e4 / (P * N), rem = e4 % (P * N), p = rem / N, j = rem % N;
        t.re4(f, p, j) = *reinterpret_cast<const V4<T>*>(in + f * fin + (size_t)(line0 + p) * N + j);
      }
    }
  }
};

// Droplet gather records.  Per node two 16-byte records, {u_x, u_y, u_z, 0}
// and {v_x, v_y, v_z, n_stage}, laid out in z-quads: the four u records of
// nodes z = 4q .. 4q+3 of an (x, y) line (64 bytes), then their four v records,
// so a 128-byte line holds both records of four consecutive z nodes and a
// droplet's 4-node z run of both records touches one or two lines instead of
// two to four.  Node index o = (x N + y) N + z.
__device__ __forceinline__ uint32_t recU(uint32_t o) { return ((o >> 2) << 3) | (o & 3u); }
__device__ __forceinline__ uint32_t recV(uint32_t o) { return recU(o) | 4u; }
// Load of record idx: a 32-bit byte offset from the uniform base where the
// array (2 N^3 records) fits in 4 GiB, else a 64-bit one (512^3 FP64).
template <typename T, int N>
__device__ __forceinline__ V4<T> ldRec(const V4<T>* base, uint32_t idx) {
  if constexpr (2ull * N * N * N * sizeof(V4<T>) <= (1ull << 32))
    return ld32(base, idx);
  else
    return base[idx];
}

// Velocity and vorticity to physical space, u stored, then the rotational
// nonlinear term u x omega transformed along z (forced-dns-sm.py:291-301, and
// the first pass of the r2c at :304-306).  Spectral input: 6 fields
// (u_x,u_y,u_z,w_x,w_y,w_z) already inverse-transformed along x and y; output:
// 3 fields written over inputs 0..2.
template <typename T, int N, int P>
struct OpZCross {
  C<T>* w; size_t fs; T* u; V4<T>* gu; T scale;   // u planar (may be null) and/or gather records (may be null)
  __device__ OpZCross shifted(size_t rs) const {
    OpZCross o = *this; o.w = rsel1(w, rs); o.u = rsel(u, rs); o.gu = rsel(gu, rs); return o;
  }
  __device__ const C<T>* specIn(int f) const { return w + f * fs; }
  __device__ C<T>* specOut(int f) const { return w + f * fs; }
  __device__ void physical(const ZTile<T, N, P>& t, int line0) const {
    constexpr size_t R = (size_t)N * N * N;
    // one element per thread and step, so consecutive lanes write
    // consecutive 16-byte records
    using Q = Quads<ZThreads<N, P, 6, 3>::v, P * N>;
#pragma unroll
    for (int it = 0; it < Q::IT; ++it) {
      if (!Q::ok(it)) break;
      const int e = Q::e(it), p = e / N, j = e % N;
      const T ux = t.re(0, p, j) * scale, uy = t.re(1, p, j) * scale, uz = t.re(2, p, j) * scale;
      const T wx = t.re(3, p, j) * scale, wy = t.re(4, p, j) * scale, wz = t.re(5, p, j) * scale;
      const size_t o = (size_t)(line0 + p) * N + j;
      if (u) { u[o] = ux; u[R + o] = uy; u[2 * R + o] = uz; }
      if (gu) { V4<T> q; q.x = ux; q.y = uy; q.z = uz; q.w = (T)0; gu[recU((uint32_t)o)] = q; }
      t.re(0, p, j) = wz * uy - wy * uz;
      t.re(1, p, j) = wx * uz - wz * ux;
      t.re(2, p, j) = wy * ux - wx * uy;
    }
  }
};

// Stage density as the stage sees it: the reference stores clip_zero(x) =
// clip(x, 0) * oldmean / newmean (forced-dns-sm.py:278-285); we keep x and
// apply the clip lazily.  clip[0] = oldmean, clip[1] = newmean, clip[2] = 1
// to apply, 0 to pass x through unchanged.
template <typename T>
__device__ __forceinline__ T nStage(T x, const double* __restrict__ clip) {
  if (clip[2] != 0.0) {
    const double xc = x > (T)0 ? (double)x : 0.0;
    return (T)((xc * clip[0]) / clip[1]);
  }
  return x;
}

// Slow-manifold velocity v to physical space (forced-dns-sm.py:319-321, last
// pass), v stored, then n*v transformed along z (:326-329, first pass).
template <typename T, int N, int P>
struct OpZVel {
  const C<T>* w; C<T>* wout; size_t fs; T* v; V4<T>* gv; const T* n; const double* clip; T scale; T tpsg;
  __device__ OpZVel shifted(size_t rs) const {
    OpZVel o = *this;
    o.w = rsel1(w, rs); o.wout = rsel1(wout, rs); o.v = rsel(v, rs); o.gv = rsel(gv, rs);
    o.n = rsel1(n, rs); o.clip = rsel1(clip, rs);
    return o;
  }
  __device__ const C<T>* specIn(int f) const { return w + f * fs; }
  __device__ C<T>* specOut(int f) const { return wout + f * fs; }
  __device__ void physical(const ZTile<T, N, P>& t, int line0) const {
    constexpr size_t R = (size_t)N * N * N;
    // one element per thread and step (consecutive lanes write consecutive
    // 16-byte records); all loads of n are issued before the first use
    using Q = Quads<ZThreads<N, P, 3, 3>::v, P * N>;
    T nq[Q::IT];
#pragma unroll
    for (int it = 0; it < Q::IT; ++it) {
      if (!Q::ok(it)) break;
      nq[it] = n[(size_t)line0 * N + Q::e(it)];
    }
#pragma unroll
    for (int it = 0; it < Q::IT; ++it) {
      if (!Q::ok(it)) break;
      const int e = Q::e(it), p = e / N, j = e % N;
      const size_t o = (size_t)(line0 + p) * N + j;
      const T vx = t.re(0, p, j) * scale, vy = t.re(1, p, j) * scale;
      const T vz = t.re(2, p, j) * scale - tpsg;
      const T nn = nStage<T>(nq[it], clip);
      if (v) { v[o] = vx; v[R + o] = vy; v[2 * R + o] = vz; }
      if (gv) { V4<T> q; q.x = vx; q.y = vy; q.z = vz; q.w = nn; gv[recV((uint32_t)o)] = q; }
      t.re(0, p, j) = nn * vx;
      t.re(1, p, j) = nn * vy;
      t.re(2, p, j) = nn * vz;
    }
  }
};

// Density RHS kn to physical space (last pass of forced-dns-sm.py:330) fused
// with the RK4 update of n (:343-360) and the two sums of the next clip_zero
// (:281-283).  Modes:
//   kn_out != null        : store kn (full_RHS semantics, no RK);
//   nacc / nnext          : nacc = (acc_init ? n0 : nacc) + wacc*kn,
//                           nnext = max(n0,0) + anext*kn   (n0 = RK base)
//   sum_target            : 0 -> sums of nnext, 1 -> sums of nacc, -1 none
template <typename T, int N, int P>
struct OpZKn {
  const C<T>* w; size_t fs; T scale;
  T* kn_out;
  const T* n0; T* nacc; T* nnext;
  T wacc, anext; int acc_init;
  int sum_target; double* partials;
  __device__ OpZKn shifted(size_t rs) const {
    OpZKn o = *this;
    o.w = rsel1(w, rs); o.kn_out = rsel(kn_out, rs); o.n0 = rsel(n0, rs); o.nacc = rsel(nacc, rs);
    o.nnext = rsel(nnext, rs); o.partials = rsel1(partials, rs);
    return o;
  }
  __device__ const C<T>* specIn(int f) const { return w + f * fs; }
  __device__ C<T>* specOut(int) const { return nullptr; }
  __device__ void physical(const ZTile<T, N, P>& t, int line0) const {
    double s_old = 0.0, s_new = 0.0;
    using Q = Quads<ZThreads<N, P, 1, 0>::v, P * N / 4>;
    const size_t o0 = (size_t)line0 * N;
    V4<T> qb[Q::IT], qa[Q::IT];
    if (nacc) {
      const T* accsrc = acc_init ? n0 : nacc;
#pragma unroll
      for (int it = 0; it < Q::IT; ++it) {
        if (!Q::ok(it)) break;
        const size_t o = o0 + 4 * Q::e(it);
        qb[it] = *reinterpret_cast<const V4<T>*>(n0 + o);
        qa[it] = *reinterpret_cast<const V4<T>*>(accsrc + o);
      }
    }
#pragma unroll
    for (int it = 0; it < Q::IT; ++it) {
      if (!Q::ok(it)) break;
      const int e4 = 4 * Q::e(it), p = e4 / N, j = e4 % N;
      const size_t o = o0 + e4;
      const V4<T> kn = scale4(t.re4(0, p, j), scale);
      if (kn_out) *reinterpret_cast<V4<T>*>(kn_out + o) = kn;
      if (nacc) {
        const T kv[4] = {kn.x, kn.y, kn.z, kn.w};
        const T bv[4] = {qb[it].x, qb[it].y, qb[it].z, qb[it].w};
        const T av[4] = {qa[it].x, qa[it].y, qa[it].z, qa[it].w};
        T ao[4], bo[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const T a = av[i] + wacc * kv[i];
          ao[i] = a;
          if (sum_target == 1) { s_old += (double)a; s_new += a > (T)0 ? (double)a : 0.0; }
          const T b = (bv[i] > (T)0 ? bv[i] : (T)0) + kv[i] * anext;
          bo[i] = b;
          if (nnext && sum_target == 0) { s_old += (double)b; s_new += b > (T)0 ? (double)b : 0.0; }
        }
        *reinterpret_cast<V4<T>*>(nacc + o) = V4<T>{ao[0], ao[1], ao[2], ao[3]};
        if (nnext) *reinterpret_cast<V4<T>*>(nnext + o) = V4<T>{bo[0], bo[1], bo[2], bo[3]};
      }
    }
    if (sum_target >= 0) {
      constexpr int NT = ZThreads<N, P, 1, 0>::v;
      __shared__ double sh[NT];
      const int blk = line0 / P;            // global tile index (launches may cover a slab)
      const double a = blockSum<NT>(s_old, sh);
      if (threadIdx.x == 0) partials[2 * blk] = a;
      __syncthreads();
      const double b = blockSum<NT>(s_new, sh);
      if (threadIdx.x == 0) partials[2 * blk + 1] = b;
    }
  }
};

// ======================================================== fused x passes
// Arguments of the fused spectral RHS pass.
template <typename T> struct RhsArgs {
  C<T>* nl;            // in: u x w, forward-transformed along z and y (3 fields); out: v_hat x-inverse
  const C<T>* us;      // stage input velocity u_hat (3)
  const C<T>* ubase;   // RK base u_hat (3)
  C<T>* acc;           // RK accumulator (3) or null
  C<T>* usnext;        // next stage input (3) or null
  C<T>* ku;            // raw RHS (3) or null
  const double* fac;   // forcing factor per shell (Nf), times isforcing*forc
  size_t fs;
  T nu_eff, tps, wacc, anext, scale;
  int acc_init;
  int kcut;            // kcut < N/2: lines (ky, kz) with ky^2 + kz^2 > kcut (kcut + 1) hold no mode
                       // inside the mask, so their nonlinear term is not read (those entries of nl
                       // may hold data of an earlier pass)
  int band;            // 1: us, ubase and acc are zero on every line outside the mask radius, so
                       // those lines need no RK update (their zero outputs are already in place)
  __device__ RhsArgs shifted(size_t rs) const {
    RhsArgs a = *this;
    a.nl = rsel1(nl, rs); a.us = rsel1(us, rs); a.ubase = rsel1(ubase, rs); a.acc = rsel(acc, rs);
    a.usnext = rsel(usnext, rs); a.ku = rsel(ku, rs); a.fac = rsel1(fac, rs);
    return a;
  }
};

// ---------------------------------------------- field-parallel x passes
// Block order of the x passes.  The hardware deals consecutive blocks to the
// 8 XCDs in turn; here logical blocks 2j and 2j + 1 run on one XCD, one right
// after the other, and the pairs are dealt in turn.  With 8 complex64 lines
// per block (64-byte row segments) the two blocks of a pair share every
// 128-byte row segment in that XCD's L2 instead of fetching it twice, while
// work stays interleaved over the XCDs (lines outside the dealiasing radius
// cost less).  Grids that are not a multiple of 16 keep the hardware order.
__device__ __forceinline__ int xPairBlock() {
  const int nb = gridDim.x, pb = blockIdx.x;
  if (nb % 16) return pb;
  const int x = pb % 8, k = pb / 8;
  return 2 * (8 * (k >> 1) + x) + (k & 1);
}
// The field-parallel x passes' lines q0 .. q0 + M - 1 ([ky][kz] order): true
// if any holds a mode inside the dealiasing radius (ky^2 + kz^2 <=
// kcut (kcut + 1)), or kcut >= N/2.  About 60 % of the blocks hold none at
// kcut = N/3; where such a block's lines are zero and store nothing it exits
// at entry instead of transforming zeros.
template <int N, int M>
__device__ __forceinline__ bool xBlockLive(int q0, int kcut) {
  constexpr int Nf = N / 2 + 1;
  if (kcut >= N / 2) return true;
  bool any = false;
#pragma unroll
  for (int j = 0; j < M; ++j) {
    const int q = q0 + j, ky = kwave(q / Nf, N), kz = q % Nf;
    any = any || ky * ky + kz * kz <= kcut * (kcut + 1);
  }
  return any;
}
// The kernels below give each of the three vector components its own thread
// group (3 * M * N1 threads): every thread transforms one field, so a thread
// holds N2 elements instead of 3 * N2, and the components meet in LDS for the
// pointwise work.  Lines per block M3 keeps the three LDS regions <= 100 KB.
template <typename T, int N, int M0> struct XM3 {
  static constexpr int v = (3L * M0 * (N + 1) * (long)sizeof(C<T>) <= 102400L) ? M0 : M0 / 2;
};

// Velocity and vorticity (forced-dns-sm.py:291-297): x-inverse of u_hat and of
// i k x u_hat, 6 output fields (u_x, u_y, u_z, w_x, w_y, w_z).  Group f
// inverse-transforms u_f and w_f.
template <typename T, int N, int M>
__global__ void __launch_bounds__(3 * M * Plan<N>::N1)
k_x_uw3(const C<T>* __restrict__ uk0, C<T>* __restrict__ out0, size_t fs, const C<T>* __restrict__ tw, T scale,
        size_t rs, int kcut) {
  const C<T>* __restrict__ uk = rsel1(uk0, rs);
  C<T>* __restrict__ out = rsel1(out0, rs);
  constexpr int N1 = Plan<N>::N1, N2 = Plan<N>::N2, Nf = N / 2 + 1, js = N * Nf, G = M * N1;
  __shared__ C<T> lds[3 * M * (N + 1)];
  if (!xBlockLive<N, M>(xPairBlock() * M, kcut)) return;   // zero lines, nothing stored
  const int f = threadIdx.x / G, loc = threadIdx.x % G;
  const int m = loc % M, n1 = loc / M;
  const Twiddles<T, N> wt(n1, tw);         // line-FFT twiddles of n1, both directions
  const int q = xPairBlock() * M + m;
  const T ky = (T)kwave(q / Nf, N), kz = (T)(q % Nf);
  C<T>* L[3] = {lds + m * (N + 1), lds + (M + m) * (N + 1), lds + (2 * M + m) * (N + 1)};
  C<T>* Lf = lds + (f * M + m) * (N + 1);
  // kcut < N/2: u_hat has no mode outside the mask, so a line with
  // ky^2 + kz^2 > kcut (kcut + 1) is zero: it is not read, and its zero
  // outputs are not written (the y pass skips those entries, the z pass takes
  // kz > kcut as zero without reading it)
  const int kyi = kwave(q / Nf, N), kzi = q % Nf;
  const bool live = kcut >= N / 2 || kyi * kyi + kzi * kzi <= kcut * (kcut + 1);
  const bool wr = live;
  C<T> u[N2], w[N2];
  if (live) {
#pragma unroll
    for (int r = 0; r < N2; ++r) u[r] = ld32(uk + f * fs, (uint32_t)(q + posOut<N>(n1, r) * js));
  } else {
#pragma unroll
    for (int r = 0; r < N2; ++r) u[r] = mkc<T>((T)0, (T)0);
  }
#pragma unroll
  for (int r = 0; r < N2; ++r) Lf[posOut<N>(n1, r)] = u[r];
  __syncthreads();
  // vorticity component f = i (k x u)_f  (forced-dns-sm.py:294-296)
#pragma unroll
  for (int r = 0; r < N2; ++r) {
    const int ix = posOut<N>(n1, r);
    const T kx = (T)kwave(ix, N);
    if (f == 0) {
      const C<T> u1 = L[1][ix], u2 = L[2][ix];
      w[r] = cmuli<T>(mkc<T>(ky * u2.x - kz * u1.x, ky * u2.y - kz * u1.y));
    } else if (f == 1) {
      const C<T> u0 = L[0][ix], u2 = L[2][ix];
      w[r] = cmuli<T>(mkc<T>(kz * u0.x - kx * u2.x, kz * u0.y - kx * u2.y));
    } else {
      const C<T> u0 = L[0][ix], u1 = L[1][ix];
      w[r] = cmuli<T>(mkc<T>(kx * u1.x - ky * u0.x, kx * u1.y - ky * u0.y));
    }
  }
  // (lineInv synchronises before it reuses the regions)
  lineInvT<T, N, 0>(w, Lf, n1, true, wt);
  C<T>* Ow = out + (3 + f) * fs;
  if (wr) {
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2) st32(Ow, (uint32_t)(q + posIn<N>(n1, n2) * js), cscale<T>(w[n2], scale));
  }
  lineInvT<T, N, 0>(u, Lf, n1, true, wt);
  C<T>* Ou = out + f * fs;
  if (wr) {
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2) st32(Ou, (uint32_t)(q + posIn<N>(n1, n2) * js), cscale<T>(u[n2], scale));
  }
}

// flag[0] <- 1 if any of the n spectral entries of u ([f][kx][ky][kz] layout)
// with |k|^2 > r2 is nonzero (or NaN); flag[0] is cleared beforehand.  With r2
// = kc (kc + 1) those are the modes outside the dealiasing mask, and a state
// without them keeps none through the RK4 stages and the step epilogue.
template <typename T, int N>
__global__ void __launch_bounds__(256) k_band_check(const C<T>* __restrict__ u, size_t n, int r2, int* flag) {
  constexpr int Nf = N / 2 + 1;
  bool any = false;
  for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (size_t)gridDim.x * 256) {
    const int kz = (int)(e % Nf);
    const size_t line = e / Nf;
    const int ky = kwave((int)(line % N), N), kx = kwave((int)((line / N) % N), N);
    if (kx * kx + ky * ky + kz * kz > r2) {
      const C<T> v = u[e];
      any = any || !(v.x == (T)0 && v.y == (T)0);
    }
  }
  if (__syncthreads_or(any) && threadIdx.x == 0) flag[0] = 1;
}

// Nonlinear term + forcing + dealiasing + pressure projection + viscosity
// (forced-dns-sm.py:239-275, 304-316), the RK4 accumulation of u_hat
// (:338-360) and the slow-manifold velocity v_hat (:319-321) x-inverse
// transformed for the next passes.
// Group f forward-transforms nl_f, the pointwise stage runs on the LDS
// spectra (modes r = f, f+3, ... of each thread's line), then group f
// inverse-transforms v_f.
template <typename T, int N, int M>
__global__ void __launch_bounds__(3 * M * Plan<N>::N1)
k_x_rhs3(RhsArgs<T> A0, const C<T>* __restrict__ tw, size_t rs) {
  const RhsArgs<T> A = A0.shifted(rs);
  constexpr int N1 = Plan<N>::N1, N2 = Plan<N>::N2, Nf = N / 2 + 1, js = N * Nf, G = M * N1;
  __shared__ C<T> lds[3 * M * (N + 1)];
  // band: lines outside the mask radius need no RK update and store nothing
  if (A.band && !xBlockLive<N, M>(xPairBlock() * M, A.kcut)) return;
  const int f = threadIdx.x / G, loc = threadIdx.x % G;
  const int m = loc % M, n1 = loc / M;
  const Twiddles<T, N> wt(n1, tw);         // line-FFT twiddles of n1, both directions
  const int q = xPairBlock() * M + m;
  const int iy = q / Nf, iz = q % Nf;
  C<T>* L0 = lds + m * (N + 1);
  C<T>* L1 = lds + (M + m) * (N + 1);
  C<T>* L2 = lds + (2 * M + m) * (N + 1);
  C<T>* Lf = lds + (f * M + m) * (N + 1);
  const size_t fs = A.fs;
  C<T>* NLf = A.nl + f * fs;
  const int kyl = kwave(iy, N);
  const bool live = A.kcut >= N / 2 || kyl * kyl + iz * iz <= A.kcut * (A.kcut + 1);
  C<T> a[N2];
  if (live) {
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2) a[n2] = ld32(NLf, (uint32_t)(q + posIn<N>(n1, n2) * js));
  } else {
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2) a[n2] = mkc<T>((T)0, (T)0);
  }
  lineFwdT<T, N, 0>(a, Lf, n1, true, wt);
  __syncthreads();
#pragma unroll
  for (int r = 0; r < N2; ++r) Lf[posOut<N>(n1, r)] = a[r];
  __syncthreads();
  const C<T>* US0 = A.us; const C<T>* US1 = A.us + fs; const C<T>* US2 = A.us + 2 * fs;
  const C<T>* UB0 = A.ubase; const C<T>* UB1 = A.ubase + fs; const C<T>* UB2 = A.ubase + 2 * fs;
  const bool idle = A.band && !live;
#pragma unroll 2
  for (int r = f; r < N2; r += 3) {
    if (idle) break;
    const int ix = posOut<N>(n1, r);
    const Mode<T> md = mode<T, N>(ix, iy, iz);
    const uint32_t o = (uint32_t)(q + ix * js);
    const C<T> us0 = ld32(US0, o), us1 = ld32(US1, o), us2 = ld32(US2, o);
    // forcing (forced-dns-sm.py:262-275)
    const T F = (T)A.fac[md.kint];
    C<T> f0 = cscale<T>(us0, F * md.mask), f1 = cscale<T>(us1, F * md.mask), f2 = cscale<T>(us2, F * md.mask);
    const C<T> dv = mkc<T>(md.kx * f0.x + md.ky * f1.x + md.kz * f2.x, md.kx * f0.y + md.ky * f1.y + md.kz * f2.y);
    const C<T> pf = cscale<T>(dv, md.invlap * md.mask);
    f0 = cscale<T>(cadd<T>(f0, cscale<T>(pf, md.kx)), md.mask);
    f1 = cscale<T>(cadd<T>(f1, cscale<T>(pf, md.ky)), md.mask);
    f2 = cscale<T>(cadd<T>(f2, cscale<T>(pf, md.kz)), md.mask);
    // dealiased nonlinear term + forcing (:304-306)
    const C<T> R0 = cscale<T>(cadd<T>(L0[ix], f0), md.mask);
    const C<T> R1 = cscale<T>(cadd<T>(L1[ix], f1), md.mask);
    const C<T> R2 = cscale<T>(cadd<T>(L2[ix], f2), md.mask);
    // pressure pk = i invlap (k . R)  (:309)
    const C<T> D = mkc<T>(md.kx * R0.x + md.ky * R1.x + md.kz * R2.x, md.kx * R0.y + md.ky * R1.y + md.kz * R2.y);
    const C<T> pk = cmuli<T>(cscale<T>(D, md.invlap));
    // i k_c pk and viscous term nu k^2 u  (:314-316)
    const C<T> ikp0 = cmuli<T>(cscale<T>(pk, md.kx)), ikp1 = cmuli<T>(cscale<T>(pk, md.ky)), ikp2 = cmuli<T>(cscale<T>(pk, md.kz));
    const T nk2 = A.nu_eff * md.k2;
    const C<T> vs0 = cscale<T>(us0, nk2), vs1 = cscale<T>(us1, nk2), vs2 = cscale<T>(us2, nk2);
    const C<T> k0 = csub<T>(csub<T>(R0, ikp0), vs0);
    const C<T> k1 = csub<T>(csub<T>(R1, ikp1), vs1);
    const C<T> k2 = csub<T>(csub<T>(R2, ikp2), vs2);
    if (A.ku) { st32(A.ku, o, k0); st32(A.ku + fs, o, k1); st32(A.ku + 2 * fs, o, k2); }
    if (A.acc) {
      C<T> c0, c1, c2;
      if (A.acc_init) { c0 = ld32(UB0, o); c1 = ld32(UB1, o); c2 = ld32(UB2, o); }
      else { c0 = ld32(A.acc, o); c1 = ld32(A.acc + fs, o); c2 = ld32(A.acc + 2 * fs, o); }
      st32(A.acc, o, cadd<T>(c0, cscale<T>(k0, A.wacc)));
      st32(A.acc + fs, o, cadd<T>(c1, cscale<T>(k1, A.wacc)));
      st32(A.acc + 2 * fs, o, cadd<T>(c2, cscale<T>(k2, A.wacc)));
    }
    if (A.usnext) {
      const C<T> b0 = ld32(UB0, o), b1 = ld32(UB1, o), b2 = ld32(UB2, o);
      st32(A.usnext, o, cadd<T>(b0, cscale<T>(k0, A.anext)));
      st32(A.usnext + fs, o, cadd<T>(b1, cscale<T>(k1, A.anext)));
      st32(A.usnext + 2 * fs, o, cadd<T>(b2, cscale<T>(k2, A.anext)));
    }
    // v_hat = (u - tps (-i k p - nu k^2 u)) * mask  (:319-321)
    L0[ix] = cscale<T>(csub<T>(us0, cscale<T>(csub<T>(cscale<T>(ikp0, (T)-1), vs0), A.tps)), md.mask);
    L1[ix] = cscale<T>(csub<T>(us1, cscale<T>(csub<T>(cscale<T>(ikp1, (T)-1), vs1), A.tps)), md.mask);
    L2[ix] = cscale<T>(csub<T>(us2, cscale<T>(csub<T>(cscale<T>(ikp2, (T)-1), vs2), A.tps)), md.mask);
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < N2; ++r) a[r] = Lf[posOut<N>(n1, r)];
  lineInvT<T, N, 0>(a, Lf, n1, true, wt);
  // v_hat x-inverse: lines outside the mask radius are zero and are not
  // written: the next y pass reads none of their entries, and the z pass takes
  // kz > kcut as zero without reading it
  if (live) {
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2) st32(NLf, (uint32_t)(q + posIn<N>(n1, n2) * js), cscale<T>(a[n2], A.scale));
  }
}

// Step epilogue in spectral space (forced-dns-sm.py:597-610, 618): the
// reality round trip u_hat <- rfftn(irfftn(u_hat)) and the projection, as one
// pointwise pass.  The round trip leaves every mode with 0 < kz < N/2 as it is
// and replaces the kz = 0 and kz = N/2 planes by their Hermitian part,
// X(kx, ky) <- (X(kx, ky) + conj(X(-kx, -ky))) / 2 (the c2r along z keeps only
// the real part of those planes' 2-D inverse transform); this kernel forms that
// result directly, so it does not add the round trip's transform rounding.
// Then u_hat += k (k . u_hat) mask / lap, and per-block maxima of Re(u_hat)
// for the blow-up test.  Blocks take runs of entry pairs grid-stride, two
// pairs of each field per thread in flight.  r2 >= 0: acc and uk are zero at
// every mode with |k|^2 > r2, so pairs of two such modes are skipped (they
// add 0 to the maximum).
template <typename T, int N>
__global__ void __launch_bounds__(256)
k_epi_proj(const C<T>* acc0, C<T>* uk0, size_t fs, double* maxpart0, size_t rs, int r2) {
  const C<T>* acc = rsel1(acc0, rs);
  C<T>* uk = rsel1(uk0, rs);
  double* maxpart = rsel1(maxpart0, rs);
  constexpr int Nf = N / 2 + 1, msk = N - 1;
  constexpr uint32_t NE2 = (uint32_t)N * N * Nf / 2;     // entry pairs (N * N is even)
  constexpr int U = 2;
  __shared__ double sh[256];
  double mx = -INFINITY;
  const uint32_t stride = gridDim.x * 256u;
  // a thread takes entry pairs (2 e2, 2 e2 + 1): 16-byte FP32 accesses
  auto inside = [&](uint32_t e) {
    const int iz = (int)(e % Nf);
    const uint32_t line = e / Nf;
    const int ky = kwave((int)(line & msk), N), kx = kwave((int)(line / N), N);
    return kx * kx + ky * ky + iz * iz <= r2;
  };
  for (uint32_t e0 = blockIdx.x * 256u + threadIdx.x; e0 < NE2; e0 += U * stride) {
    C2<T> a[U][3];
    bool kp[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t e2 = e0 + u * stride;
      kp[u] = e2 < NE2 && (r2 < 0 || inside(2 * e2) || inside(2 * e2 + 1));
      if (kp[u]) {
#pragma unroll
        for (int c = 0; c < 3; ++c) a[u][c] = ld32(reinterpret_cast<const C2<T>*>(acc + c * fs), e2);
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t e2 = e0 + u * stride;
      if (e2 >= NE2) break;
      if (!kp[u]) { mx = fmax(mx, 0.0); continue; }
      C2<T> o[3];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const uint32_t e = 2 * e2 + h;
        const int iz = (int)(e % Nf), iy = (int)((e / Nf) & msk), ix = (int)(e / ((uint32_t)N * Nf));
        C<T> v[3] = {h ? a[u][0].b : a[u][0].a, h ? a[u][1].b : a[u][1].a, h ? a[u][2].b : a[u][2].a};
        if (iz == 0 || iz == N / 2) {
          const uint32_t ep = ((uint32_t)((-ix) & msk) * N + (uint32_t)((-iy) & msk)) * Nf + iz;
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            const C<T> p = ld32(acc + c * fs, ep);
            v[c] = mkc<T>((v[c].x + p.x) * (T)0.5, (v[c].y - p.y) * (T)0.5);
          }
        }
        const Mode<T> md = mode<T, N>(ix, iy, iz);
        const C<T> d = mkc<T>(md.kx * v[0].x + md.ky * v[1].x + md.kz * v[2].x,
                              md.kx * v[0].y + md.ky * v[1].y + md.kz * v[2].y);
        const C<T> pk = cscale<T>(d, md.invlap);
        const C<T> b0 = cadd<T>(v[0], cscale<T>(pk, md.kx));
        const C<T> b1 = cadd<T>(v[1], cscale<T>(pk, md.ky));
        const C<T> b2 = cadd<T>(v[2], cscale<T>(pk, md.kz));
        if (h) { o[0].b = b0; o[1].b = b1; o[2].b = b2; } else { o[0].a = b0; o[1].a = b1; o[2].a = b2; }
        mx = fmax(mx, fmax((double)b0.x, fmax((double)b1.x, (double)b2.x)));
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) st32(reinterpret_cast<C2<T>*>(uk + c * fs), e2, o[c]);
    }
  }
  const double bm = blockMax<256>(mx, sh);
  if (threadIdx.x == 0) maxpart[blockIdx.x] = bm;
}

// Density RHS in spectral space (forced-dns-sm.py:325-330):
//   kn_hat = (-i k.(nv)_hat*mask - fc_hat*mask) * mask, then x-inverse.
// In: 4 fields (nv_x, nv_y, nv_z, fc) forward-transformed along z and y;
// out: field 0 (x-inverse of kn_hat).
// Four field groups (4 * M * N1 threads) forward-transform one input each;
// the pointwise stage runs on the LDS spectra (modes r = g, g+4, ... of each
// thread's line) and group 0 inverse-transforms the result.  Lines per block
// keep the four LDS regions <= 132 KB.
template <typename T, int N, int M0> struct XM4 {
  static constexpr int v = (4L * M0 * (N + 1) * (long)sizeof(C<T>) <= 135168L) ? M0 : M0 / 2;
};
template <typename T, int N, int M>
__global__ void __launch_bounds__(4 * M * Plan<N>::N1)
k_x_kn4(C<T>* w0, size_t fs, const C<T>* __restrict__ tw, T scale, size_t rs, int kcut) {
  C<T>* w = rsel1(w0, rs);
  constexpr int N1 = Plan<N>::N1, N2 = Plan<N>::N2, Nf = N / 2 + 1, js = N * Nf, G = M * N1;
  __shared__ C<T> lds[4 * M * (N + 1)];
  if (!xBlockLive<N, M>(xPairBlock() * M, kcut)) return;   // zero lines, nothing stored
  const int g = threadIdx.x / G, loc = threadIdx.x % G;
  const int m = loc % M, n1 = loc / M;
  const Twiddles<T, N> wt(n1, tw);         // line-FFT twiddles of n1, both directions
  const int q = xPairBlock() * M + m;
  const int iy = q / Nf, iz = q % Nf;
  C<T>* L0 = lds + m * (N + 1);
  C<T>* L1 = lds + (M + m) * (N + 1);
  C<T>* L2 = lds + (2 * M + m) * (N + 1);
  C<T>* L3 = lds + (3 * M + m) * (N + 1);
  C<T>* Lg = lds + (g * M + m) * (N + 1);
  // kcut < N/2: kn_hat is masked, so lines with ky^2 + kz^2 > kcut (kcut + 1)
  // read no input (the previous y pass wrote none there) and give zero, which
  // is not written (the y pass skips those entries, the z pass takes kz > kcut
  // as zero without reading it)
  const int kyl = kwave(iy, N);
  const bool live = kcut >= N / 2 || kyl * kyl + iz * iz <= kcut * (kcut + 1);
  C<T> a[N2];
  const C<T>* W = w + g * fs;
  if (live) {
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2) a[n2] = ld32(W, (uint32_t)(q + posIn<N>(n1, n2) * js));
  } else {
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2) a[n2] = mkc<T>((T)0, (T)0);
  }
  lineFwdT<T, N, 0>(a, Lg, n1, true, wt);
  __syncthreads();
#pragma unroll
  for (int r = 0; r < N2; ++r) Lg[posOut<N>(n1, r)] = a[r];
  __syncthreads();
  for (int r = g; r < N2; r += 4) {
    const int ix = posOut<N>(n1, r);
    const Mode<T> md = mode<T, N>(ix, iy, iz);
    C<T> acc = cscale<T>(cscale<T>(L0[ix], md.mask), md.kx);
    acc = cadd<T>(acc, cscale<T>(cscale<T>(L1[ix], md.mask), md.ky));
    acc = cadd<T>(acc, cscale<T>(cscale<T>(L2[ix], md.mask), md.kz));
    const C<T> mi = mkc<T>(acc.y, -acc.x);   // -i * (k . nv)
    L0[ix] = cscale<T>(csub<T>(mi, cscale<T>(L3[ix], md.mask)), md.mask);
  }
  __syncthreads();
  if (g == 0) {
#pragma unroll
    for (int r = 0; r < N2; ++r) a[r] = L0[posOut<N>(n1, r)];
  }
  lineInvT<T, N, 0>(a, L0, n1, g == 0, wt);
  if (g == 0 && live) {
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2) st32(w, (uint32_t)(q + posIn<N>(n1, n2) * js), cscale<T>(a[n2], scale));
  }
}


// ======================================================== small kernels
// Shell energies of the forced shells and the constant-power factors
// (forced-dns-sm.py:243-254).  One block; loops over the modes with every
// |component| <= smax, which contains every shell s <= smax.
//   ek[s]  = sum over kint == s of 0.5 |u_hat|^2 normalize
//   fac[s] = f0 / (2 ek[s]) for s in shells (0 if |ek[s]| < 1e-10), times fscale
template <typename T, int N>
__global__ void __launch_bounds__(256)
k_shell(const C<T>* __restrict__ uk0, size_t fs, int smax, const int* __restrict__ shells, int nsh,
        double f0, double fscale, double norm0, double* __restrict__ ek0, double* __restrict__ fac0, size_t rs) {
  const C<T>* __restrict__ uk = rsel1(uk0, rs);
  double* __restrict__ ek = rsel1(ek0, rs);
  double* __restrict__ fac = rsel1(fac0, rs);
  constexpr int NT = 256;
  __shared__ double sh[NT];
  constexpr int Nf = N / 2 + 1;
  const int w = 2 * smax + 1;
  const int nmodes = w * w * (smax + 1);
  double part[9];
#pragma unroll
  for (int s = 0; s < 9; ++s) part[s] = 0.0;
  for (int e = threadIdx.x; e < nmodes; e += NT) {
    const int kz = e % (smax + 1), rem = e / (smax + 1);
    const int ky = rem % w - smax, kx = rem / w - smax;
    const int k2 = kx * kx + ky * ky + kz * kz;
    const int ki = kround(k2);
    if (ki > smax || ki >= 9) continue;
    const int ix = kx < 0 ? kx + N : kx, iy = ky < 0 ? ky + N : ky;
    const size_t o = ((size_t)ix * N + iy) * Nf + kz;
    const double nrm = (kz == 0 || kz == N / 2) ? norm0 : 2.0 * norm0;
    double e3 = 0.0;
    for (int c = 0; c < 3; ++c) {
      const C<T> v = uk[c * fs + o];
      e3 += (double)v.x * (double)v.x + (double)v.y * (double)v.y;
    }
    const double val = 0.5 * e3 * nrm;
#pragma unroll
    for (int s = 0; s < 9; ++s) if (s == ki) part[s] += val;
  }
  for (int s = 0; s <= smax && s < 9; ++s) {
    const double tot = blockSum<NT>(part[s], sh);
    if (threadIdx.x == 0) ek[s] = tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    for (int i = 0; i < nsh; ++i) {
      const int s = shells[i];
      const double e = ek[s];
      fac[s] = (fabs(e) < 1e-10) ? 0.0 : fscale * (f0 / (2.0 * e));
    }
  }
}

// Two sums of clip_zero over a real field: sum(x) and sum(max(x, 0))
// (forced-dns-sm.py:281, :283), per block.
template <typename T>
__global__ void __launch_bounds__(256)
k_sum2(const T* __restrict__ x0, size_t n, double* __restrict__ partials0, size_t rs) {
  const T* __restrict__ x = rsel1(x0, rs);
  double* __restrict__ partials = rsel1(partials0, rs);
  constexpr int NT = 256;
  __shared__ double sh[NT];
  // block b sums quads [qb, qe) of 4 consecutive entries (16-byte FP32
  // loads), four quads per thread in flight; the last block adds the n % 4 tail
  const size_t nq = n / 4, chunk = (nq + gridDim.x - 1) / gridDim.x;
  const size_t qb = blockIdx.x * chunk, qe = (qb + chunk < nq) ? qb + chunk : nq;
  const V4<T>* xq = reinterpret_cast<const V4<T>*>(x);
  double s0 = 0.0, s1 = 0.0;
  constexpr int U = 4;
  for (size_t i0 = qb + threadIdx.x; i0 < qe; i0 += U * NT) {
    V4<T> v[U];
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (i0 + u * NT < qe) v[u] = xq[i0 + u * NT];
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (i0 + u * NT < qe) {
        const double a[4] = {(double)v[u].x, (double)v[u].y, (double)v[u].z, (double)v[u].w};
#pragma unroll
        for (int k = 0; k < 4; ++k) { s0 += a[k]; s1 += a[k] > 0.0 ? a[k] : 0.0; }
      }
  }
  if (blockIdx.x == gridDim.x - 1)
    for (size_t i = 4 * nq + threadIdx.x; i < n; i += NT) {
      const double v = (double)x[i];
      s0 += v; s1 += v > 0.0 ? v : 0.0;
    }
  const double a = blockSum<NT>(s0, sh);
  if (threadIdx.x == 0) partials[2 * blockIdx.x] = a;
  __syncthreads();
  const double c = blockSum<NT>(s1, sh);
  if (threadIdx.x == 0) partials[2 * blockIdx.x + 1] = c;
}

// Final clip_zero means: clip = {sum/N^3, sum+/N^3, 1}.
static __global__ void __launch_bounds__(256)
k_clip_final(const double* __restrict__ partials0, int nblk, double inv_n3, double* __restrict__ clip0, size_t rs) {
  const double* __restrict__ partials = rsel1(partials0, rs);
  double* __restrict__ clip = rsel1(clip0, rs);
  constexpr int NT = 256;
  __shared__ double sh[NT];
  double s0 = 0.0, s1 = 0.0;
  for (int i = threadIdx.x; i < nblk; i += NT) { s0 += partials[2 * i]; s1 += partials[2 * i + 1]; }
  const double a = blockSum<NT>(s0, sh);
  __syncthreads();
  const double b = blockSum<NT>(s1, sh);
  if (threadIdx.x == 0) { clip[0] = a * inv_n3; clip[1] = b * inv_n3; clip[2] = 1.0; }
}

// out = clip_zero(x) with x clipped in place when clip_x (forced-dns-sm.py:282-285).
template <typename T>
__global__ void k_clip_apply(T* x0, T* out0, size_t n, const double* __restrict__ clip0, int clip_x, size_t rs) {
  T* x = rsel1(x0, rs);
  T* out = rsel1(out0, rs);
  const double* __restrict__ clip = rsel1(clip0, rs);
  const double m0 = clip[0], m1 = clip[1];
  // quads of 4 entries (16-byte FP32 accesses), four per thread in flight
  const size_t nq = n / 4, stride = (size_t)gridDim.x * blockDim.x;
  V4<T>* xq = reinterpret_cast<V4<T>*>(x);
  V4<T>* oq = reinterpret_cast<V4<T>*>(out);
  constexpr int U = 4;
  for (size_t i0 = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i0 < nq; i0 += U * stride) {
    V4<T> v[U];
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (i0 + u * stride < nq) v[u] = xq[i0 + u * stride];
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (i0 + u * stride < nq) {
        V4<T> c;
        c.x = v[u].x > (T)0 ? v[u].x : (T)0; c.y = v[u].y > (T)0 ? v[u].y : (T)0;
        c.z = v[u].z > (T)0 ? v[u].z : (T)0; c.w = v[u].w > (T)0 ? v[u].w : (T)0;
        if (clip_x) xq[i0 + u * stride] = c;
        V4<T> r;
        r.x = (T)(((double)c.x * m0) / m1); r.y = (T)(((double)c.y * m0) / m1);
        r.z = (T)(((double)c.z * m0) / m1); r.w = (T)(((double)c.w * m0) / m1);
        oq[i0 + u * stride] = r;
      }
  }
  for (size_t i = 4 * nq + blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += stride) {
    const T v = x[i];
    const T c = v > (T)0 ? v : (T)0;
    if (clip_x) x[i] = c;
    out[i] = (T)(((double)c * m0) / m1);
  }
}

// Blow-up test: flag |= max Re(u_hat) > thresh (forced-dns-sm.py:618); the
// per-step maxima are kept for diagnostics.
static __global__ void k_max_final(const double* __restrict__ part0, int nblk, double thresh, int* flag0,
                                   double* lastmax0, size_t rs) {
  const double* __restrict__ part = rsel1(part0, rs);
  int* flag = rsel1(flag0, rs);
  double* lastmax = rsel1(lastmax0, rs);
  constexpr int NT = 256;
  __shared__ double sh[NT];
  double m = -INFINITY;
  for (int i = threadIdx.x; i < nblk; i += NT) m = fmax(m, part[i]);
  const double b = blockMax<NT>(m, sh);
  if (threadIdx.x == 0) { *lastmax = b; if (b > thresh) *flag = 1; }
}

// forcing(uk, fk) as a standalone pass (forced-dns-sm.py:262-275), given the
// per-shell factors from k_shell.
template <typename T, int N>
__global__ void k_forcing_apply(const C<T>* __restrict__ uk, C<T>* __restrict__ fk, size_t fs,
                                const double* __restrict__ fac) {
  constexpr int Nf = N / 2 + 1;
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= fs) return;
  const int iz = (int)(i % Nf), rem = (int)(i / Nf), iy = rem % N, ix = rem / N;
  const Mode<T> md = mode<T, N>(ix, iy, iz);
  const T F = (T)fac[md.kint];
  C<T> f0 = cscale<T>(uk[i], F * md.mask), f1 = cscale<T>(uk[fs + i], F * md.mask), f2 = cscale<T>(uk[2 * fs + i], F * md.mask);
  const C<T> dv = mkc<T>(md.kx * f0.x + md.ky * f1.x + md.kz * f2.x, md.kx * f0.y + md.ky * f1.y + md.kz * f2.y);
  const C<T> pf = cscale<T>(dv, md.invlap * md.mask);
  fk[i] = cscale<T>(cadd<T>(f0, cscale<T>(pf, md.kx)), md.mask);
  fk[fs + i] = cscale<T>(cadd<T>(f1, cscale<T>(pf, md.ky)), md.mask);
  fk[2 * fs + i] = cscale<T>(cadd<T>(f2, cscale<T>(pf, md.kz)), md.mask);
}

template <typename T>
__global__ void k_fill(T* x0, size_t n, T v, size_t rs) {
  T* x = rsel1(x0, rs);
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) x[i] = v;
}
template <typename T>
__global__ void k_relu(T* x, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const T v = x[i];
    if (!(v > (T)0)) x[i] = (T)0;
  }
}

}  // namespace ld
