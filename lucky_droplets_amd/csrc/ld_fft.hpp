// Line-FFT building blocks for CDNA4 (gfx950), wavefront = 64 lanes.
//
// A length-N line (N = 2^p, 16 <= N <= 512) is split N = N1 * N2 with
// N1 = 2^floor(p/2) threads per line, each holding N2 = N / N1 elements in
// registers.  One transform is
//     step 1: N2-point DFT in registers (thread n1 owns x[n1 + N1*n2]),
//     inter-step twiddle W_N^(n1*k2),
//     one exchange through LDS,
//     step 2: N2/N1 DFTs of N1 points in registers.
// Forward output is left in "pos_out" order, and the inverse consumes that
// order directly (the steps run backwards), so a forward transform, a
// pointwise spectral operation and an inverse transform need no extra
// reordering.  See posIn()/posOut().
#pragma once
#include "ld_common.hpp"
#include <type_traits>

namespace ld {

// Q = 0: the default split N1 = 2^floor(p/2); Q > 0: N1 = Q threads per line
// (a power of two with Q * Q <= N, so that N2 >= N1).
template <int N, int Q = 0> struct Plan {
  static constexpr int lg = (N == 16) ? 4 : (N == 32) ? 5 : (N == 64) ? 6 : (N == 128) ? 7
                          : (N == 256) ? 8 : (N == 512) ? 9 : -1;
  static_assert(lg > 0, "N must be a power of two in [16, 512]");
  static constexpr int N1 = Q ? Q : 1 << (lg / 2);
  static constexpr int N2 = N / N1;
  static_assert(N1 >= 2 && N2 % N1 == 0, "split N = N1 * N2 with N1 <= N2");
  static constexpr int S = N2 / N1;     // step-2 DFTs per thread
  static constexpr int Nf = N / 2 + 1;
};

// ---------------------------------------------------------------- compile-time loop
template <int I, int E> struct SFor {
  template <class F> static __device__ __forceinline__ void run(F&& f) {
    f(std::integral_constant<int, I>{});
    SFor<I + 1, E>::run(f);
  }
};
template <int E> struct SFor<E, E> {
  template <class F> static __device__ __forceinline__ void run(F&&) {}
};

// v * exp(DIR * 2*pi*i * K / R), R <= 32, 0 <= K < R/2, all compile-time.
template <typename T, int R, int DIR, int K>
__device__ __forceinline__ C<T> twiddleR(C<T> v) {
  constexpr int j = K * (32 / R);
  if constexpr (j == 0) {
    return v;
  } else if constexpr (j == 8) {
    // exp(DIR * i*pi/2) = DIR * i
    if constexpr (DIR < 0) return mkc<T>(v.y, -v.x);
    else return mkc<T>(-v.y, v.x);
  } else {
    constexpr T c = (T)cos32(j);
    constexpr T s = (T)(DIR * sin32(j));
    return mkc<T>(v.x * c - v.y * s, v.x * s + v.y * c);
  }
}

// In-register R-point DFT, natural order in and out (radix-2 decimation in time).
template <typename T, int R, int DIR> struct Dft {
  static __device__ __forceinline__ void run(C<T>* a) {
    C<T> e[R / 2], o[R / 2];
#pragma unroll
    for (int i = 0; i < R / 2; ++i) { e[i] = a[2 * i]; o[i] = a[2 * i + 1]; }
    Dft<T, R / 2, DIR>::run(e);
    Dft<T, R / 2, DIR>::run(o);
    SFor<0, R / 2>::run([&](auto kc) {
      constexpr int k = decltype(kc)::value;
      C<T> t = twiddleR<T, R, DIR, k>(o[k]);
      a[k] = cadd<T>(e[k], t);
      a[k + R / 2] = csub<T>(e[k], t);
    });
  }
};
template <typename T, int DIR> struct Dft<T, 1, DIR> {
  static __device__ __forceinline__ void run(C<T>*) {}
};
template <typename T, int DIR> struct Dft<T, 2, DIR> {
  static __device__ __forceinline__ void run(C<T>* a) {
    C<T> t = a[1];
    a[1] = csub<T>(a[0], t);
    a[0] = cadd<T>(a[0], t);
  }
};

// Element index held in register r by thread t after a forward transform.
template <int N, int Q = 0> __device__ __forceinline__ int posOut(int t, int r) {
  constexpr int N1 = Plan<N, Q>::N1, N2 = Plan<N, Q>::N2;
  return (t + N1 * (r / N1)) + N2 * (r % N1);
}
// Element index held in register n2 by thread n1 before a forward transform.
template <int N, int Q = 0> __device__ __forceinline__ int posIn(int n1, int n2) {
  return n1 + Plan<N, Q>::N1 * n2;
}

// LDS swizzle of the exchange: element (k2, n1) of a line lives at
// k2*N1 + ((n1 + k2) mod N1).
template <int N, int Q = 0> __device__ __forceinline__ int xchg(int k2, int n1) {
  constexpr int N1 = Plan<N, Q>::N1;
  return k2 * N1 + ((n1 + k2) & (N1 - 1));
}

// Forward DFT (exp(-2 pi i jk/N)) of one line.  On entry register n2 holds
// x[posIn(n1, n2)]; on exit register r holds X[posOut(n1, r)].  L is this
// line's LDS scratch (N elements).  Every thread of the block must call it
// (it contains block barriers); threads with active == false only sync.
// The inter-step twiddle W_N^(n1 k2) comes from twd(k2): a load of the table
// tw[m] = exp(-2 pi i m / N) (lineFwd), or a Twiddles source (below).
template <typename T, int N, int Q, class TW>
__device__ __forceinline__ void lineFwdT(C<T>* a, C<T>* L, int n1, bool active, TW&& twd) {
  constexpr int N1 = Plan<N, Q>::N1, N2 = Plan<N, Q>::N2, S = Plan<N, Q>::S;
  if (active) {
    Dft<T, N2, -1>::run(a);
#pragma unroll
    for (int k2 = 1; k2 < N2; ++k2) {
      if (n1 != 0) a[k2] = cmul<T>(a[k2], twd(k2));
    }
  }
  __syncthreads();
  if (active) {
#pragma unroll
    for (int k2 = 0; k2 < N2; ++k2) L[xchg<N, Q>(k2, n1)] = a[k2];
  }
  __syncthreads();
  if (active) {
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const int k2 = n1 + N1 * s;
#pragma unroll
      for (int j = 0; j < N1; ++j) a[s * N1 + j] = L[xchg<N, Q>(k2, j)];
      Dft<T, N1, -1>::run(a + s * N1);
    }
  }
}
template <typename T, int N, int Q = 0>
__device__ __forceinline__ void lineFwd(C<T>* a, C<T>* L, int n1, bool active, const C<T>* __restrict__ tw) {
  lineFwdT<T, N, Q>(a, L, n1, active, [&](int k2) { return tw[n1 * k2]; });
}
// Twiddles of thread n1 for both directions, loaded into registers.
template <typename T, int N, int Q = 0>
__device__ __forceinline__ void loadTwiddles(C<T>* w, int n1, const C<T>* __restrict__ tw) {
#pragma unroll
  for (int k2 = 1; k2 < Plan<N, Q>::N2; ++k2) w[k2] = tw[n1 * k2];
}

// Inverse DFT (exp(+2 pi i jk/N), unnormalised) of one line: consumes the
// posOut order, produces the posIn order.  Same barrier rules as lineFwd;
// twd(k2) as there (the conjugate is taken here).
template <typename T, int N, int Q, class TW>
__device__ __forceinline__ void lineInvT(C<T>* a, C<T>* L, int n1, bool active, TW&& twd) {
  constexpr int N1 = Plan<N, Q>::N1, N2 = Plan<N, Q>::N2, S = Plan<N, Q>::S;
  if (active) {
#pragma unroll
    for (int s = 0; s < S; ++s) Dft<T, N1, +1>::run(a + s * N1);
  }
  __syncthreads();
  if (active) {
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const int k2 = n1 + N1 * s;
#pragma unroll
      for (int j = 0; j < N1; ++j) L[xchg<N, Q>(k2, j)] = a[s * N1 + j];
    }
  }
  __syncthreads();
  if (active) {
#pragma unroll
    for (int k2 = 0; k2 < N2; ++k2) a[k2] = L[xchg<N, Q>(k2, n1)];
#pragma unroll
    for (int k2 = 1; k2 < N2; ++k2) {
      if (n1 != 0) a[k2] = cmul<T>(a[k2], cconj<T>(twd(k2)));
    }
    Dft<T, N2, +1>::run(a);
  }
}
// PRE: the twiddles are loaded on entry, so their latency overlaps the first
// step and the exchange instead of following the second barrier (2 N2 more
// registers live across the exchange).
template <typename T, int N, int Q = 0, bool PRE = false>
__device__ __forceinline__ void lineInv(C<T>* a, C<T>* L, int n1, bool active, const C<T>* __restrict__ tw) {
  if constexpr (PRE) {
    C<T> w[Plan<N, Q>::N2];
    loadTwiddles<T, N, Q>(w, active ? n1 : 0, tw);
    lineInvT<T, N, Q>(a, L, n1, active, [&](int k2) { return w[k2]; });
  } else {
    lineInvT<T, N, Q>(a, L, n1, active, [&](int k2) { return tw[n1 * k2]; });
  }
}

// Thread n1's twiddles for a kernel that runs several line transforms: held
// in registers, loaded once at kernel entry so their latency hides behind the
// first global loads, when the set fits in 32 VGPRs (N2 complex values of at
// most 128 bytes); above that the registers would cost occupancy or spill,
// and each step reads the table instead.  Pass to lineFwdT / lineInvT.
template <typename T, int N, bool REG = (Plan<N>::N2 * sizeof(C<T>) <= 128)>
struct Twiddles;
template <typename T, int N> struct Twiddles<T, N, true> {
  C<T> w[Plan<N>::N2];
  __device__ __forceinline__ Twiddles(int n1, const C<T>* __restrict__ tw) { loadTwiddles<T, N>(w, n1, tw); }
  __device__ __forceinline__ C<T> operator()(int k2) const { return w[k2]; }
};
template <typename T, int N> struct Twiddles<T, N, false> {
  const C<T>* __restrict__ tw;
  int n1;
  __device__ __forceinline__ Twiddles(int n1_, const C<T>* __restrict__ tw_) : tw(tw_), n1(n1_) {}
  __device__ __forceinline__ C<T> operator()(int k2) const { return tw[n1 * k2]; }
};

}  // namespace ld
