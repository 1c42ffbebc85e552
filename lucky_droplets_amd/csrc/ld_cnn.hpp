// Learned subgrid form: a small periodic 3-D CNN on a coarse grid, trained and
// run with hand-written bf16 MFMA kernels (FP32 accumulation).  No
// counterpart in the reference.
//
// Layout: activations channels-last, A[node][C] in bf16, node = (x N + y) N + z,
// C a multiple of 16; weights per layer W[tap][co][ci] (and the transposed
// copy Wt[tap][ci][co] for the data gradient), tap = (dx+1) 9 + (dy+1) 3 +
// (dz+1) for offsets in {-1, 0, 1}; periodic neighbours.
//
// A 3x3x3 convolution is an implicit GEMM: M = nodes, K = 27 taps x CI,
// N = CO.  One wavefront computes 16 nodes x CO outputs with
// __builtin_amdgcn_mfma_f32_16x16x16bf16_1k, whose operand layout is
//   A (16x16, M x K): lane l holds A[l % 16][4 (l / 16) + i], i = 0..3
//   B (16x16, K x N): lane l holds B[4 (l / 16) + i][l % 16]
//   D (16x16):        lane l holds D[4 (l / 16) + i][l % 16]
// so a lane reads 4 consecutive channels (8 bytes) of one node.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace ld {
namespace cnn {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint16_t bf16;

// FP32 -> bf16, round to nearest even (NaN kept quiet).
__device__ __forceinline__ bf16 f2b(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (bf16)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (bf16)(u >> 16);
}
__device__ __forceinline__ float b2f(bf16 b) { return __uint_as_float((uint32_t)b << 16); }

__device__ __forceinline__ int nbr(int node, int tap, int N, bool flip) {
  int dx = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dz = tap % 3 - 1;
  if (flip) { dx = -dx; dy = -dy; dz = -dz; }
  const int z = node & (N - 1), y = (node / N) & (N - 1), x = node / (N * N);
  return (((x + dx) & (N - 1)) * N + ((y + dy) & (N - 1))) * N + ((z + dz) & (N - 1));
}

// out[node][co] = bias[co] + sum_{tap, ci} X[nbr(node, tap)][ci] W[tap][co][ci]
// FLIP: neighbours at -offset (the data gradient, with Wt);
// RELU: max(., 0); MASK: times (mask[node][co] > 0) (ReLU derivative of the
// layer below); Yb / Yf: bf16 / FP32 outputs (either may be null).
// 256 threads = 4 wavefronts x 16 nodes; N^3 must be a multiple of 64.
template <int CI, int CO, bool FLIP, bool RELU, bool MASK>
__global__ void __launch_bounds__(256)
k_conv(const bf16* __restrict__ X, const bf16* __restrict__ W, const float* __restrict__ bias,
       const bf16* __restrict__ mask, bf16* __restrict__ Yb, float* __restrict__ Yf, int N) {
  static_assert(CI % 16 == 0 && CO % 16 == 0, "channels: multiples of 16");
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int node0 = blockIdx.x * 64 + wv * 16;
  const int row = lane & 15, kq = 4 * (lane >> 4);
  f32x4 acc[CO / 16];
#pragma unroll
  for (int t = 0; t < CO / 16; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int tap = 0; tap < 27; ++tap) {
    const bf16* xa = X + (size_t)nbr(node0 + row, tap, N, FLIP) * CI + kq;
    const bf16* wb = W + ((size_t)tap * CO + row) * CI + kq;
#pragma unroll
    for (int cb = 0; cb < CI / 16; ++cb) {
      const bf16x4 a = *reinterpret_cast<const bf16x4*>(xa + cb * 16);
#pragma unroll
      for (int t = 0; t < CO / 16; ++t) {
        const bf16x4 b = *reinterpret_cast<const bf16x4*>(wb + (size_t)t * 16 * CI + cb * 16);
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[t], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int t = 0; t < CO / 16; ++t) {
    const int co = t * 16 + row;
    const float bv = bias ? bias[co] : 0.f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const size_t o = (size_t)(node0 + kq + i) * CO + co;
      float v = acc[t][i] + bv;
      if (RELU) v = v > 0.f ? v : 0.f;
      if (MASK) v = b2f(mask[o]) > 0.f ? v : 0.f;
      if (Yb) Yb[o] = f2b(v);
      if (Yf) Yf[o] = v;
    }
  }
}

// The same convolution with the gfx950 double-rate instruction
// __builtin_amdgcn_mfma_f32_16x16x32_bf16, oriented M = output channels,
// N = nodes, K = (tap, input channel) pairs in steps of 32:
//   A (16x32): lane l holds W[tap][co0 + l % 16][ci0 + i]       (8 bf16, 16 B)
//   B (32x16): lane l holds X[nbr(node0 + l % 16, tap)][ci0 + i] (8 bf16, 16 B)
//   D (16x16): lane l holds Y[node0 + l % 16][co0 + 4 (l / 16) + i], i < 4
// with (tap, ci0) = divmod(32 ks + 8 (l / 16), CI) for K-step ks (CI = 16: a
// step covers two taps; the 28th tap of the last step has zero weights).
// The layer's weights are staged in LDS once per block; a wavefront computes
// NT tiles of 16 consecutive nodes along z (one z-line segment) for all CO,
// so each weight fragment feeds NT instructions and each input fragment CO/16.
// Blocks loop over segments (persistent grid), XCD-aware: each XCD takes a
// contiguous run of blocks, so neighbouring lines share its L2.
// Needs N >= 16 * NT and N a multiple of 16 * NT.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// Writes a wavefront's 16-node x 16-channel output tile (lane l: node n0 +
// l % 16, channels c0 + 4 (l / 16) + i, i < 4, in v) to the node-blocked
// channel-major layout of k_cblock through 512 bytes of LDS owned by the
// wavefront: 32 runs of 8 nodes x 1 channel, one 16-byte store each.  n0 is a
// multiple of 16.  The LDS writes and reads are issued by one wavefront, whose
// LDS operations complete in order.
template <int CO>
__device__ __forceinline__ void store_blocked(bf16* __restrict__ Yblk, bf16* sT, size_t n0, int c0, const uint32_t pk[2]) {
  const int lane = threadIdx.x & 63, n16 = lane & 15, q = lane >> 4;
  sT[(4 * q + 0) * 16 + n16] = (bf16)(pk[0] & 0xffff);
  sT[(4 * q + 1) * 16 + n16] = (bf16)(pk[0] >> 16);
  sT[(4 * q + 2) * 16 + n16] = (bf16)(pk[1] & 0xffff);
  sT[(4 * q + 3) * 16 + n16] = (bf16)(pk[1] >> 16);
  __builtin_amdgcn_wave_barrier();
  if (lane < 32) {
    const int cl = lane >> 1, half = lane & 1;
    const uint4 v = *reinterpret_cast<const uint4*>(sT + cl * 16 + half * 8);
    *reinterpret_cast<uint4*>(Yblk + (((n0 >> 3) + half) * CO + c0 + cl) * 8) = v;
  }
  __builtin_amdgcn_wave_barrier();
}

template <int CI, int CO, bool FLIP, bool RELU, bool MASK, int NT, int WPB>
__global__ void __launch_bounds__(64 * WPB)
k_conv2(const bf16* __restrict__ X, const bf16* __restrict__ W, const float* __restrict__ bias,
        const bf16* __restrict__ mask, bf16* __restrict__ Yb, float* __restrict__ Yf, int N, int nseg,
        bf16* __restrict__ Yblk) {
  static_assert(CI % 16 == 0 && CO % 16 == 0, "channels: multiples of 16");
  constexpr int MT = CO / 16;
  constexpr int KS = (27 * CI + 31) / 32;
  constexpr int WN = 27 * CO * CI;
  __shared__ __attribute__((aligned(16))) bf16 sW[WN];
  __shared__ __attribute__((aligned(16))) bf16 sT[WPB][256];   // per-wavefront tile for the blocked copy
  for (int e = threadIdx.x; e < WN / 8; e += 64 * WPB)
    reinterpret_cast<uint4*>(sW)[e] = reinterpret_cast<const uint4*>(W)[e];
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int n16 = lane & 15, q = lane >> 4;
  const int G = gridDim.x;
  const int lb = (G % 8 == 0) ? (blockIdx.x % 8) * (G / 8) + blockIdx.x / 8 : blockIdx.x;
  const int spl = N / (16 * NT);                  // segments per z-line
  const int msk = N - 1;
  for (int seg = lb * WPB + wv; seg < nseg; seg += G * WPB) {
    const int line = seg / spl, z0 = (seg % spl) * 16 * NT;
    const int x = line / N, y = line & msk;
    f32x4 acc[NT][MT];
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int m = 0; m < MT; ++m) acc[t][m] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 2
    for (int ks = 0; ks < KS; ++ks) {
      const int k = 32 * ks + 8 * q;
      const int tap0 = k / CI, ci0 = k % CI;
      const bool valid = tap0 < 27;
      const int tap = valid ? tap0 : 26;
      int dx = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dz = tap % 3 - 1;
      if (FLIP) { dx = -dx; dy = -dy; dz = -dz; }
      bf16x8 wf[MT], xf[NT];
      // opaque to the compiler, so the weight fragments are re-read from LDS
      // per segment instead of being hoisted into registers for all K-steps
      int wofs = (tap * CO + n16) * CI + ci0;
      asm volatile("" : "+v"(wofs));
#pragma unroll
      for (int m = 0; m < MT; ++m) {
        const uint4 v = *reinterpret_cast<const uint4*>(sW + wofs + m * 16 * CI);
        wf[m] = valid ? __builtin_bit_cast(bf16x8, v) : bf16x8{};
      }
      const size_t row = ((size_t)((x + dx) & msk) * N + ((y + dy) & msk)) * N;
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const int zz = (z0 + 16 * t + n16 + dz) & msk;
        const uint4 v = *reinterpret_cast<const uint4*>(X + (row + zz) * CI + ci0);
        xf[t] = valid ? __builtin_bit_cast(bf16x8, v) : bf16x8{};
      }
#pragma unroll
      for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int m = 0; m < MT; ++m) acc[t][m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf[m], xf[t], acc[t][m], 0, 0, 0);
    }
    const size_t nbase = (size_t)line * N + z0 + n16;
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int m = 0; m < MT; ++m) {
        const int co = m * 16 + 4 * q;
        const size_t o = (nbase + 16 * t) * CO + co;
        float v[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          v[i] = acc[t][m][i] + (bias ? bias[co + i] : 0.f);
          if (RELU) v[i] = v[i] > 0.f ? v[i] : 0.f;
        }
        if (MASK) {
          const uint2 mk = *reinterpret_cast<const uint2*>(mask + o);
          const bf16 mb[4] = {(bf16)(mk.x & 0xffff), (bf16)(mk.x >> 16), (bf16)(mk.y & 0xffff), (bf16)(mk.y >> 16)};
#pragma unroll
          for (int i = 0; i < 4; ++i) v[i] = b2f(mb[i]) > 0.f ? v[i] : 0.f;
        }
        uint32_t pk[2];
        pk[0] = (uint32_t)f2b(v[0]) | ((uint32_t)f2b(v[1]) << 16);
        pk[1] = (uint32_t)f2b(v[2]) | ((uint32_t)f2b(v[3]) << 16);
        if (Yb) *reinterpret_cast<uint2*>(Yb + o) = make_uint2(pk[0], pk[1]);
        if (Yf) *reinterpret_cast<float4*>(Yf + o) = make_float4(v[0], v[1], v[2], v[3]);
        if (Yblk) store_blocked<CO>(Yblk, sT[wv], nbase - n16 + 16 * t, m * 16, pk);
      }
  }
}

// CI = 32 form of k_conv2 that reuses input fragments across the dz taps: per
// neighbour line (dx, dy) a wavefront loads its NT tiles once (plus the node
// before and after the segment) and forms the dz = +1 and -1 operands by
// shifting lanes within each 16-lane row (DPP row_shl:1 / row_shr:1, with the
// row's end lane filled from the neighbouring tile by row_shr:15 / row_shl:15),
// so one 16-byte load feeds 3 MT instructions instead of MT.
template <unsigned CTRL>
__device__ __forceinline__ uint32_t dpp_u(uint32_t old, uint32_t src) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)src, CTRL, 0xf, 0xf, false);
}
// lanes of a 16-lane row take the fragment of the next node (dz = +1): lane i
// <- lane i + 1 of `a`, lane 15 <- lane 0 of `next`
__device__ __forceinline__ uint4 shift_up(const uint4& a, const uint4& next) {
  uint4 r;
  r.x = dpp_u<0x101>(dpp_u<0x11f>(0u, next.x), a.x);
  r.y = dpp_u<0x101>(dpp_u<0x11f>(0u, next.y), a.y);
  r.z = dpp_u<0x101>(dpp_u<0x11f>(0u, next.z), a.z);
  r.w = dpp_u<0x101>(dpp_u<0x11f>(0u, next.w), a.w);
  return r;
}
// dz = -1: lane i <- lane i - 1 of `a`, lane 0 <- lane 15 of `prev`
__device__ __forceinline__ uint4 shift_down(const uint4& a, const uint4& prev) {
  uint4 r;
  r.x = dpp_u<0x111>(dpp_u<0x10f>(0u, prev.x), a.x);
  r.y = dpp_u<0x111>(dpp_u<0x10f>(0u, prev.y), a.y);
  r.z = dpp_u<0x111>(dpp_u<0x10f>(0u, prev.z), a.z);
  r.w = dpp_u<0x111>(dpp_u<0x10f>(0u, prev.w), a.w);
  return r;
}

template <int CO, bool FLIP, bool RELU, bool MASK, int NT, int WPB>
__global__ void __launch_bounds__(64 * WPB)
k_conv3(const bf16* __restrict__ X, const bf16* __restrict__ W, const float* __restrict__ bias,
        const bf16* __restrict__ mask, bf16* __restrict__ Yb, float* __restrict__ Yf, int N, int nseg,
        bf16* __restrict__ Yblk) {
  constexpr int CI = 32, MT = CO / 16;
  constexpr int WN = 27 * CO * CI;
  __shared__ __attribute__((aligned(16))) bf16 sW[WN];
  __shared__ __attribute__((aligned(16))) bf16 sT[WPB][256];   // per-wavefront tile for the blocked copy
  for (int e = threadIdx.x; e < WN / 8; e += 64 * WPB)
    reinterpret_cast<uint4*>(sW)[e] = reinterpret_cast<const uint4*>(W)[e];
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int n16 = lane & 15, q = lane >> 4, ci0 = 8 * q;
  const int G = gridDim.x;
  const int lb = (G % 8 == 0) ? (blockIdx.x % 8) * (G / 8) + blockIdx.x / 8 : blockIdx.x;
  const int spl = N / (16 * NT);
  const int msk = N - 1;
  for (int seg = lb * WPB + wv; seg < nseg; seg += G * WPB) {
    const int line = seg / spl, z0 = (seg % spl) * 16 * NT;
    const int x = line / N, y = line & msk;
    f32x4 acc[NT][MT];
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int m = 0; m < MT; ++m) acc[t][m] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 1
    for (int dxy = 0; dxy < 9; ++dxy) {
      const int ndx = dxy / 3 - 1, ndy = dxy % 3 - 1;     // neighbour line offset
      const size_t row = ((size_t)((x + ndx) & msk) * N + ((y + ndy) & msk)) * N;
      uint4 xc[NT];
#pragma unroll
      for (int t = 0; t < NT; ++t)
        xc[t] = *reinterpret_cast<const uint4*>(X + (row + ((z0 + 16 * t + n16) & msk)) * CI + ci0);
      uint4 blo = make_uint4(0u, 0u, 0u, 0u), bhi = make_uint4(0u, 0u, 0u, 0u);
      if (n16 == 15) blo = *reinterpret_cast<const uint4*>(X + (row + ((z0 - 1) & msk)) * CI + ci0);
      if (n16 == 0) bhi = *reinterpret_cast<const uint4*>(X + (row + ((z0 + 16 * NT) & msk)) * CI + ci0);
#pragma unroll
      for (int ndz = -1; ndz <= 1; ++ndz) {
        // weight tap of this neighbour offset (mirrored for the data gradient)
        const int tap = FLIP ? (1 - ndx) * 9 + (1 - ndy) * 3 + (1 - ndz) : (ndx + 1) * 9 + (ndy + 1) * 3 + (ndz + 1);
        int wofs = (tap * CO + n16) * CI + ci0;
        asm volatile("" : "+v"(wofs));
        bf16x8 wf[MT];
#pragma unroll
        for (int m = 0; m < MT; ++m)
          wf[m] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(sW + wofs + m * 16 * CI));
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          uint4 v;
          if (ndz == 0) v = xc[t];
          else if (ndz > 0) v = shift_up(xc[t], t + 1 < NT ? xc[t + 1] : bhi);
          else v = shift_down(xc[t], t > 0 ? xc[t - 1] : blo);
          const bf16x8 xf = __builtin_bit_cast(bf16x8, v);
#pragma unroll
          for (int m = 0; m < MT; ++m)
            acc[t][m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf[m], xf, acc[t][m], 0, 0, 0);
        }
      }
    }
    const size_t nbase = (size_t)line * N + z0 + n16;
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int m = 0; m < MT; ++m) {
        const int co = m * 16 + 4 * q;
        const size_t o = (nbase + 16 * t) * CO + co;
        float v[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          v[i] = acc[t][m][i] + (bias ? bias[co + i] : 0.f);
          if (RELU) v[i] = v[i] > 0.f ? v[i] : 0.f;
        }
        if (MASK) {
          const uint2 mk = *reinterpret_cast<const uint2*>(mask + o);
          const bf16 mb[4] = {(bf16)(mk.x & 0xffff), (bf16)(mk.x >> 16), (bf16)(mk.y & 0xffff), (bf16)(mk.y >> 16)};
#pragma unroll
          for (int i = 0; i < 4; ++i) v[i] = b2f(mb[i]) > 0.f ? v[i] : 0.f;
        }
        uint32_t pk[2];
        pk[0] = (uint32_t)f2b(v[0]) | ((uint32_t)f2b(v[1]) << 16);
        pk[1] = (uint32_t)f2b(v[2]) | ((uint32_t)f2b(v[3]) << 16);
        if (Yb) *reinterpret_cast<uint2*>(Yb + o) = make_uint2(pk[0], pk[1]);
        if (Yf) *reinterpret_cast<float4*>(Yf + o) = make_float4(v[0], v[1], v[2], v[3]);
        if (Yblk) store_blocked<CO>(Yblk, sT[wv], nbase - n16 + 16 * t, m * 16, pk);
      }
  }
}

// Weight and bias gradients of one layer, per block over `chunk` nodes:
//   part[blk][(tap CO + co) CI + ci] = sum_node dZ[node][co] X[nbr(node, tap)][ci]
//   part[blk][27 CO CI + co]         = sum_node dZ[node][co]
// (M = co, N = ci, K = nodes).  The dZ rows of 64 nodes are staged in LDS for
// the block, the X neighbour rows per wavefront (wavefront w does taps
// w, w + 4, ...).  chunk is a multiple of 64.
template <int CI, int CO>
__global__ void __launch_bounds__(256)
k_wgrad(const bf16* __restrict__ X, const bf16* __restrict__ dZ, float* __restrict__ part, int N, int chunk) {
  constexpr int NT = 7;                       // taps per wavefront (27 = 7 + 7 + 7 + 6)
  __shared__ bf16 sdz[64 * CO];
  __shared__ bf16 sx[4][64 * CI];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int row = lane & 15, kq = 4 * (lane >> 4);
  f32x4 acc[NT][CO / 16][CI / 16];
#pragma unroll
  for (int j = 0; j < NT; ++j)
#pragma unroll
    for (int a = 0; a < CO / 16; ++a)
#pragma unroll
      for (int b = 0; b < CI / 16; ++b) acc[j][a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
  float db = 0.f;
  const size_t nbase = (size_t)blockIdx.x * chunk;
  for (int s = 0; s < chunk; s += 64) {
    __syncthreads();
    for (int e = threadIdx.x; e < 64 * CO; e += 256) sdz[e] = dZ[(nbase + s) * CO + e];
    __syncthreads();
    if (threadIdx.x < CO) {
      for (int n = 0; n < 64; ++n) db += b2f(sdz[n * CO + threadIdx.x]);
    }
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      const int tap = wv + 4 * j;
      if (tap >= 27) break;
      // this lane's node: neighbour row of X
      const int node = (int)(nbase + s) + lane;
      const bf16* src = X + (size_t)nbr(node, tap, N, false) * CI;
#pragma unroll
      for (int c = 0; c < CI; c += 4)
        *reinterpret_cast<bf16x4*>(&sx[wv][lane * CI + c]) = *reinterpret_cast<const bf16x4*>(src + c);
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) {
#pragma unroll
        for (int a = 0; a < CO / 16; ++a) {
          bf16x4 fa;
#pragma unroll
          for (int i = 0; i < 4; ++i) fa[i] = (short)sdz[(kk * 16 + kq + i) * CO + a * 16 + row];
#pragma unroll
          for (int b = 0; b < CI / 16; ++b) {
            bf16x4 fb;
#pragma unroll
            for (int i = 0; i < 4; ++i) fb[i] = (short)sx[wv][(kk * 16 + kq + i) * CI + b * 16 + row];
            acc[j][a][b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(fa, fb, acc[j][a][b], 0, 0, 0);
          }
        }
      }
    }
  }
  float* out = part + (size_t)blockIdx.x * (27 * CO * CI + CO);
#pragma unroll
  for (int j = 0; j < NT; ++j) {
    const int tap = wv + 4 * j;
    if (tap >= 27) break;
#pragma unroll
    for (int a = 0; a < CO / 16; ++a)
#pragma unroll
      for (int b = 0; b < CI / 16; ++b)
#pragma unroll
        for (int i = 0; i < 4; ++i)
          out[((size_t)tap * CO + a * 16 + kq + i) * CI + b * 16 + row] = acc[j][a][b][i];
  }
  if (threadIdx.x < CO) out[27 * CO * CI + threadIdx.x] = db;
}

// Channels-last A[node][C] -> node-blocked channel-major AB (bf16): entry
// (node, c) at ((node / 8) C + c) 8 + node % 8, i.e. for each group of 8
// nodes, C runs of 8 consecutive nodes (16 bytes each).  One thread per group
// and 8 channels: eight 16-byte loads, an 8 x 8 transpose in registers,
// eight 16-byte stores.
template <int C>
__global__ void __launch_bounds__(256)
k_cblock(const bf16* __restrict__ A, bf16* __restrict__ AB, size_t nodes) {
  constexpr int CG = C / 8;
  const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (e >= nodes / 8 * CG) return;
  const size_t g = e / CG;
  const int c8 = (int)(e % CG) * 8;
  uint32_t r[8][4];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const uint4 v = *reinterpret_cast<const uint4*>(A + (g * 8 + i) * C + c8);
    r[i][0] = v.x; r[i][1] = v.y; r[i][2] = v.z; r[i][3] = v.w;
  }
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    uint32_t o[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint32_t a = (r[2 * k][j / 2] >> (16 * (j & 1))) & 0xffffu;
      const uint32_t b = (r[2 * k + 1][j / 2] >> (16 * (j & 1))) & 0xffffu;
      o[k] = a | (b << 16);
    }
    *reinterpret_cast<uint4*>(AB + (g * C + c8 + j) * 8) = make_uint4(o[0], o[1], o[2], o[3]);
  }
}

// Weight and bias gradients from node-blocked channel-major operands
// (k_cblock layout) on the 16x16x32 bf16 MFMA (M = output channels,
// N = input channels, K = nodes):
//   A (16x32): lane l holds dZ[node k0 + 8 (l / 16) + i][co0 + l % 16]  (one 16-byte run)
//   B (32x16): lane l holds X[nbr(node k0 + 8 (l / 16) + i, tap)][ci0 + l % 16]
// so the 16 lanes of a row read 256 consecutive bytes.
// Eight consecutive nodes of a lane lie on one z-line, so their neighbours for
// tap (dx, dy, dz) are eight consecutive z on the neighbour line, shifted by
// dz: the lane loads the aligned 16 bytes at z plus one dword on each side
// (periodic in z) and forms the dz = -1 and +1 operands with v_alignbyte.
// Neighbour line l = 3 (dx + 1) + (dy + 1) carries taps 3 l .. 3 l + 2; a
// wavefront owns LPW consecutive lines.  Block b sums nodes
// [b chunk, (b + 1) chunk) in K-steps of 32 and writes its partial
//   part[b][(tap CO + co) CI + ci],  part[b][27 CO CI + co] (bias, wavefront 0).
// LPW neighbour lines per wavefront (9 / LPW wavefronts per block): LPW = 3
// gives wavefront w the lines (w - 1, -1..1), so the dZ fragments are read by
// three wavefronts instead of nine and each feeds 9 taps.
template <int CI, int CO, int LPW>
__global__ void __launch_bounds__(64 * 9 / LPW)
k_wgrad2(const bf16* __restrict__ XB, const bf16* __restrict__ dZB, float* __restrict__ part, int N, int chunk) {
  constexpr int MT = CO / 16, NT = CI / 16;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int n16 = lane & 15, q = lane >> 4, msk = N - 1;
  f32x4 acc[3 * LPW][MT][NT];
#pragma unroll
  for (int d = 0; d < 3 * LPW; ++d)
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
      for (int t = 0; t < NT; ++t) acc[d][m][t] = f32x4{0.f, 0.f, 0.f, 0.f};
  float db[MT];
#pragma unroll
  for (int m = 0; m < MT; ++m) db[m] = 0.f;
  // one K-step: the lane's 8 nodes start at z = zq on line (x, y) (8-node
  // group gq of the dZ copy); A from dZ, B from the neighbour lines' X copy
  auto kstep = [&](size_t gq, int zq, int x, int y) {
    bf16x8 a[MT];
#pragma unroll
    for (int m = 0; m < MT; ++m) {
      const uint4 v = *reinterpret_cast<const uint4*>(dZB + (gq * CO + m * 16 + n16) * 8);
      a[m] = __builtin_bit_cast(bf16x8, v);
      if (w == 0) {
        const uint32_t u[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int i = 0; i < 4; ++i) db[m] += b2f((bf16)(u[i] & 0xffff)) + b2f((bf16)(u[i] >> 16));
      }
    }
    const int G = N >> 3;                          // 8-node groups per z-line
#pragma unroll
    for (int li = 0; li < LPW; ++li) {
      const int wl = w * LPW + li;                 // neighbour line index 0..8
      const int dx = wl / 3 - 1, dy = wl % 3 - 1;
      const size_t gb = ((size_t)((x + dx) & msk) * N + ((y + dy) & msk)) * G;   // its first group
      const size_t cg = gb + (zq >> 3);
      const size_t pg = zq ? cg - 1 : gb + G - 1;                                // periodic in z
      const size_t ng = zq + 8 < N ? cg + 1 : gb;
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const int ci = t * 16 + n16;
        const uint4 c = *reinterpret_cast<const uint4*>(XB + (cg * CI + ci) * 8);
        const uint32_t pv = *reinterpret_cast<const uint32_t*>(XB + (pg * CI + ci) * 8 + 6);
        const uint32_t nv = *reinterpret_cast<const uint32_t*>(XB + (ng * CI + ci) * 8);
        const uint32_t s1 = __builtin_amdgcn_alignbyte(c.y, c.x, 2), s2 = __builtin_amdgcn_alignbyte(c.z, c.y, 2),
                       s3 = __builtin_amdgcn_alignbyte(c.w, c.z, 2);
        const bf16x8 fm = __builtin_bit_cast(bf16x8, make_uint4(__builtin_amdgcn_alignbyte(c.x, pv, 2), s1, s2, s3));
        const bf16x8 f0 = __builtin_bit_cast(bf16x8, c);
        const bf16x8 fp = __builtin_bit_cast(bf16x8, make_uint4(s1, s2, s3, __builtin_amdgcn_alignbyte(nv, c.w, 2)));
#pragma unroll
        for (int m = 0; m < MT; ++m) {
          acc[3 * li][m][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[m], fm, acc[3 * li][m][t], 0, 0, 0);
          acc[3 * li + 1][m][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[m], f0, acc[3 * li + 1][m][t], 0, 0, 0);
          acc[3 * li + 2][m][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[m], fp, acc[3 * li + 2][m][t], 0, 0, 0);
        }
      }
    }
  };
  const size_t kb = (size_t)blockIdx.x * chunk;
  if (N >= 32) {
    // the chunk is whole z-lines: line coordinates once per line, K-steps
    // along z with plain increments
    const int lines = chunk / N;
    const size_t l0 = kb / N;
    for (int li0 = 0; li0 < lines; ++li0) {
      const size_t line = l0 + li0;
      const int x = (int)(line / N), y = (int)(line & msk);
      const size_t gl = line * (size_t)(N >> 3);
#pragma unroll 2
      for (int z0 = 0; z0 < N; z0 += 32) {
        const int zq = z0 + 8 * q;
        kstep(gl + (zq >> 3), zq, x, y);
      }
    }
  } else {
    for (int ks = 0; ks < chunk / 32; ++ks) {
      const size_t nq = kb + 32 * ks + 8 * q;
      const size_t line = nq / N;
      kstep(nq >> 3, (int)(nq & msk), (int)(line / N), (int)(line & msk));
    }
  }
  float* out = part + (size_t)blockIdx.x * (27 * CO * CI + CO);
#pragma unroll
  for (int d = 0; d < 3 * LPW; ++d) {
    const int tap = 3 * (w * LPW) + d;
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
      for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i)
          out[((size_t)tap * CO + m * 16 + 4 * q + i) * CI + t * 16 + n16] = acc[d][m][t][i];
  }
  if (w == 0) {
#pragma unroll
    for (int m = 0; m < MT; ++m) {
      float v = db[m];
      v += __shfl_xor(v, 16, 64);
      v += __shfl_xor(v, 32, 64);
      if (q == 0) out[27 * CO * CI + m * 16 + n16] = v;
    }
  }
}

// out[i] = sum over blocks of part[b][i] (fixed order), i < n.
static __global__ void __launch_bounds__(256)
k_reduce_parts(const float* __restrict__ part, int nblk, int n, float* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  // loads in batches of 16 ahead of their (ordered) additions
  float s = 0.f;
  int b = 0;
  for (; b + 16 <= nblk; b += 16) {
    float v[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) v[u] = part[(size_t)(b + u) * n + i];
#pragma unroll
    for (int u = 0; u < 16; ++u) s += v[u];
  }
  for (; b < nblk; ++b) s += part[(size_t)b * n + i];
  out[i] = s;
}

// x[node][cin] FP32 -> A[node][CP] bf16 (channels cin..CP-1 zero).
// One node per thread and step, 16-byte stores (CP is 16 or 32).
static __global__ void k_pack_input(const float* __restrict__ x, int cin, int CP, size_t nodes, bf16* __restrict__ A) {
  for (size_t node = blockIdx.x * (size_t)blockDim.x + threadIdx.x; node < nodes;
       node += (size_t)gridDim.x * blockDim.x) {
    for (int c8 = 0; c8 < CP; c8 += 8) {
      uint32_t w[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int c = c8 + 2 * i;
        const uint32_t lo = c < cin ? f2b(x[node * cin + c]) : 0u;
        const uint32_t hi = c + 1 < cin ? f2b(x[node * cin + c + 1]) : 0u;
        w[i] = lo | (hi << 16);
      }
      *reinterpret_cast<uint4*>(A + node * CP + c8) = make_uint4(w[0], w[1], w[2], w[3]);
    }
  }
}

// Master FP32 weights w[tap][co][ci] -> bf16 W[tap][co][ci] and Wt[tap][ci][co].
static __global__ void k_pack_weights(const float* __restrict__ w, int CI, int CO, bf16* __restrict__ Wb,
                                      bf16* __restrict__ Wt) {
  const int n = 27 * CO * CI;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    const int tap = e / (CO * CI), co = (e / CI) % CO, ci = e % CI;
    const bf16 b = f2b(w[e]);
    Wb[e] = b;
    Wt[((size_t)tap * CI + ci) * CO + co] = b;
  }
}

// Mean squared error over nodes x cout: dY[node][co] = 2 (y - t) / (nodes cout)
// for co < cout (0 for padding channels), per-block partial sums of (y - t)^2.
static __global__ void __launch_bounds__(256)
k_mse(const float* __restrict__ y, const float* __restrict__ t, int cout, int CP, size_t nodes,
      bf16* __restrict__ dY, double* __restrict__ part) {
  __shared__ double sh[256];
  double s = 0.0;
  const double scale = 2.0 / ((double)nodes * cout);
  // one node per thread and step, 16-byte stores of dY (CP is 16 or 32)
  for (size_t node = blockIdx.x * (size_t)blockDim.x + threadIdx.x; node < nodes;
       node += (size_t)gridDim.x * blockDim.x) {
    for (int c8 = 0; c8 < CP; c8 += 8) {
      uint32_t w[4];
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const int c = c8 + i;
        float g = 0.f;
        if (c < cout) {
          const double d = (double)y[node * CP + c] - (double)t[node * cout + c];
          s += d * d;
          g = (float)(scale * d);
        }
        const uint32_t b = f2b(g);
        if (i & 1) w[i / 2] |= b << 16; else w[i / 2] = b;
      }
      *reinterpret_cast<uint4*>(dY + node * CP + c8) = make_uint4(w[0], w[1], w[2], w[3]);
    }
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// Adam on the FP32 master parameters (Kingma & Ba 2015), step t >= 1.
static __global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                              float* __restrict__ v, int n, float lr, float b1, float b2, float eps, int t) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float mh = mi / (1.f - powf(b1, (float)t)), vh = vi / (1.f - powf(b2, (float)t));
  w[i] -= lr * mh / (sqrtf(vh) + eps);
}

// Output rows [node][CP] FP32 -> [node][cout].
static __global__ void k_unpack_output(const float* __restrict__ yp, int CP, int cout, size_t nodes, float* __restrict__ y) {
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < nodes * cout; e += (size_t)gridDim.x * blockDim.x)
    y[e] = yp[(e / cout) * CP + e % cout];
}

}  // namespace cnn
}  // namespace ld
