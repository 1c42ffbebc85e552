// Shared device/host helpers for the lucky-droplets MI355X step library.
//
// Complex numbers are float2/double2 (interleaved re, im), matching numpy's
// complex64/complex128 memory layout so buffers cross the C-ABI unchanged.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>

namespace ld {

template <typename T> struct cvec;
template <> struct cvec<float> { using type = float2; };
template <> struct cvec<double> { using type = double2; };
template <typename T> using C = typename cvec<T>::type;

template <typename T>
__host__ __device__ __forceinline__ C<T> mkc(T x, T y) { C<T> r; r.x = x; r.y = y; return r; }
template <typename T>
__device__ __forceinline__ C<T> cadd(C<T> a, C<T> b) { return mkc<T>(a.x + b.x, a.y + b.y); }
template <typename T>
__device__ __forceinline__ C<T> csub(C<T> a, C<T> b) { return mkc<T>(a.x - b.x, a.y - b.y); }
template <typename T>
__device__ __forceinline__ C<T> cmul(C<T> a, C<T> b) {
  return mkc<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T>
__device__ __forceinline__ C<T> cconj(C<T> a) { return mkc<T>(a.x, -a.y); }
template <typename T>
__device__ __forceinline__ C<T> cscale(C<T> a, T s) { return mkc<T>(a.x * s, a.y * s); }
// multiply by +i
template <typename T>
__device__ __forceinline__ C<T> cmuli(C<T> a) { return mkc<T>(-a.y, a.x); }

// cos(2*pi*j/32) for j = 0..8; sin(2*pi*j/32) = cos(2*pi*(8-j)/32).
__host__ __device__ constexpr double kCos32(int j) {
  return j == 0 ? 1.0 : j == 1 ? 0.9807852804032304 : j == 2 ? 0.9238795325112867
       : j == 3 ? 0.8314696123025452 : j == 4 ? 0.7071067811865476 : j == 5 ? 0.5555702330196022
       : j == 6 ? 0.3826834323650898 : j == 7 ? 0.19509032201612828 : 0.0;
}
// cos and sin of 2*pi*j/32 for 0 <= j < 16.
__host__ __device__ constexpr double cos32(int j) { return j <= 8 ? kCos32(j) : -kCos32(16 - j); }
__host__ __device__ constexpr double sin32(int j) { return j <= 8 ? kCos32(8 - j) : kCos32(j - 8); }

// Wavenumber of FFT index i on an N-point axis (numpy fftfreq(N, 1/N)).
__device__ __forceinline__ int kwave(int i, int N) { return i < N / 2 ? i : i - N; }

// round(|k|) for integer |k|^2 < 2^24: m = floor(sqrt(k2)) from sqrtf with one
// integer correction each way (sqrtf is within one unit of the exact root
// there), then |k| rounds up iff k2 > m^2 + m (|k| is never m + 1/2 exactly
// when |k|^2 is an integer).
__device__ __forceinline__ int kround(int k2) {
  int m = (int)sqrtf((float)k2);
  m -= (m * m > k2) ? 1 : 0;
  m += ((m + 1) * (m + 1) <= k2) ? 1 : 0;
  return (k2 > m * m + m) ? m + 1 : m;
}

}  // namespace ld
