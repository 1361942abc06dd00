// stencil.h -- periodic central finite-difference stencils for gfx950.
//
// Coefficients are those of the reference's compile-time-order stencils
// (utils/math.h): first derivative (math.h:74-100, 231-263, 412-450,
// 554-586), pure second derivative (math.h:873-901, 958-1085), diagonal-only
// mixed second derivative (math.h:676-855) and Kreiss-Oliger dissipation
// (math.h:36-72, orders 2 and 8 only).  Here they are applied to an LDS tile
// with compile-time strides; 1/dx and 1/dx^2 are applied by the caller.
#pragma once

#include <hip/hip_runtime.h>

namespace cg {

// first derivative: sum_m c1(m) * (f(+m) - f(-m))
template <int R>
__host__ __device__ constexpr double c1(int m) {
  return R == 1 ? 0.5
       : R == 2 ? (m == 1 ? 2.0 / 3.0 : -1.0 / 12.0)
       : R == 3 ? (m == 1 ? 3.0 / 4.0 : m == 2 ? -3.0 / 20.0 : 1.0 / 60.0)
                : (m == 1 ? 4.0 / 5.0 : m == 2 ? -1.0 / 5.0 : m == 3 ? 4.0 / 105.0 : -1.0 / 280.0);
}
// pure second derivative: c2(0) f(0) + sum_m c2(m) * (f(+m) + f(-m))
template <int R>
__host__ __device__ constexpr double c2(int m) {
  return R == 1 ? (m == 0 ? -2.0 : 1.0)
       : R == 2 ? (m == 0 ? -5.0 / 2.0 : m == 1 ? 4.0 / 3.0 : -1.0 / 12.0)
       : R == 3 ? (m == 0 ? -49.0 / 18.0 : m == 1 ? 3.0 / 2.0 : m == 2 ? -3.0 / 20.0 : 1.0 / 90.0)
                : (m == 0 ? -205.0 / 72.0 : m == 1 ? 8.0 / 5.0 : m == 2 ? -1.0 / 5.0
                   : m == 3 ? 8.0 / 315.0 : -1.0 / 560.0);
}
// diagonal mixed derivative: sum_m cm(m) * Q(m),
// Q(m) = f(+m,+m) - f(+m,-m) - f(-m,+m) + f(-m,-m)
template <int R>
__host__ __device__ constexpr double cm(int m) {
  return R == 1 ? 0.25
       : R == 2 ? (m == 1 ? 1.0 / 3.0 : -1.0 / 48.0)
       : R == 3 ? (m == 1 ? 135.0 / 360.0 : m == 2 ? -13.5 / 360.0 : 1.0 / 360.0)
                : (m == 1 ? 2.0 / 5.0 : m == 2 ? -1.0 / 20.0 : m == 3 ? 2.0 / 315.0 : -1.0 / 2240.0);
}

template <int R, typename T>
__device__ __forceinline__ T fd1(const T* s, int c, int st) {
  T acc = T(c1<R>(1)) * (s[c + st] - s[c - st]);
#pragma unroll
  for (int m = 2; m <= R; ++m) acc += T(c1<R>(m)) * (s[c + m * st] - s[c - m * st]);
  return acc;
}

template <int R, typename T>
__device__ __forceinline__ T fd2(const T* s, int c, int st) {
  T acc = T(c2<R>(0)) * s[c];
#pragma unroll
  for (int m = 1; m <= R; ++m) acc += T(c2<R>(m)) * (s[c + m * st] + s[c - m * st]);
  return acc;
}

template <int R, typename T>
__device__ __forceinline__ T fdm(const T* s, int c, int sa, int sb) {
  T acc = T(0);
#pragma unroll
  for (int m = 1; m <= R; ++m) {
    const int p = m * (sa + sb), q = m * (sa - sb);
    acc += T(cm<R>(m)) * ((s[c + p] - s[c + q]) - (s[c - q] - s[c - p]));
  }
  return acc;
}

// The same stencils on a tile of 32-bit fixed-point words (the conformal
// metric of the fixed-point storage mode), for the FP32 algebra: first
// differences q(+m) - q(-m), second differences about the centre
// q(+m) + q(-m) - 2 q(0) (the stencil weights sum to zero) and the mixed
// cross differences are formed exactly in 32-bit wrap-around arithmetic
// (exact whenever the difference itself fits in int32, which it does for a
// field sampled finely enough to be differentiated), converted, and summed
// with FP32 FMA; the caller applies the field's scale.  The FP32 sums keep
// the relative precision of the derivative itself; FP32 storage of the
// values would instead lose it to the cancellation of neighbouring values.
__device__ __forceinline__ float qd(int a, int b) { return (float)(int)((unsigned)a - (unsigned)b); }
// (a - c0) + (b - c0) as (a + b) - 2 c0: the same value in wrap-around
// arithmetic, one operation fewer (c2 = 2 c0 once per stencil)
__device__ __forceinline__ float qdd(int a, int b, unsigned c2) {
  return (float)(int)(((unsigned)a + (unsigned)b) - c2);
}

template <int R>
__device__ __forceinline__ float fd1qf(const int* s, int c, int st) {
  float acc = float(c1<R>(1)) * qd(s[c + st], s[c - st]);
#pragma unroll
  for (int m = 2; m <= R; ++m) acc += float(c1<R>(m)) * qd(s[c + m * st], s[c - m * st]);
  return acc;
}

template <int R>
__device__ __forceinline__ float fd2qf(const int* s, int c, int st) {
  const unsigned q2 = 2u * (unsigned)s[c];
  float acc = float(c2<R>(1)) * qdd(s[c + st], s[c - st], q2);
#pragma unroll
  for (int m = 2; m <= R; ++m) acc += float(c2<R>(m)) * qdd(s[c + m * st], s[c - m * st], q2);
  return acc;
}

template <int R>
__device__ __forceinline__ float fdmqf(const int* s, int c, int sa, int sb) {
  float acc = 0.0f;
#pragma unroll
  for (int m = 1; m <= R; ++m) {
    const int p = m * (sa + sb), q = m * (sa - sb);
    const unsigned d = ((unsigned)s[c + p] - (unsigned)s[c + q]) - ((unsigned)s[c - q] - (unsigned)s[c - p]);
    acc += float(cm<R>(m)) * (float)(int)d;
  }
  return acc;
}

// Kreiss-Oliger stencil sum (without prefactor) over the three axes.
// order 2: [1,-4,6,-4,1];  order 8: [1,-10,45,-120,210,-252,...]
template <int ORDER, typename T>
__device__ __forceinline__ T ko_sum(const T* s, int c, int si, int sj, int sk) {
  if constexpr (ORDER == 2) {
    constexpr double w[3] = {6.0, -4.0, 1.0};
    T acc = T(3.0 * w[0]) * s[c];
#pragma unroll
    for (int m = 1; m <= 2; ++m)
      acc += T(w[m]) * ((s[c + m * si] + s[c - m * si]) + (s[c + m * sj] + s[c - m * sj]) +
                        (s[c + m * sk] + s[c - m * sk]));
    return acc;
  } else if constexpr (ORDER == 8) {
    constexpr double w[6] = {-252.0, 210.0, -120.0, 45.0, -10.0, 1.0};
    T acc = T(3.0 * w[0]) * s[c];
#pragma unroll
    for (int m = 1; m <= 5; ++m)
      acc += T(w[m]) * ((s[c + m * si] + s[c - m * si]) + (s[c + m * sj] + s[c - m * sj]) +
                        (s[c + m * sk] + s[c - m * sk]));
    return acc;
  } else {
    return T(0);
  }
}

// ko_sum on a fixed-point tile, in FP64
template <int ORDER>
__device__ __forceinline__ double ko_sumq(const int* s, int c, int si, int sj, int sk) {
  if constexpr (ORDER == 2 || ORDER == 8) {
    constexpr int NM = ORDER == 2 ? 2 : 5;
    constexpr double w2[3] = {6.0, -4.0, 1.0};
    constexpr double w8[6] = {-252.0, 210.0, -120.0, 45.0, -10.0, 1.0};
    auto w = [&](int m) { return ORDER == 2 ? w2[m] : w8[m]; };
    double acc = 3.0 * w(0) * (double)s[c];
#pragma unroll
    for (int m = 1; m <= NM; ++m)
      acc += w(m) * ((((double)s[c + m * si] + (double)s[c - m * si]) + ((double)s[c + m * sj] + (double)s[c - m * sj])) +
                     ((double)s[c + m * sk] + (double)s[c - m * sk]));
    return acc;
  } else {
    return 0.0;
  }
}

}  // namespace cg
