// Element types of the supernodal LU: real FP64, real FP32 (mixed precision)
// and complex FP64 (small-signal AC systems G + j w C).
//
// Num<T> gives, for factor element type T, the type of the matrix values the
// factorization reads (D), the type the triangular solves accumulate in (A),
// and the real type of pivot magnitudes (R).  Complex pivots are ranked by
// |re| + |im|, the measure LAPACK's zgetrf uses (cabs1).
#pragma once
#include <hip/hip_runtime.h>

namespace vxm {

struct cplx {
  double re, im;
  cplx() = default;
  __host__ __device__ constexpr cplx(double r, double i = 0.0) : re(r), im(i) {}
};

__device__ __forceinline__ cplx operator+(cplx a, cplx b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cplx operator-(cplx a, cplx b) { return {a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ cplx operator-(cplx a) { return {-a.re, -a.im}; }
__device__ __forceinline__ cplx operator*(cplx a, cplx b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
// Smith's algorithm (Smith 1962): scales by the larger part of the divisor.
__device__ __forceinline__ cplx operator/(cplx a, cplx b) {
  if (fabs(b.re) >= fabs(b.im)) {
    const double r = b.im / b.re, d = b.re + b.im * r;
    return {(a.re + a.im * r) / d, (a.im - a.re * r) / d};
  }
  const double r = b.re / b.im, d = b.im + b.re * r;
  return {(a.re * r + a.im) / d, (a.im * r - a.re) / d};
}
__device__ __forceinline__ cplx& operator+=(cplx& a, cplx b) { return a = a + b; }
__device__ __forceinline__ cplx& operator-=(cplx& a, cplx b) { return a = a - b; }
__device__ __forceinline__ cplx& operator/=(cplx& a, cplx b) { return a = a / b; }

template <typename T>
struct Num;
template <>
struct Num<double> {
  using D = double;
  using A = double;
  using R = double;
};
template <>
struct Num<float> {
  using D = double;
  using A = double;
  using R = float;
};
template <>
struct Num<cplx> {
  using D = cplx;
  using A = cplx;
  using R = double;
};

// pivot magnitude, NaN mapped to +inf so a NaN column is picked (and propagates)
__device__ __forceinline__ double pmag(double x) { return x != x ? INFINITY : fabs(x); }
__device__ __forceinline__ float pmag(float x) { return x != x ? INFINITY : fabsf(x); }
__device__ __forceinline__ double pmag(cplx z) {
  const double m = fabs(z.re) + fabs(z.im);
  return m != m ? INFINITY : m;
}

// value conversions between element, data and accumulation types
template <typename To>
__device__ __forceinline__ To cvt(double x) { return static_cast<To>(x); }
template <typename To>
__device__ __forceinline__ To cvt(float x) { return static_cast<To>(x); }
template <typename To>
__device__ __forceinline__ To cvt(cplx x) { return x; }

__device__ __forceinline__ double shfl(double v, int lane, int width) {
  return __shfl(v, lane, width);
}
__device__ __forceinline__ float shfl(float v, int lane, int width) {
  return __shfl(v, lane, width);
}
__device__ __forceinline__ cplx shfl(cplx v, int lane, int width) {
  return {__shfl(v.re, lane, width), __shfl(v.im, lane, width)};
}

}  // namespace vxm
