// Forward-mode dual numbers for device code: a value and N tangents.
//
// The potential's device functions are templates on their scalar type; with
// double they are the plain value path, with Dual<N> they also carry the
// derivatives along N seeded directions (a direction in position, or a
// potential parameter).  Comparisons and branches use the value only, which
// is how jax.grad treats where/clip/max: the branch taken at the value point
// is the one differentiated.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>

namespace smx {

template <int N>
struct Dual {
  double v;
  double d[N];
};

using ::sqrt;
using ::log;
using ::exp;
using ::pow;
using ::fabs;

__device__ __forceinline__ double val(double a) { return a; }
template <int N> __device__ __forceinline__ double val(const Dual<N>& a) { return a.v; }

template <int N> __device__ __forceinline__ Dual<N> dconst(double a) {
  Dual<N> r;
  r.v = a;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = 0.0;
  return r;
}
// value a with tangent 1 in slot k
template <int N> __device__ __forceinline__ Dual<N> dseed(double a, int k) {
  Dual<N> r = dconst<N>(a);
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = (i == k) ? 1.0 : 0.0;
  return r;
}

// a constant of scalar type S
template <class S> struct Lift;
template <> struct Lift<double> {
  __device__ static __forceinline__ double c(double a) { return a; }
};
template <int N> struct Lift<Dual<N>> {
  __device__ static __forceinline__ Dual<N> c(double a) { return dconst<N>(a); }
};

// chain rule helper: f(a) with f'(a) = g
// f(a) with f'(a) = g.  A tangent that is exactly zero stays zero even where
// g is infinite (sqrt or log at 0): that is the derivative reverse mode
// (jax.grad) takes of sqrt(maximum(x, 0)) at a clamped x, where maximum's
// select passes no cotangent, whereas 0 * inf would make the tangent NaN.
template <int N> __device__ __forceinline__ Dual<N> dchain(const Dual<N>& a, double f, double g) {
  Dual<N> r;
  r.v = f;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] == 0.0) ? 0.0 : g * a.d[i];
  return r;
}

template <int N> __device__ __forceinline__ Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r;
  r.v = a.v + b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i];
  return r;
}
template <int N> __device__ __forceinline__ Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r;
  r.v = a.v - b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i];
  return r;
}
template <int N> __device__ __forceinline__ Dual<N> operator-(const Dual<N>& a) {
  Dual<N> r;
  r.v = -a.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = -a.d[i];
  return r;
}
template <int N> __device__ __forceinline__ Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r;
  r.v = a.v * b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
  return r;
}
template <int N> __device__ __forceinline__ Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r;
  const double ib = 1.0 / b.v;
  r.v = a.v / b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) * ib;
  return r;
}
// mixed with double
template <int N> __device__ __forceinline__ Dual<N> operator+(const Dual<N>& a, double b) {
  Dual<N> r = a;
  r.v = a.v + b;
  return r;
}
template <int N> __device__ __forceinline__ Dual<N> operator+(double a, const Dual<N>& b) {
  Dual<N> r = b;
  r.v = a + b.v;
  return r;
}
template <int N> __device__ __forceinline__ Dual<N> operator-(const Dual<N>& a, double b) {
  Dual<N> r = a;
  r.v = a.v - b;
  return r;
}
template <int N> __device__ __forceinline__ Dual<N> operator-(double a, const Dual<N>& b) {
  Dual<N> r;
  r.v = a - b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = -b.d[i];
  return r;
}
template <int N> __device__ __forceinline__ Dual<N> operator*(const Dual<N>& a, double b) {
  Dual<N> r;
  r.v = a.v * b;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b;
  return r;
}
template <int N> __device__ __forceinline__ Dual<N> operator*(double a, const Dual<N>& b) {
  Dual<N> r;
  r.v = a * b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a * b.d[i];
  return r;
}
template <int N> __device__ __forceinline__ Dual<N> operator/(const Dual<N>& a, double b) {
  Dual<N> r;
  r.v = a.v / b;
  const double ib = 1.0 / b;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * ib;
  return r;
}
template <int N> __device__ __forceinline__ Dual<N> operator/(double a, const Dual<N>& b) {
  Dual<N> r;
  r.v = a / b.v;
  const double g = -r.v / b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = g * b.d[i];
  return r;
}
template <int N, class T> __device__ __forceinline__ Dual<N>& operator+=(Dual<N>& a, const T& b) {
  a = a + b;
  return a;
}
template <int N, class T> __device__ __forceinline__ Dual<N>& operator-=(Dual<N>& a, const T& b) {
  a = a - b;
  return a;
}
template <int N, class T> __device__ __forceinline__ Dual<N>& operator*=(Dual<N>& a, const T& b) {
  a = a * b;
  return a;
}

template <int N> __device__ __forceinline__ bool operator>(const Dual<N>& a, double b) { return a.v > b; }
template <int N> __device__ __forceinline__ bool operator<(const Dual<N>& a, double b) { return a.v < b; }
template <int N> __device__ __forceinline__ bool operator>=(const Dual<N>& a, double b) { return a.v >= b; }
template <int N> __device__ __forceinline__ bool operator<=(const Dual<N>& a, double b) { return a.v <= b; }

template <int N> __device__ __forceinline__ Dual<N> sqrt(const Dual<N>& a) {
  const double s = ::sqrt(a.v);
  return dchain(a, s, 0.5 / s);
}
template <int N> __device__ __forceinline__ Dual<N> log(const Dual<N>& a) {
  return dchain(a, ::log(a.v), 1.0 / a.v);
}
template <int N> __device__ __forceinline__ Dual<N> exp(const Dual<N>& a) {
  const double e = ::exp(a.v);
  return dchain(a, e, e);
}
template <int N> __device__ __forceinline__ Dual<N> pow(const Dual<N>& a, double p) {
  const double f = ::pow(a.v, p);
  return dchain(a, f, p * ::pow(a.v, p - 1.0));
}
// |a| with derivative sign(a) (0 at 0)
template <int N> __device__ __forceinline__ Dual<N> fabs(const Dual<N>& a) {
  const double s = a.v > 0.0 ? 1.0 : (a.v < 0.0 ? -1.0 : 0.0);
  return dchain(a, ::fabs(a.v), s);
}

}  // namespace smx
