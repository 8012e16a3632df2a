// Forward-mode dual numbers for device code: value plus K tangents.
//
// The reference obtains forces as -jax.grad(potential) (model.py:152-157), the
// exact derivative of the potential's expression graph.  Evaluating the same
// templated potential with Dual<3> seeded on (x, y, z) gives that derivative
// in one forward pass, with no tape.
#pragma once
#include <hip/hip_runtime.h>

namespace smx {

template <int K>
struct Dual {
  double v;
  double g[K];
};

template <int K>
__device__ __forceinline__ Dual<K> dconst(double c) {
  Dual<K> r;
  r.v = c;
#pragma unroll
  for (int i = 0; i < K; ++i) r.g[i] = 0.0;
  return r;
}

#define SMX_DUAL_BIN(OP, VAL, GRAD)                                                   \
  template <int K>                                                                   \
  __device__ __forceinline__ Dual<K> operator OP(const Dual<K>& a, const Dual<K>& b) { \
    Dual<K> r;                                                                       \
    r.v = VAL;                                                                       \
    _Pragma("unroll") for (int i = 0; i < K; ++i) r.g[i] = GRAD;                     \
    return r;                                                                        \
  }
SMX_DUAL_BIN(+, a.v + b.v, a.g[i] + b.g[i])
SMX_DUAL_BIN(-, a.v - b.v, a.g[i] - b.g[i])
SMX_DUAL_BIN(*, a.v * b.v, a.g[i] * b.v + a.v * b.g[i])
#undef SMX_DUAL_BIN

template <int K>
__device__ __forceinline__ Dual<K> operator/(const Dual<K>& a, const Dual<K>& b) {
  Dual<K> r;
  r.v = a.v / b.v;
  const double ib = 1.0 / b.v;
#pragma unroll
  for (int i = 0; i < K; ++i) r.g[i] = (a.g[i] - r.v * b.g[i]) * ib;
  return r;
}

template <int K>
__device__ __forceinline__ Dual<K> operator-(const Dual<K>& a) {
  Dual<K> r;
  r.v = -a.v;
#pragma unroll
  for (int i = 0; i < K; ++i) r.g[i] = -a.g[i];
  return r;
}

// mixed with plain doubles
template <int K>
__device__ __forceinline__ Dual<K> operator+(const Dual<K>& a, double b) {
  Dual<K> r = a;
  r.v = a.v + b;
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator+(double b, const Dual<K>& a) {
  Dual<K> r = a;
  r.v = b + a.v;
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator-(const Dual<K>& a, double b) {
  Dual<K> r = a;
  r.v = a.v - b;
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator-(double b, const Dual<K>& a) {
  Dual<K> r;
  r.v = b - a.v;
#pragma unroll
  for (int i = 0; i < K; ++i) r.g[i] = -a.g[i];
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator*(const Dual<K>& a, double b) {
  Dual<K> r;
  r.v = a.v * b;
#pragma unroll
  for (int i = 0; i < K; ++i) r.g[i] = a.g[i] * b;
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator*(double b, const Dual<K>& a) {
  Dual<K> r;
  r.v = b * a.v;
#pragma unroll
  for (int i = 0; i < K; ++i) r.g[i] = b * a.g[i];
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator/(const Dual<K>& a, double b) {
  Dual<K> r;
  r.v = a.v / b;
#pragma unroll
  for (int i = 0; i < K; ++i) r.g[i] = a.g[i] / b;
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator/(double b, const Dual<K>& a) {
  Dual<K> r;
  r.v = b / a.v;
  const double f = -r.v / a.v;
#pragma unroll
  for (int i = 0; i < K; ++i) r.g[i] = f * a.g[i];
  return r;
}

// elementary functions; plain-double overloads keep templated code uniform
__device__ __forceinline__ double dsqrt(double a) { return sqrt(a); }
__device__ __forceinline__ double dlog(double a) { return log(a); }
__device__ __forceinline__ double dval(double a) { return a; }

template <int K>
__device__ __forceinline__ Dual<K> dsqrt(const Dual<K>& a) {
  Dual<K> r;
  r.v = sqrt(a.v);
  const double f = 0.5 / r.v;
#pragma unroll
  for (int i = 0; i < K; ++i) r.g[i] = a.g[i] * f;
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> dlog(const Dual<K>& a) {
  Dual<K> r;
  r.v = log(a.v);
#pragma unroll
  for (int i = 0; i < K; ++i) r.g[i] = a.g[i] / a.v;
  return r;
}
template <int K>
__device__ __forceinline__ double dval(const Dual<K>& a) { return a.v; }

}  // namespace smx
