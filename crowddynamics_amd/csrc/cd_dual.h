// Forward-mode dual numbers for the adjoint: a value and K directional
// derivatives carried through +, -, *, /, sqrt, exp and hypot.
// The physics functions of cd_device.h are templates on their scalar type,
// so evaluating one of them on Dual<K> yields the term and its Jacobian
// along K input directions with the very branches, cut-offs and
// truncations of the forward step.  Comparisons act on the value, which is
// computed by the forward's own operations (an IEEE quotient, the library
// hypot); the derivative parts of a quotient or hypot multiply by one
// reciprocal of the divisor rather than dividing K times.
#pragma once

#include "cd_device.h"

namespace cd {

template <int K>
struct Dual {
  double v;
  double d[K];
  __device__ __forceinline__ Dual() : v(0.0) {
#pragma unroll
    for (int k = 0; k < K; ++k) d[k] = 0.0;
  }
  __device__ __forceinline__ Dual(double c) : v(c) {
#pragma unroll
    for (int k = 0; k < K; ++k) d[k] = 0.0;
  }
  // the value c seeded along direction `dir`
  __device__ static __forceinline__ Dual seed(double c, int dir) {
    Dual r(c);
#pragma unroll
    for (int k = 0; k < K; ++k) r.d[k] = (k == dir) ? 1.0 : 0.0;
    return r;
  }
  __device__ __forceinline__ Dual &operator+=(const Dual &b) {
    v += b.v;
#pragma unroll
    for (int k = 0; k < K; ++k) d[k] += b.d[k];
    return *this;
  }
  __device__ __forceinline__ Dual &operator*=(const Dual &b) {
#pragma unroll
    for (int k = 0; k < K; ++k) d[k] = d[k] * b.v + v * b.d[k];
    v *= b.v;
    return *this;
  }
};

template <int K>
__device__ __forceinline__ Dual<K> operator+(const Dual<K> &a, const Dual<K> &b) {
  Dual<K> r;
  r.v = a.v + b.v;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = a.d[k] + b.d[k];
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator-(const Dual<K> &a, const Dual<K> &b) {
  Dual<K> r;
  r.v = a.v - b.v;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = a.d[k] - b.d[k];
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator-(const Dual<K> &a) {
  Dual<K> r;
  r.v = -a.v;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = -a.d[k];
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator*(const Dual<K> &a, const Dual<K> &b) {
  Dual<K> r;
  r.v = a.v * b.v;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator/(const Dual<K> &a, const Dual<K> &b) {
  Dual<K> r;
  r.v = a.v / b.v;
  const double inv = 1.0 / b.v;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) * inv;
  return r;
}

template <int K>
__device__ __forceinline__ bool operator<(const Dual<K> &a, const Dual<K> &b) { return a.v < b.v; }
template <int K>
__device__ __forceinline__ bool operator>(const Dual<K> &a, const Dual<K> &b) { return a.v > b.v; }
template <int K>
__device__ __forceinline__ bool operator<=(const Dual<K> &a, const Dual<K> &b) {
  return a.v <= b.v;
}
template <int K>
__device__ __forceinline__ bool operator==(const Dual<K> &a, const Dual<K> &b) {
  return a.v == b.v;
}
template <int K>
__device__ __forceinline__ bool operator!=(const Dual<K> &a, const Dual<K> &b) {
  return a.v != b.v;
}

template <int K>
__device__ __forceinline__ bool is_nan(const Dual<K> &a) {
  return __builtin_isnan(a.v);
}
template <int K>
__device__ __forceinline__ Dual<K> dsqrt(const Dual<K> &a) {
  Dual<K> r;
  r.v = sqrt(a.v);
  const double s = 0.5 / r.v;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = a.d[k] * s;
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> dexp(const Dual<K> &a) {
  Dual<K> r;
  r.v = exp(a.v);
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = a.d[k] * r.v;
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> dhypot(const Dual<K> &a, const Dual<K> &b) {
  Dual<K> r;
  r.v = hypot_ref(a.v, b.v);
  const double inv = 1.0 / r.v;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = (a.v * a.d[k] + b.v * b.d[k]) * inv;
  return r;
}

}  // namespace cd
