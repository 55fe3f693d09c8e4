// fr_dual.hpp -- forward-mode dual numbers for the adjoint kernels.
//
// Dual<N> carries a value and N partial derivatives.  The pointwise physics
// (fluxes, common fluxes, ghost states) are templates on the scalar type, so
// instantiating them with Dual<N> gives their exact Jacobians with respect
// to N seeded inputs from the same code the residual kernels run.  The
// adjoint kernels contract those Jacobians with cotangents (J^T lambda).
// Branches (max, abs, the characteristic boundary's cases) follow the value,
// so the derivative is the one of the branch the residual takes.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace frdev {

// keep the scalar overloads visible next to the Dual ones in this namespace
using ::fabs;
using ::fmax;
using ::fmin;
using ::pow;
using ::sqrt;

template <int N>
struct Dual {
  double v;
  double d[N];
  __host__ __device__ Dual() : v(0.0) {
#pragma unroll
    for (int i = 0; i < N; i++) d[i] = 0.0;
  }
  __host__ __device__ Dual(double x) : v(x) {
#pragma unroll
    for (int i = 0; i < N; i++) d[i] = 0.0;
  }
  // the variable x with derivative 1 in direction `seed`
  __host__ __device__ static Dual var(double x, int seed) {
    Dual r(x);
#pragma unroll
    for (int i = 0; i < N; i++) r.d[i] = (i == seed) ? 1.0 : 0.0;
    return r;
  }
  __host__ __device__ Dual &operator+=(const Dual &b) {
    v += b.v;
#pragma unroll
    for (int i = 0; i < N; i++) d[i] += b.d[i];
    return *this;
  }
  __host__ __device__ Dual &operator-=(const Dual &b) {
    v -= b.v;
#pragma unroll
    for (int i = 0; i < N; i++) d[i] -= b.d[i];
    return *this;
  }
  __host__ __device__ Dual &operator*=(const Dual &b) {
#pragma unroll
    for (int i = 0; i < N; i++) d[i] = d[i] * b.v + v * b.d[i];
    v *= b.v;
    return *this;
  }
  __host__ __device__ Dual &operator/=(const Dual &b) {
    const double inv = 1.0 / b.v, q = v * inv;
#pragma unroll
    for (int i = 0; i < N; i++) d[i] = (d[i] - q * b.d[i]) * inv;
    v = q;
    return *this;
  }
};

template <int N>
__host__ __device__ inline Dual<N> operator-(const Dual<N> &a) {
  Dual<N> r;
  r.v = -a.v;
#pragma unroll
  for (int i = 0; i < N; i++) r.d[i] = -a.d[i];
  return r;
}
template <int N>
__host__ __device__ inline Dual<N> operator+(Dual<N> a, const Dual<N> &b) {
  return a += b;
}
template <int N>
__host__ __device__ inline Dual<N> operator-(Dual<N> a, const Dual<N> &b) {
  return a -= b;
}
template <int N>
__host__ __device__ inline Dual<N> operator*(Dual<N> a, const Dual<N> &b) {
  return a *= b;
}
template <int N>
__host__ __device__ inline Dual<N> operator/(Dual<N> a, const Dual<N> &b) {
  return a /= b;
}
template <int N>
__host__ __device__ inline Dual<N> operator+(Dual<N> a, double b) {
  a.v += b;
  return a;
}
template <int N>
__host__ __device__ inline Dual<N> operator+(double b, Dual<N> a) {
  a.v += b;
  return a;
}
template <int N>
__host__ __device__ inline Dual<N> operator-(Dual<N> a, double b) {
  a.v -= b;
  return a;
}
template <int N>
__host__ __device__ inline Dual<N> operator-(double b, const Dual<N> &a) {
  return Dual<N>(b) - a;
}
template <int N>
__host__ __device__ inline Dual<N> operator*(Dual<N> a, double b) {
  a.v *= b;
#pragma unroll
  for (int i = 0; i < N; i++) a.d[i] *= b;
  return a;
}
template <int N>
__host__ __device__ inline Dual<N> operator*(double b, Dual<N> a) {
  return a * b;
}
template <int N>
__host__ __device__ inline Dual<N> operator/(Dual<N> a, double b) {
  return a / Dual<N>(b);
}
template <int N>
__host__ __device__ inline Dual<N> operator/(double b, const Dual<N> &a) {
  return Dual<N>(b) / a;
}

#define FR_DUAL_CMP(op)                                                                        \
  template <int N>                                                                             \
  __host__ __device__ inline bool operator op(const Dual<N> &a, const Dual<N> &b) {            \
    return a.v op b.v;                                                                         \
  }                                                                                            \
  template <int N>                                                                             \
  __host__ __device__ inline bool operator op(const Dual<N> &a, double b) {                    \
    return a.v op b;                                                                           \
  }                                                                                            \
  template <int N>                                                                             \
  __host__ __device__ inline bool operator op(double a, const Dual<N> &b) {                    \
    return a op b.v;                                                                           \
  }
FR_DUAL_CMP(<)
FR_DUAL_CMP(>)
FR_DUAL_CMP(<=)
FR_DUAL_CMP(>=)
#undef FR_DUAL_CMP

template <int N>
__host__ __device__ inline Dual<N> sqrt(const Dual<N> &a) {
  Dual<N> r;
  r.v = ::sqrt(a.v);
  const double h = (r.v > 0.0) ? 0.5 / r.v : 0.0;
#pragma unroll
  for (int i = 0; i < N; i++) r.d[i] = a.d[i] * h;
  return r;
}
template <int N>
__host__ __device__ inline Dual<N> fabs(const Dual<N> &a) {
  return (a.v < 0.0) ? -a : a;
}
template <int N>
__host__ __device__ inline Dual<N> fmax(const Dual<N> &a, const Dual<N> &b) {
  return (b.v > a.v) ? b : a;
}
template <int N>
__host__ __device__ inline Dual<N> fmin(const Dual<N> &a, const Dual<N> &b) {
  return (b.v < a.v) ? b : a;
}
// a^b = exp(b ln a): d = a^b (b' ln a + b a' / a)
template <int N>
__host__ __device__ inline Dual<N> pow(const Dual<N> &a, const Dual<N> &b) {
  Dual<N> r;
  r.v = ::pow(a.v, b.v);
  const double la = (a.v > 0.0) ? ::log(a.v) : 0.0;
  const double ea = (a.v != 0.0) ? b.v / a.v : 0.0;
#pragma unroll
  for (int i = 0; i < N; i++) r.d[i] = r.v * (b.d[i] * la + ea * a.d[i]);
  return r;
}

}  // namespace frdev
