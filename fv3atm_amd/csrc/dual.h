// dual.h -- forward-mode derivative numbers for the tangent-linear and
// adjoint kernels.  Dual<N> carries a value and N partial derivatives.
// Comparisons, min/max selections and sign tests act on the value only, so
// a piecewise-linear limiter evaluated on Dual<N> follows the branches of
// the forward trajectory and returns their (frozen-branch) derivatives.
#pragma once
#include <hip/hip_runtime.h>

namespace fv3 {

// keep the double overloads visible next to the Dual ones below
using ::copysign;
using ::fabs;

template <int N>
struct Dual {
  double v;
  double d[N];
  __host__ __device__ constexpr Dual() : v(0.0), d{} {}
  __host__ __device__ constexpr Dual(double x) : v(x), d{} {}
};

#define FV3_DUAL_LOOP for (int n_ = 0; n_ < N; n_++)

template <int N>
__host__ __device__ __forceinline__ Dual<N> operator+(const Dual<N> &a, const Dual<N> &b) {
  Dual<N> r(a.v + b.v);
  FV3_DUAL_LOOP r.d[n_] = a.d[n_] + b.d[n_];
  return r;
}
template <int N>
__host__ __device__ __forceinline__ Dual<N> operator-(const Dual<N> &a, const Dual<N> &b) {
  Dual<N> r(a.v - b.v);
  FV3_DUAL_LOOP r.d[n_] = a.d[n_] - b.d[n_];
  return r;
}
template <int N>
__host__ __device__ __forceinline__ Dual<N> operator-(const Dual<N> &a) {
  Dual<N> r(-a.v);
  FV3_DUAL_LOOP r.d[n_] = -a.d[n_];
  return r;
}
template <int N>
__host__ __device__ __forceinline__ Dual<N> operator*(const Dual<N> &a, const Dual<N> &b) {
  Dual<N> r(a.v * b.v);
  FV3_DUAL_LOOP r.d[n_] = a.d[n_] * b.v + a.v * b.d[n_];
  return r;
}
template <int N>
__host__ __device__ __forceinline__ Dual<N> operator/(const Dual<N> &a, const Dual<N> &b) {
  const double q = a.v / b.v;
  Dual<N> r(q);
  FV3_DUAL_LOOP r.d[n_] = (a.d[n_] - q * b.d[n_]) / b.v;
  return r;
}
template <int N> __host__ __device__ __forceinline__ Dual<N> operator+(const Dual<N> &a, double b) { return a + Dual<N>(b); }
template <int N> __host__ __device__ __forceinline__ Dual<N> operator+(double a, const Dual<N> &b) { return Dual<N>(a) + b; }
template <int N> __host__ __device__ __forceinline__ Dual<N> operator-(const Dual<N> &a, double b) { return a - Dual<N>(b); }
template <int N> __host__ __device__ __forceinline__ Dual<N> operator-(double a, const Dual<N> &b) { return Dual<N>(a) - b; }
template <int N> __host__ __device__ __forceinline__ Dual<N> operator*(const Dual<N> &a, double b) {
  Dual<N> r(a.v * b);
  FV3_DUAL_LOOP r.d[n_] = a.d[n_] * b;
  return r;
}
template <int N> __host__ __device__ __forceinline__ Dual<N> operator*(double a, const Dual<N> &b) { return b * a; }
template <int N> __host__ __device__ __forceinline__ Dual<N> operator/(const Dual<N> &a, double b) {
  Dual<N> r(a.v / b);
  FV3_DUAL_LOOP r.d[n_] = a.d[n_] / b;
  return r;
}
template <int N> __host__ __device__ __forceinline__ Dual<N> operator/(double a, const Dual<N> &b) { return Dual<N>(a) / b; }

template <int N> __host__ __device__ __forceinline__ bool operator<(const Dual<N> &a, const Dual<N> &b) { return a.v < b.v; }
template <int N> __host__ __device__ __forceinline__ bool operator>(const Dual<N> &a, const Dual<N> &b) { return a.v > b.v; }
template <int N> __host__ __device__ __forceinline__ bool operator<=(const Dual<N> &a, const Dual<N> &b) { return a.v <= b.v; }
template <int N> __host__ __device__ __forceinline__ bool operator>=(const Dual<N> &a, const Dual<N> &b) { return a.v >= b.v; }

template <int N>
__host__ __device__ __forceinline__ Dual<N> fabs(const Dual<N> &a) {
  return a.v < 0.0 ? -a : a;
}
// Fortran SIGN semantics on the value: |a| with the sign of b
template <int N>
__host__ __device__ __forceinline__ Dual<N> copysign(const Dual<N> &a, const Dual<N> &b) {
  const Dual<N> m = fabs(a);
  return b.v < 0.0 || (b.v == 0.0 && __builtin_signbit(b.v)) ? -m : m;
}

#undef FV3_DUAL_LOOP

// value and single tangent
using D1 = Dual<1>;

}  // namespace fv3
