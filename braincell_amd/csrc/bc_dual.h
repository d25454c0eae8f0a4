// Forward-mode dual numbers for the Jacobians of the linearly implicit
// point-neuron steps.  The rate templates of bc_channels.h run unchanged on
// bc_dual<real> (value v, derivative d with respect to one seed variable),
// so every rate's derivative with respect to V comes from the same formula
// that gives its value.
#pragma once

#include "bc_channels.h"

template <typename T>
struct bc_dual {
  T v, d;
  __device__ __forceinline__ bc_dual() : v(T(0)), d(T(0)) {}
  // constants (literals in the rate formulas and parameters) have d = 0
  __device__ __forceinline__ explicit bc_dual(double c) : v(T(c)), d(T(0)) {}
  static __device__ __forceinline__ bc_dual of(T v, T d) {
    bc_dual r;
    r.v = v;
    r.d = d;
    return r;
  }
};

template <typename T>
__device__ __forceinline__ bc_dual<T> operator+(bc_dual<T> a, bc_dual<T> b) {
  return bc_dual<T>::of(a.v + b.v, a.d + b.d);
}
template <typename T>
__device__ __forceinline__ bc_dual<T> operator-(bc_dual<T> a, bc_dual<T> b) {
  return bc_dual<T>::of(a.v - b.v, a.d - b.d);
}
template <typename T>
__device__ __forceinline__ bc_dual<T> operator-(bc_dual<T> a) {
  return bc_dual<T>::of(-a.v, -a.d);
}
template <typename T>
__device__ __forceinline__ bc_dual<T> operator*(bc_dual<T> a, bc_dual<T> b) {
  return bc_dual<T>::of(a.v * b.v, a.d * b.v + a.v * b.d);
}
template <typename T>
__device__ __forceinline__ bool operator<(bc_dual<T> a, bc_dual<T> b) { return a.v < b.v; }
template <typename T>
__device__ __forceinline__ bool operator<=(bc_dual<T> a, bc_dual<T> b) { return a.v <= b.v; }
template <typename T>
__device__ __forceinline__ bool operator>(bc_dual<T> a, bc_dual<T> b) { return a.v > b.v; }
template <typename T>
__device__ __forceinline__ bool operator>=(bc_dual<T> a, bc_dual<T> b) { return a.v >= b.v; }
template <typename T>
__device__ __forceinline__ bc_dual<T> fabs(bc_dual<T> a) { return a.v < T(0) ? -a : a; }

// Math policy for the rate templates: the exponential of the value (the
// library one by default, or that of the policy B), chained into the
// derivative.
template <typename T, typename B = bc_math<T>>
struct bc_dmath {
  static __device__ __forceinline__ bc_dual<T> ex(bc_dual<T> x) {
    const T e = B::ex(x.v);
    return bc_dual<T>::of(e, e * x.d);
  }
};

template <>
__device__ __forceinline__ bc_dual<float> bc_rcp<bc_dual<float>>(bc_dual<float> x) {
  const float r = bc_rcp<float>(x.v);
  return bc_dual<float>::of(r, -x.d * r * r);
}
template <>
__device__ __forceinline__ bc_dual<double> bc_rcp<bc_dual<double>>(bc_dual<double> x) {
  const double r = 1.0 / x.v;
  return bc_dual<double>::of(r, -x.d * r * r);
}
