// Value representations ("semirings") of the inside/outside tables.
//
// LogSR -- the plain FP64 path: natural-log values, (+) = log-sum-exp with the
//          reference's exact -inf guards (src/CommonUtils.jl:145-155), (x) = +.
// XF    -- the scaled FP32 path: value = m * 2^e with a float mantissa m in
//          [0.5, 1) and an int32 exponent e per cell (power-of-two scaling per
//          cell), (+) = exponent-aligned float add, (x) = float multiply and
//          exponent add.  No transcendental per term.
//
// Both store 8 bytes per cell.  A bifurcation sum over split points is done in
// two passes: pass 1 takes the maximum "key" of the terms (log value, or
// exponent sum), pass 2 sums term contributions relative to that key.
#pragma once

#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>

namespace messi {

struct LogSR {
  typedef double V;     // stored cell
  typedef double K;     // key of a term
  typedef double S;     // partial sum relative to the key
  typedef double C;     // rule constant (log probability)
  static constexpr bool scaled = false;
  static __device__ __forceinline__ V zero() { return -HUGE_VAL; }
  static __device__ __forceinline__ bool is_zero(V a) { return a == -HUGE_VAL; }
  static __device__ __forceinline__ K kmin() { return -HUGE_VAL; }
  static __device__ __forceinline__ K key(V a, V b) { return a + b; }
  static __device__ __forceinline__ K kmax(K a, K b) { return fmax(a, b); }
  static __device__ __forceinline__ S contrib(V a, V b, K km) { return exp(a + b - km); }
  static __device__ __forceinline__ V finish(K km, S s) { return s > 0.0 ? km + log(s) : -HUGE_VAL; }
  static __device__ __forceinline__ V add(V a, V b) {
    if (b == -HUGE_VAL) return a;
    if (a == -HUGE_VAL) return b;
    return a < b ? b + log1p(exp(a - b)) : a + log1p(exp(b - a));
  }
  static __device__ __forceinline__ V mul(V a, V b) {
    return (a == -HUGE_VAL || b == -HUGE_VAL) ? -HUGE_VAL : a + b;
  }
  static __device__ __forceinline__ V mulc(V a, C c) { return a == -HUGE_VAL ? a : a + c; }
  // natural-log value of a cell (for export); baseline added by the caller
  static __device__ __forceinline__ double to_log(V a) { return a; }
  static __device__ __forceinline__ V from_log(double x) { return x; }
  static __host__ __device__ inline C rule(double logp) { return logp; }
  static __device__ __forceinline__ K shfl_k(K k, int o) { return __shfl_xor(k, o); }
  static __device__ __forceinline__ S shfl_s(S s, int o) { return __shfl_xor(s, o); }
};

struct XFv {
  float m;
  int e;
};

struct XF {
  typedef XFv V;
  typedef int K;
  typedef float S;
  typedef float C;  // rule probability (linear)
  static constexpr bool scaled = true;
  static constexpr int EZ = -(1 << 28);   // exponent of zero; sums of a few stay far from INT_MIN
  static __device__ __forceinline__ V zero() { return V{0.0f, EZ}; }
  static __device__ __forceinline__ bool is_zero(V a) { return a.m == 0.0f; }
  static __device__ __forceinline__ K kmin() { return INT_MIN / 2; }
  static __device__ __forceinline__ K key(V a, V b) { return a.e + b.e; }
  static __device__ __forceinline__ K kmax(K a, K b) { return max(a, b); }
  static __device__ __forceinline__ S contrib(V a, V b, K km) {
    int d = a.e + b.e - km;
    d = d < -200 ? -200 : d;
    return ldexpf(a.m * b.m, d);
  }
  static __device__ __forceinline__ V norm(float m, int e) {
    if (m == 0.0f) return zero();
    int x;
    float f = frexpf(m, &x);
    return V{f, e + x};
  }
  static __device__ __forceinline__ V finish(K km, S s) { return norm(s, km); }
  static __device__ __forceinline__ V add(V a, V b) {
    if (b.m == 0.0f) return a;
    if (a.m == 0.0f) return b;
    if (a.e >= b.e) {
      int d = b.e - a.e;
      d = d < -200 ? -200 : d;
      return norm(a.m + ldexpf(b.m, d), a.e);
    }
    int d = a.e - b.e;
    d = d < -200 ? -200 : d;
    return norm(b.m + ldexpf(a.m, d), b.e);
  }
  static __device__ __forceinline__ V mul(V a, V b) {
    if (a.m == 0.0f || b.m == 0.0f) return zero();
    return norm(a.m * b.m, a.e + b.e);
  }
  static __device__ __forceinline__ V mulc(V a, C c) {
    if (a.m == 0.0f) return a;
    return norm(a.m * c, a.e);
  }
  static __device__ __forceinline__ double to_log(V a) {
    return a.m == 0.0f ? -HUGE_VAL : log((double)a.m) + (double)a.e * 0.69314718055994530942;
  }
  // natural-log x -> m * 2^e
  static __device__ __forceinline__ V from_log(double x) {
    if (x == -HUGE_VAL) return zero();
    double t = x * 1.4426950408889634;  // bits
    double fl = floor(t);
    return norm((float)exp2(t - fl), (int)fl);
  }
  static __host__ __device__ inline C rule(double logp) { return (C)exp(logp); }
  static __device__ __forceinline__ K shfl_k(K k, int o) { return __shfl_xor(k, o); }
  static __device__ __forceinline__ S shfl_s(S s, int o) { return __shfl_xor(s, o); }
};

}  // namespace messi
