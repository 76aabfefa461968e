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
  static __device__ __forceinline__ V mul_raw(V a, V b) { return a + b; }
  static __device__ __forceinline__ V keep(V t, bool c) { return c ? t : -HUGE_VAL; }
  // natural-log value of a cell (for export); baseline added by the caller
  static __device__ __forceinline__ double to_log(V a) { return a; }
  static __device__ __forceinline__ V from_log(double x) { return x; }
  static __host__ __device__ inline C rule(double logp) { return logp; }
  static __device__ __forceinline__ K shfl_k(K k, int o) { return __shfl_xor(k, o); }
  static __device__ __forceinline__ S shfl_s(S s, int o) { return __shfl_xor(s, o); }
  // Online accumulator of a split-point sum: value = m + log(s), one exp per term.
  struct Acc {
    double m, s;
  };
  static __device__ __forceinline__ Acc acc(V v) { return Acc{v, v == -HUGE_VAL ? 0.0 : 1.0}; }
  static __device__ __forceinline__ void acc_add(Acc &a, V t) {
    if (t == -HUGE_VAL) return;
    if (t > a.m) {
      a.s = (a.s == 0.0 ? 0.0 : a.s * exp(a.m - t)) + 1.0;
      a.m = t;
    } else {
      a.s += exp(t - a.m);
    }
  }
  static __device__ __forceinline__ V acc_value(const Acc &a) { return a.s > 0.0 ? a.m + log(a.s) : -HUGE_VAL; }
  // Add two terms at once; the three exponentials are independent.
  static __device__ __forceinline__ void acc_add2(Acc &a, V t1, V t2) {
    const double m = fmax(a.m, fmax(t1, t2));
    if (m == -HUGE_VAL) return;
    a.s = a.s * exp(a.m - m) + exp(t1 - m) + exp(t2 - m);
    a.m = m;
  }
  // Rule weights used when a cell is finalised: log p6 and linear p0, p2, p3, p4.
  struct FinRules {
    double l6, p0, p2, p3, p4;
  };
  static __device__ __forceinline__ FinRules fin_rules(double lp6, double p0, double p2, double p3, double p4,
                                                       double) {
    return FinRules{lp6, p0, p2, p3, p4};
  }
  // S = log(p0 e^bif + p2 e^pf), F = log(p4 e^bif + p3 e^pf), L = log p6 + pf, with
  // bif = a.m + log(a.s).
  static __device__ __forceinline__ void finalize(const Acc &a, V pf, const FinRules &r, V &S, V &F, V &Lv) {
    const bool hb = a.s > 0.0, hp = pf != -HUGE_VAL;
    Lv = hp ? pf + r.l6 : -HUGE_VAL;
    if (!hb && !hp) {
      S = F = -HUGE_VAL;
      return;
    }
    double m = hb ? (hp ? fmax(a.m, pf) : a.m) : pf;
    double ex = hb ? a.s * exp(a.m - m) : 0.0;
    double ey = hp ? exp(pf - m) : 0.0;
    S = m + log(r.p0 * ex + r.p2 * ey);
    F = m + log(r.p4 * ex + r.p3 * ey);
  }
};

// 8-byte aligned, so a cell moves with one 64-bit load or store (ds_read_b64
// takes a 16-bit immediate offset; two 32-bit halves would not).
struct alignas(8) XFv {
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
  // frexp maps 0 to 0, so only the exponent of a zero needs a select
  static __device__ __forceinline__ V norm(float m, int e) {
    int x;
    const float f = frexpf(m, &x);
    return V{f, m == 0.0f ? EZ : e + x};
  }
  static __device__ __forceinline__ V finish(K km, S s) { return norm(s, km); }
  // Both operands aligned to the larger exponent and added, with no branch: the
  // operand with that exponent is scaled by 2^0 (exact), the other by at most
  // 2^-200 (far below a float's precision), and a zero (mantissa 0, exponent EZ)
  // adds nothing, so this equals the case analysis "zero operand: return the
  // other; else align the smaller to the larger" (up to the normalisation of an
  // unnormalised operand, which changes no value) without the divergent paths.
  static __device__ __forceinline__ V add(V a, V b) {
    const int ne = max(a.e, b.e);
    return norm(ldexpf(a.m, max(a.e - ne, -200)) + ldexpf(b.m, max(b.e - ne, -200)), ne);
  }
  // acc + t1 + t2 with one alignment and one normalisation: the split-point
  // products that one launch makes ready are folded into a cell's sum in a
  // single step (terms may be unnormalised; zero terms carry exponent EZ).
  static __device__ __forceinline__ V add3(V acc, V t1, V t2) {
    const int ne = max(acc.e, max(t1.e, t2.e));
    return norm(ldexpf(acc.m, acc.e - ne) + ldexpf(t1.m, t1.e - ne) + ldexpf(t2.m, t2.e - ne), ne);
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
  // Unnormalised accumulator: terms are raw products (mantissa in [0.25, 1)),
  // aligned to the larger exponent and added; normalised once when finalised.
  typedef V Acc;
  static __device__ __forceinline__ Acc acc(V v) { return v; }
  static __device__ __forceinline__ V acc_value(const Acc &a) { return norm(a.m, a.e); }
  static __device__ __forceinline__ V mul_raw(V a, V b) { return V{a.m * b.m, a.e + b.e}; }
  static __device__ __forceinline__ V keep(V t, bool c) { return c ? t : V{0.0f, EZ}; }
  static __device__ __forceinline__ void acc_add(Acc &a, V t) {
    const int ne = max(a.e, t.e);
    a.m = ldexpf(a.m, max(a.e - ne, -200)) + ldexpf(t.m, max(t.e - ne, -200));
    a.e = ne;
  }
  struct FinRules {
    float p0, p2, p3, p4, p6;
  };
  static __device__ __forceinline__ FinRules fin_rules(float, float p0, float p2, float p3, float p4, float p6) {
    return FinRules{p0, p2, p3, p4, p6};
  }
  // In acc_add2 and finalize the exponent differences are <= 0 and the mantissas
  // are far below 2^49, so a difference below -200 already scales to zero:
  // ldexp takes it unclamped (any int32) with the same result.
  static __device__ __forceinline__ void acc_add2(Acc &a, V t1, V t2) {
    const int ne = max(a.e, max(t1.e, t2.e));
    a.m = ldexpf(a.m, a.e - ne) + ldexpf(t1.m, t1.e - ne) + ldexpf(t2.m, t2.e - ne);
    a.e = ne;
  }
  static __device__ __forceinline__ void finalize(const Acc &bif, V pf, const FinRules &r, V &S, V &F, V &Lv) {
    // pf is a raw product; bif an unnormalised sum.
    const int ne = max(bif.e, pf.e);
    const float xb = ldexpf(bif.m, bif.e - ne), xp = ldexpf(pf.m, pf.e - ne);
    // explicit fused forms: the rounding does not depend on how the compiler
    // contracts the sums at each inlined call site
    S = norm(fmaf(r.p0, xb, r.p2 * xp), ne);
    F = norm(fmaf(r.p4, xb, r.p3 * xp), ne);
    // L = p6 * pf is kept unnormalised: pf's mantissa is a product of two
    // normalised ones, in [1/4, 1), so L's lies in [p6/4, p6) and a zero pf already
    // carries a zero's exponent (the sum of a zero's and a finite one's)
    Lv = V{r.p6 * pf.m, pf.e};
  }
};

// Representation of each storage precision: double -> LogSR, float -> XF.
template <typename T> struct SRof;
template <> struct SRof<double> { typedef LogSR type; };
template <> struct SRof<float> { typedef XF type; };

}  // namespace messi
