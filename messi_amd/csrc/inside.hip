// KH99 inside / outside / base-pair posterior kernels for gfx950.
//
// Semantics follow the reference CPU path:
//   inside  -- computeinsideparallel / computeblock (src/InsideOutsideAlgorithm.jl:648-749)
//   outside -- computeoutside (src/InsideOutsideAlgorithm.jl:447-483)
//   bpp     -- computebasepairprobs (src/InsideOutsideAlgorithm.jl:486-533)
//
// Representation.  Tables hold log values in "units" relative to a per-position
// baseline: v_hat[i,j] = v[i,j] * toU - (C[j+1] - C[i]), C the prefix sum of
// lambda_k.  Because every derivation of a span emits each of its positions
// exactly once, the baseline is additive over bifurcations and cancels out of
// every recursion; only the pair term becomes ptv[j,y] = paired[j,y]*B*toU -
// lambda_j - lambda_y.  The FP64 path uses natural logs with lambda = 0 (the
// reference's own numbers); the FP32 path uses base-2 logs with lambda_k =
// (u_k + log p(S->LS) + log p(L->s)) * B / ln 2, which removes the per-column
// trend of about -20 nats so FP32 keeps its precision, and exp2/log2 map to
// single gfx950 instructions.
//
// Log-sum-exp over a bifurcation is computed as max + log(sum exp(t - max)),
// two passes over the split point, with the split range shared by a group of
// G lanes of one wavefront.
#include "common.h"

namespace messi {

template <typename T> struct Lg;
template <> struct Lg<double> {
  static __device__ __forceinline__ double ex(double x) { return exp(x); }
  static __device__ __forceinline__ double lg(double x) { return log(x); }
  static __device__ __forceinline__ double ninf() { return -HUGE_VAL; }
  static __device__ __forceinline__ double lse(double a, double b) {
    if (a == -HUGE_VAL) return b;
    if (b == -HUGE_VAL) return a;
    return a < b ? b + log1p(exp(a - b)) : a + log1p(exp(b - a));
  }
};
template <> struct Lg<float> {
  static __device__ __forceinline__ float ex(float x) { return exp2f(x); }
  static __device__ __forceinline__ float lg(float x) { return log2f(x); }
  static __device__ __forceinline__ float ninf() { return -HUGE_VALF; }
  static __device__ __forceinline__ float lse(float a, float b) {
    if (a == -HUGE_VALF) return b;
    if (b == -HUGE_VALF) return a;
    float m = fmaxf(a, b), d = fabsf(a - b);
    return m + log2f(1.0f + exp2f(-d));
  }
};

template <typename T> struct Rules {
  T p[7];
};

template <typename T>
__device__ __forceinline__ T group_max(T v, int G) {
  for (int o = G >> 1; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}
template <typename T>
__device__ __forceinline__ T group_sum(T v, int G) {
  for (int o = G >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

static inline int pow2ceil(int x) {
  int g = 1;
  while (g < x) g <<= 1;
  return g;
}
static inline int lanes_for(int terms) {
  int g = pow2ceil(terms / 8 > 1 ? terms / 8 : 1);
  return g > 64 ? 64 : g;
}

// ---------------------------------------------------------------------------
// Preparation: baseline lambda (units), its prefix sums C, u_hat and ptv.
// ---------------------------------------------------------------------------
__global__ void baseline_kernel(int L, const double *u, double B, double toU, double lamconst,
                                int scaled, double *lam, double *C, double *uhat) {
  // One workgroup of 1024 threads: chunked serial sums + shared scan.
  __shared__ double part[1024];
  int t = threadIdx.x, nt = blockDim.x;
  int chunk = (L + nt - 1) / nt;
  int b = t * chunk, e = min(L, b + chunk);
  double s = 0.0;
  for (int k = b; k < e; ++k) {
    double lk = scaled ? (u[k] + lamconst) * B * toU : 0.0;
    lam[k] = lk;
    uhat[k] = u[k] * toU - lk;
    s += lk;
  }
  part[t] = s;
  __syncthreads();
  if (t == 0) {
    double acc = 0.0;
    for (int i = 0; i < nt; ++i) {
      double v = part[i];
      part[i] = acc;
      acc += v;
    }
  }
  __syncthreads();
  double acc = part[t];
  if (t == 0) C[0] = 0.0;
  for (int k = b; k < e; ++k) {
    acc += lam[k];
    C[k + 1] = acc;
  }
}

template <typename T>
__global__ void ptv_kernel(int L, const double *paired, const double *lam, double BtoU, T *ptv) {
  int y = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y;
  if (y >= L) return;
  // Julia column-major element (j, y); only the upper triangle (j < y) is used.
  double v = paired[(size_t)y * L + j];
  T out = (j == y || v == -HUGE_VAL) ? Lg<T>::ninf() : (T)(v * BtoU - lam[j] - lam[y]);
  ptv[(size_t)j * L + y] = out;
}

template <typename T>
__global__ void inside_init_kernel(int L, const double *u, const double *lam, double BtoU,
                                   Rules<T> r0, T *Lr, T *Lc, T *Sr, T *Sc, T *F) {
  // r0.p[1] = log p(S->s) * B * toU, r0.p[5] = log p(L->s) * B * toU.
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L) return;
  size_t c = (size_t)i * L + i;
  double ub = u[i] * BtoU - lam[i];
  T s = (T)(ub + (double)r0.p[1]);
  T l = (T)(ub + (double)r0.p[5]);
  Sr[c] = s;
  Sc[c] = s;
  Lr[c] = l;
  Lc[c] = l;
  F[c] = Lg<T>::ninf();
}

// One anti-diagonal d of the inside recursion; G lanes per cell.
template <typename T>
__global__ __launch_bounds__(256) void inside_diag_kernel(int L, int d, int G, T *Lr, T *Lc, T *Sr,
                                                          T *Sc, T *F, const T *ptv, Rules<T> r) {
  long tid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  int cell = (int)(tid / G), sub = (int)(tid % G);
  bool valid = cell < L - d;
  int j = valid ? cell : 0;
  int y = j + d;
  const T NINF = Lg<T>::ninf();
  const T *Lrow = Lr + (size_t)j * L;          // L[j, h]
  const T *Scol = Sc + (size_t)y * L + 1;      // S[h+1, y]
  T m = NINF;
  if (valid)
    for (int h = j + sub; h < y; h += G) m = fmax(m, Lrow[h] + Scol[h]);
  m = group_max(m, G);
  T s = 0;
  if (valid && m != NINF)
    for (int h = j + sub; h < y; h += G) s += Lg<T>::ex(Lrow[h] + Scol[h] - m);
  s = group_sum(s, G);
  if (valid && sub == 0) {
    T bif = s > (T)0 ? m + Lg<T>::lg(s) : NINF;
    T pt = ptv[(size_t)j * L + y];
    T f = d >= 2 ? F[(size_t)(j + 1) * L + (y - 1)] : NINF;
    T pb = pt + f;
    T b0 = bif == NINF ? NINF : r.p[0] + bif;
    T b4 = bif == NINF ? NINF : r.p[4] + bif;
    T q2 = pb == NINF ? NINF : r.p[2] + pb;
    T q3 = pb == NINF ? NINF : r.p[3] + pb;
    T q6 = pb == NINF ? NINF : r.p[6] + pb;
    T vS = Lg<T>::lse(b0, q2);
    T vF = Lg<T>::lse(b4, q3);
    size_t rc = (size_t)j * L + y, cc = (size_t)y * L + j;
    Sr[rc] = vS;
    Sc[cc] = vS;
    F[rc] = vF;
    Lr[rc] = q6;
    Lc[cc] = q6;
  }
}

// Outside, one anti-diagonal b (descending); G lanes per cell.
template <typename T>
__global__ __launch_bounds__(256) void outside_diag_kernel(int L, int b, int G, const T *iSr,
                                                           const T *iLc, T *oSr, T *oSc, T *oFr,
                                                           T *oFc, T *oL, const T *ptv, Rules<T> r) {
  long tid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  int cell = (int)(tid / G), sub = (int)(tid % G);
  bool valid = cell < L - b;
  int j = valid ? cell : 0;
  int y = j + b;
  const T NINF = Lg<T>::ninf();
  // right extension: k in (y, L-1]: out[A, j, k] + in[S, y+1, k]
  const T *oSrow = oSr + (size_t)j * L, *oFrow = oFr + (size_t)j * L;
  const T *iSrow = iSr + (size_t)(y + 1 < L ? y + 1 : y) * L;
  // left extension: k in [0, j-1]: out[A, k, y] + in[L, k, j-1]
  const T *oScol = oSc + (size_t)y * L, *oFcol = oFc + (size_t)y * L;
  const T *iLcol = iLc + (size_t)(j >= 1 ? j - 1 : 0) * L;
  T mrS = NINF, mrF = NINF, mlS = NINF, mlF = NINF;
  if (valid) {
    for (int k = y + 1 + sub; k < L; k += G) {
      T is = iSrow[k];
      mrS = fmax(mrS, oSrow[k] + is);
      mrF = fmax(mrF, oFrow[k] + is);
    }
    for (int k = sub; k < j; k += G) {
      T il = iLcol[k];
      mlS = fmax(mlS, oScol[k] + il);
      mlF = fmax(mlF, oFcol[k] + il);
    }
  }
  mrS = group_max(mrS, G);
  mrF = group_max(mrF, G);
  mlS = group_max(mlS, G);
  mlF = group_max(mlF, G);
  T srS = 0, srF = 0, slS = 0, slF = 0;
  if (valid) {
    for (int k = y + 1 + sub; k < L; k += G) {
      T is = iSrow[k];
      if (mrS != NINF) srS += Lg<T>::ex(oSrow[k] + is - mrS);
      if (mrF != NINF) srF += Lg<T>::ex(oFrow[k] + is - mrF);
    }
    for (int k = sub; k < j; k += G) {
      T il = iLcol[k];
      if (mlS != NINF) slS += Lg<T>::ex(oScol[k] + il - mlS);
      if (mlF != NINF) slF += Lg<T>::ex(oFcol[k] + il - mlF);
    }
  }
  srS = group_sum(srS, G);
  srF = group_sum(srF, G);
  slS = group_sum(slS, G);
  slF = group_sum(slF, G);
  if (valid && sub == 0) {
    T rS = srS > (T)0 ? mrS + Lg<T>::lg(srS) : NINF;
    T rF = srF > (T)0 ? mrF + Lg<T>::lg(srF) : NINF;
    T lS = slS > (T)0 ? mlS + Lg<T>::lg(slS) : NINF;
    T lF = slF > (T)0 ? mlF + Lg<T>::lg(slF) : NINF;
    // rule S->LS then F->LS: L gets the right extensions, S the left ones.
    T vL = Lg<T>::lse(rS == NINF ? NINF : rS + r.p[0], rF == NINF ? NINF : rF + r.p[4]);
    T vS = Lg<T>::lse(lS == NINF ? NINF : lS + r.p[0], lF == NINF ? NINF : lF + r.p[4]);
    T vF = NINF;
    if (j >= 1 && y + 1 < L) {
      size_t pc = (size_t)(j - 1) * L + (y + 1);
      T pt = ptv[pc];
      if (pt != NINF) {
        T a = oSr[pc], f = oFr[pc], l = oL[pc];
        vF = Lg<T>::lse(vF, a == NINF ? NINF : r.p[2] + pt + a);
        vF = Lg<T>::lse(vF, f == NINF ? NINF : r.p[3] + pt + f);
        vF = Lg<T>::lse(vF, l == NINF ? NINF : r.p[6] + pt + l);
      }
    }
    size_t rc = (size_t)j * L + y, cc = (size_t)y * L + j;
    oSr[rc] = vS;
    oSc[cc] = vS;
    oFr[rc] = vF;
    oFc[cc] = vF;
    oL[rc] = vL;
  }
}

template <typename T>
__global__ void outside_init_kernel(int L, T *oSr, T *oSc, T *oFr, T *oFc, T *oL) {
  size_t rc = (size_t)0 * L + (L - 1), cc = (size_t)(L - 1) * L + 0;
  oSr[rc] = (T)0;
  oSc[cc] = (T)0;
  oFr[rc] = Lg<T>::ninf();
  oFc[cc] = Lg<T>::ninf();
  oL[rc] = Lg<T>::ninf();
}

// Export a table triple to the Julia (3, L, L) layout in natural-log absolute
// values; cells below the diagonal are -inf.
template <typename T>
__global__ void export3_kernel(int L, const T *Sc, const T *Fr, const T *Lrow, const T *Lcol,
                               const double *C, double fromU, int outside, double CL, double *o3) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y;
  if (i >= L) return;
  size_t base = 3 * ((size_t)i + (size_t)L * j);
  if (i > j) {
    o3[base] = o3[base + 1] = o3[base + 2] = -HUGE_VAL;
    return;
  }
  // inside baseline of span [i, j]: C[j+1] - C[i]; outside: C[L] - C[j+1] + C[i].
  double bl = outside ? (CL - C[j + 1] + C[i]) : (C[j + 1] - C[i]);
  T s = Sc[(size_t)j * L + i];
  T f = Fr[(size_t)i * L + j];
  T l = Lcol ? Lcol[(size_t)j * L + i] : Lrow[(size_t)i * L + j];
  o3[base] = s == Lg<T>::ninf() ? -HUGE_VAL : ((double)s + bl) * fromU;
  o3[base + 1] = f == Lg<T>::ninf() ? -HUGE_VAL : ((double)f + bl) * fromU;
  o3[base + 2] = l == Lg<T>::ninf() ? -HUGE_VAL : ((double)l + bl) * fromU;
}

// Import a Julia (3, L, L) table into row/column-major scaled tables.
template <typename T>
__global__ void import3_kernel(int L, const double *i3, const double *C, double toU, int outside,
                               double CL, T *Sr, T *Sc, T *Fr, T *Fc, T *Lr, T *Lc) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y;
  if (i >= L || i > j) return;
  size_t base = 3 * ((size_t)i + (size_t)L * j);
  double bl = outside ? (CL - C[j + 1] + C[i]) : (C[j + 1] - C[i]);
  double s = i3[base], f = i3[base + 1], l = i3[base + 2];
  T ts = s == -HUGE_VAL ? Lg<T>::ninf() : (T)(s * toU - bl);
  T tf = f == -HUGE_VAL ? Lg<T>::ninf() : (T)(f * toU - bl);
  T tl = l == -HUGE_VAL ? Lg<T>::ninf() : (T)(l * toU - bl);
  size_t rc = (size_t)i * L + j, cc = (size_t)j * L + i;
  Sr[rc] = ts;
  Sc[cc] = ts;
  Fr[rc] = tf;
  if (Fc) Fc[cc] = tf;
  if (Lr) Lr[rc] = tl;
  if (Lc) Lc[cc] = tl;
}

// Base-pair posteriors: log table (symmetric) and unpaired logs, scaled units.
template <typename T>
__global__ void bpp_logs_kernel(int L, const T *oSr, const T *oFr, const T *oL, const T *iF,
                                const T *ptv, const double *uhat, Rules<T> r, const T *Zhat, T *lp,
                                T *res) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y;
  if (k >= L) return;
  const T NINF = Lg<T>::ninf();
  T z = Zhat[0];
  if (k == j) {
    lp[(size_t)j * L + j] = NINF;
    T a = oSr[(size_t)j * L + j], l = oL[(size_t)j * L + j];
    T t = Lg<T>::lse(a == NINF ? NINF : a + r.p[1], l == NINF ? NINF : l + r.p[5]);
    res[j] = t == NINF ? NINF : (T)((double)t + uhat[j] - (double)z);
    return;
  }
  if (k < j) return;
  size_t c = (size_t)j * L + k;
  T inres = (j + 1 <= k - 1) ? iF[(size_t)(j + 1) * L + (k - 1)] : NINF;
  T pt = ptv[c];
  T v = NINF;
  if (inres != NINF && pt != NINF) {
    T a = oSr[c], f = oFr[c], l = oL[c];
    T t = NINF;
    t = Lg<T>::lse(t, a == NINF ? NINF : a + r.p[2] + inres);
    t = Lg<T>::lse(t, f == NINF ? NINF : f + r.p[3] + inres);
    t = Lg<T>::lse(t, l == NINF ? NINF : l + r.p[6] + inres);
    v = t == NINF ? NINF : t + pt - z;
  }
  lp[c] = v;
  lp[(size_t)k * L + j] = v;
}

// Per-row maximum and normaliser (one workgroup per row).
template <typename T>
__global__ __launch_bounds__(256) void bpp_rownorm_kernel(int L, const T *lp, const T *res, T *rmax,
                                                          T *rnorm) {
  __shared__ T sh[256];
  int x = blockIdx.x, t = threadIdx.x;
  const T *row = lp + (size_t)x * L;
  T m = res[x];
  for (int y = t; y < L; y += 256) m = fmax(m, row[y]);
  sh[t] = m;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o) sh[t] = fmax(sh[t], sh[t + o]);
    __syncthreads();
  }
  m = sh[0];
  __syncthreads();
  T s = 0;
  for (int y = t; y < L; y += 256) s += Lg<T>::ex(row[y] - m);
  sh[t] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o) sh[t] += sh[t + o];
    __syncthreads();
  }
  if (t == 0) {
    rmax[x] = m;
    rnorm[x] = sh[0] + Lg<T>::ex(res[x] - m);
  }
}

template <typename T>
__global__ void bpp_write_kernel(int L, const T *lp, const T *res, const T *rmax, const T *rnorm,
                                 double *unpaired, double *pair) {
  int x = blockIdx.x * blockDim.x + threadIdx.x;
  int y = blockIdx.y;
  if (x >= L) return;
  T m = rmax[x], n = rnorm[x];
  pair[(size_t)x + (size_t)L * y] = (double)(Lg<T>::ex(lp[(size_t)y * L + x] - m) / n);
  if (y == 0 && unpaired) unpaired[x] = (double)(Lg<T>::ex(res[x] - m) / n);
}

template <typename T>
__global__ void zhat_kernel(int L, const T *Sc, T *z) {
  z[0] = Sc[(size_t)(L - 1) * L + 0];
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
template <typename T> struct Units;
template <> struct Units<double> {
  static constexpr double toU = 1.0;
  static constexpr int scaled = 0;
};
template <> struct Units<float> {
  static constexpr double toU = 1.4426950408889634;  // 1 / ln 2
  static constexpr int scaled = 1;
};

template <typename T>
static Rules<T> make_rules(double B) {
  double lp[7];
  kh99_logprobs(lp);
  Rules<T> r;
  for (int i = 0; i < 7; ++i) r.p[i] = (T)(lp[i] * B * Units<T>::toU);
  return r;
}

template <typename T>
void inside_alloc(InsideWork &w, int L, bool outside) {
  size_t n = (size_t)L * L * sizeof(T);
  w.L = L;
  w.Lr.ensure(n);
  w.Lc.ensure(n);
  w.Sr.ensure(n);
  w.Sc.ensure(n);
  w.F.ensure(n);
  w.ptv.ensure(n);
  w.lam.ensure(sizeof(double) * L);
  w.C.ensure(sizeof(double) * (L + 1));
  w.uhat.ensure(sizeof(double) * L);
  w.u.ensure(sizeof(double) * L);
  w.paired.ensure(sizeof(double) * L * L);
  w.red.ensure(sizeof(double) * 4 * (L + 16));
  if (outside) {
    w.oSr.ensure(n);
    w.oSc.ensure(n);
    w.oFr.ensure(n);
    w.oFc.ensure(n);
    w.oL.ensure(n);
    w.lp.ensure(n);
  }
}

// Baseline, ptv and the diagonal of the inside tables from w.u and w.paired (device).
template <typename T>
void inside_prepare(hipStream_t s, InsideWork &w, int L, double B) {
  double lp[7];
  kh99_logprobs(lp);
  double toU = Units<T>::toU;
  hipLaunchKernelGGL(baseline_kernel, dim3(1), dim3(1024), 0, s, L, w.u.as<double>(), B, toU,
                     lp[0] + lp[5], Units<T>::scaled, w.lam.as<double>(), w.C.as<double>(),
                     w.uhat.as<double>());
  hipLaunchKernelGGL(ptv_kernel<T>, dim3((L + 255) / 256, L), dim3(256), 0, s, L,
                     w.paired.as<double>(), w.lam.as<double>(), B * toU, w.ptv.as<T>());
}

template <typename T>
void inside_run(hipStream_t s, InsideWork &w, int L, double B) {
  Rules<T> r = make_rules<T>(B);
  hipLaunchKernelGGL(inside_init_kernel<T>, dim3((L + 255) / 256), dim3(256), 0, s, L,
                     w.u.as<double>(), w.lam.as<double>(), B * Units<T>::toU, r, w.Lr.as<T>(),
                     w.Lc.as<T>(), w.Sr.as<T>(), w.Sc.as<T>(), w.F.as<T>());
  for (int d = 1; d < L; ++d) {
    int G = lanes_for(d);
    long threads = (long)(L - d) * G;
    int blocks = (int)((threads + 255) / 256);
    hipLaunchKernelGGL(inside_diag_kernel<T>, dim3(blocks), dim3(256), 0, s, L, d, G, w.Lr.as<T>(),
                       w.Lc.as<T>(), w.Sr.as<T>(), w.Sc.as<T>(), w.F.as<T>(), w.ptv.as<T>(), r);
  }
  MESSI_HIP(hipGetLastError());
}

// log Z in natural units (host value).
template <typename T>
double inside_logZ(hipStream_t s, InsideWork &w, int L) {
  T z;
  double CL;
  MESSI_HIP(hipMemcpyAsync(&z, w.Sc.as<T>() + (size_t)(L - 1) * L, sizeof(T), hipMemcpyDeviceToHost, s));
  MESSI_HIP(hipMemcpyAsync(&CL, w.C.as<double>() + L, sizeof(double), hipMemcpyDeviceToHost, s));
  MESSI_HIP(hipStreamSynchronize(s));
  if (z == -INFINITY) return -INFINITY;
  return ((double)z + CL) / Units<T>::toU;
}

template <typename T>
void inside_export(hipStream_t s, InsideWork &w, int L, double *d_in3) {
  hipLaunchKernelGGL(export3_kernel<T>, dim3((L + 255) / 256, L), dim3(256), 0, s, L,
                     (const T *)w.Sc.as<T>(), (const T *)w.F.as<T>(), (const T *)nullptr,
                     (const T *)w.Lc.as<T>(), (const double *)w.C.as<double>(),
                     1.0 / Units<T>::toU, 0, 0.0, d_in3);
}

template <typename T>
void inside_import(hipStream_t s, InsideWork &w, int L, const double *d_in3) {
  hipLaunchKernelGGL(import3_kernel<T>, dim3((L + 255) / 256, L), dim3(256), 0, s, L, d_in3,
                     (const double *)w.C.as<double>(), Units<T>::toU, 0, 0.0, w.Sr.as<T>(),
                     w.Sc.as<T>(), w.F.as<T>(), (T *)nullptr, w.Lr.as<T>(), w.Lc.as<T>());
}

template <typename T>
void outside_run(hipStream_t s, InsideWork &w, int L) {
  Rules<T> r = make_rules<T>(1.0);
  hipLaunchKernelGGL(outside_init_kernel<T>, dim3(1), dim3(1), 0, s, L, w.oSr.as<T>(), w.oSc.as<T>(),
                     w.oFr.as<T>(), w.oFc.as<T>(), w.oL.as<T>());
  for (int b = L - 2; b >= 0; --b) {
    int G = lanes_for(L - b);
    long threads = (long)(L - b) * G;
    int blocks = (int)((threads + 255) / 256);
    hipLaunchKernelGGL(outside_diag_kernel<T>, dim3(blocks), dim3(256), 0, s, L, b, G,
                       (const T *)w.Sr.as<T>(), (const T *)w.Lc.as<T>(), w.oSr.as<T>(),
                       w.oSc.as<T>(), w.oFr.as<T>(), w.oFc.as<T>(), w.oL.as<T>(),
                       (const T *)w.ptv.as<T>(), r);
  }
  MESSI_HIP(hipGetLastError());
}

template <typename T>
void outside_export(hipStream_t s, InsideWork &w, int L, double *d_out3) {
  double CL;
  MESSI_HIP(hipMemcpyAsync(&CL, w.C.as<double>() + L, sizeof(double), hipMemcpyDeviceToHost, s));
  MESSI_HIP(hipStreamSynchronize(s));
  hipLaunchKernelGGL(export3_kernel<T>, dim3((L + 255) / 256, L), dim3(256), 0, s, L,
                     (const T *)w.oSc.as<T>(), (const T *)w.oFr.as<T>(), (const T *)w.oL.as<T>(),
                     (const T *)nullptr, (const double *)w.C.as<double>(), 1.0 / Units<T>::toU, 1,
                     CL, d_out3);
}

template <typename T>
void outside_import(hipStream_t s, InsideWork &w, int L, const double *d_out3) {
  double CL;
  MESSI_HIP(hipMemcpyAsync(&CL, w.C.as<double>() + L, sizeof(double), hipMemcpyDeviceToHost, s));
  MESSI_HIP(hipStreamSynchronize(s));
  hipLaunchKernelGGL(import3_kernel<T>, dim3((L + 255) / 256, L), dim3(256), 0, s, L, d_out3,
                     (const double *)w.C.as<double>(), Units<T>::toU, 1, CL, w.oSr.as<T>(),
                     w.oSc.as<T>(), w.oFr.as<T>(), w.oFc.as<T>(), w.oL.as<T>(), (T *)nullptr);
}

// Posteriors into device arrays (unpaired may be null).
template <typename T>
void bpp_run(hipStream_t s, InsideWork &w, int L, double *d_unpaired, double *d_pair) {
  Rules<T> r = make_rules<T>(1.0);
  T *z = reinterpret_cast<T *>(w.red.as<double>());
  T *res = z + 16;
  T *rmax = res + L;
  T *rnorm = rmax + L;
  hipLaunchKernelGGL(zhat_kernel<T>, dim3(1), dim3(1), 0, s, L, (const T *)w.Sc.as<T>(), z);
  hipLaunchKernelGGL(bpp_logs_kernel<T>, dim3((L + 255) / 256, L), dim3(256), 0, s, L,
                     (const T *)w.oSr.as<T>(), (const T *)w.oFr.as<T>(), (const T *)w.oL.as<T>(),
                     (const T *)w.F.as<T>(), (const T *)w.ptv.as<T>(),
                     (const double *)w.uhat.as<double>(), r, (const T *)z, w.lp.as<T>(), res);
  hipLaunchKernelGGL(bpp_rownorm_kernel<T>, dim3(L), dim3(256), 0, s, L, (const T *)w.lp.as<T>(),
                     (const T *)res, rmax, rnorm);
  hipLaunchKernelGGL(bpp_write_kernel<T>, dim3((L + 255) / 256, L), dim3(256), 0, s, L,
                     (const T *)w.lp.as<T>(), (const T *)res, (const T *)rmax, (const T *)rnorm,
                     d_unpaired, d_pair);
  MESSI_HIP(hipGetLastError());
}

#define MESSI_INST(T)                                                                   \
  template void inside_alloc<T>(InsideWork &, int, bool);                               \
  template void inside_prepare<T>(hipStream_t, InsideWork &, int, double);              \
  template void inside_run<T>(hipStream_t, InsideWork &, int, double);                  \
  template double inside_logZ<T>(hipStream_t, InsideWork &, int);                       \
  template void inside_export<T>(hipStream_t, InsideWork &, int, double *);             \
  template void inside_import<T>(hipStream_t, InsideWork &, int, const double *);       \
  template void outside_run<T>(hipStream_t, InsideWork &, int);                         \
  template void outside_export<T>(hipStream_t, InsideWork &, int, double *);            \
  template void outside_import<T>(hipStream_t, InsideWork &, int, const double *);      \
  template void bpp_run<T>(hipStream_t, InsideWork &, int, double *, double *);
MESSI_INST(double)
MESSI_INST(float)

}  // namespace messi
