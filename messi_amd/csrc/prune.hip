// Felsenstein pruning kernels for gfx950: 16-state column-pair pruning on MFMA
// and 4-state site pruning.
//
// Both run level by level over the tree (all internal nodes of equal height in
// one launch) on per-node deduplicated entries, the reference's pattern dedup
// (getindexdict, src/FelsensteinCUDA.jl:77-142, and the FastCache of
// src/FARCE.jl:2182-2254): an entry of node v is one distinct pair (or single)
// of subtree column patterns at v, and carries the 16 (or 4) partial
// likelihoods plus an integer power-of-two exponent.  An internal entry is
//   part_v = (P_l * part_l[idx_l]) .* (P_r * part_r[idx_r]),
// rescaled so its maximum lies in [0.5, 1) (exact power-of-two scaling; the
// reference divides by the maximum when it falls below 1e-20,
// src/FARCE.jl:2225-2238, and always on the GPU, src/cuda/felsenstein.cu:193-198).
// At the root the log-likelihood log(sum_a pi_a part_a) + exponent*ln 2 is mixed
// into a running log-sum-exp over components (src/FelsensteinCUDA.jl:400-419).
//
// MFMA mapping (V_MFMA_F64_16X16X4_F64 / V_MFMA_F32_16X16X4_F32, one wavefront
// per block of 16 entries): the message P_c * part_c is a 16x16 (states) times
// 16x16 (states x entries) product done as four K=4 steps.  Both instructions
// take A[i][k] from lane i + 16k and B[k][j] from lane j + 16k, and return
// D[i][j] in lane j + 16g, register r, where the row is i = state(g, r):
// g + 4r for f64 and 4g + r for f32 (checked by messi_selftest).  Step s of a
// node therefore uses register s of the child's accumulator as B, contracting
// the states {state(g, s)}, with A = P[i][state(g, s)]: a node's output feeds
// its parent's MFMA with no data movement.  Partials are stored in lane order
// (position 4g + r holds state(g, r)) so every lane loads and stores 4
// consecutive values of one entry.
#include <type_traits>

#include "common.h"
#include "prune.h"

namespace messi {

template <typename T> struct Mfma;
template <> struct Mfma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 op(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __host__ __device__ __forceinline__ int state(int g, int r) { return g + 4 * r; }
};
template <> struct Mfma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 op(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __host__ __device__ __forceinline__ int state(int g, int r) { return 4 * g + r; }
};

__device__ __forceinline__ double lse_d(double a, double b) {
  if (a == -HUGE_VAL) return b;
  if (b == -HUGE_VAL) return a;
  return a < b ? b + log1p(exp(a - b)) : a + log1p(exp(b - a));
}

__device__ __forceinline__ void load4(const double *p, double v[4]) {
  const double2 *q = reinterpret_cast<const double2 *>(p);
  double2 a = q[0], b = q[1];
  v[0] = a.x;
  v[1] = a.y;
  v[2] = b.x;
  v[3] = b.y;
}
__device__ __forceinline__ void load4(const float *p, float v[4]) {
  float4 a = *reinterpret_cast<const float4 *>(p);
  v[0] = a.x;
  v[1] = a.y;
  v[2] = a.z;
  v[3] = a.w;
}
__device__ __forceinline__ void store4(double *p, const double v[4]) {
  double2 *q = reinterpret_cast<double2 *>(p);
  q[0] = make_double2(v[0], v[1]);
  q[1] = make_double2(v[2], v[3]);
}
__device__ __forceinline__ void store4(float *p, const float v[4]) {
  *reinterpret_cast<float4 *>(p) = make_float4(v[0], v[1], v[2], v[3]);
}

__device__ __forceinline__ int frexp_exp(double m) {
  int e;
  (void)frexp(m, &e);
  return e;
}
__device__ __forceinline__ int frexp_exp(float m) {
  int e;
  (void)frexpf(m, &e);
  return e;
}

// A stored 16-state entry carries its int32 power-of-two exponent inside its
// values: every value is nonnegative and below 2 (the maximum is scaled into
// [0.5, 1)), so the sign bit and the top exponent bit (bits 31 and 30 of a
// float, 63 and 62 of a double) are zero, and value k of the entry (position
// 4g + r, held by lane group g) holds exponent bits 2k and 2k+1 there.  One
// random read then brings an entry's values and exponent together; clearing the
// two bits gives back the values exactly.
template <typename T> struct TopBits;
template <> struct TopBits<float> {
  static __device__ __forceinline__ unsigned get(float x) { return __float_as_uint(x) >> 30; }
  static __device__ __forceinline__ float clear(float x) { return __uint_as_float(__float_as_uint(x) & 0x3FFFFFFFu); }
  static __device__ __forceinline__ float put(float x, unsigned b) { return __uint_as_float(__float_as_uint(x) | (b << 30)); }
};
template <> struct TopBits<double> {
  static __device__ __forceinline__ unsigned get(double x) {
    return (unsigned)((unsigned long long)__double_as_longlong(x) >> 62);
  }
  static __device__ __forceinline__ double clear(double x) {
    return __longlong_as_double((long long)((unsigned long long)__double_as_longlong(x) & 0x3FFFFFFFFFFFFFFFull));
  }
  static __device__ __forceinline__ double put(double x, unsigned b) {
    return __longlong_as_double(
        (long long)((unsigned long long)__double_as_longlong(x) | ((unsigned long long)b << 62)));
  }
};
// This lane's 4 values of a loaded entry: their exponent bits, placed at bits
// 8g .. 8g+7, and the values with those bits cleared.
template <typename T>
__device__ __forceinline__ unsigned take_exp_bits(T v[4], int g) {
  unsigned w = 0;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    w |= TopBits<T>::get(v[r]) << (2 * r);
    v[r] = TopBits<T>::clear(v[r]);
  }
  return w << (8 * g);
}
// Lanes col + 16g (g = 0..3) of an entry sit in the four 16-lane rows.  A
// gfx950 row swap of a register with itself returns its even and odd rows side
// by side (v_permlane16_swap: rows 0/1 and 2/3; v_permlane32_swap: rows 0/2 and
// 1/3), so two swaps combine the four rows on the VALU, with no LDS round trip.
__device__ __forceinline__ unsigned rows_or(unsigned w) {
  const auto a = __builtin_amdgcn_permlane16_swap(w, w, false, false);
  w = a[0] | a[1];
  const auto b = __builtin_amdgcn_permlane32_swap(w, w, false, false);
  return b[0] | b[1];
}
template <typename T> __device__ __forceinline__ T rows_max(T m);
template <> __device__ __forceinline__ float rows_max(float m) {
  const auto a = __builtin_amdgcn_permlane16_swap(__float_as_uint(m), __float_as_uint(m), false, false);
  m = fmaxf(__uint_as_float(a[0]), __uint_as_float(a[1]));
  const auto b = __builtin_amdgcn_permlane32_swap(__float_as_uint(m), __float_as_uint(m), false, false);
  return fmaxf(__uint_as_float(b[0]), __uint_as_float(b[1]));
}
template <> __device__ __forceinline__ double rows_max(double m) {
  m = fmax(m, __shfl_xor(m, 16));
  return fmax(m, __shfl_xor(m, 32));
}
// The entry's exponent from the four lane groups' bit fields.
__device__ __forceinline__ int join_exp_bits(unsigned w) { return (int)rows_or(w); }
// This lane's 4 values with exponent E's bits 8g .. 8g+7 put in.
template <typename T>
__device__ __forceinline__ void put_exp_bits(T v[4], int E, int g) {
  const unsigned w = ((unsigned)E >> (8 * g)) & 0xFFu;
#pragma unroll
  for (int r = 0; r < 4; ++r) v[r] = TopBits<T>::put(v[r], (w >> (2 * r)) & 3u);
}

// p advanced by a 32-bit byte offset: with p uniform, the access becomes a
// scalar-base, 32-bit-lane-offset load or store (no 64-bit address per lane).
template <class X> __device__ __forceinline__ X *boff(X *p, unsigned bytes) {
  return reinterpret_cast<X *>(reinterpret_cast<typename std::conditional<std::is_const<X>::value, const char, char>::type *>(p) + bytes);
}

// This lane's 4 values of one child entry in the operand lane order: a leaf's
// two codes expanded to indicator values, or a stored entry gathered (its
// exponent bits still in place).
template <typename T>
__device__ __forceinline__ void child_values(bool leaf, const uint16_t *lc, const T *pp, unsigned ci, int g,
                                             T out[4]) {
  if (leaf) {
    const unsigned cc = *boff(lc, 2u * ci);
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int a = Mfma<T>::state(g, s);
      out[s] = (T)((cc >> (a >> 2)) & 1u) * (T)((cc >> (8 + (a & 3))) & 1u);
    }
  } else {
    load4(boff(pp, (ci * 16u + 4u * g) * (unsigned)sizeof(T)), out);
  }
}

// One node entry from its two children's values xl, xr (exponents El, Er): the
// products with the children's transition matrices on MFMA, their elementwise
// product, the maximum scaled into [0.5, 1) by an exact power of two.  The
// values come out in the accumulator order, which is the operand order of the
// parent's product; the entry's exponent is returned.
template <typename T>
__device__ __forceinline__ int node_entry(const T pl[4], const T xl[4], int El, const T pr[4], const T xr[4],
                                          int Er, T v[4]) {
  typedef typename Mfma<T>::v4 v4;
  v4 ml = {0, 0, 0, 0}, mr = {0, 0, 0, 0};
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    ml = Mfma<T>::op(pl[s], xl[s], ml);
    mr = Mfma<T>::op(pr[s], xr[s], mr);
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) v[r] = ml[r] * mr[r];
  const T m = rows_max<T>(fmax(fmax(v[0], v[1]), fmax(v[2], v[3])));
  int E = El + Er;
  if (m > (T)0) {
    const int ex = frexp_exp(m);
    const T sc = ldexp((T)1, -ex);
#pragma unroll
    for (int r = 0; r < 4; ++r) v[r] *= sc;
    E += ex;
  }
  return E;
}

// One wavefront per item of 16*NB consecutive entries of one node.  The kernel
// gathers child entries at random: its speed follows the number of loads in
// flight, so the register budget is set for 6 wavefronts per SIMD in FP32 (80
// VGPRs, a few spilled) and 4 in FP64 (123 VGPRs, none spilled).  With VL the
// node's first child is not stored: its entry is computed in registers from
// its own two children's entries (the same steps as its own launch would take,
// so the result is the same to the bit), which saves that child's store and
// the reread of it when most of its entries serve one entry of the node.
template <typename T, int NB, bool VL>
__global__ __launch_bounds__(256, sizeof(T) == 4 ? (VL ? 4 : 6) : 4) void pair_level_kernel(
    const int2 *__restrict__ items, int nitems, const NodeDesc *__restrict__ nodes,
    const int *__restrict__ idx, const uint16_t *__restrict__ leafc, T *__restrict__ part,
    const T *__restrict__ P, const double *__restrict__ pi16,
    const double *__restrict__ logws, int first, double *__restrict__ rootacc,
    double *__restrict__ rootcomp, CompBatch cb) {
  const int lane = threadIdx.x & 63;
  const bool comp = rootcomp != nullptr;  // keep each component's root values
  {
    const int64_t c = blockIdx.y;  // component of this grid row
    part += c * cb.part;
    P += c * cb.P;
    pi16 += c * cb.pi;
    logws += c * cb.logw;
    rootacc += c * cb.root;
    if (comp) rootcomp += c * cb.root;
  }
  // the wavefront's item, uniform: the item and its node descriptor come through
  // scalar loads, and every gather below is a scalar base plus a 32-bit lane offset
  const int wave = __builtin_amdgcn_readfirstlane(
      (int)((xcd_chunk_id(blockIdx.x, gridDim.x) * (unsigned)blockDim.x + threadIdx.x) >> 6));
  if (wave >= nitems) return;
  const int2 it = items[wave];
  const NodeDesc nd = nodes[it.x];
  const int col = lane & 15, g = lane >> 4;
  T al[4], ar[4], aa[4], ab[4];
  const T *Pl = P + (size_t)nd.node_l * 256;
  const T *Pr = P + (size_t)nd.node_r * 256;
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    al[s] = Pl[col + 16 * Mfma<T>::state(g, s)];
    ar[s] = Pr[col + 16 * Mfma<T>::state(g, s)];
  }
  if constexpr (VL) {
    const T *Pa = P + (size_t)nd.node_va * 256;
    const T *Pb = P + (size_t)nd.node_vb * 256;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      aa[s] = Pa[col + 16 * Mfma<T>::state(g, s)];
      ab[s] = Pb[col + 16 * Mfma<T>::state(g, s)];
    }
  }
  const bool ROOT = nd.root != 0;
  double pis[4] = {0, 0, 0, 0};
  if (ROOT) {
#pragma unroll
    for (int r = 0; r < 4; ++r) pis[r] = pi16[16 * nd.member + Mfma<T>::state(g, r)];
  }
  const unsigned U = (unsigned)nd.U, e0 = (unsigned)it.y;  // entries of one node fit in 32 bits
  const int *__restrict__ il = idx + nd.idx_l;
  const int *__restrict__ ir = idx + nd.idx_r;
  const int *__restrict__ ix = idx + nd.idx_v2;
  const T *pl = part + (size_t)nd.cpart_l * 16, *pr = part + (size_t)nd.cpart_r * 16;
  const T *px = part + (size_t)nd.cpart_v2 * 16;
  const uint16_t *lcl = leafc + nd.cpart_l, *lcr = leafc + nd.cpart_r, *lcx = leafc + nd.cpart_v2;
  // Every load of the NB blocks is issued before the first store: part is both
  // read (children) and written (this node), so loads after a store could not
  // be hoisted above it.
  unsigned cl[NB], cr[NB], cx[NB];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const unsigned e = e0 + b * 16 + col;
    const unsigned ec = e < U ? e : U - 1;
    cl[b] = (unsigned)*boff(il, 4u * ec);
    cr[b] = (unsigned)*boff(ir, 4u * ec);
    if constexpr (VL) cx[b] = (unsigned)*boff(ix, 4u * ec);
  }
  // this lane's exponent bit fields of the children (0 for a leaf)
  T bl[NB][4], br[NB][4], bx[NB][4];
  unsigned wl[NB], wr[NB], wx[NB];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    child_values<T>(nd.leaf_l, lcl, pl, cl[b], g, bl[b]);
    child_values<T>(nd.leaf_r, lcr, pr, cr[b], g, br[b]);
    if constexpr (VL) child_values<T>(nd.leaf_v2, lcx, px, cx[b], g, bx[b]);
  }
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    wl[b] = nd.leaf_l ? 0u : take_exp_bits<T>(bl[b], g);
    wr[b] = nd.leaf_r ? 0u : take_exp_bits<T>(br[b], g);
    if constexpr (VL) wx[b] = nd.leaf_v2 ? 0u : take_exp_bits<T>(bx[b], g);
  }
  T *po = part + (size_t)nd.part * 16;
  double *rc = rootcomp + nd.out, *ra = rootacc + nd.out;
  // root: lane group b keeps block b's sum and exponent, so the logarithms of the
  // wavefront's 16 NB entries run once, on all lanes, after the blocks
  static_assert(NB <= 4, "one lane group per block at the root");
  double rs = 1.0;
  int rE = 0;
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const unsigned e = e0 + b * 16 + col;
    const bool valid = e < U;
    T v[4];
    int E;
    if constexpr (VL) {
      T vc[4];
      const int Ec = node_entry<T>(aa, bl[b], join_exp_bits(wl[b]), ab, bx[b], join_exp_bits(wx[b]), vc);
      E = node_entry<T>(al, vc, Ec, ar, br[b], join_exp_bits(wr[b]), v);
    } else {
      E = node_entry<T>(al, bl[b], join_exp_bits(wl[b]), ar, br[b], join_exp_bits(wr[b]), v);
    }
    if (!ROOT) {
      put_exp_bits<T>(v, E, g);
      if (valid) store4(boff(po, (e * 16u + 4u * g) * (unsigned)sizeof(T)), v);
    } else {
      double s = pis[0] * (double)v[0] + pis[1] * (double)v[1] + pis[2] * (double)v[2] +
                 pis[3] * (double)v[3];
      s += __shfl_xor(s, 16);
      s += __shfl_xor(s, 32);
      if (g == b) {
        rs = s;
        rE = E;
      }
    }
  }
  if (ROOT) {
    const unsigned e = e0 + g * 16 + col;
    if (g < NB && e < U) {
      const double ll = log(rs) + (double)rE * 0.69314718055994530942;
      if (comp) *boff(rc, 8u * e) = ll;
      const double t = logws[nd.member] + ll;
      double *const pa = boff(ra, 8u * e);
      *pa = first ? t : lse_d(*pa, t);
    }
  }
}
// One thread per entry of a node; 256 entries per item.
template <typename T>
__global__ __launch_bounds__(256) void site_level_kernel(
    const int2 *__restrict__ items, int nitems, const NodeDesc *__restrict__ nodes,
    const int *__restrict__ idx, const uint8_t *__restrict__ leafc, T *__restrict__ part,
    int *__restrict__ expo, const T *__restrict__ P4, const double *__restrict__ pi4,
    const double *__restrict__ logws, int first, double *__restrict__ rootacc, CompBatch cb) {
  if ((int)blockIdx.x >= nitems) return;
  {
    const int64_t c = blockIdx.y;  // component of this grid row
    part += c * cb.part;
    expo += c * cb.expo;
    P4 += c * cb.P;
    pi4 += c * cb.pi;
    logws += c * cb.logw;
    rootacc += c * cb.root;
  }
  const int2 it = items[blockIdx.x];
  const NodeDesc nd = nodes[it.x];
  const int64_t e = (int64_t)it.y + threadIdx.x;
  if (e >= nd.U) return;
  T vl[4], vr[4];
  int el = 0, er = 0;
  const int cl = idx[nd.idx_l + e], cr = idx[nd.idx_r + e];
  if (nd.leaf_l) {
    const unsigned c = leafc[nd.cpart_l + cl];
    for (int a = 0; a < 4; ++a) vl[a] = (T)((c >> a) & 1u);
  } else {
    load4(part + (size_t)(nd.cpart_l + cl) * 4, vl);
    el = expo[nd.cpart_l + cl];
  }
  if (nd.leaf_r) {
    const unsigned c = leafc[nd.cpart_r + cr];
    for (int a = 0; a < 4; ++a) vr[a] = (T)((c >> a) & 1u);
  } else {
    load4(part + (size_t)(nd.cpart_r + cr) * 4, vr);
    er = expo[nd.cpart_r + cr];
  }
  const T *Pl = P4 + (size_t)nd.node_l * 16;
  const T *Pr = P4 + (size_t)nd.node_r * 16;
  T v[4];
  T m = 0;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    T sl = 0, sr = 0;
#pragma unroll
    for (int d = 0; d < 4; ++d) {
      sl += Pl[a + 4 * d] * vl[d];
      sr += Pr[a + 4 * d] * vr[d];
    }
    v[a] = sl * sr;
    m = fmax(m, v[a]);
  }
  int E = el + er;
  if (m > (T)0) {
    const int ex = frexp_exp(m);
    const T sc = ldexp((T)1, -ex);
    for (int a = 0; a < 4; ++a) v[a] *= sc;
    E += ex;
  }
  if (!nd.root) {
    store4(part + (size_t)(nd.part + e) * 4, v);
    expo[nd.part + e] = E;
  } else {
    double s = 0.0;
    for (int a = 0; a < 4; ++a) s += pi4[4 * nd.member + a] * (double)v[a];
    const double t = logws[nd.member] + log(s) + (double)E * 0.69314718055994530942;
    const int64_t o = nd.out + e;
    rootacc[o] = first ? t : lse_d(rootacc[o], t);
  }
}

// Site expansion: u[c] = root entry of column c's root pattern (one grid row per member).
__global__ void site_expand_kernel(const ExpandDesc *__restrict__ dd, const double *__restrict__ acc) {
  const ExpandDesc d = dd[blockIdx.y];
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < d.L) d.u[c] = acc[d.rootoff + d.pid0[c]];
}

// Pair expansion into a symmetric row-major L x L matrix; grid (columns, rows, members).
__global__ void pair_expand_kernel(const ExpandDesc *__restrict__ dd, const double *__restrict__ acc,
                                   double *const *__restrict__ outs) {
  const ExpandDesc d = dd[blockIdx.z];
  const int L = d.L;
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  int r = blockIdx.y;
  if (c >= L || r >= L) return;
  int x = min(r, c), y = max(r, c);
  double v = -HUGE_VAL;
  if (x != y && y - x <= d.maxbp && (!d.mask || d.mask[(size_t)x + (size_t)L * y])) {
    int k = d.T0[(size_t)d.pid0[x] * d.n0 + d.pid0[y]];
    if (k >= 0) v = acc[d.rootoff + k];
  }
  double *out = outs ? outs[blockIdx.z] : d.paired;
  out[(size_t)r * L + c] = v;
}

// Fixed-structure log-likelihood of one member from its site values (d.u) and its
// pair root entries: out[0] = sum of u over unpaired columns, out[1] = sum of the
// pair log-likelihoods of the structure's pairs, out[2] = number of pairs that the
// member's pair list does not hold.  partner is 1-based (0 unpaired).  One
// workgroup; the fixed per-thread order and tree make the sums deterministic.
__global__ __launch_bounds__(1024) void structure_sum_kernel(const ExpandDesc *__restrict__ dd,
                                                             const double *__restrict__ acc,
                                                             const int *__restrict__ partner, double *out) {
  const ExpandDesc d = *dd;
  const int L = d.L, t = threadIdx.x;
  __shared__ double su[1024], sp[1024];
  __shared__ int sm[1024];
  double a = 0.0, b = 0.0;
  int miss = 0;
  for (int x = t; x < L; x += 1024) {
    const int p = partner[x];
    if (p == 0) {
      a += d.u[x];
    } else if (p - 1 > x) {
      const int y = p - 1;
      int k = -1;
      if (y - x <= d.maxbp && (!d.mask || d.mask[(size_t)x + (size_t)L * y]))
        k = d.T0[(size_t)d.pid0[x] * d.n0 + d.pid0[y]];
      if (k >= 0) b += acc[d.rootoff + k];
      else ++miss;
    }
  }
  su[t] = a;
  sp[t] = b;
  sm[t] = miss;
  __syncthreads();
  for (int o = 512; o > 0; o >>= 1) {
    if (t < o) {
      su[t] += su[t + o];
      sp[t] += sp[t + o];
      sm[t] += sm[t + o];
    }
    __syncthreads();
  }
  if (t == 0) {
    out[0] = su[0];
    out[1] = sp[0];
    out[2] = (double)sm[0];
  }
}

void launch_structure_sum(hipStream_t s, const ExpandDesc *d, const double *acc, const int *partner, double *out) {
  hipLaunchKernelGGL(structure_sum_kernel, dim3(1), dim3(1024), 0, s, d, acc, partner, out);
}

__global__ void site_gather_kernel(const ExpandDesc *__restrict__ dd, int64_t rstride, const double *__restrict__ acc,
                                   double *__restrict__ out) {
  const ExpandDesc d = *dd;
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  acc += blockIdx.y * rstride;
  out += (size_t)blockIdx.y * d.L;
  if (x < d.L) out[x] = acc[d.rootoff + d.pid0[x]];
}

__global__ void pair_gather_kernel(const ExpandDesc *__restrict__ dd, int64_t rstride, const double *__restrict__ acc,
                                   const int *__restrict__ partner, double *__restrict__ out, int *missing) {
  const ExpandDesc d = *dd;
  const int L = d.L, x = blockIdx.x * blockDim.x + threadIdx.x;
  acc += blockIdx.y * rstride;
  out += (size_t)blockIdx.y * L;
  if (x >= L) return;
  const int p = partner[x];
  double v = -HUGE_VAL;
  if (p - 1 > x) {
    const int y = p - 1;
    int k = -1;
    if (y - x <= d.maxbp && (!d.mask || d.mask[(size_t)x + (size_t)L * y]))
      k = d.T0[(size_t)d.pid0[x] * d.n0 + d.pid0[y]];
    if (k >= 0) v = acc[d.rootoff + k];
    else atomicAdd(missing, 1);
  }
  out[x] = v;
}

// One workgroup per column x; each thread takes the partners y = t, t + 256, ...
// and keeps an online (maximum, sum) per category, then the workgroup reduces
// them in a fixed tree.  The reference adds the terms one by one in column
// order; the sum is the same up to rounding.
constexpr int SRP_THREADS = 256, SRP_MAXCATS = 8;
__global__ __launch_bounds__(SRP_THREADS) void site_rate_posterior_kernel(
    const ExpandDesc *__restrict__ dd, int64_t rstride, const double *__restrict__ paircomp,
    const double *__restrict__ logw16, int K16, int cats, const double *__restrict__ sitec,
    const double *__restrict__ single, const double *__restrict__ pairprob, double *__restrict__ out) {
  const ExpandDesc d = *dd;
  const int L = d.L, x = blockIdx.x, t = threadIdx.x;
  const int M = K16 / (cats * cats);
  // per-thread accumulators (maximum, sum) of each category, in LDS (the category
  // of a term is known only at run time)
  __shared__ double am[SRP_MAXCATS][SRP_THREADS], as[SRP_MAXCATS][SRP_THREADS];
  for (int c = 0; c < SRP_MAXCATS; ++c) {
    am[c][t] = -HUGE_VAL;
    as[c][t] = 0.0;
  }
  auto add = [&](int c, double v) {
    if (v == -HUGE_VAL) return;
    const double mc = am[c][t];
    if (v > mc) {
      as[c][t] = as[c][t] * exp(mc - v) + 1.0;
      am[c][t] = v;
    } else {
      as[c][t] += exp(v - mc);
    }
  };
  if (t == 0) {
    const double ls = log(single[x]);
    for (int c = 0; c < cats; ++c) {
      const double u = sitec[(size_t)c * L + x];
      if (u != -HUGE_VAL) add(c, u + ls);
    }
  }
  for (int y = t; y < L; y += SRP_THREADS) {
    if (y == x) continue;
    const int a = min(x, y), b = max(x, y);
    if (b - a > d.maxbp || (d.mask && !d.mask[(size_t)a + (size_t)L * b])) continue;
    const int k0 = d.T0[(size_t)d.pid0[a] * d.n0 + d.pid0[b]];
    if (k0 < 0) continue;
    const double lpp = log(pairprob[(size_t)a + (size_t)L * b]);
    if (lpp == -HUGE_VAL) continue;
    const bool first = x == a;
    for (int k = 0; k < K16; ++k) {
      const int c1 = k / (cats * M), c2 = (k / M) % cats;
      add(first ? c1 : c2, lpp + logw16[k] + paircomp[(size_t)k * rstride + d.rootoff + k0]);
    }
  }
  __syncthreads();
  for (int c = 0; c < cats; ++c) {
    for (int o = SRP_THREADS / 2; o > 0; o >>= 1) {
      if (t < o) {
        const double m1 = am[c][t], m2 = am[c][t + o];
        const double mm = fmax(m1, m2);
        if (mm != -HUGE_VAL) {
          as[c][t] = (m1 == -HUGE_VAL ? 0.0 : as[c][t] * exp(m1 - mm)) +
                     (m2 == -HUGE_VAL ? 0.0 : as[c][t + o] * exp(m2 - mm));
          am[c][t] = mm;
        }
      }
      __syncthreads();
    }
    if (t == 0) out[(size_t)c + (size_t)cats * x] = as[c][0] > 0.0 ? am[c][0] + log(as[c][0]) : -HUGE_VAL;
  }
}

void launch_site_rate_posterior(hipStream_t s, int L, int cats, int K16, int64_t root_stride,
                                const ExpandDesc *dp, const double *paircomp, const double *logw16,
                                const double *sitec, const double *single, const double *pairprob, double *out) {
  hipLaunchKernelGGL(site_rate_posterior_kernel, dim3(L), dim3(SRP_THREADS), 0, s, dp, root_stride, paircomp,
                     logw16, K16, cats, sitec, single, pairprob, out);
}

void launch_site_gather(hipStream_t s, int L, int K, int64_t root_stride, const ExpandDesc *ds, const double *acc,
                        double *out) {
  hipLaunchKernelGGL(site_gather_kernel, dim3((L + 255) / 256, K), dim3(256), 0, s, ds, root_stride, acc, out);
}

void launch_pair_gather(hipStream_t s, int L, int K, int64_t root_stride, const ExpandDesc *dp, const double *acc,
                        const int *partner, double *out, int *missing) {
  hipLaunchKernelGGL(pair_gather_kernel, dim3((L + 255) / 256, K), dim3(256), 0, s, dp, root_stride, acc, partner,
                     out, missing);
}

template <typename T, int NB, bool VL>
void launch_pair_level(hipStream_t s, const int2 *items, int nitems, const NodeDesc *nodes, const int *idx,
                       const uint16_t *leafc, T *part, const T *P, const double *pi16,
                       const double *logw, int first, double *rootacc, double *rootcomp, const CompBatch &cb) {
  if (nitems <= 0) return;
  int blocks = (nitems + 3) / 4;
  hipLaunchKernelGGL((pair_level_kernel<T, NB, VL>), dim3(blocks, cb.K), dim3(256), 0, s, items, nitems, nodes, idx,
                     leafc, part, P, pi16, logw, first, rootacc, rootcomp, cb);
}

template <typename T>
void launch_site_level(hipStream_t s, const int2 *items, int nitems, const NodeDesc *nodes, const int *idx,
                       const uint8_t *leafc, T *part, int *expo, const T *P4, const double *pi4,
                       const double *logw, int first, double *rootacc, const CompBatch &cb) {
  if (nitems <= 0) return;
  hipLaunchKernelGGL((site_level_kernel<T>), dim3(nitems, cb.K), dim3(256), 0, s, items, nitems, nodes, idx, leafc,
                     part, expo, P4, pi4, logw, first, rootacc, cb);
}

void launch_site_expand(hipStream_t s, int n, int Lmax, const ExpandDesc *d, const double *acc) {
  hipLaunchKernelGGL(site_expand_kernel, dim3((Lmax + 255) / 256, n), dim3(256), 0, s, d, acc);
}

void launch_pair_expand(hipStream_t s, int n, int Lmax, const ExpandDesc *d, const double *acc,
                        double *const *outs) {
  hipLaunchKernelGGL(pair_expand_kernel, dim3((Lmax + 255) / 256, Lmax, n), dim3(256), 0, s, d, acc, outs);
}

#define MESSI_PAIR_LEVEL(T, NB, VL)                                                                           \
  template void launch_pair_level<T, NB, VL>(hipStream_t, const int2 *, int, const NodeDesc *, const int *,        \
                                             const uint16_t *, T *, const T *, const double *, const double *, int, \
                                             double *, double *, const CompBatch &);
MESSI_PAIR_LEVEL(double, PAIR_NB, false)
MESSI_PAIR_LEVEL(float, PAIR_NB, false)
MESSI_PAIR_LEVEL(double, PAIR_VNB, true)
MESSI_PAIR_LEVEL(float, PAIR_VNB, true)
#undef MESSI_PAIR_LEVEL
template void launch_site_level<double>(hipStream_t, const int2 *, int, const NodeDesc *, const int *,
                                        const uint8_t *, double *, int *, const double *, const double *,
                                        const double *, int, double *, const CompBatch &);
template void launch_site_level<float>(hipStream_t, const int2 *, int, const NodeDesc *, const int *,
                                       const uint8_t *, float *, int *, const float *, const double *,
                                       const double *, int, double *, const CompBatch &);

// ---------------------------------------------------------------------------
// Self-test of the MFMA operand layouts the pair kernel relies on.
// D = A (16x4) * B (4x16); A[i][k] at lane i + 16k, B[k][j] at lane j + 16k,
// D[state(g, r)][j] in register r of lane j + 16g.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void mfma_probe_kernel(const T *A, const T *Bm, T *D) {
  typedef typename Mfma<T>::v4 v4;
  int l = threadIdx.x;
  T a = A[(l % 16) * 4 + l / 16];
  T b = Bm[(l / 16) * 16 + l % 16];
  v4 acc = {0, 0, 0, 0};
  acc = Mfma<T>::op(a, b, acc);
  for (int r = 0; r < 4; ++r) D[Mfma<T>::state(l / 16, r) * 16 + l % 16] = acc[r];
}

template <typename T>
bool mfma_probe(hipStream_t s) {
  std::vector<T> A(64), Bm(64), D(256), R(256, 0);
  for (int i = 0; i < 64; ++i) {
    A[i] = (T)(1 + (i * 7) % 13);
    Bm[i] = (T)(1 + (i * 5) % 11);
  }
  for (int i = 0; i < 16; ++i)
    for (int j = 0; j < 16; ++j)
      for (int k = 0; k < 4; ++k) R[i * 16 + j] += A[i * 4 + k] * Bm[k * 16 + j];
  DevBuf dA, dB, dD;
  upload(dA, A, s);
  upload(dB, Bm, s);
  dD.ensure(256 * sizeof(T));
  hipLaunchKernelGGL(mfma_probe_kernel<T>, dim3(1), dim3(64), 0, s, dA.as<T>(), dB.as<T>(), dD.as<T>());
  MESSI_HIP(hipGetLastError());
  MESSI_HIP(hipMemcpyAsync(D.data(), dD.p, 256 * sizeof(T), hipMemcpyDeviceToHost, s));
  MESSI_HIP(hipStreamSynchronize(s));
  for (int i = 0; i < 256; ++i)
    if (D[i] != R[i]) return false;
  return true;
}
template bool mfma_probe<double>(hipStream_t);
template bool mfma_probe<float>(hipStream_t);

}  // namespace messi
