// KH99 inside / outside / base-pair posterior kernels for gfx950.
//
// Semantics follow the reference CPU path:
//   inside  -- computeinsideparallel / computeblock (src/InsideOutsideAlgorithm.jl:648-749)
//   outside -- computeoutside (src/InsideOutsideAlgorithm.jl:447-483)
//   bpp     -- computebasepairprobs (src/InsideOutsideAlgorithm.jl:486-533)
//
// Two value representations (sr.h): the plain FP64 path keeps natural-log
// values with exact log-sum-exp (LogSR); the scaled FP32 path keeps m * 2^e per
// cell (XF) relative to a per-position baseline lambda_k = (u_k + log p(S->LS) +
// log p(L->s)) * B, so a bifurcation term is one float multiply plus exponent
// arithmetic and the bulk split-point sums run on MFMA.  Every derivation of a
// span emits each of its positions once, so the baseline is additive over
// bifurcations and cancels out of every recursion; only the pair term becomes
// ptv[j,y] = paired[j,y]*B - lambda_j - lambda_y.
//
// Inside schedule: the upper triangle is cut into T x T tiles.  Launch t
// finalises every tile (I, I+t) with an in-LDS wavefront over its cells, from
// (a) split points h inside tile I, (c) split points inside tile J and (b) the
// bulk of split points in the tiles strictly between, accumulated earlier: the
// tile product (I,K) x (K,J) becomes ready in launch max(K-I, J-K) + 1, when
// both operand tiles are final, and runs either in that launch inside the
// workgroup that finalises (I,J) (K = I+1 and K = J-1) or in the tile's next
// bulk visit (a workgroup of its own every W launches, see bulk_items).  A
// launch therefore carries the cell wavefronts of one tile-diagonal plus the
// split-point products that became ready since the tiles' previous visits.
#include <algorithm>
#include <cstring>

#include "common.h"
#include "prune.h"
#include "sr.h"
#include "inside_sched.h"
#include "wave_layout.h"

namespace messi {

// Tile edge T of the inside schedule; a workgroup has T*T threads (one per cell).

template <class SR> struct Rules {
  typename SR::C p[7];    // rule weights in the representation's own form (log or linear)
  typename SR::C lin[7];  // linear rule probabilities p^B
};

template <class SR> static Rules<SR> make_rules(double B) {
  double lp[7];
  kh99_logprobs(lp);
  Rules<SR> r;
  for (int i = 0; i < 7; ++i) {
    r.p[i] = SR::rule(lp[i] * B);
    r.lin[i] = (typename SR::C)std::exp(lp[i] * B);
  }
  return r;
}

// ---------------------------------------------------------------------------
// Instances: one inside problem each (own length and tables); a launch serves all.
// ---------------------------------------------------------------------------
template <class V> struct InstDesc {
  int L, NT, T;  // length, tiles per side, tile edge
  int KB;  // largest tile offset K - I whose L tile (I, K) can hold a finite value
  V *Lr, *Lc, *Sr, *Sc, *F, *acc, *ptv;
  // acc is tile-major: tile (I, J) holds T*T cells at (I*NT + J)*T*T, row-major.
  // Scaled path only, tile-major too: L as a product's A operand, each row of
  // tile (I, K) aligned to its largest exponent (Lp block (I, K) row-major, LpE
  // at (I*NT + K)*T + row), and S as the B operand, each column segment of rows
  // KT+1 .. KT+T of tile column J aligned to its largest exponent (Sp block
  // (K, J) column-major, SpE at (K*NT + J)*T + column).
  float *Lp, *Sp;
  int *LpE, *SpE;
  const double *u, *paired;
  double *lam, *C, *uhat, *z;
  const ExpandDesc *exp;  // pair term from the pair root entries when non-null
  double *zout;           // second destination of z (the batch's gathered log Z), or null
  int *pflag;             // dependency flags of the single-launch form (pflag_ints(NT)), or null
};

// Baseline lambda, its prefix sums C and u - lambda; one workgroup per instance.
template <class V>
__global__ void baseline_kernel(const InstDesc<V> *__restrict__ insts, double B, double lamconst, int scaled) {
  const InstDesc<V> d = insts[blockIdx.x];
  const int L = d.L;
  __shared__ double part[1024];
  int t = threadIdx.x, nt = blockDim.x;
  int chunk = (L + nt - 1) / nt;
  int b = t * chunk, e = min(L, b + chunk);
  double s = 0.0;
  for (int k = b; k < e; ++k) {
    double lk = scaled ? (d.u[k] + lamconst) * B : 0.0;
    d.lam[k] = lk;
    d.uhat[k] = d.u[k] - lk;
    s += lk;
  }
  part[t] = s;
  __syncthreads();
  // Hillis-Steele inclusive scan of the chunk sums.
  for (int o = 1; o < nt; o <<= 1) {
    double v = t >= o ? part[t - o] : 0.0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  double acc = t > 0 ? part[t - 1] : 0.0;
  if (t == 0) d.C[0] = 0.0;
  for (int k = b; k < e; ++k) {
    acc += d.lam[k];
    d.C[k + 1] = acc;
  }
}

// ptv (row-major) from the Julia column-major paired matrix (upper triangle);
// grid (columns, rows, instances).
template <class SR>
__global__ void ptv_kernel(const InstDesc<typename SR::V> *__restrict__ insts, double B) {
  const InstDesc<typename SR::V> d = insts[blockIdx.z];
  const int L = d.L;
  int y = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y;
  if (y >= L || j >= L) return;
  double v = d.paired[(size_t)y * L + j];
  double x = (j >= y || v == -HUGE_VAL) ? -HUGE_VAL : v * B - d.lam[j] - d.lam[y];
  d.ptv[(size_t)j * L + y] = SR::from_log(x);
}

// The same pair term gathered straight from the pair pruning's root entries
// (no L x L pair matrix): the pair (j, y) maps to root entry
// T0[pid0[j], pid0[y]] when y - j <= maxbp and the mask allows it.  One grid
// z-slice serves a run of members on one dataset (groups[z] = first member,
// member count): the entry of a cell is looked up once and each member's
// value gathered from its own root entries.  Only cells y > j are written;
// every reader of the pair term takes zero at y <= j without reading it.
template <class SR>
__global__ void ptv_group_kernel(const InstDesc<typename SR::V> *__restrict__ insts,
                                 const int2 *__restrict__ groups, const double *__restrict__ rootacc, double B) {
  constexpr int MU = 4;  // members whose gathers are in flight together
  const int2 g = groups[blockIdx.z];
  const ExpandDesc e = *insts[g.x].exp;
  const int L = e.L;
  const int j = blockIdx.y;
  const int y = j + 1 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (y >= L) return;
  int k = -1;
  if (y - j <= e.maxbp && (!e.mask || gp(e.mask)[(size_t)j + (size_t)L * y]))
    k = gp(e.T0)[(size_t)gp(e.pid0)[j] * e.n0 + gp(e.pid0)[y]];
  const size_t c = (size_t)j * L + y;
  for (int m0 = 0; m0 < g.y; m0 += MU) {
    double v[MU], lj[MU], ly[MU];
#pragma unroll
    for (int u = 0; u < MU; ++u) {
      const InstDesc<typename SR::V> &d = insts[g.x + min(m0 + u, g.y - 1)];
      v[u] = k >= 0 ? gp(rootacc)[d.exp->rootoff + k] : -HUGE_VAL;
      lj[u] = gp(d.lam)[j];
      ly[u] = gp(d.lam)[y];
    }
#pragma unroll
    for (int u = 0; u < MU; ++u)
      if (m0 + u < g.y) {
        const double x = v[u] == -HUGE_VAL ? -HUGE_VAL : v[u] * B - lj[u] - ly[u];
        gp(insts[g.x + m0 + u].ptv)[c] = SR::from_log(x);
      }
  }
}

// Set the diagonal of S and L (F has no s rule).  The bulk accumulator needs no
// clearing: it is read only after a bulk visit wrote it (had_visit).
template <class SR>
__global__ void inside_init_kernel(const InstDesc<typename SR::V> *__restrict__ insts, double B, double lp1,
                                   double lp5) {
  const InstDesc<typename SR::V> d = insts[blockIdx.z];
  const int L = d.L;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= L) return;
  const size_t c = (size_t)j * L + j;
  const double ub = d.u[j] * B - d.lam[j];
  const typename SR::V sv = SR::from_log(ub + lp1 * B), lv = SR::from_log(ub + lp5 * B);
  d.Sr[c] = sv;
  d.Sc[c] = sv;
  d.Lr[c] = lv;
  d.Lc[c] = lv;
  d.F[c] = SR::zero();
}

// log Z (natural, absolute) = S[0, L-1] plus the baseline of the whole span, per instance.
template <class SR>
__global__ void zlog_kernel(const InstDesc<typename SR::V> *__restrict__ insts, int n, const int *gave_up) {
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  const InstDesc<typename SR::V> d = insts[m];
  // a single-launch pass that gave up waiting leaves no log Z: NaN
  double v = gave_up && *gave_up ? __builtin_nan("") : SR::to_log(d.Sc[(size_t)(d.L - 1) * d.L]);
  d.z[0] = v == -HUGE_VAL ? v : v + d.C[d.L];
  d.z[1] = v;  // baseline-relative
  if (d.zout) {
    d.zout[0] = d.z[0];
    d.zout[1] = v;
  }
}

// ---------------------------------------------------------------------------
// Group reductions of the two-pass split-point sums
// ---------------------------------------------------------------------------
template <class SR>
__device__ __forceinline__ typename SR::K grp_kmax(typename SR::K k, int G) {
  for (int o = G >> 1; o > 0; o >>= 1) k = SR::kmax(k, SR::shfl_k(k, o));
  return k;
}
template <class SR>
__device__ __forceinline__ typename SR::S grp_sum(typename SR::S s, int G) {
  for (int o = G >> 1; o > 0; o >>= 1) s += SR::shfl_s(s, o);
  return s;
}

// ---------------------------------------------------------------------------
// Shared memory of one step workgroup.
// ---------------------------------------------------------------------------
// Staging of one T x T tile product sum_c A[a][c] * B[c][b].  Log values are
// staged as they are; scaled values as exponent-aligned float mantissas with
// their row / column exponent maxima, plus the MFMA result.
template <class SR, int T, bool S = SR::scaled> struct ProdLDS;
template <class SR, int T> struct ProdLDS<SR, T, false> {
  typedef typename SR::V V;
  V A[T][T + 1];
  V B[T][T + 1];
};
template <class SR, int T> struct ProdLDS<SR, T, true> {
  XFv A[T][T + 1];     // the operands as they are (for the exact fallback)
  XFv B[T][T + 1];
  union {
    float Ap[T][T + 1];  // exponent-aligned mantissas ...
    float R[T][T + 1];   // ... overwritten by the MFMA results once all are computed
  };
  float Bp[T][T + 1];
  int rA[T], cB[T];
};

// Staging of one product-ready (scaled) tile product: aligned mantissas, the
// row / column exponents, and the MFMA result over A's mantissas.
template <int T> struct PackLDS {
  union {
    float Ap[T][T + 1];
    float R[T][T + 1];
  };
  float Bp[T][T + 1];
  int rA[T], cB[T];
};
template <class SR, int T, bool S = SR::scaled> struct PackArea {
  char unused[4];
};
template <class SR, int T> struct PackArea<SR, T, true> {
  PackLDS<T> q[4];  // the four products of one bulk call
};

// The finalising workgroup's tables in one flat array laid out by wave::Layout
// (diagonal tiles, pair term, then the tile's S, F and L cells with their zero
// regions); the part after the staged tables shares its storage with the
// split-point products, which are done before the wavefront starts.
template <class SR, int T> struct StepProducts {
  union {
    ProdLDS<SR, T> p[2];  // the (at most two) split-point products of a launch
    PackArea<SR, T> pk;   // scaled path: up to four products of a bulk window
  };
};
template <class SR, int T> struct StepStaged {
  typename SR::V staged[wave::Layout<T>::STAGED];
  StepProducts<SR, T> q;
};
template <class SR, int T> struct StepLDS {
  typedef typename SR::V V;
  typedef wave::Layout<T> X;
  union {
    V all[X::END];
    StepStaged<SR, T> s;
  } m;
  V Sb[T], Fbot[T], Fleft[T];
  double Fbd[T], Fld[T];  // scaled path, double wavefront: Fbot and Fleft at the tile's scale
  // scaled path, T = 16: the finalising workgroup's two products (MFMA results)
  float Rd[2][(SR::scaled && T == 16) ? 16 : 1][17];
  int flag;    // a dependency wait's outcome (single-launch form)
  int escale;  // scaled path, double wavefront: the tile's power-of-two scale
  int bad;     // scaled path, double wavefront: a value left the safe range
  __device__ V &at(int i) { return m.all[i]; }
};

// Scaled path, off-diagonal tiles: the in-tile wavefront runs in plain double
// precision.  FP64 FMAs issue at the unpacked FP32 rate on CDNA4, and a
// double's range (2^+-1022) holds a tile's values at one power-of-two scale, so a
// step is two fused multiply-adds instead of per-cell exponent alignment and
// normalisation.  The diagonal tiles (I, I), (J, J) and the pair term enter
// unscaled; the tile's own cells, its bulk sums and its border S / F values at
// 2^-escale, with escale the largest exponent of the border values.  Any value
// that would leave [2^-1000, 2^1000] (or a zero) sends the tile through the
// per-cell-exponent wavefront instead, so the double path never loses range.
typedef double __attribute__((may_alias)) adbl;
constexpr double D_LO = 0x1p-1000, D_HI = 0x1p1000;

__device__ __forceinline__ bool d_in_range(double x) {
  const double ax = fabs(x);
  return ax == 0.0 || (ax >= D_LO && ax <= D_HI);  // false for NaN and infinities
}
// m * 2^(e - E) as a double (0 for a zero mantissa); clears ok when a nonzero
// value lands outside [2^-1000, 2^1000].
__device__ __forceinline__ double xf_to_d(XFv v, int E, bool &ok) {
  if (v.m == 0.0f) return 0.0;
  const double x = ldexp((double)v.m, max(min(v.e - E, 1100), -1100));
  const double ax = fabs(x);
  ok = ok && ax >= D_LO && ax <= D_HI;
  return x;
}
// A double at scale 2^E as a normalised XF cell (mantissa rounded to float).
__device__ __forceinline__ XFv d_to_xf(double x, int E) {
  if (x == 0.0) return XF::zero();
  int e;
  float f = (float)frexp(x, &e);
  if (f == 1.0f) {  // the mantissa rounded up to 1
    f = 0.5f;
    ++e;
  }
  return XFv{f, e + E};
}
// Largest exponent of the nonzero values held by the first T lanes of a
// wavefront (INT_MIN / 2 when all are zero), uniform across the wavefront: a
// max-butterfly over each 16-lane row (DPP quad permutes, then row rotations
// by 4 and 8), then the rows' results read from lanes 0 and 16.
__device__ __forceinline__ int row16_max(int e) {
  e = max(e, __builtin_amdgcn_update_dpp(e, e, 0xB1, 0xF, 0xF, false));   // quad_perm [1,0,3,2]
  e = max(e, __builtin_amdgcn_update_dpp(e, e, 0x4E, 0xF, 0xF, false));   // quad_perm [2,3,0,1]
  e = max(e, __builtin_amdgcn_update_dpp(e, e, 0x124, 0xF, 0xF, false));  // row_ror:4
  e = max(e, __builtin_amdgcn_update_dpp(e, e, 0x128, 0xF, 0xF, false));  // row_ror:8
  return e;
}
template <int T>
__device__ __forceinline__ int border_emax(int e) {
  e = row16_max(e);
  int m = __builtin_amdgcn_readlane(e, 0);
  if constexpr (T > 16) m = max(m, __builtin_amdgcn_readlane(e, 16));
  return m;
}
// Maximum over each aligned group of T consecutive lanes (T = 16 or 32), in
// every lane of the group; the whole wavefront must be active.
template <int T>
__device__ __forceinline__ int group_max(int e) {
  static_assert(T == 16 || T == 32, "tile edge 16 or 32");
  e = row16_max(e);
  if constexpr (T == 32) e = max(e, __shfl_xor(e, 16));
  return e;
}

// Where the two operand tiles of a product live: element (r, c) of A is
// A[(ar0 + r) * L + ac0 + c] and element (c, r) of B is B[(br0 + r) * L + bc0 + c],
// each only where it names a cell of the upper triangle (row <= column in a
// row-major table: up == true; column <= row index in a column-major one: up ==
// false), zero elsewhere.
template <class SR> struct OpSrc {
  const MESSI_GLOBAL typename SR::V *A;
  int ar0, ac0;
  bool aup;
  const MESSI_GLOBAL typename SR::V *B;
  int br0, bc0;
  bool bup;
};

template <class SR>
__device__ __forceinline__ typename SR::V op_a(const OpSrc<SR> &s, int L, int r, int c) {
  const int ar = s.ar0 + r, ac = s.ac0 + c;
  return (ar >= 0 && ac >= 0 && ar < L && ac < L && (s.aup ? ar <= ac : ac <= ar)) ? typename SR::V(s.A[(size_t)ar * L + ac])
                                                                                 : SR::zero();
}
template <class SR>
__device__ __forceinline__ typename SR::V op_b(const OpSrc<SR> &s, int L, int r, int c) {
  const int br = s.br0 + r, bc = s.bc0 + c;
  return (br >= 0 && bc >= 0 && br < L && bc < L && (s.bup ? br <= bc : bc <= br)) ? typename SR::V(s.B[(size_t)br * L + bc])
                                                                                 : SR::zero();
}

// Operands of one product, one element of each tile per thread (r = tid / T,
// c = tid % T): a = A(r, c), b = B(c, r).
template <class SR> struct Operands {
  typename SR::V a, b;
};

template <class SR, int T>
__device__ __forceinline__ Operands<SR> fetch_operands(const OpSrc<SR> &s, int L) {
  const int tid = threadIdx.x, r = tid / T, c = tid % T;
  Operands<SR> o;
  o.a = op_a(s, L, r, c);
  o.b = op_b(s, L, r, c);
  return o;
}

// Inside split-point product (I,K) x (K,J): L[IT+a, KT+c] * S[KT+c+1, JT+b].
template <class SR, int T>
__device__ __forceinline__ OpSrc<SR> inside_src(int I, int K, int J, const MESSI_GLOBAL typename SR::V *Lr,
                                               const MESSI_GLOBAL typename SR::V *Sc) {
  return OpSrc<SR>{Lr, I * T, K * T, true, Sc, J * T, K * T + 1, false};
}

// Exact sum over c of A[a][c] * B[c][b] of a staged product (the rare cell whose
// exponent-aligned sum lost the significant terms).
template <int T>
__device__ __forceinline__ XFv exact_cell_xf(const XFv (*A)[T + 1], const XFv (*B)[T + 1], int a, int b) {
  int km = XF::kmin();
  for (int c = 0; c < T; ++c) km = max(km, XF::key(A[a][c], B[c][b]));
  float sum = 0.0f;
  for (int c = 0; c < T; ++c) sum += XF::contrib(A[a][c], B[c][b], km);
  return XF::finish(km, sum);
}

// acc0 (+)= product 0 and acc1 (+)= product 1 (each if do_k) for this thread's
// cell (a, b) = (tid / T, tid % T), from operands already fetched into registers.
// Both products are staged together; the MFMAs of the two run on different
// wavefronts; each thread then folds its own cell.  Ends with a barrier.
template <class SR, int T>
__device__ __forceinline__ void tile_products(const Operands<SR> &o0, bool do0, const Operands<SR> &o1, bool do1,
                                              ProdLDS<SR, T> *pp, typename SR::V &acc0, typename SR::V &acc1) {
  typedef typename SR::V V;
  const int tid = threadIdx.x, r = tid / T, c = tid % T;
  if constexpr (SR::scaled) {
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const Operands<SR> &o = k ? o1 : o0;
      if (!(k ? do1 : do0)) continue;
      // Row maxima of A's exponents and column maxima of B's (over the split
      // points c, held by T consecutive lanes), then exponent-aligned floats.
      int ea = o.a.m != 0.0f ? o.a.e : XF::EZ;
      int eb = o.b.m != 0.0f ? o.b.e : XF::EZ;
#pragma unroll
      for (int off = T / 2; off > 0; off >>= 1) {
        ea = max(ea, __shfl_xor(ea, off));
        eb = max(eb, __shfl_xor(eb, off));
      }
      pp[k].A[r][c] = o.a;
      pp[k].B[c][r] = o.b;
      pp[k].Ap[r][c] = o.a.m == 0.0f ? 0.0f : ldexpf(o.a.m, max(o.a.e - ea, -200));
      pp[k].Bp[c][r] = o.b.m == 0.0f ? 0.0f : ldexpf(o.b.m, max(o.b.e - eb, -200));
      if (c == 0) {
        pp[k].rA[r] = ea;
        pp[k].cB[r] = eb;
      }
    }
    __syncthreads();
    // One MFMA 16x16x4 f32 chain per 16x16 block; product k's blocks on waves
    // k * NB^2 .. (k + 1) * NB^2 - 1.
    typedef float v4f __attribute__((ext_vector_type(4)));
    const int wave = tid >> 6, lane = tid & 63;
    constexpr int NB = T / 16, NB2 = NB * NB;
    const int k = wave / NB2, blk = wave % NB2;
    const bool mf = wave < 2 * NB2 && (k == 0 ? do0 : do1);
    const int bi = blk / NB, bj = blk % NB;
    v4f acc = {0.f, 0.f, 0.f, 0.f};
    if (mf) {
#pragma unroll
      for (int q = 0; q < T / 4; ++q) {
        float av = pp[k].Ap[bi * 16 + (lane & 15)][4 * q + (lane >> 4)];
        float bv = pp[k].Bp[4 * q + (lane >> 4)][bj * 16 + (lane & 15)];
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc, 0, 0, 0);
      }
    }
    // R shares storage with Ap: with several wavefronts per product (T = 32),
    // block (bi, bj)'s result overwrites mantissas that block (bi, bj') still
    // reads, so no result is stored before every wavefront has read its operands.
    if constexpr (NB > 1) __syncthreads();
    if (mf) {
#pragma unroll
      for (int q = 0; q < 4; ++q) pp[k].R[bi * 16 + 4 * (lane >> 4) + q][bj * 16 + (lane & 15)] = acc[q];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      if (!(k ? do1 : do0)) continue;
      const int ra = pp[k].rA[r], cb = pp[k].cB[c];
      V contrib = SR::zero();
      if (ra != XF::EZ && cb != XF::EZ) {
        const float cv = pp[k].R[r][c];
        contrib = cv >= 0x1p-100f ? XF::norm(cv, ra + cb) : exact_cell_xf<T>(pp[k].A, pp[k].B, r, c);
      }
      V &acc = k ? acc1 : acc0;
      acc = SR::add(acc, contrib);
    }
  } else {
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      if (!(k ? do1 : do0)) continue;
      const Operands<SR> &o = k ? o1 : o0;
      pp[k].A[r][c] = o.a;
      pp[k].B[c][r] = o.b;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      if (!(k ? do1 : do0)) continue;
      typename SR::K km = SR::kmin();
      for (int q = 0; q < T; ++q) km = SR::kmax(km, SR::key(pp[k].A[r][q], pp[k].B[q][c]));
      typename SR::S sum = 0;
      if (km != SR::kmin())
        for (int q = 0; q < T; ++q) sum += SR::contrib(pp[k].A[r][q], pp[k].B[q][c], km);
      V &acc = k ? acc1 : acc0;
      acc = SR::add(acc, SR::finish(km, sum));
    }
  }
  __syncthreads();
}

// Cell positions inside a tile of the product-ready copies and of the bulk
// accumulator.  With T = 16 they follow V_MFMA_F32_16X16X4_F32's lane order, so
// a wavefront computing a product loads each operand and its accumulator with
// one contiguous vector per lane: lane l, register q takes A[l & 15][4q + l/16]
// and B[4q + l/16][l & 15], and holds result row 4(l/16) + q of column l & 15.
// With T = 32 the copies stay row-major (A) / column-major (B), the accumulator
// row-major.
template <int T> __host__ __device__ constexpr int op_a_at(int i, int k) {
  return T == 16 ? (i + 16 * (k & 3)) * 4 + (k >> 2) : i * T + k;
}
template <int T> __host__ __device__ constexpr int op_b_at(int k, int j) {
  return T == 16 ? (j + 16 * (k & 3)) * 4 + (k >> 2) : j * T + k;
}
template <int T, bool MF> __host__ __device__ constexpr int acc_at(int a, int b) {
  return (MF && T == 16) ? (b + 16 * (a >> 2)) * 4 + (a & 3) : a * T + b;
}

// Scaled-path operands of one product read from the product-ready copies: this
// thread's (r, c) = (tid / T, tid % T) element of each, as aligned float mantissas,
// and the exponent of its row of A and of its column r of B.
struct PackedOps {
  float a, b;
  int ea, eb;
};

// The product-ready copies hold whole tiles (zero mantissas and EZ exponents
// beyond the table), so the loads need no bounds checks; tile offsets fit in 32
// bits (NT^2 T^2 < 2^31 for every length the tables fit in memory).
template <int T>
__device__ __forceinline__ PackedOps fetch_packed(const InstDesc<XFv> &d, int I, int K, int J) {
  const int tid = threadIdx.x, r = tid / T;
  const unsigned ta = (unsigned)(I * d.NT + K), tb = (unsigned)(K * d.NT + J);
  PackedOps o;
  o.a = gp(d.Lp)[ta * (T * T) + op_a_at<T>(r, tid % T)];
  o.b = gp(d.Sp)[tb * (T * T) + op_b_at<T>(tid % T, r)];
  o.ea = gp(d.LpE)[ta * T + r];
  o.eb = gp(d.SpE)[tb * T + r];
  return o;
}

// Exact sum of one cell of product (I, K) x (K, J) from the tables themselves, for
// the rare cell whose exponent-aligned sum lost its significant terms.
template <int T>
__device__ __forceinline__ XFv exact_cell_global(const InstDesc<XFv> &d, int I, int K, int J, int r, int c) {
  const int L = d.L, row = I * T + r, col = J * T + c;
  auto av = [&](int q) {
    const int h = K * T + q;
    return (row < L && h < L && row <= h) ? gp(d.Lr)[(size_t)row * L + h] : XF::zero();
  };
  auto bv = [&](int q) {
    const int s = K * T + 1 + q;
    return (col < L && s < L && s <= col) ? gp(d.Sc)[(size_t)col * L + s] : XF::zero();
  };
  int km = XF::kmin();
#pragma unroll 1
  for (int q = 0; q < T; ++q) km = max(km, XF::key(av(q), bv(q)));
  float sum = 0.0f;
#pragma unroll 1
  for (int q = 0; q < T; ++q) sum += XF::contrib(av(q), bv(q), km);
  return XF::finish(km, sum);
}

// Cell (r, c) of a staged product as a raw term: the MFMA sum of aligned
// mantissas with the row and column exponents (zero when either tile row or
// column is all zero), or, when that sum lost its significant terms (below
// 2^-100), the product recomputed exactly from the tables.
template <int T, class PL>
__device__ __forceinline__ XFv packed_term(const InstDesc<XFv> &d, const PL &p, int I, int K, int J, int r, int c) {
  const int ra = p.rA[r], cb = p.cB[c];
  const bool live = ra != XF::EZ && cb != XF::EZ;
  const float cv = p.R[r][c];
  XFv v = live ? XFv{cv, ra + cb} : XF::zero();
  if (live && !(cv >= 0x1p-100f)) v = exact_cell_global<T>(d, I, K, J, r, c);
  return v;
}

// tile_products for the scaled inside path from product-ready operands: the
// same MFMA, with no exponent search and half the operand bytes; sum (+)= both
// products (each if do_k), folded in one step (XF::add3).
template <int T>
__device__ __forceinline__ void tile_products_packed(const InstDesc<XFv> &d, int I, int J, const PackedOps &o0,
                                                     int K0, bool do0, const PackedOps &o1, int K1, bool do1,
                                                     ProdLDS<XF, T> *pp, XFv &sum) {
  const int tid = threadIdx.x, r = tid / T, c = tid % T;
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const PackedOps &o = k ? o1 : o0;
    if (!(k ? do1 : do0)) continue;
    pp[k].Ap[r][c] = o.a;
    pp[k].Bp[c][r] = o.b;
    if (c == 0) {
      pp[k].rA[r] = o.ea;
      pp[k].cB[r] = o.eb;
    }
  }
  __syncthreads();
  typedef float v4f __attribute__((ext_vector_type(4)));
  const int wave = tid >> 6, lane = tid & 63;
  constexpr int NB = T / 16, NB2 = NB * NB;
  const int k = wave / NB2, blk = wave % NB2;
  const bool mf = wave < 2 * NB2 && (k == 0 ? do0 : do1);
  const int bi = blk / NB, bj = blk % NB;
  v4f acc = {0.f, 0.f, 0.f, 0.f};
  if (mf) {
#pragma unroll
    for (int q = 0; q < T / 4; ++q) {
      float av = pp[k].Ap[bi * 16 + (lane & 15)][4 * q + (lane >> 4)];
      float bv = pp[k].Bp[4 * q + (lane >> 4)][bj * 16 + (lane & 15)];
      acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc, 0, 0, 0);
    }
  }
  if constexpr (NB > 1) __syncthreads();  // R shares storage with Ap
  if (mf) {
#pragma unroll
    for (int q = 0; q < 4; ++q) pp[k].R[bi * 16 + 4 * (lane >> 4) + q][bj * 16 + (lane & 15)] = acc[q];
  }
  __syncthreads();
  const XFv t0 = do0 ? packed_term<T>(d, pp[0], I, K0, J, r, c) : XF::zero();
  const XFv t1 = do1 ? packed_term<T>(d, pp[1], I, K1, J, r, c) : XF::zero();
  sum = XF::add3(sum, t0, t1);
  __syncthreads();
}

// Bulk window W: a bulk tile at distance d is visited in the launches t = d-1,
// d-1-W, d-1-2W, ... (those with 2t - 2 >= d), each visit adding the products
// that became ready since the previous one, so its accumulator is read and
// written once per W launches (inside_sched.h: item counts, decode, visits).
using sched::bulk_items;
using sched::bulk_per;
using sched::had_visit;
using sched::last_visit;

// Dependency flags of the single-launch form (InstDesc::pflag, per instance):
// tile_done (NT x NT, 1 once the tile is final), visit_done (NT x NT, t+1 after
// the tile's bulk visit of launch t) and diag_done (NT, finalised tiles per
// diagonal); ctl[0] the ticket counter, ctl[1] the give-up flag.  All start at
// zero for each evaluation.
__host__ __device__ constexpr size_t pflag_ints(int NT) { return 2 * (size_t)NT * NT + NT; }

// Bound on a dependency wait (polls of about a microsecond each): a wait that
// runs out sets the give-up flag, every workgroup then leaves, and the host
// reports the evaluation as failed.
constexpr int PERSIST_SPIN = 1 << 22;

// Thread 0 waits until *p[i] >= v[i] for the given flags (null: none) and
// acquires what the signalling workgroups wrote (the invalidation covers the
// CU's L1 and the XCD's L2 that the other waves read through after the
// barrier).  False for every thread when the wait ran out or another workgroup
// gave up.
__device__ __forceinline__ bool wg_wait(int *ctl, int *lds_flag, const int *p0, int v0, const int *p1, int v1,
                                        const int *p2, int v2) {
  if (threadIdx.x == 0) {
    const int *ps[3] = {p0, p1, p2};
    const int vs[3] = {v0, v1, v2};
    int ok = 1;
    for (int i = 0; i < 3 && ok; ++i) {
      if (!ps[i]) continue;
      int n = 0;
      while (__hip_atomic_load(ps[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < vs[i]) {
        if (++n >= PERSIST_SPIN) {
          __hip_atomic_store(ctl + 1, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          ok = 0;
          break;
        }
        if ((n & 255) == 0 && __hip_atomic_load(ctl + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
          ok = 0;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    *lds_flag = ok;
  }
  __syncthreads();
  return *lds_flag != 0;
}

// Every thread's global writes are released (agent scope), then thread 0 sets
// flag *p = v (and adds 1 to *q when q is non-null).
__device__ __forceinline__ void wg_signal(int *p, int v, int *p2, int v2, int *q) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  if (threadIdx.x == 0) {
    if (p) __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    if (p2) __hip_atomic_store(p2, v2, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    if (q) __hip_atomic_fetch_add(q, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Cell of a T x T off-diagonal tile that thread k takes in the in-tile wavefront:
// the cells in decreasing order of their finalisation step own = b + T-1 - a
// (ties by row).  Each wavefront then holds cells of neighbouring steps: it has
// nothing left to do after its first cell's step, and the cells finalised in one
// step sit in one or two wavefronts.
template <int T> __device__ __forceinline__ void cell_by_step(int k, int &a, int &b) {
  a = 0;
  b = T - 1;
  for (int o = 2 * T - 2; o >= 0; --o) {
    const int alo = max(0, T - 1 - o), ahi = min(T - 1, 2 * T - 2 - o), n = ahi - alo + 1;
    if (k < n) {
      a = alo + k;
      b = o - (T - 1) + a;
      return;
    }
    k -= n;
  }
}

// The outside's cell order: its cells finalise at step (T-1 - b) + a.
template <int T> __device__ __forceinline__ void out_cell_by_step(int k, int &a, int &b) {
  a = T - 1;
  b = 0;
  for (int o = 2 * T - 2; o >= 0; --o) {
    const int alo = max(0, o - (T - 1)), ahi = min(T - 1, o), n = ahi - alo + 1;
    if (k < n) {
      a = alo + k;
      b = a - o + (T - 1);
      return;
    }
    k -= n;
  }
}

// Up to four product-ready products k = 0..3 (pairs (0, 1) and (2, 3) of two
// launches) folded into acc, one MFMA chain set per product on its own
// wavefronts.  Each pair is folded with the same XF::add3 as in
// tile_products_packed, so the result has the same bits whatever the window.
template <int T>
__device__ __forceinline__ void tile_products_packed4(const InstDesc<XFv> &d, int I, int J, const PackedOps *o,
                                                      const int *K, const bool *dok, PackLDS<T> *pp, XFv &acc) {
  const int tid = threadIdx.x, r = tid / T, c = tid % T;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (!dok[k]) continue;
    pp[k].Ap[r][c] = o[k].a;
    pp[k].Bp[c][r] = o[k].b;
    if (c == 0) {
      pp[k].rA[r] = o[k].ea;
      pp[k].cB[r] = o[k].eb;
    }
  }
  __syncthreads();
  typedef float v4f __attribute__((ext_vector_type(4)));
  const int wave = tid >> 6, lane = tid & 63;
  constexpr int NB = T / 16, NB2 = NB * NB;
  static_assert(T * T / 64 >= 4 * NB2, "one wavefront per 16 x 16 block of four products");
  const int k = wave / NB2, blk = wave % NB2;
  const bool mf = k < 4 && dok[k < 4 ? k : 0];
  const int bi = blk / NB, bj = blk % NB;
  v4f m = {0.f, 0.f, 0.f, 0.f};
  if (mf) {
#pragma unroll
    for (int q = 0; q < T / 4; ++q) {
      float av = pp[k].Ap[bi * 16 + (lane & 15)][4 * q + (lane >> 4)];
      float bv = pp[k].Bp[4 * q + (lane >> 4)][bj * 16 + (lane & 15)];
      m = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, m, 0, 0, 0);
    }
  }
  if constexpr (NB > 1) __syncthreads();  // R shares storage with Ap
  if (mf) {
#pragma unroll
    for (int q = 0; q < 4; ++q) pp[k].R[bi * 16 + 4 * (lane >> 4) + q][bj * 16 + (lane & 15)] = m[q];
  }
  __syncthreads();
#pragma unroll
  for (int h = 0; h < 4; h += 2) {
    if (!(dok[h] || dok[h + 1])) continue;
    const XFv t0 = dok[h] ? packed_term<T>(d, pp[h], I, K[h], J, r, c) : XF::zero();
    const XFv t1 = dok[h + 1] ? packed_term<T>(d, pp[h + 1], I, K[h + 1], J, r, c) : XF::zero();
    acc = XF::add3(acc, t0, t1);
  }
  __syncthreads();
}

// Bulk items handled by one workgroup of the scaled T = 16 launch sequence: one
// per wavefront.
constexpr int WAVE_ITEMS = 4;

// One bulk visit of the scaled path with T = 16, done by one wavefront in
// registers: bulk item idx of launch t (the same tile, products and order as
// the workgroup form in inside_item).  Each lane loads its operand vectors of a
// product (op_a_at / op_b_at order), runs the four MFMA steps, and folds the
// product pairs of each launch into its four accumulator cells (acc_at order:
// rows 4(l/16) .. 4(l/16)+3 of column l & 15) with the workgroup form's
// packed_term / XF::add3 arithmetic, so the results are bit-identical.  No LDS,
// no barrier.  The accumulator and product-ready tiles are whole tiles, so all
// 256 cells are read and written with no bounds checks (cells beyond the table
// are never used).
template <int W>
__device__ __forceinline__ void bulk_wave_item(const InstDesc<XFv> &ins, int idx, int t) {
  constexpr int T = 16;
  typedef float v4f __attribute__((ext_vector_type(4)));
  const int lane = threadIdx.x & 63, g = lane >> 4, col = lane & 15;
  const int NT = ins.NT;
  const int nd = sched::bulk_dists(NT, t, W, ins.KB);
  if (idx >= bulk_items(NT, t, W, nd)) return;
  int I, J;
  sched::bulk_tile(NT, t, W, nd, idx, I, J);
  const int smin = max(t - W + 1, (J - I + 3) / 2);  // a product of launch s needs dist <= 2s - 2
  MESSI_GLOBAL XFv *acc = gp(ins.acc) + (unsigned)(I * NT + J) * (T * T) + lane * 4;
  XFv av[4];
  if (had_visit(t, J - I, W, ins.KB, false)) {
#pragma unroll
    for (int q = 0; q < 4; ++q) av[q] = XFv(acc[q]);
  } else {
#pragma unroll
    for (int q = 0; q < 4; ++q) av[q] = XF::zero();
  }
  // the operands of product (I, K) x (K, J) for this lane: its float4 of each
  // tile and the exponents of its rows of A and its column of B
  struct Ops {
    float4 A, B;
    int4 ra;
    int cb;
  };
  auto load = [&](int K) {
    const unsigned ta = (unsigned)(I * NT + K), tb = (unsigned)(K * NT + J);
    Ops o;
    o.A = *reinterpret_cast<const MESSI_GLOBAL float4 *>(gp(ins.Lp) + ta * (T * T) + lane * 4);
    o.B = *reinterpret_cast<const MESSI_GLOBAL float4 *>(gp(ins.Sp) + tb * (T * T) + lane * 4);
    o.ra = *reinterpret_cast<const MESSI_GLOBAL int4 *>(gp(ins.LpE) + ta * T + 4 * g);
    o.cb = gp(ins.SpE)[tb * T + col];
    return o;
  };
  // the product for this lane's cells, as raw XF terms
  auto product = [&](const Ops &o, int K, XFv out[4]) {
    v4f m = {0.f, 0.f, 0.f, 0.f};
    m = __builtin_amdgcn_mfma_f32_16x16x4f32(o.A.x, o.B.x, m, 0, 0, 0);
    m = __builtin_amdgcn_mfma_f32_16x16x4f32(o.A.y, o.B.y, m, 0, 0, 0);
    m = __builtin_amdgcn_mfma_f32_16x16x4f32(o.A.z, o.B.z, m, 0, 0, 0);
    m = __builtin_amdgcn_mfma_f32_16x16x4f32(o.A.w, o.B.w, m, 0, 0, 0);
    const int ras[4] = {o.ra.x, o.ra.y, o.ra.z, o.ra.w};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const bool live = ras[q] != XF::EZ && o.cb != XF::EZ;
      const float cv = m[q];
      out[q] = live ? XFv{cv, ras[q] + o.cb} : XF::zero();
      if (live && !(cv >= 0x1p-100f)) out[q] = exact_cell_global<T>(ins, I, K, J, 4 * g + q, col);
    }
  };
  // each launch's pair of products; both products' loads issue before either
  // is used, so a launch's products cost one memory round trip
#pragma unroll
  for (int i = 0; i < W; ++i) {
    const int s = t - W + 1 + i, k1 = I + s - 1, k2 = J - (s - 1);
    const bool d1 = s >= smin && k1 - I <= ins.KB;
    const bool d2 = s >= smin && k2 != k1 && k2 - I <= ins.KB;
    if (!(d1 || d2)) continue;
    // one launch's operands in registers at a time (the next launch's loads are
    // not hoisted above this one's products)
    __builtin_amdgcn_sched_barrier(0);
    Ops o1, o2;
    if (d1) o1 = load(k1);
    if (d2) o2 = load(k2);
    XFv t1[4], t2[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) t1[q] = t2[q] = XF::zero();
    if (d1) product(o1, k1, t1);
    if (d2) product(o2, k2, t2);
#pragma unroll
    for (int q = 0; q < 4; ++q) av[q] = XF::add3(av[q], t1[q], t2[q]);
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) acc[q] = av[q];  // every visit writes
}

// One work item of the inside schedule (see the file header): the finalisation
// of tile (item, item+t) of launch t, or bulk workgroup `item` of launch t.
// With P (the single-launch form) the item first waits for what it reads and
// afterwards signals what it wrote (pflag).
template <class SR, int T, int W, bool P>
__device__ __forceinline__ void inside_item(const InstDesc<typename SR::V> &ins, bool fin_wg, int item, int t,
                                            const Rules<SR> &r, StepLDS<SR, T> &sm, int *ctl) {
  typedef typename SR::V V;
  typedef wave::Layout<T> WL;
  const int tid = threadIdx.x;
  constexpr bool by_step = !SR::scaled;  // in-tile wavefront cells in step order (see below)
  constexpr int BP = bulk_per(W);
  const int L = ins.L, NT = ins.NT;
  MESSI_GLOBAL V *__restrict__ Lr = gp(ins.Lr);
  MESSI_GLOBAL V *__restrict__ Lc = gp(ins.Lc);
  MESSI_GLOBAL V *__restrict__ Sr = gp(ins.Sr);
  MESSI_GLOBAL V *__restrict__ Sc = gp(ins.Sc);
  MESSI_GLOBAL V *__restrict__ F = gp(ins.F);
  MESSI_GLOBAL V *__restrict__ acc = gp(ins.acc);
  const MESSI_GLOBAL V *__restrict__ ptv = gp(ins.ptv);
  int *const tile_done = ins.pflag, *const visit_done = ins.pflag + (size_t)NT * NT;
  int *const diag_done = ins.pflag + 2 * (size_t)NT * NT;
  if (!fin_wg) {
    // The tiles (I, I+d) at the k-th distance d = t+1+kW gather the products of
    // the launches s = max(t-W+1, smin(d)) .. t in increasing s, the two of each
    // s in the order a window of one launch adds them.
    const int nd = sched::bulk_dists(NT, t, W, ins.KB);
    const int nbulk = bulk_items(NT, t, W, nd);
    const int first = BP * item;
    int Ib[BP], Jb[BP], smin[BP];
    bool has[BP], d1[BP][W], d2[BP][W];
    V av[BP];
    Operands<SR> ob1[BP][W], ob2[BP][W];
    PackedOps qb1[BP][W], qb2[BP][W];
    const int a = tid / T, b = tid % T;
#pragma unroll
    for (int u = 0; u < BP; ++u) {
      const int idx = first + u;
      has[u] = idx < nbulk;
      sched::bulk_tile(NT, t, W, nd, idx, Ib[u], Jb[u]);
      smin[u] = max(t - W + 1, (Jb[u] - Ib[u] + 3) / 2);  // a product of launch s needs dist <= 2s - 2
    }
    if constexpr (P) {
      // every tile of the diagonals before t is final, and the tiles' previous
      // visits have written their accumulators
      int pv[2] = {-1, -1};
#pragma unroll
      for (int u = 0; u < BP; ++u) pv[u] = has[u] ? last_visit(t, Jb[u] - Ib[u], W, ins.KB, false) : -1;
      if (!wg_wait(ctl, &sm.flag, diag_done + (t - 1), NT - (t - 1),
                   pv[0] >= 0 ? visit_done + (size_t)Ib[0] * NT + Jb[0] : nullptr, pv[0] + 1,
                   BP > 1 && pv[BP - 1] >= 0 ? visit_done + (size_t)Ib[BP - 1] * NT + Jb[BP - 1] : nullptr,
                   pv[BP - 1] + 1))
        return;
    }
#pragma unroll
    for (int u = 0; u < BP; ++u) {
      const int j = Ib[u] * T + a, y = Jb[u] * T + b;
      av[u] = (has[u] && had_visit(t, Jb[u] - Ib[u], W, ins.KB, false) && j < L && y < L)
                  ? typename SR::V(acc[(unsigned)(Ib[u] * NT + Jb[u]) * (T * T) + acc_at<T, SR::scaled>(a, b)])
                  : SR::zero();
#pragma unroll
      for (int i = 0; i < W; ++i) {
        const int s = t - W + 1 + i, k1 = Ib[u] + s - 1, k2 = Jb[u] - (s - 1);
        d1[u][i] = has[u] && s >= smin[u] && k1 - Ib[u] <= ins.KB;
        d2[u][i] = has[u] && s >= smin[u] && k2 != k1 && k2 - Ib[u] <= ins.KB;
        if constexpr (SR::scaled) {
          if (d1[u][i]) qb1[u][i] = fetch_packed<T>(ins, Ib[u], k1, Jb[u]);
          if (d2[u][i]) qb2[u][i] = fetch_packed<T>(ins, Ib[u], k2, Jb[u]);
        } else {
          if (d1[u][i]) ob1[u][i] = fetch_operands<SR, T>(inside_src<SR, T>(Ib[u], k1, Jb[u], Lr, Sc), L);
          if (d2[u][i]) ob2[u][i] = fetch_operands<SR, T>(inside_src<SR, T>(Ib[u], k2, Jb[u], Lr, Sc), L);
        }
      }
    }
#pragma unroll
    for (int u = 0; u < BP; ++u) {
      if (!has[u]) continue;
      if constexpr (SR::scaled && W >= 2) {
        // two launches' products per call, four wavefronts on MFMA
#pragma unroll
        for (int i = 0; i < W; i += 2) {
          const bool dk[4] = {d1[u][i], d2[u][i], d1[u][i + 1], d2[u][i + 1]};
          if (!(dk[0] || dk[1] || dk[2] || dk[3])) continue;
          const int s = t - W + 1 + i;
          const int Ks[4] = {Ib[u] + s - 1, Jb[u] - (s - 1), Ib[u] + s, Jb[u] - s};
          const PackedOps ops[4] = {qb1[u][i], qb2[u][i], qb1[u][i + 1], qb2[u][i + 1]};
          tile_products_packed4<T>(ins, Ib[u], Jb[u], ops, Ks, dk, sm.m.s.q.pk.q, av[u]);
        }
      } else {
#pragma unroll
        for (int i = 0; i < W; ++i) {
          if (!(d1[u][i] || d2[u][i])) continue;
          const int s = t - W + 1 + i, k1 = Ib[u] + s - 1, k2 = Jb[u] - (s - 1);
          if constexpr (SR::scaled) {
            tile_products_packed<T>(ins, Ib[u], Jb[u], qb1[u][i], k1, d1[u][i], qb2[u][i], k2, d2[u][i],
                                    sm.m.s.q.p, av[u]);
          } else {
            V acc2 = SR::zero();
            tile_products<SR, T>(ob1[u][i], d1[u][i], ob2[u][i], d2[u][i], sm.m.s.q.p, av[u], acc2);
            av[u] = SR::add(av[u], acc2);
          }
        }
      }
      const int j = Ib[u] * T + a, y = Jb[u] * T + b;
      if (j < L && y < L) acc[(unsigned)(Ib[u] * NT + Jb[u]) * (T * T) + acc_at<T, SR::scaled>(a, b)] = av[u];  // every visit writes
    }
    if constexpr (P)
      wg_signal(has[0] ? visit_done + (size_t)Ib[0] * NT + Jb[0] : nullptr, t + 1,
                BP > 1 && has[BP - 1] ? visit_done + (size_t)Ib[BP - 1] * NT + Jb[BP - 1] : nullptr, t + 1,
                nullptr);
    return;
  }
  const bool fin = true;
  const int I = item, J = item + t;
  const int IT = I * T, JT = J * T;
  if constexpr (P) {
    // the two neighbouring tiles of the previous diagonal are final (and with
    // them every tile this one reads), and the tile's last bulk visit is done
    const int lv = last_visit(t, t, W, ins.KB, true);
    if (!wg_wait(ctl, &sm.flag, t >= 1 ? tile_done + (size_t)I * NT + J - 1 : nullptr, 1,
                 t >= 1 ? tile_done + (size_t)(I + 1) * NT + J : nullptr, 1,
                 lv >= 0 ? visit_done + (size_t)I * NT + J : nullptr, lv + 1))
      return;
  }
  // ---- split-point products that became ready with tile-diagonal t-1: every
  // global load (accumulator and both products' operands) is issued up front.
  // L[j, h] is -inf for h - j beyond the pairing band: such L tiles contribute nothing.
  const int K1 = I + t - 1, K2 = J - (t - 1);
  const bool do1 = t >= 2 && K1 - I <= ins.KB;
  const bool do2 = t >= 2 && K2 != K1 && K2 - I <= ins.KB;
  V accv = SR::zero();
  {
    const int a = tid / T, b = tid % T, j = IT + a, y = JT + b;
    if (had_visit(t, t, W, ins.KB, true) && j < L && y < L)
      accv = acc[(unsigned)(I * NT + J) * (T * T) + acc_at<T, SR::scaled>(a, b)];
  }
  Operands<SR> o1, o2;
  PackedOps q1, q2;
  // Scaled path with T = 16: the two products run on wavefronts 0 and 1 straight
  // from the product-ready tiles (MFMA lane order: one float4 per operand and
  // lane), with no LDS staging of operands; every thread loads the row and column
  // exponents of its own cell.
  constexpr bool direct = SR::scaled && T == 16;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  float4 dA = {0.f, 0.f, 0.f, 0.f}, dB = {0.f, 0.f, 0.f, 0.f};
  int dra[2] = {0, 0}, dcb[2] = {0, 0};
  const bool dmine = wv == 0 ? do1 : (wv == 1 ? do2 : false);  // this wavefront runs a product
  if constexpr (direct) {
    const int r = tid / T, c = tid % T;
    if (dmine) {
      const int K = wv == 0 ? K1 : K2;
      const unsigned ta = (unsigned)(I * NT + K), tb = (unsigned)(K * NT + J);
      dA = *reinterpret_cast<const MESSI_GLOBAL float4 *>(gp(ins.Lp) + ta * (T * T) + (tid & 63) * 4);
      dB = *reinterpret_cast<const MESSI_GLOBAL float4 *>(gp(ins.Sp) + tb * (T * T) + (tid & 63) * 4);
    }
    if (do1) {
      dra[0] = gp(ins.LpE)[(unsigned)(I * NT + K1) * T + r];
      dcb[0] = gp(ins.SpE)[(unsigned)(K1 * NT + J) * T + c];
    }
    if (do2) {
      dra[1] = gp(ins.LpE)[(unsigned)(I * NT + K2) * T + r];
      dcb[1] = gp(ins.SpE)[(unsigned)(K2 * NT + J) * T + c];
    }
  } else if constexpr (SR::scaled) {
    if (do1) q1 = fetch_packed<T>(ins, I, K1, J);
    if (do2) q2 = fetch_packed<T>(ins, I, K2, J);
  } else {
    const OpSrc<SR> s1 = inside_src<SR, T>(I, K1, J, Lr, Sc), s2 = inside_src<SR, T>(I, K2, J, Lr, Sc);
    if (do1) o1 = fetch_operands<SR, T>(s1, L);
    if (do2) o2 = fetch_operands<SR, T>(s2, L);
  }
  // ---- finalising workgroup: stage the tile's inputs now, while the product
  // operands are in flight (these LDS arrays are not shared with the products).
  // Off-diagonal tiles of the scaled path stage the diagonal tiles and the pair
  // term as unscaled doubles for the double wavefront (dbl); dok turns false when
  // a value the double wavefront uses is out of its range.
  const bool dbl = SR::scaled && t > 0;
  adbl *const D = reinterpret_cast<adbl *>(sm.m.all);
  bool dok = true;
  // The staged tables of this thread's cell: the diagonal tiles (I, I) and
  // (J, J) and the pair term.  The loads use clamped indices (element 0 of the
  // table for a cell outside it) and select zero afterwards, so they issue with
  // no branch or wait between them and their latencies overlap; the LDS writes
  // come after every load.
  struct StagedCell {
    V ld, cp, sd;
  };
  auto load_tables = [&]() {
    const int a = tid / T, b = tid % T, j = IT + a, y = JT + b, jd = IT + b, ys = JT + a;
    const bool k1 = j < L && jd < L && jd >= j, k2 = j < y && y < L, k3 = ys < L && b <= a;
    const V x1 = typename SR::V(Lr[k1 ? (size_t)j * L + jd : 0]);   // L[j, jd] of tile (I, I)
    const V x2 = typename SR::V(ptv[k2 ? (size_t)j * L + y : 0]);   // pair term of the cell
    const V x3 = typename SR::V(Sc[k3 ? (size_t)ys * L + y : 0]);   // S[JT+b, JT+a] of tile (J, J)
    return StagedCell{k1 ? x1 : SR::zero(), k2 ? x2 : SR::zero(), k3 ? x3 : SR::zero()};
  };
  auto put_tables = [&](const StagedCell &g, bool as_d) {
    auto put = [&](int i, V v) {
      if constexpr (SR::scaled) {
        if (as_d) {
          D[i] = xf_to_d(v, 0, dok);
          return;
        }
      }
      sm.at(i) = v;
    };
    const int a = tid / T, b = tid % T;
    put(WL::ld(a, b), g.ld);
    put(WL::cp(a, b), g.cp);
    put(WL::sd(b, a), g.sd);
    for (int k = tid; k < WL::N_ZERO_STAGED; k += T * T) put(WL::zero_staged(k), SR::zero());
  };
  double sbd = 0.0;  // double wavefront: this lane's bottom-border S at the tile's scale (lanes < T)
  if (fin) {
    const StagedCell g = load_tables();
    // the border values (lanes < T): S of the tile below, F below and to the left
    V sb, fb, fl;
    {
      const int b = tid, y = JT + b, rb = (I + 1) * T, yc = JT + b - 1, ja = IT + b + 1, yl = JT - 1;
      const bool lane = tid < T && t >= 1;
      const bool k1 = lane && y < L && rb <= y, k2 = lane && rb < L && yc < L && rb <= yc,
                 k3 = lane && ja < L && yl >= 0 && ja <= yl;
      const V x1 = typename SR::V(Sc[k1 ? (size_t)y * L + rb : 0]);
      const V x2 = typename SR::V(F[k2 ? (size_t)rb * L + yc : 0]);
      const V x3 = typename SR::V(F[k3 ? (size_t)ja * L + yl : 0]);
      sb = k1 ? x1 : SR::zero();
      fb = k2 ? x2 : SR::zero();
      fl = k3 ? x3 : SR::zero();
    }
    put_tables(g, dbl);
    if (tid < T) {
      sm.Sb[tid] = sb;
      sm.Fbot[tid] = fb;
      sm.Fleft[tid] = fl;
    }
    if (tid == 0) sm.bad = 0;
    if constexpr (SR::scaled) {
      if (dbl && tid < 64) {
        // the tile's scale: the largest exponent of its nonzero border values
        // (S of the tile below, F below and to the left), 0 when all are zero;
        // then the borders at that scale (the S row goes into the tile's LDS
        // area once the products are done)
        int e = INT_MIN / 2;
        if (tid < T) {
          if (sb.m != 0.0f) e = max(e, sb.e);
          if (fb.m != 0.0f) e = max(e, fb.e);
          if (fl.m != 0.0f) e = max(e, fl.e);
        }
        int E = border_emax<T>(e);
        if (E < -(1 << 27)) E = 0;
        if (tid < T) {
          sbd = xf_to_d(sb, E, dok);
          sm.Fbd[tid] = xf_to_d(fb, E, dok);
          sm.Fld[tid] = xf_to_d(fl, E, dok);
        }
        if (tid == 0) sm.escale = E;
      }
    }
  }
  if constexpr (direct) {
    // nothing shares the wavefront's LDS area with the products here, so its zero
    // regions and bottom border row are written now, under the loads' latency
    if (t > 0) {
      if (!dbl) {
        for (int k = tid; k < WL::N_ZERO_WAVE; k += T * T) sm.at(WL::zero_wave(k)) = SR::zero();
        if (tid < T) sm.at(WL::cs(T, tid)) = sm.Sb[tid];
      } else {
        for (int k = tid; k < WL::N_ZERO_WAVE; k += T * T) D[WL::zero_wave(k)] = 0.0;
        if (tid < T) D[WL::cs(T, tid)] = sbd;
      }
    }
    if (dmine) {
      typedef float v4f __attribute__((ext_vector_type(4)));
      const int lane = tid & 63;
      v4f mm = {0.f, 0.f, 0.f, 0.f};
      mm = __builtin_amdgcn_mfma_f32_16x16x4f32(dA.x, dB.x, mm, 0, 0, 0);
      mm = __builtin_amdgcn_mfma_f32_16x16x4f32(dA.y, dB.y, mm, 0, 0, 0);
      mm = __builtin_amdgcn_mfma_f32_16x16x4f32(dA.z, dB.z, mm, 0, 0, 0);
      mm = __builtin_amdgcn_mfma_f32_16x16x4f32(dA.w, dB.w, mm, 0, 0, 0);
#pragma unroll
      for (int q = 0; q < 4; ++q) sm.Rd[wv][4 * (lane >> 4) + q][lane & 15] = mm[q];
    }
    __syncthreads();
    if (do1 || do2) {
      // each thread folds its cell as tile_products_packed does (packed_term, XF::add3)
      const int r = tid / T, c = tid % T;
      XFv tk[2];
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        tk[k] = XF::zero();
        if (k == 0 ? do1 : do2) {
          const bool live = dra[k] != XF::EZ && dcb[k] != XF::EZ;
          const float cv = sm.Rd[k][r][c];
          tk[k] = live ? XFv{cv, dra[k] + dcb[k]} : XF::zero();
          if (live && !(cv >= 0x1p-100f)) tk[k] = exact_cell_global<T>(ins, I, k == 0 ? K1 : K2, J, r, c);
        }
      }
      accv = XF::add3(accv, tk[0], tk[1]);
    }
  } else if (do1 || do2) {
    if constexpr (SR::scaled) {
      tile_products_packed<T>(ins, I, J, q1, K1, do1, q2, K2, do2, sm.m.s.q.p, accv);
    } else {
      V acc2 = SR::zero();
      tile_products<SR, T>(o1, do1, o2, do2, sm.m.s.q.p, accv, acc2);
      accv = SR::add(accv, acc2);
    }
  }
  if (!fin) {
    const int a = tid / T, b = tid % T, j = IT + a, y = JT + b;
    if (j < L && y < L) acc[(size_t)j * L + y] = accv;
    return;
  }
  if (t == 0) {
    for (int idx = tid; idx < T * T; idx += T * T) {
      int a = idx / T, b = idx % T;
      int j = IT + a;
      bool d = (a == b) && j < L;
      sm.at(WL::cs(a, b)) = d ? typename SR::V(Sr[(size_t)j * L + j]) : SR::zero();
      sm.at(WL::cl(a, b)) = d ? typename SR::V(Lr[(size_t)j * L + j]) : SR::zero();
      sm.at(WL::cf(a, b)) = SR::zero();
    }
  } else if constexpr (!direct) {
    if (by_step) sm.at(WL::cf(tid / T, tid % T)) = accv;  // to the thread that takes the cell below
    // the zero regions of the wavefront's reads and the bottom border row as row T of cS
    if (!dbl) {
      for (int k = tid; k < WL::N_ZERO_WAVE; k += T * T) sm.at(WL::zero_wave(k)) = SR::zero();
      if (tid < T) sm.at(WL::cs(T, tid)) = sm.Sb[tid];
    } else if constexpr (SR::scaled) {
      for (int k = tid; k < WL::N_ZERO_WAVE; k += T * T) D[WL::zero_wave(k)] = 0.0;
      if (tid < T) D[WL::cs(T, tid)] = sbd;
    }
  }
  // (the direct form's staging was ordered by its barrier above, except the
  // diagonal tiles' cells just written)
  if (!direct || t == 0) __syncthreads();
  // One thread per cell.  A cell's split-point sum is accumulated online: at
  // step s it adds the (at most two) terms whose newest operand was finalised at
  // step s-1, and it finalises itself at its own step.  In the FP64 path
  // off-diagonal tiles give the cells to the threads in step order
  // (cell_by_step): a wavefront's cells then finish together and the costly
  // finalisations of one step fall in one or two wavefronts; each thread reads its
  // accumulated products back from the cell's F slot, which only the cell's own
  // finalisation overwrites.  (In the FP32 path step order scatters a
  // wavefront's LDS reads over many rows; row order, whose lanes read
  // neighbouring cells, measured faster.)
  int a = tid / T, b = tid % T;
  if (by_step && t > 0) cell_by_step<T>(tid, a, b);
  const int j = IT + a, y = JT + b;
  const bool mine = j < L && y < L && (t > 0 || a < b);
  const typename SR::FinRules rk = SR::fin_rules(r.p[6], r.lin[0], r.lin[2], r.lin[3], r.lin[4], r.lin[6]);
  typename SR::Acc ac = SR::acc(by_step && t > 0 ? sm.at(WL::cf(a, b)) : accv);
  if (t == 0) {
    const int own = b - a;
    for (int st = 1; st < T; ++st) {
      const bool act = mine && st <= own, fin = mine && st == own;
      V l1 = SR::zero(), s1 = SR::zero(), l2 = SR::zero(), s2 = SR::zero(), f = SR::zero(), pt = SR::zero();
      if (act) {
        const bool ok = own - st <= st - 1;
        const int c1 = a + st - 1, c2 = b - st;
        if (ok && c1 <= b - 1) {
          l1 = sm.at(WL::cl(a, c1));
          s1 = sm.at(WL::cs(c1 + 1, b));
        }
        if (ok && c2 >= a && c2 != c1) {
          l2 = sm.at(WL::cl(a, c2));
          s2 = sm.at(WL::cs(c2 + 1, b));
        }
      }
      if (fin) {
        pt = sm.at(WL::cp(a, b));
        if (own >= 2) f = sm.at(WL::cf(a + 1, b - 1));
      }
      if (act) SR::acc_add2(ac, SR::mul_raw(l1, s1), SR::mul_raw(l2, s2));
      if (fin) {
        V vs, vf, vl;
        SR::finalize(ac, SR::mul_raw(pt, f), rk, vs, vf, vl);
        sm.at(WL::cs(a, b)) = vs;
        sm.at(WL::cf(a, b)) = vf;
        sm.at(WL::cl(a, b)) = vl;
      }
      __syncthreads();
    }
  } else if constexpr (SR::scaled) {
    // Off-diagonal tile, scaled path.  Every step reads the lane's four cells at
    // fixed offsets from per-lane bases (wave::Layout): terms that must be zero
    // read a zero cell, terms after the lane's finalisation feed a dead
    // accumulator, so a step has no index arithmetic and no masks.  The
    // bottom-border split point h = IT+T-1 is the column term of step b (row T
    // of cS holds S[(I+1)T, y]).  Lanes of columns y >= L compute values nothing
    // reads (the tile's in-range cells read only in-range cells) and are not
    // written back.
    const int own = b + T - 1 - a;
    const int bcl = WL::base_col_l(a, b), bcs = WL::base_col_s(a, b);
    const int brl = WL::base_row_l(a, b), brs = WL::base_row_s(a, b);
    // F[j+1, y-1] for the pair term (inner cell, finalised two steps earlier, or
    // the bottom / left border), read one step ahead at a fixed position
    const bool fin_in = a + 1 < T && b >= 1;
    const int wave_last = __builtin_amdgcn_readfirstlane(2 * T - 2 - (tid & ~63) / T);
    bool xf_wave = !dbl;
    if (dbl) {
      // Double wavefront (see D_LO / D_HI): the same steps on doubles, the cells at
      // 2^-E.  S and F of a span of two or more are positive, so a zero there, like
      // any value out of range, or a pair term that underflowed, hands the tile to
      // the per-cell-exponent wavefront below, which recomputes it from the tables.
      const int E = sm.escale;
      double acd = xf_to_d(accv, E, dok);
      const adbl *fsrc = fin_in ? &D[WL::cf(a + 1, b - 1)]
                                : (a + 1 == T ? reinterpret_cast<const adbl *>(&sm.Fbd[b])
                                              : reinterpret_cast<const adbl *>(&sm.Fld[a]));
      const double pterm = j + 1 > y - 1 ? 0.0 : D[WL::cp(a, b)];
      const double p0 = rk.p0, p2 = rk.p2, p3 = rk.p3, p4 = rk.p4, p6 = rk.p6;
      double fin_next = *fsrc;
      double vs = 0.0, vf = 0.0, vl = 0.0, vpf = 0.0, vfin = 0.0;
#pragma unroll
      for (int st = 0; st <= 2 * T - 2; ++st) {
        if (st <= wave_last) {
          const double fin_f = fin_next;
          fin_next = *fsrc;
          acd = fma(D[bcl + WL::off_col_l(st)], D[bcs + WL::off_col_s(st)], acd);
          acd = fma(D[brl + WL::off_row_l(st)], D[brs + WL::off_row_s(st)], acd);
          if (st == own) {
            vfin = fin_f;
            vpf = pterm * fin_f;
            vs = fma(p0, acd, p2 * vpf);
            vf = fma(p4, acd, p3 * vpf);
            vl = p6 * vpf;
            D[WL::cs(a, b)] = vs;
            D[WL::cf(a, b)] = vf;
            D[WL::cl(a, b)] = vl;
          }
        }
        __syncthreads();
      }
      if (mine) {
        const bool pf_ok = pterm == 0.0 || vfin == 0.0 || (vpf != 0.0 && d_in_range(vpf));
        dok = dok && vs != 0.0 && vf != 0.0 && d_in_range(vs) && d_in_range(vf) && d_in_range(vl) && pf_ok;
      }
      if (!dok) atomicOr(&sm.bad, 1);
      // the cell as an XF value in its LDS slots for the write-back (no other
      // thread reads the doubles after the last step's barrier)
      sm.at(WL::cs(a, b)) = d_to_xf(vs, E);
      sm.at(WL::cf(a, b)) = d_to_xf(vf, E);
      sm.at(WL::cl(a, b)) = d_to_xf(vl, E);
      __syncthreads();
      xf_wave = sm.bad != 0;
      if (xf_wave) {
        // restage the tables, zero regions and bottom border as XF values
        put_tables(load_tables(), false);
        for (int k = tid; k < WL::N_ZERO_WAVE; k += T * T) sm.at(WL::zero_wave(k)) = SR::zero();
        if (tid < T) sm.at(WL::cs(T, tid)) = sm.Sb[tid];
        __syncthreads();
      }
    }
    if (xf_wave) {
      // Per-cell-exponent wavefront: each cell keeps its own exponent, so no
      // range limit applies.
      const V *fsrc = fin_in ? &sm.at(WL::cf(a + 1, b - 1)) : (a + 1 == T ? &sm.Fbot[b] : &sm.Fleft[a]);
      const V pterm = j + 1 > y - 1 ? SR::zero() : sm.at(WL::cp(a, b));
      V fin_next = *fsrc;
#pragma unroll
      for (int st = 0; st <= 2 * T - 2; ++st) {
        if (st <= wave_last) {
          const V fin_f = fin_next;
          fin_next = *fsrc;
          const V t1 = SR::mul_raw(sm.at(bcl + WL::off_col_l(st)), sm.at(bcs + WL::off_col_s(st)));
          const V t2 = SR::mul_raw(sm.at(brl + WL::off_row_l(st)), sm.at(brs + WL::off_row_s(st)));
          SR::acc_add2(ac, t1, t2);
          if (st == own) {
            V vs, vf, vl;
            SR::finalize(ac, SR::mul_raw(pterm, fin_f), rk, vs, vf, vl);
            sm.at(WL::cs(a, b)) = vs;
            sm.at(WL::cf(a, b)) = vf;
            sm.at(WL::cl(a, b)) = vl;
          }
        }
        __syncthreads();
      }
    }
  } else {
    const int own = b + T - 1 - a;
    // split point h = IT+T-1: S[(I+1)T, y] lies in the tile below.
    if (mine) SR::acc_add2(ac, SR::mul_raw(sm.at(WL::ld(a, T - 1)), sm.Sb[b]), SR::zero());
    // F[j+1, y-1] for the pair term: inside this tile, or on the tile's bottom /
    // left border.  The inner cell (a+1, b-1) is finalised two steps before this
    // cell (its step is own - 2, and own >= 2 for it).  Every step reads the
    // lane's F cell (inner or border, a fixed LDS position) for use in the next
    // step, so the read overlaps a whole step instead of delaying the
    // finalisation; with no inner pair (j + 1 > y - 1) the pair term is zero.
    const bool fin_in = a + 1 < T && b >= 1;
    const V *fsrc = fin_in ? &sm.at(WL::cf(a + 1, b - 1)) : (a + 1 == T ? &sm.Fbot[b] : &sm.Fleft[a]);
    const V pterm = j + 1 > y - 1 ? SR::zero() : sm.at(WL::cp(a, b));
    // primed before step 0: a border cell may finalise at step 0 (own = 0) or 1;
    // an inner cell's first reads are stale but unused (own >= 2)
    V fin_next = *fsrc;
    auto work = [&](int st) {
      const V fin_f = fin_next;
      fin_next = *fsrc;
      const bool act = mine && st >= 1 && st <= own;
      // column producer (ap, b) and row producer (a, bp), finalised at step st-1;
      // both terms are read at clamped (always valid) positions and masked, so
      // the step has no divergent branch but the finalisation
      const int ap = b + T - st, bp = st - T + a;
      const int apc = min(max(ap, 1), T - 1), bpc = min(max(bp, 0), T - 2);
      const V t1 = SR::mul_raw(sm.at(WL::ld(a, apc - 1)), sm.at(WL::cs(apc, b)));
      const V t2 = SR::mul_raw(sm.at(WL::cl(a, bpc)), sm.at(WL::sd(bpc + 1, b)));
      SR::acc_add2(ac, SR::keep(t1, act && ap > a && ap < T), SR::keep(t2, act && bp >= 0 && bp < b));
      if (mine && st == own) {
        V vs, vf, vl;
        SR::finalize(ac, SR::mul_raw(pterm, fin_f), rk, vs, vf, vl);
        sm.at(WL::cs(a, b)) = vs;
        sm.at(WL::cf(a, b)) = vf;
        sm.at(WL::cl(a, b)) = vl;
      }
    };
    // the last step with work for this wavefront (a scalar, so the skip below is
    // a uniform branch and every step has one barrier): that of its first cell in
    // step order, or of the last cell of its first row in row order
    int wlast;
    if constexpr (by_step) {
      int wa, wb;
      cell_by_step<T>(tid & ~63, wa, wb);
      wlast = wb + T - 1 - wa;
    } else {
      wlast = 2 * T - 2 - (tid & ~63) / T;
    }
    const int wave_last = __builtin_amdgcn_readfirstlane(wlast);
    auto step = [&](int st) {
      if (st <= wave_last) work(st);
      __syncthreads();
    };
    // The FP64 step (exponentials and logarithms) stays a rolled loop: unrolled,
    // the kernel outgrows the instruction cache.
#pragma unroll 1
    for (int st = 0; st <= 2 * T - 2; ++st) step(st);
  }
  // ---- write the tile back (row-major copies with b fastest, column-major with a fastest)
  for (int idx = tid; idx < T * T; idx += T * T) {
    int a = idx / T, b = idx % T, j = IT + a, y = JT + b;
    if (j < L && y < L && j < y) {
      size_t rc = (size_t)j * L + y;
      Sr[rc] = sm.at(WL::cs(a, b));
      F[rc] = sm.at(WL::cf(a, b));
      Lr[rc] = sm.at(WL::cl(a, b));
    }
  }
  for (int idx = tid; idx < T * T; idx += T * T) {
    int b = idx / T, a = idx % T, j = IT + a, y = JT + b;
    if (j < L && y < L && j < y) {
      size_t cc = (size_t)y * L + j;
      Sc[cc] = sm.at(WL::cs(a, b));
      Lc[cc] = sm.at(WL::cl(a, b));
    }
  }
  if constexpr (SR::scaled) {
    // Product-ready copies of the finished tile.  A operand: each row segment of
    // L aligned to its largest exponent (the T lanes of a row reduce by DPP).
    {
      const int a = tid / T, b = tid % T, j = IT + a, y = JT + b;
      const bool ok = j < L && y < L;
      const XFv v = sm.at(WL::cl(a, b));
      const int e = group_max<T>(ok && v.m != 0.0f ? v.e : XF::EZ);
      // every lane writes (zero beyond the table), so products read whole tiles
      // with no bounds checks
      gp(ins.Lp)[((size_t)I * NT + J) * (T * T) + op_a_at<T>(a, b)] = !ok || v.m == 0.0f ? 0.0f : ldexpf(v.m, max(v.e - e, -200));
      if (b == 0) gp(ins.LpE)[((size_t)I * NT + J) * T + a] = e;
    }
    // B operand: the column segment of S over rows IT+1 .. IT+T (the last one is
    // row (I+1)T of the tile below, staged in Sb), aligned to its largest exponent.
    {
      const int c = tid % T, b = tid / T, row = IT + 1 + c, y = JT + b;
      XFv v = XF::zero();
      if (row < L && y < L && row <= y) v = c + 1 < T ? sm.at(WL::cs(c + 1, b)) : sm.Sb[b];
      const int e = group_max<T>(v.m != 0.0f ? v.e : XF::EZ);
      gp(ins.Sp)[((size_t)I * NT + J) * (T * T) + op_b_at<T>(c, b)] = v.m == 0.0f ? 0.0f : ldexpf(v.m, max(v.e - e, -200));
      if (c == 0) gp(ins.SpE)[((size_t)I * NT + J) * T + b] = e;
    }
  }
  if constexpr (P) wg_signal(tile_done + (size_t)I * NT + J, 1, nullptr, 0, diag_done + t);
}

// One launch of the inside schedule (see the file header).  The scaled T = 16
// form holds seven workgroups per CU (its LDS allows seven; the register budget
// is set to match, 72 per lane).
template <class SR, int T, int W>
__global__ __launch_bounds__(T * T, (SR::scaled && T == 16) ? 7 : 1) void inside_step_kernel(
    const InstDesc<typename SR::V> *__restrict__ insts, const int *__restrict__ prefix, int n, int t, Rules<SR> r) {
  __shared__ StepLDS<SR, T> sm;
  static_assert(sizeof(StepStaged<SR, T>) == sizeof(typename SR::V) * wave::Layout<T>::STAGED +
                                                 sizeof(StepProducts<SR, T>),
                "the products follow the staged tables");
  // ---- work item.  The first workgroups (in dispatch order, so they start
  // first, round-robin over the XCDs) finalise the tiles of every instance:
  // prefix[0..n] counts them per instance; a few padding workgroups round their
  // number up to a multiple of 8.  The rest are bulk workgroups, prefix[n+1 ..
  // 2n+1] per instance, renumbered so that each XCD takes a contiguous run.
  const int *fpre = prefix, *bpre = prefix + (n + 1);
  const int nfin_all = fpre[n], fpad = (nfin_all + 7) & ~7;
  const bool fin_wg = (int)blockIdx.x < fpad;
  if (fin_wg && (int)blockIdx.x >= nfin_all) return;  // padding
  const int bx = fin_wg ? (int)blockIdx.x : xcd_chunk_id((int)blockIdx.x - fpad, (int)gridDim.x - fpad);
  const int *pre = fin_wg ? fpre : bpre;
  int lo = 0, hi = n;  // instance: pre[lo] <= bx < pre[lo+1]
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (pre[mid] <= bx) lo = mid;
    else hi = mid;
  }
  if constexpr (SR::scaled && T == 16) {
    // bulk items one per wavefront (WAVE_ITEMS per workgroup), in registers
    if (!fin_wg) {
      bulk_wave_item<W>(insts[lo], WAVE_ITEMS * (bx - pre[lo]) + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)), t);
      return;
    }
  }
  inside_item<SR, T, W, false>(insts[lo], fin_wg, bx - pre[lo], t, r, sm, nullptr);
}

// The whole inside pass of a batch in one launch (single-launch form): one
// workgroup per work item of every launch of the schedule.  A workgroup takes
// its item from a ticket counter when it starts, in the order of the ticket
// table (inside_sched.h), and waits on the flags of the items it reads.  An
// item only waits for items of smaller tickets, which workgroups that already
// started hold, so the oldest unfinished item can always proceed whatever order
// the workgroups are dispatched in; the waits are bounded (PERSIST_SPIN).
template <class SR, int T, int W>
__global__ __launch_bounds__(T * T) void inside_persist_kernel(const InstDesc<typename SR::V> *__restrict__ insts,
                                                               const int *__restrict__ tab, int n, int NP,
                                                               Rules<SR> r, int *ctl) {
  typedef typename SR::V V;
  __shared__ StepLDS<SR, T> sm;
  if (threadIdx.x == 0)
    sm.flag = __hip_atomic_load(ctl + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                  ? -1
                  : __hip_atomic_fetch_add(ctl, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  const int k = __builtin_amdgcn_readfirstlane(sm.flag);  // uniform: the decode stays scalar
  __syncthreads();  // the item's waits reuse sm.flag
  if (k < 0 || k >= tab[NP]) return;  // another workgroup gave up
  const sched::Ticket w = sched::decode_ticket(tab, n, NP, k);
  const InstDesc<V> ins = insts[w.m];
  const int item = w.kind == 2 ? w.item + sched::crit_groups(ins.NT, w.t, W, ins.KB) : w.item;
  inside_item<SR, T, W, true>(ins, w.kind == 0, item, w.t, r, sm, ctl);
}

// ---------------------------------------------------------------------------
// Tiled outside.  With W = p(S->LS) oS + p(F->LS) oF (the weight of a cell as
// the parent of a bifurcation),
//   oL[j,y] = sum_{k>y} W[j,k] iS[y+1,k]        (L is the left child)
//   oS[j,y] = sum_{k<j} W[k,y] iL[k,j-1]        (S is the right child)
//   oF[j,y] = ptv[j-1,y+1] Q[j-1,y+1],  Q = p2 oS + p3 oF + p6 oL.
// Tiles are finalised by descending tile distance d = J - I (launch u finalises
// d = NT-1-u).  The sums over k outside the tile's own row / column tiles are
// tile products: row products W(I, J+m) x iS(rows JT+1.., cols (J+m)T..)^T into
// the L accumulator, column products iL(rows (I-m)T.., cols IT-1..)^T x W(I-m, J)
// into the S accumulator; the pair with offset m becomes ready in launch
// NT - d - m (its W tile has distance d + m).  m = 1 runs in the finalising
// workgroup, m >= 2 in bulk workgroups.  The remaining terms (k inside tile J,
// resp. I) are pulled by an in-LDS wavefront over the tile's cells in
// descending span order, at most one row and one column term per step.
// ---------------------------------------------------------------------------
struct OutDesc {
  int L, NT, band, lband;
  const void *iSr, *iLc, *ptv;              // inside S (row-major), L (column-major), pair term
  void *oSr, *oSc, *oFr, *oL, *Wr, *Wc;     // outside tables and W (row- and column-major)
  void *accL, *accS;                        // bulk accumulators (row-major)
  // Scaled path with T = 16: product-ready operands of the bulk products, tile-major
  // in MFMA lane order (op_a_at / op_b_at), one exponent per A row or B column.
  //   WpA (I, K): W tile (I, K) as the A operand of a row product, rows aligned
  //   WpB (K, J): W tile (K, J) as the B operand of a column product, columns aligned
  //   iSpB (J, K): B[c][b] = iS[JT+b+1, KT+c] (row product), columns b aligned
  //   iLpA (K, I): A[a][c] = iL[KT+c, IT+a-1] (column product), rows a aligned
  float *WpA, *WpB, *iSpB, *iLpA;
  int *WpAE, *WpBE, *iSpBE, *iLpAE;
};

// Per launch u and tile distance d < NT-1-u: the bulk tiles (I, I+d) with a
// ready product are I in [0, rend) and [cbeg, cend); cum counts the items of
// the smaller distances.
struct OutRange {
  int cum, rend, cbeg, cend;
};

__host__ __device__ inline bool out_row_ok(int d, int T, int lband) {
  return lband < 0 || (d == 0 ? 0 : d * T - (T - 1)) <= lband;
}
__host__ __device__ inline bool out_col_ok(int m, int T, int band) { return band < 0 || (m - 1) * T <= band; }

template <class SR, int T> struct OutLDS {
  typedef typename SR::V V;
  union {
    ProdLDS<SR, T> p[2];  // product phase
    struct {
      V W[T][T + 1], Q[T][T + 1];
      V iSd[T][T + 1];  // iS[JT+b+1, JT+c] as [b][c]
      V iLd[T][T + 1];  // iL[IT+c, IT+a-1] as [c][a]
      V cP[T][T + 1];   // ptv[j-1, y+1]
    } w;   // wavefront phase
  } u;
  V Qtop[T], Qright[T];
  double Qtd[T], Qrd[T];  // scaled path, double wavefront: Qtop and Qright at the tile's scale
  int escale;             // scaled path, double wavefront: the tile's power-of-two scale
  int bad;                // scaled path, double wavefront: a value left the safe range
};

// Product-ready copies of the inside tables for the outside's bulk products
// (scaled path, T = 16), one workgroup per tile pair p < q: iSpB (p, q) and
// iLpA (p, q) (see OutDesc), aligned as tile_products aligns the operands it
// stages (largest nonzero exponent of a B column / an A row), so a product from
// them has the same bits.
__global__ __launch_bounds__(256) void out_operands_kernel(OutDesc od) {
  constexpr int T = 16;
  const int p = blockIdx.y, q = blockIdx.x;
  if (p >= q) return;
  const int L = od.L, NT = od.NT, tid = threadIdx.x, hi = tid / T, lo = tid % T;
  const MESSI_GLOBAL XFv *iSr = gp(static_cast<const XFv *>(od.iSr));
  const MESSI_GLOBAL XFv *iLc = gp(static_cast<const XFv *>(od.iLc));
  const unsigned tile = (unsigned)(p * NT + q);
  {
    // B[c][b] = iS[pT + b + 1, qT + c] (c = lo, b = hi), columns b aligned
    const int c = lo, b = hi, row = p * T + b + 1, col = q * T + c;
    const XFv v = (row < L && col < L && row <= col) ? XFv(iSr[(size_t)row * L + col]) : XF::zero();
    const int e = group_max<T>(v.m != 0.0f ? v.e : XF::EZ);
    gp(od.iSpB)[tile * (T * T) + op_b_at<T>(c, b)] = v.m == 0.0f ? 0.0f : ldexpf(v.m, max(v.e - e, -200));
    if (c == 0) gp(od.iSpBE)[tile * T + b] = e;
  }
  {
    // A[a][c] = iL[pT + c, qT + a - 1] (a = hi, c = lo), rows a aligned
    const int a = hi, c = lo, row = p * T + c, col = q * T + a - 1;
    const XFv v = (col >= 0 && row < L && col < L && row <= col) ? XFv(iLc[(size_t)col * L + row]) : XF::zero();
    const int e = group_max<T>(v.m != 0.0f ? v.e : XF::EZ);
    gp(od.iLpA)[tile * (T * T) + op_a_at<T>(a, c)] = v.m == 0.0f ? 0.0f : ldexpf(v.m, max(v.e - e, -200));
    if (c == 0) gp(od.iLpAE)[tile * T + a] = e;
  }
}

// Exact sums of one cell of an outside bulk product from the tables (the rare
// cell whose aligned MFMA sum lost its significant terms), as exact_cell_xf
// computes them from staged operands: the row product (I, K) x (J, K)^T,
//   sum_c W[IT+a, KT+c] iS[JT+b+1, KT+c],
// and the column product,
//   sum_c iL[KT+c, IT+a-1] W[KT+c, JT+b].
template <int T>
__device__ __forceinline__ XFv exact_out_cell(const OutDesc &od, bool row_product, int I, int K, int J, int a, int b) {
  const int L = od.L;
  const MESSI_GLOBAL XFv *Wr = gp(static_cast<const XFv *>(od.Wr));
  const MESSI_GLOBAL XFv *Wc = gp(static_cast<const XFv *>(od.Wc));
  const MESSI_GLOBAL XFv *iSr = gp(static_cast<const XFv *>(od.iSr));
  const MESSI_GLOBAL XFv *iLc = gp(static_cast<const XFv *>(od.iLc));
  auto av = [&](int c) -> XFv {
    if (row_product) {
      const int r = I * T + a, k = K * T + c;
      return (r < L && k < L && r <= k) ? XFv(Wr[(size_t)r * L + k]) : XF::zero();
    }
    const int r = K * T + c, k = I * T + a - 1;  // iL[r, k], column-major
    return (k >= 0 && r < L && k < L && r <= k) ? XFv(iLc[(size_t)k * L + r]) : XF::zero();
  };
  auto bv = [&](int c) -> XFv {
    if (row_product) {
      const int r = J * T + b + 1, k = K * T + c;  // iS[r, k]
      return (r < L && k < L && r <= k) ? XFv(iSr[(size_t)r * L + k]) : XF::zero();
    }
    const int r = K * T + c, k = J * T + b;  // W[r, k], column-major copy
    return (r < L && k < L && r <= k) ? XFv(Wc[(size_t)k * L + r]) : XF::zero();
  };
  int km = XF::kmin();
#pragma unroll 1
  for (int c = 0; c < T; ++c) km = max(km, XF::key(av(c), bv(c)));
  float sum = 0.0f;
#pragma unroll 1
  for (int c = 0; c < T; ++c) sum += XF::contrib(av(c), bv(c), km);
  return XF::finish(km, sum);
}

// One bulk visit of the outside's scaled path with T = 16, by one wavefront in
// registers (the tile, products and order of the workgroup form in
// outside_step_kernel, with tile_products' arithmetic, so the results are
// bit-identical): bulk item idx of launch u.  Lane l holds cells (4(l/16) + q,
// l & 15), q = 0..3, of both accumulators.
template <int W>
__device__ __forceinline__ void out_bulk_wave_item(const OutDesc &od, const OutRange *rg, int u, int idx) {
  constexpr int T = 16;
  typedef float v4f __attribute__((ext_vector_type(4)));
  const int lane = threadIdx.x & 63, g = lane >> 4, col = lane & 15;
  const int L = od.L, NT = od.NT, du = NT - 1 - u;
  if (du < 1) return;
  const OutRange last = rg[du - 1];
  if (idx >= last.cum + last.rend + (last.cend - last.cbeg)) return;
  int lo = 0, hi = du - 1;  // largest distance with cum <= idx
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (rg[mid].cum <= idx) lo = mid;
    else hi = mid - 1;
  }
  const int d = lo, i0 = idx - rg[d].cum;
  const int I = i0 < rg[d].rend ? i0 : rg[d].cbeg + (i0 - rg[d].rend);
  const int m = du + 1 - d, J = I + d;
  unsigned rows = 0, cols = 0;  // bit i: offset m + W - 1 - i has a row / column product
#pragma unroll
  for (int i = 0; i < W; ++i) {
    const int mi = m + W - 1 - i;
    if (out_row_ok(d, T, od.lband) && J + mi <= NT - 1) rows |= 1u << i;
    if (out_col_ok(mi, T, od.band) && I - mi >= 0) cols |= 1u << i;
  }
  MESSI_GLOBAL XFv *accL = gp(static_cast<XFv *>(od.accL));
  MESSI_GLOBAL XFv *accS = gp(static_cast<XFv *>(od.accS));
  const int y = J * T + col;
  XFv vL[4], vS[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int j = I * T + 4 * g + q;
    const bool cell = j < L && y < L;
    vL[q] = (cell && rows) ? XFv(accL[(size_t)j * L + y]) : XF::zero();
    vS[q] = (cell && cols) ? XFv(accS[(size_t)j * L + y]) : XF::zero();
  }
  // the product of A (I-side) and B (J-side) tiles, folded into acc (XF::add per cell)
  auto product = [&](const float *Ab, const int *AE, const float *Bb, const int *BE, unsigned ta, unsigned tb,
                     bool row_product, int K, XFv acc[4]) {
    const float4 A = *reinterpret_cast<const MESSI_GLOBAL float4 *>(gp(Ab) + ta * (T * T) + lane * 4);
    const float4 B = *reinterpret_cast<const MESSI_GLOBAL float4 *>(gp(Bb) + tb * (T * T) + lane * 4);
    const int4 ra = *reinterpret_cast<const MESSI_GLOBAL int4 *>(gp(AE) + ta * T + 4 * g);
    const int cb = gp(BE)[tb * T + col];
    v4f mm = {0.f, 0.f, 0.f, 0.f};
    mm = __builtin_amdgcn_mfma_f32_16x16x4f32(A.x, B.x, mm, 0, 0, 0);
    mm = __builtin_amdgcn_mfma_f32_16x16x4f32(A.y, B.y, mm, 0, 0, 0);
    mm = __builtin_amdgcn_mfma_f32_16x16x4f32(A.z, B.z, mm, 0, 0, 0);
    mm = __builtin_amdgcn_mfma_f32_16x16x4f32(A.w, B.w, mm, 0, 0, 0);
    const int ras[4] = {ra.x, ra.y, ra.z, ra.w};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      XFv contrib = XF::zero();
      if (ras[q] != XF::EZ && cb != XF::EZ) {
        const float cv = mm[q];
        contrib = cv >= 0x1p-100f ? XF::norm(cv, ras[q] + cb)
                                  : exact_out_cell<T>(od, row_product, I, K, J, 4 * g + q, col);
      }
      acc[q] = XF::add(acc[q], contrib);
    }
  };
#pragma unroll
  for (int i = 0; i < W; ++i) {
    const int mi = m + W - 1 - i;
    if ((rows >> i) & 1u) {
      const int K = J + mi;
      product(od.WpA, od.WpAE, od.iSpB, od.iSpBE, (unsigned)(I * NT + K), (unsigned)(J * NT + K), true, K, vL);
    }
    if ((cols >> i) & 1u) {
      const int K = I - mi;
      product(od.iLpA, od.iLpAE, od.WpB, od.WpBE, (unsigned)(K * NT + I), (unsigned)(K * NT + J), false, K, vS);
    }
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int j = I * T + 4 * g + q;
    if (j < L && y < L) {
      if (rows) accL[(size_t)j * L + y] = vL[q];
      if (cols) accS[(size_t)j * L + y] = vS[q];
    }
  }
}

template <class SR, int T, int W>
__global__ __launch_bounds__(T * T) void outside_step_kernel(OutDesc od, const OutRange *__restrict__ ranges,
                                                             int u, Rules<SR> r) {
  typedef typename SR::V V;
  __shared__ OutLDS<SR, T> sm;
  const int tid = threadIdx.x, L = od.L, NT = od.NT;
  const int du = NT - 1 - u;
  const MESSI_GLOBAL V *__restrict__ iSr = gp(static_cast<const V *>(od.iSr));
  const MESSI_GLOBAL V *__restrict__ iLc = gp(static_cast<const V *>(od.iLc));
  const MESSI_GLOBAL V *__restrict__ ptv = gp(static_cast<const V *>(od.ptv));
  MESSI_GLOBAL V *__restrict__ oSr = gp(static_cast<V *>(od.oSr));
  MESSI_GLOBAL V *__restrict__ oSc = gp(static_cast<V *>(od.oSc));
  MESSI_GLOBAL V *__restrict__ oFr = gp(static_cast<V *>(od.oFr));
  MESSI_GLOBAL V *__restrict__ oL = gp(static_cast<V *>(od.oL));
  MESSI_GLOBAL V *__restrict__ Wr = gp(static_cast<V *>(od.Wr));
  MESSI_GLOBAL V *__restrict__ Wc = gp(static_cast<V *>(od.Wc));
  MESSI_GLOBAL V *__restrict__ accL = gp(static_cast<V *>(od.accL));
  MESSI_GLOBAL V *__restrict__ accS = gp(static_cast<V *>(od.accS));
  // ---- decode: u+1 finalising tiles first, then the bulk tiles by distance,
  // in plain dispatch order (round-robin over the XCDs: the bulk tiles of one
  // distance read disjoint operands, and grouping them per XCD measured slower)
  int I, d, m;
  bool fin;
  const int item = (int)blockIdx.x;
  if constexpr (SR::scaled && T == 16) {
    // bulk items one per wavefront (WAVE_ITEMS per workgroup), in registers
    if (item > u) {
      out_bulk_wave_item<W>(od, ranges + (size_t)u * NT, u,
                            WAVE_ITEMS * (item - (u + 1)) + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)));
      return;
    }
  }
  if (item <= u) {
    fin = true;
    I = item;
    d = du;
    m = 1;
  } else {
    fin = false;
    const int idx = item - (u + 1);
    const OutRange *rg = ranges + (size_t)u * NT;
    int lo = 0, hi = du - 1;  // largest distance with cum <= idx
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (rg[mid].cum <= idx) lo = mid;
      else hi = mid - 1;
    }
    d = lo;
    const int i = idx - rg[d].cum;
    I = i < rg[d].rend ? i : rg[d].cbeg + (i - rg[d].rend);
    m = du + 1 - d;
  }
  const int J = I + d, IT = I * T, JT = J * T;
  const int a0 = tid / T, b0 = tid % T, j0 = IT + a0, y0 = JT + b0;
  const bool cell0 = j0 < L && y0 < L;
  V vL = SR::zero(), vS = SR::zero();
  if (!fin) {
    // A bulk tile is visited once per W launches (the host table lists it only
    // in the launches with (m - 2) % W == 0) and adds the products of the
    // offsets m + W - 1 .. m, in that order: the order of the launches that made
    // them ready.
    unsigned rows = 0, cols = 0;  // bit i: offset m + W - 1 - i has a row / column product
#pragma unroll
    for (int i = 0; i < W; ++i) {
      const int mi = m + W - 1 - i;
      if (out_row_ok(d, T, od.lband) && J + mi <= NT - 1) rows |= 1u << i;
      if (out_col_ok(mi, T, od.band) && I - mi >= 0) cols |= 1u << i;
    }
    if (cell0 && rows) vL = accL[(size_t)j0 * L + y0];
    if (cell0 && cols) vS = accS[(size_t)j0 * L + y0];
#pragma unroll
    for (int i = 0; i < W; ++i) {
      const bool dr = (rows >> i) & 1u, dc = (cols >> i) & 1u;
      if (!(dr || dc)) continue;
      const int mi = m + W - 1 - i;
      const OpSrc<SR> srow{Wr, IT, (J + mi) * T, true, iSr, JT + 1, (J + mi) * T, true};
      const OpSrc<SR> scol{iLc, IT - 1, (I - mi) * T, false, Wc, JT, (I - mi) * T, false};
      Operands<SR> orow, ocol;
      if (dr) orow = fetch_operands<SR, T>(srow, L);
      if (dc) ocol = fetch_operands<SR, T>(scol, L);
      tile_products<SR, T>(orow, dr, ocol, dc, sm.u.p, vL, vS);
    }
    if (cell0) {
      if (rows) accL[(size_t)j0 * L + y0] = vL;
      if (cols) accS[(size_t)j0 * L + y0] = vS;
    }
    return;
  }
  // ---- the finalising workgroup's own products (offset 1), loads issued first
  {
    const bool dorow = out_row_ok(d, T, od.lband) && J + 1 <= NT - 1;
    const bool docol = out_col_ok(1, T, od.band) && I - 1 >= 0;
    if (cell0) {
      vL = accL[(size_t)j0 * L + y0];
      vS = accS[(size_t)j0 * L + y0];
    }
    const OpSrc<SR> srow{Wr, IT, (J + 1) * T, true, iSr, JT + 1, (J + 1) * T, true};
    const OpSrc<SR> scol{iLc, IT - 1, (I - 1) * T, false, Wc, JT, (I - 1) * T, false};
    Operands<SR> orow, ocol;
    if (dorow) orow = fetch_operands<SR, T>(srow, L);
    if (docol) ocol = fetch_operands<SR, T>(scol, L);
    if (dorow || docol) tile_products<SR, T>(orow, dorow, ocol, docol, sm.u.p, vL, vS);
  }
  // ---- finalise tile (I, J)
  // The cells are given to the threads in order of their finalisation step (as
  // in the inside step); the products' sums reach the cell's new thread through
  // its W and Q slots, which only the cell's own finalisation overwrites (W and
  // Q of a cell are read only once it is final).  The scaled path runs the
  // wavefront in double precision at one power-of-two scale per tile (the
  // largest exponent of the tile's border Q values), as the inside step does;
  // a tile whose values would leave [2^-1000, 2^1000], or one of whose nonzero
  // product terms would fall below 2^-1000, is recomputed by the
  // per-cell-exponent wavefront.
  constexpr bool by_step = true;
  adbl *const DW = reinterpret_cast<adbl *>(&sm.u.w.W[0][0]);
  adbl *const DQ = reinterpret_cast<adbl *>(&sm.u.w.Q[0][0]);
  adbl *const DiS = reinterpret_cast<adbl *>(&sm.u.w.iSd[0][0]);
  adbl *const DiL = reinterpret_cast<adbl *>(&sm.u.w.iLd[0][0]);
  adbl *const DcP = reinterpret_cast<adbl *>(&sm.u.w.cP[0][0]);
  constexpr int RSO = T + 1;  // row stride of the wavefront arrays
  const bool dbl = SR::scaled;
  bool dok = true;
  // iS / iL of the tile's own columns and rows and the pair term, loaded with
  // clamped indices (element 0 of the table for a cell outside it) and zero
  // selected afterwards, so that these loads and the border loads below issue
  // together, before the first LDS write; staged with the products' sums handed
  // to the cells' new threads, as unscaled doubles for the double wavefront
  // (as_d), else in the representation itself
  struct StagedCell {
    V iS, iL, cP;
  };
  auto load_tables = [&]() {
    const int j = j0, y = y0;
    const int rr = JT + a0 + 1, cc = JT + b0;  // iSd[a][b] = iS[JT+a+1, JT+b]
    const int lr = IT + a0, lc = IT + b0 - 1;  // iLd[a][b] = iL[IT+a, IT+b-1]
    const bool k1 = rr < L && cc < L && rr <= cc, k2 = lc >= 0 && lr < L && lc < L && lr <= lc,
               k3 = j >= 1 && j <= y && y + 1 < L;
    const V x1 = typename SR::V(iSr[k1 ? (size_t)rr * L + cc : 0]);
    const V x2 = typename SR::V(iLc[k2 ? (size_t)lc * L + lr : 0]);
    const V x3 = typename SR::V(ptv[k3 ? (size_t)(j - 1) * L + (y + 1) : 0]);
    return StagedCell{k1 ? x1 : SR::zero(), k2 ? x2 : SR::zero(), k3 ? x3 : SR::zero()};
  };
  auto put_tables = [&](const StagedCell &g, bool as_d) {
    const int a = a0, b = b0;
    auto put = [&](V &slot, adbl *darr, V v) {
      if constexpr (SR::scaled) {
        if (as_d) {
          darr[a * RSO + b] = xf_to_d(v, 0, dok);
          return;
        }
      }
      slot = v;
    };
    put(sm.u.w.iSd[a][b], DiS, g.iS);
    put(sm.u.w.iLd[a][b], DiL, g.iL);
    put(sm.u.w.cP[a][b], DcP, g.cP);
    // the hand-over stays in the representation (converted by the new thread)
    sm.u.w.W[a][b] = by_step ? vL : SR::zero();
    sm.u.w.Q[a][b] = by_step ? vS : SR::zero();
  };
  const StagedCell g = load_tables();
  // Q on the row above the tile (IT-1, JT+b+1) and the column right of it
  // (IT+b-1, JT+T), lanes b < T
  V qt = SR::zero(), qr = SR::zero();
  {
    const int bq = tid, tr = IT - 1, tc = JT + bq + 1, rr = IT + bq - 1, rc = JT + T;
    const bool kt = tid < T && tr >= 0 && tc < L, kr = tid < T && bq >= 1 && rr < L && rc < L;
    const size_t ct = kt ? (size_t)tr * L + tc : 0, cr = kr ? (size_t)rr * L + rc : 0;
    const V s1 = typename SR::V(oSr[ct]), f1 = typename SR::V(oFr[ct]), l1 = typename SR::V(oL[ct]);
    const V s2 = typename SR::V(oSr[cr]), f2 = typename SR::V(oFr[cr]), l2 = typename SR::V(oL[cr]);
    if (kt) qt = SR::add(SR::add(SR::mulc(s1, r.p[2]), SR::mulc(f1, r.p[3])), SR::mulc(l1, r.p[6]));
    if (kr) qr = SR::add(SR::add(SR::mulc(s2, r.p[2]), SR::mulc(f2, r.p[3])), SR::mulc(l2, r.p[6]));
  }
  put_tables(g, dbl);
  if (tid == 0) sm.bad = 0;
  if (tid < T) {
    sm.Qtop[tid] = qt;
    sm.Qright[tid] = qr;
  }
  if constexpr (SR::scaled) {
    if (tid < 64) {
      // the tile's scale from its nonzero border values (0 when all are zero)
      int e = INT_MIN / 2;
      if (tid < T) {
        if (qt.m != 0.0f) e = max(e, qt.e);
        if (qr.m != 0.0f) e = max(e, qr.e);
      }
      int E = border_emax<T>(e);
      if (E < -(1 << 27)) E = 0;
      if (tid < T) {
        sm.Qtd[tid] = xf_to_d(qt, E, dok);
        sm.Qrd[tid] = xf_to_d(qr, E, dok);
      }
      if (tid == 0) sm.escale = E;
    }
  }
  __syncthreads();
  int a = a0, b = b0;
  if constexpr (by_step) out_cell_by_step<T>(tid, a, b);
  const int j = IT + a, y = JT + b;
  const bool mine = j < L && y < L && j <= y;
  const bool lkeep = out_row_ok(d, T, od.lband);
  const int own = (T - 1 - b) + a;
  // Q[j-1, y+1] for the pair term: inside this tile, or on the row above / the
  // column right of it.  The inner cell (a-1, b+1) is finalised two steps before
  // this cell (own - 2; own >= 2 for it): every step reads the lane's Q cell
  // (inner or border, a fixed LDS position) for use in the next step, off the
  // finalisation's path.
  const bool q_in = a >= 1 && b + 1 < T;
  // the last step with work for this wavefront (scalar: a uniform skip, one barrier a step)
  int wlast;
  if constexpr (by_step) {
    int wa, wb;
    out_cell_by_step<T>(tid & ~63, wa, wb);
    wlast = (T - 1 - wb) + wa;
  } else {
    wlast = 2 * T - 2;
  }
  const int wave_last = __builtin_amdgcn_readfirstlane(wlast);
  V fS = SR::zero(), fF = SR::zero(), fL = SR::zero(), fW = SR::zero();
  bool xf_wave = !dbl;
  if constexpr (SR::scaled) {
    // ---- double wavefront
    const int E = sm.escale;
    double acL = xf_to_d(sm.u.w.W[a][b], E, dok), acS = xf_to_d(sm.u.w.Q[a][b], E, dok);
    if (!mine) dok = true;  // only cells of the table are checked (the others read zeros)
    bool ok = true;
    const adbl *qsrc = q_in ? &DQ[(a - 1) * RSO + b + 1]
                            : (a == 0 ? reinterpret_cast<const adbl *>(&sm.Qtd[b])
                                      : reinterpret_cast<const adbl *>(&sm.Qrd[a]));
    const double pterm = DcP[a * RSO + b];
    const double p0 = r.p[0], p2 = r.p[2], p3 = r.p[3], p4 = r.p[4], p6 = r.p[6];
    double q_next = *qsrc;
    double dS = 0.0, dF = 0.0, dL = 0.0, dW = 0.0, dQ = 0.0;
    // a nonzero term of nonzero factors below 2^-1000 (or flushed to zero)
    auto tiny = [](double x, double y, double t) { return x != 0.0 && y != 0.0 && !(t >= D_LO); };
    for (int st = 0; st <= 2 * T - 2; ++st) {
      if (st <= wave_last) {
        const double q_f = q_next;
        q_next = *qsrc;
        const bool act = mine && st >= 1 && st <= own, fn = mine && st == own;
        if (act) {
          const int bs = T - st + a;  // row producer (a, bs), finalised at step st-1
          if (bs > b && bs < T) {
            const double w1 = DW[a * RSO + bs], s1 = DiS[b * RSO + bs], t1 = w1 * s1;
            ok = ok && !tiny(w1, s1, t1);
            acL += t1;
          }
          const int as = st - T + b;  // column producer (as, b), finalised at step st-1
          if (as >= 0 && as < a) {
            const double l2 = DiL[as * RSO + a], w2 = DW[as * RSO + b], t2 = l2 * w2;
            ok = ok && !tiny(l2, w2, t2);
            acS += t2;
          }
        }
        if (fn) {
          dL = lkeep ? acL : 0.0;
          dS = acS;
          dF = pterm * q_f;
          ok = ok && !tiny(pterm, q_f, dF);
          if (j == 0 && y == L - 1) dS = ldexp(1.0, -E);  // the start symbol
          dW = p0 * dS + p4 * dF;
          dQ = p2 * dS + p3 * dF + p6 * dL;
          DW[a * RSO + b] = dW;
          DQ[a * RSO + b] = dQ;
        }
      }
      __syncthreads();
    }
    if (mine)
      dok = dok && ok && d_in_range(dS) && d_in_range(dF) && d_in_range(dL) && d_in_range(dW) && d_in_range(dQ);
    if (!dok) atomicOr(&sm.bad, 1);
    fS = d_to_xf(dS, E);
    fF = d_to_xf(dF, E);
    fL = d_to_xf(dL, E);
    fW = d_to_xf(dW, E);
    __syncthreads();
    xf_wave = sm.bad != 0;
    if (xf_wave) {
      // restage the tables and the hand-over in the representation itself
      put_tables(load_tables(), false);
      __syncthreads();
    }
  }
  if (xf_wave) {
    // ---- per-cell-exponent wavefront
    typename SR::Acc acL = SR::acc(by_step ? sm.u.w.W[a][b] : vL);
    typename SR::Acc acS = SR::acc(by_step ? sm.u.w.Q[a][b] : vS);
    fS = fF = fL = SR::zero();
    const V *qsrc = q_in ? &sm.u.w.Q[a - 1][b + 1] : (a == 0 ? &sm.Qtop[b] : &sm.Qright[a]);
    const V pterm = sm.u.w.cP[a][b];
    V q_next = *qsrc;  // primed: a border cell may finalise at step 0 or 1 (see the inside step)
    auto work = [&](int st) {
      const V q_f = q_next;
      q_next = *qsrc;
      const bool act = mine && st >= 1 && st <= own, fn = mine && st == own;
      V w1 = SR::zero(), s1 = SR::zero(), l2 = SR::zero(), w2 = SR::zero();
      if (act) {
        const int bs = T - st + a;  // row producer (a, bs), finalised at step st-1
        if (bs > b && bs < T) {
          w1 = sm.u.w.W[a][bs];
          s1 = sm.u.w.iSd[b][bs];
        }
        const int as = st - T + b;  // column producer (as, b), finalised at step st-1
        if (as >= 0 && as < a) {
          l2 = sm.u.w.iLd[as][a];
          w2 = sm.u.w.W[as][b];
        }
        SR::acc_add2(acL, SR::mul_raw(w1, s1), SR::zero());
        SR::acc_add2(acS, SR::mul_raw(l2, w2), SR::zero());
      }
      if (fn) {
        fL = lkeep ? SR::acc_value(acL) : SR::zero();
        fS = SR::acc_value(acS);
        fF = SR::mul(pterm, q_f);
        if (j == 0 && y == L - 1) fS = SR::from_log(0.0);  // the start symbol
        fW = SR::add(SR::mulc(fS, r.p[0]), SR::mulc(fF, r.p[4]));
        sm.u.w.W[a][b] = fW;
        sm.u.w.Q[a][b] = SR::add(SR::add(SR::mulc(fS, r.p[2]), SR::mulc(fF, r.p[3])), SR::mulc(fL, r.p[6]));
      }
    };
    auto step = [&](int st) {
      if (st <= wave_last) work(st);
      __syncthreads();
    };
    // as in the inside step: the FP64 step stays a rolled loop
    if constexpr (SR::scaled) {
      for (int st = 0; st <= 2 * T - 2; ++st) step(st);
    } else {
#pragma unroll 1
      for (int st = 0; st <= 2 * T - 2; ++st) step(st);
    }
  }
  // ---- write back: row-major from registers, column-major through LDS
  if (mine) {
    const size_t rc = (size_t)j * L + y;
    oSr[rc] = fS;
    oFr[rc] = fF;
    oL[rc] = fL;
    Wr[rc] = fW;
  }
  sm.u.w.Q[a][b] = fS;
  sm.u.w.W[a][b] = fW;
  __syncthreads();
  {
    const int bb = tid / T, aa = tid % T, jj = IT + aa, yy = JT + bb;
    if (jj < L && yy < L && jj <= yy) {
      const size_t cc = (size_t)yy * L + jj;
      oSc[cc] = sm.u.w.Q[aa][bb];
      Wc[cc] = sm.u.w.W[aa][bb];
    }
  }
  if constexpr (SR::scaled && T == 16) {
    // product-ready copies of the finished W tile (see OutDesc), aligned as
    // tile_products aligns the W operands it stages; whole tiles, zero beyond
    // the table and below the diagonal
    const unsigned tile = (unsigned)(I * NT + J);
    {
      const int ra = tid / T, rc = tid % T;  // A: element (ra, rc), rows aligned
      const XFv v = (IT + ra < L && JT + rc < L && IT + ra <= JT + rc) ? sm.u.w.W[ra][rc] : XF::zero();
      const int e = group_max<T>(v.m != 0.0f ? v.e : XF::EZ);
      gp(od.WpA)[tile * (T * T) + op_a_at<T>(ra, rc)] = v.m == 0.0f ? 0.0f : ldexpf(v.m, max(v.e - e, -200));
      if (rc == 0) gp(od.WpAE)[tile * T + ra] = e;
    }
    {
      const int c = tid % T, bb = tid / T;  // B: element (c, bb), columns aligned
      const XFv v = (IT + c < L && JT + bb < L && IT + c <= JT + bb) ? sm.u.w.W[c][bb] : XF::zero();
      const int e = group_max<T>(v.m != 0.0f ? v.e : XF::EZ);
      gp(od.WpB)[tile * (T * T) + op_b_at<T>(c, bb)] = v.m == 0.0f ? 0.0f : ldexpf(v.m, max(v.e - e, -200));
      if (c == 0) gp(od.WpBE)[tile * T + bb] = e;
    }
  }
}

// Zero the outside accumulators (upper triangle).
template <class SR>
__global__ void outside_tiled_init_kernel(int L, typename SR::V *accL, typename SR::V *accS) {
  int y = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y;
  if (y >= L || j > y) return;
  const size_t c = (size_t)j * L + y;
  accL[c] = SR::zero();
  accS[c] = SR::zero();
}

// ---------------------------------------------------------------------------
// Export / import of the Julia (3, L, L) layout (natural-log absolute values).
// ---------------------------------------------------------------------------
template <class SR>
__global__ void export3_kernel(int L, const typename SR::V *Scol, const typename SR::V *Frow,
                               const typename SR::V *Lrow, const typename SR::V *Lcol, const double *C,
                               int outside, double *o3) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y;
  if (i >= L) return;
  size_t base = 3 * ((size_t)i + (size_t)L * j);
  if (i > j) {
    o3[base] = o3[base + 1] = o3[base + 2] = -HUGE_VAL;
    return;
  }
  double bl = outside ? (C[L] - C[j + 1] + C[i]) : (C[j + 1] - C[i]);
  double s = SR::to_log(Scol[(size_t)j * L + i]);
  double f = SR::to_log(Frow[(size_t)i * L + j]);
  double l = SR::to_log(Lcol ? Lcol[(size_t)j * L + i] : Lrow[(size_t)i * L + j]);
  o3[base] = s == -HUGE_VAL ? s : s + bl;
  o3[base + 1] = f == -HUGE_VAL ? f : f + bl;
  o3[base + 2] = l == -HUGE_VAL ? l : l + bl;
}

template <class SR>
__global__ void import3_kernel(int L, const double *i3, const double *C, int outside, typename SR::V *Sr,
                               typename SR::V *Sc, typename SR::V *Fr, typename SR::V *Fc, typename SR::V *Lr,
                               typename SR::V *Lc) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y;
  if (i >= L || i > j) return;
  size_t base = 3 * ((size_t)i + (size_t)L * j);
  double bl = outside ? (C[L] - C[j + 1] + C[i]) : (C[j + 1] - C[i]);
  double s = i3[base], f = i3[base + 1], l = i3[base + 2];
  typename SR::V ts = SR::from_log(s == -HUGE_VAL ? s : s - bl);
  typename SR::V tf = SR::from_log(f == -HUGE_VAL ? f : f - bl);
  typename SR::V tl = SR::from_log(l == -HUGE_VAL ? l : l - bl);
  size_t rc = (size_t)i * L + j, cc = (size_t)j * L + i;
  Sr[rc] = ts;
  Sc[cc] = ts;
  Fr[rc] = tf;
  if (Fc) Fc[cc] = tf;
  if (Lr) Lr[rc] = tl;
  if (Lc) Lc[cc] = tl;
}


// ---------------------------------------------------------------------------
// Base-pair posteriors (natural logs in double, then per-row normalisation).
// ---------------------------------------------------------------------------
template <class SR>
__global__ void bpp_logs_kernel(int L, const typename SR::V *oSr, const typename SR::V *oFr,
                                const typename SR::V *oL, const typename SR::V *iF, const typename SR::V *ptv,
                                const double *uhat, Rules<SR> r, const double *zz, double *lp, double *res) {
  typedef typename SR::V V;
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y;
  if (k >= L) return;
  const double zhat = zz[1];
  if (k == j) {
    lp[(size_t)j * L + j] = -HUGE_VAL;
    V t = SR::add(SR::mulc(oSr[(size_t)j * L + j], r.p[1]), SR::mulc(oL[(size_t)j * L + j], r.p[5]));
    double tl = SR::to_log(t);
    res[j] = tl == -HUGE_VAL ? tl : tl + uhat[j] - zhat;
    return;
  }
  if (k < j) return;
  size_t c = (size_t)j * L + k;
  V inres = (j + 1 <= k - 1) ? iF[(size_t)(j + 1) * L + (k - 1)] : SR::zero();
  V t = SR::add(SR::add(SR::mulc(oSr[c], r.p[2]), SR::mulc(oFr[c], r.p[3])), SR::mulc(oL[c], r.p[6]));
  double v = SR::to_log(SR::mul(SR::mul(t, inres), ptv[c]));
  v = v == -HUGE_VAL ? v : v - zhat;
  lp[c] = v;
  lp[(size_t)k * L + j] = v;
}

__global__ __launch_bounds__(256) void bpp_rownorm_kernel(int L, const double *lp, const double *res,
                                                          double *rmax, double *rnorm) {
  __shared__ double sh[256];
  int x = blockIdx.x, t = threadIdx.x;
  const double *row = lp + (size_t)x * L;
  double m = res[x];
  for (int y = t; y < L; y += 256) m = fmax(m, row[y]);
  sh[t] = m;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o) sh[t] = fmax(sh[t], sh[t + o]);
    __syncthreads();
  }
  m = sh[0];
  __syncthreads();
  double s = 0;
  for (int y = t; y < L; y += 256) s += exp(row[y] - m);
  sh[t] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o) sh[t] += sh[t + o];
    __syncthreads();
  }
  if (t == 0) {
    rmax[x] = m;
    rnorm[x] = sh[0] + exp(res[x] - m);
  }
}

__global__ void bpp_write_kernel(int L, const double *lp, const double *res, const double *rmax,
                                 const double *rnorm, double *unpaired, double *pair) {
  int x = blockIdx.x * blockDim.x + threadIdx.x;
  int y = blockIdx.y;
  if (x >= L) return;
  double m = rmax[x], n = rnorm[x];
  pair[(size_t)x + (size_t)L * y] = exp(lp[(size_t)y * L + x] - m) / n;
  if (y == 0 && unpaired) unpaired[x] = exp(res[x] - m) / n;
}

// ---------------------------------------------------------------------------
// Host side (T = double -> LogSR, T = float -> XF)
// ---------------------------------------------------------------------------

template <typename T>
void inside_alloc(InsideWork &w, int L, bool outside) {
  typedef typename SRof<T>::type SR;
  size_t n = (size_t)L * L * sizeof(typename SR::V);
  w.L = L;
  w.Lr.ensure(n);
  w.Lc.ensure(n);
  w.Sr.ensure(n);
  w.Sc.ensure(n);
  w.F.ensure(n);
  // tile-major buffers: room for the padded square at either tile edge
  const size_t n16 = (size_t)(L + 15) / 16, n32 = (size_t)(L + 31) / 32;
  const size_t cells = std::max(n16 * n16 * 256, n32 * n32 * 1024), rows = std::max(n16 * n16 * 16, n32 * n32 * 32);
  w.acc.ensure(cells * sizeof(typename SR::V));
  w.ptv.ensure(n);
  if (SR::scaled) {
    w.Lp.ensure(sizeof(float) * cells);
    w.Sp.ensure(sizeof(float) * cells);
    w.LpE.ensure(sizeof(int) * rows);
    w.SpE.ensure(sizeof(int) * rows);
  }
  w.lam.ensure(sizeof(double) * L);
  w.C.ensure(sizeof(double) * (L + 1));
  w.uhat.ensure(sizeof(double) * L);
  w.u.ensure(sizeof(double) * L);
  w.paired.ensure(sizeof(double) * L * L);
  w.red.ensure(sizeof(double) * (3 * L + 16));
  if (outside) {
    w.oSr.ensure(n);
    w.oSc.ensure(n);
    w.oFr.ensure(n);
    w.oL.ensure(n);
    w.oWr.ensure(n);
    w.oWc.ensure(n);
    w.oaL.ensure(n);
    w.oaS.ensure(n);
    w.lp.ensure(sizeof(double) * L * L);
    if (SR::scaled) {
      const size_t c16 = n16 * n16 * 256, r16 = n16 * n16 * 16;
      w.oWpA.ensure(sizeof(float) * c16);
      w.oWpB.ensure(sizeof(float) * c16);
      w.iSpB.ensure(sizeof(float) * c16);
      w.iLpA.ensure(sizeof(float) * c16);
      w.oWpAE.ensure(sizeof(int) * r16);
      w.oWpBE.ensure(sizeof(int) * r16);
      w.iSpBE.ensure(sizeof(int) * r16);
      w.iLpAE.ensure(sizeof(int) * r16);
    }
  }
}

template <class SR>
static InstDesc<typename SR::V> inst_of(InsideWork &w) {
  typedef typename SR::V V;
  InstDesc<V> d;
  d.L = w.L;
  d.NT = 0;
  d.T = 16;
  d.KB = 0;
  d.Lr = w.Lr.as<V>();
  d.Lc = w.Lc.as<V>();
  d.Sr = w.Sr.as<V>();
  d.Sc = w.Sc.as<V>();
  d.F = w.F.as<V>();
  d.acc = w.acc.as<V>();
  d.ptv = w.ptv.as<V>();
  d.Lp = SR::scaled ? w.Lp.as<float>() : nullptr;
  d.Sp = SR::scaled ? w.Sp.as<float>() : nullptr;
  d.LpE = SR::scaled ? w.LpE.as<int>() : nullptr;
  d.SpE = SR::scaled ? w.SpE.as<int>() : nullptr;
  d.u = w.u.as<double>();
  d.paired = w.paired.as<double>();
  d.lam = w.lam.as<double>();
  d.C = w.C.as<double>();
  d.uhat = w.uhat.as<double>();
  d.z = w.red.as<double>();
  d.exp = static_cast<const ExpandDesc *>(w.exp);
  d.zout = w.zout;
  d.pflag = nullptr;
  return d;
}

// Baseline and pair term of every instance (u and paired already on the device;
// with root entries, groups holds the ngroups runs of members on one dataset).
template <class SR>
static void prepare_batch(hipStream_t s, const InstDesc<typename SR::V> *d_inst, int n, int Lmax, double B,
                          const double *rootacc = nullptr, const int2 *groups = nullptr, int ngroups = 0) {
  double lp[7];
  kh99_logprobs(lp);
  hipLaunchKernelGGL(baseline_kernel<typename SR::V>, dim3(n), dim3(1024), 0, s, d_inst, B, lp[0] + lp[5],
                     (int)SR::scaled);
  if (rootacc) {
    if (Lmax > 1)
      hipLaunchKernelGGL(ptv_group_kernel<SR>, dim3((Lmax - 1 + 255) / 256, Lmax, ngroups), dim3(256), 0, s, d_inst,
                         groups, rootacc, B);
  } else {
    hipLaunchKernelGGL(ptv_kernel<SR>, dim3((Lmax + 255) / 256, Lmax, n), dim3(256), 0, s, d_inst, B);
  }
}

// Clears n ints (the single-launch form's counter and flags).
__global__ void zero_ints_kernel(int *p, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0;
}

template <typename T>
void inside_prepare(hipStream_t s, InsideWork &w, int L, double B) {
  typedef typename SRof<T>::type SR;
  std::vector<InstDesc<typename SR::V>> d(1, inst_of<SR>(w));
  upload(w.inst1, d, s);
  prepare_batch<SR>(s, w.inst1.as<InstDesc<typename SR::V>>(), 1, L, B);
}

// Inside pass (preparation included) for a batch of instances; each work's red[0]
// receives its log Z.
template <typename T>
int inside_batch_run(hipStream_t s, InsideWork *const *ws, int n, InsideBatch &ib, double B,
                     const double *rootacc, hipEvent_t steps_begin, hipEvent_t steps_end) {
  typedef typename SRof<T>::type SR;
  typedef typename SR::V V;
  if (n <= 0) return 0;
  const int TE = ws[0]->tile == 32 ? 32 : 16;
  const int W = ws[0]->window;
  MESSI_REQUIRE(W == 1 || W == 2 || W == 4, "bulk window must be 1, 2 or 4");
  std::vector<InstDesc<V>> d(n);
  int Lmax = 0, NTmax = 0;
  const bool persist = ws[0]->persist;
  size_t pf_total = 2;  // ticket counter and give-up flag, then the instances' flags
  for (int m = 0; m < n; ++m) {
    d[m] = inst_of<SR>(*ws[m]);
    d[m].NT = (d[m].L + TE - 1) / TE;
    d[m].T = TE;
    // tile (I, K) holds an L cell with h - j <= band iff (K - I - 1) T + 1 <= band
    const int band = ws[m]->band;
    d[m].KB = (band < 0 || band >= d[m].L) ? d[m].NT : (band == 0 ? 0 : (band - 1) / TE + 1);
    MESSI_REQUIRE((size_t)d[m].NT * d[m].NT * TE * TE < ((size_t)1 << 31),
                  "inside: length too large for the 32-bit tile offsets");
    Lmax = std::max(Lmax, d[m].L);
    NTmax = std::max(NTmax, d[m].NT);
    d[m].pflag = nullptr;
    if (persist) {
      d[m].pflag = (int *)(uintptr_t)pf_total;  // offset for now
      pf_total += pflag_ints(d[m].NT);
    }
  }
  if (persist) {
    ib.pflags.ensure(sizeof(int) * pf_total);
    for (int m = 0; m < n; ++m) d[m].pflag = ib.pflags.as<int>() + (size_t)(uintptr_t)d[m].pflag;
  }
  // Work items of launch t: the tiles (I, I+t) and the ready split-point products
  // (I, K) x (K, J) with K - I = t - 1 or J - K = t - 1; a bulk tile (I, I+dist)
  // is scheduled only when one of its two products lies within the band.
  // per launch: finalising workgroups per instance (prefix, n+1 entries), then
  // bulk workgroups per instance (prefix, n+1 entries)
  const size_t stride = 2 * (size_t)(n + 1);
  std::vector<int> prefix((size_t)NTmax * stride);
  std::vector<int> total(NTmax);
  for (int t = 0; t < NTmax; ++t) {
    int fa = 0, ba = 0;
    int *fp = prefix.data() + (size_t)t * stride, *bp = fp + (n + 1);
    for (int m = 0; m < n; ++m) {
      fp[m] = fa;
      bp[m] = ba;
      const int NT = d[m].NT, KB = d[m].KB;
      if (t < NT) {
        int nbulk = 0;
        int dmax = std::min(2 * t - 2, NT - 1);
        if (t - 1 > KB) dmax = std::min(dmax, t - 1 + KB);
        for (int dist = t + 1; dist <= dmax; dist += W) nbulk += NT - dist;
        fa += NT - t;
        // bulk workgroups: WAVE_ITEMS items each in the scaled T = 16 launch
        // sequence, bulk_per(W) in the other forms
        const int per = (SR::scaled && TE == 16 && !persist) ? WAVE_ITEMS : bulk_per(W);
        ba += (nbulk + per - 1) / per;
      }
    }
    fp[n] = fa;
    bp[n] = ba;
    total[t] = ((fa + 7) & ~7) + ba;
  }
  // the tables change only when a member's buffers move or its shape changes
  auto refresh = [&](DevBuf &dev, std::vector<unsigned char> &host, const void *p, size_t bytes) {
    if (host.size() == bytes && (bytes == 0 || std::memcmp(host.data(), p, bytes) == 0)) return;
    host.assign((const unsigned char *)p, (const unsigned char *)p + bytes);
    dev.ensure(bytes);
    if (bytes) MESSI_HIP(hipMemcpyAsync(dev.p, host.data(), bytes, hipMemcpyHostToDevice, s));
  };
  refresh(ib.inst, ib.h_inst, d.data(), d.size() * sizeof(InstDesc<V>));
  refresh(ib.prefix, ib.h_prefix, prefix.data(), prefix.size() * sizeof(int));
  // runs of consecutive members on one dataset share their pair-term lookups
  std::vector<int2> groups;
  for (int m = 0; m < n; ++m)
    if (m > 0 && ws[m]->dset && ws[m]->dset == ws[m - 1]->dset && d[m].L == d[m - 1].L)
      groups.back().y++;
    else
      groups.push_back(make_int2(m, 1));
  refresh(ib.groups, ib.h_groups, groups.data(), groups.size() * sizeof(int2));
  const InstDesc<V> *di = ib.inst.as<InstDesc<V>>();
  double lp[7];
  kh99_logprobs(lp);
  Rules<SR> r = make_rules<SR>(B);
  for (int m = 0; m < n; ++m)
    MESSI_REQUIRE((rootacc != nullptr) == (ws[m]->exp != nullptr), "pair term source differs between members");
  prepare_batch<SR>(s, di, n, Lmax, B, rootacc, ib.groups.as<int2>(), (int)groups.size());
  hipLaunchKernelGGL(inside_init_kernel<SR>, dim3((Lmax + 255) / 256, 1, n), dim3(256), 0, s, di, B, lp[1],
                     lp[5]);
  if (persist) {
    // single-launch form: the ticket table, the flags cleared, one launch
    const int NP = NTmax + 1;
    std::vector<int> NTs(n), KBs(n);
    for (int m = 0; m < n; ++m) {
      NTs[m] = d[m].NT;
      KBs[m] = d[m].KB;
    }
    std::vector<int> tab(sched::table_size(n, NP));
    sched::build_table(n, NTs.data(), KBs.data(), W, NP, tab.data());
    refresh(ib.ptab, ib.h_ptab, tab.data(), tab.size() * sizeof(int));
    int *ctl = ib.pflags.as<int>();
    hipLaunchKernelGGL(zero_ints_kernel, dim3((unsigned)std::min<size_t>((pf_total + 255) / 256, 4096)), dim3(256), 0,
                       s, ctl, pf_total);
    if (steps_begin) MESSI_HIP(hipEventRecord(steps_begin, s));
    const int *dt = ib.ptab.as<int>();
    const int tickets = tab[NP];
    auto go = [&](auto kern, int threads) {
      hipLaunchKernelGGL(kern, dim3(std::max(1, tickets)), dim3(threads), 0, s, di, dt, n, NP, r, ctl);
    };
    if (TE == 32) {
      if (W == 4) go(inside_persist_kernel<SR, 32, 4>, 32 * 32);
      else if (W == 2) go(inside_persist_kernel<SR, 32, 2>, 32 * 32);
      else go(inside_persist_kernel<SR, 32, 1>, 32 * 32);
    } else {
      if (W == 4) go(inside_persist_kernel<SR, 16, 4>, 16 * 16);
      else if (W == 2) go(inside_persist_kernel<SR, 16, 2>, 16 * 16);
      else go(inside_persist_kernel<SR, 16, 1>, 16 * 16);
    }
    if (steps_end) MESSI_HIP(hipEventRecord(steps_end, s));
    hipLaunchKernelGGL(zlog_kernel<SR>, dim3((n + 63) / 64), dim3(64), 0, s, di, n, (const int *)(ctl + 1));
    MESSI_HIP(hipGetLastError());
    return 1;
  }
  if (steps_begin) MESSI_HIP(hipEventRecord(steps_begin, s));
  for (int t = 0; t < NTmax; ++t) {
    const int *pf = ib.prefix.as<int>() + (size_t)t * stride;
    auto go = [&](auto kern, int threads) {
      hipLaunchKernelGGL(kern, dim3(total[t]), dim3(threads), 0, s, di, pf, n, t, r);
    };
    if (TE == 32) {
      if (W == 4) go(inside_step_kernel<SR, 32, 4>, 32 * 32);
      else if (W == 2) go(inside_step_kernel<SR, 32, 2>, 32 * 32);
      else go(inside_step_kernel<SR, 32, 1>, 32 * 32);
    } else {
      if (W == 4) go(inside_step_kernel<SR, 16, 4>, 16 * 16);
      else if (W == 2) go(inside_step_kernel<SR, 16, 2>, 16 * 16);
      else go(inside_step_kernel<SR, 16, 1>, 16 * 16);
    }
  }
  if (steps_end) MESSI_HIP(hipEventRecord(steps_end, s));
  hipLaunchKernelGGL(zlog_kernel<SR>, dim3((n + 63) / 64), dim3(64), 0, s, di, n, (const int *)nullptr);
  MESSI_HIP(hipGetLastError());
  return NTmax;
}

template <typename T>
void inside_run(hipStream_t s, InsideWork &w, int L, double B) {
  InsideWork *ws[1] = {&w};
  (void)L;
  inside_batch_run<T>(s, ws, 1, w.ib1, B);
}

template <typename T>
double inside_logZ(hipStream_t s, InsideWork &w, int L) {
  double z[2];
  MESSI_HIP(hipMemcpyAsync(z, w.red.p, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  MESSI_HIP(hipStreamSynchronize(s));
  return z[0];
}

template <typename T>
void inside_export(hipStream_t s, InsideWork &w, int L, double *d_in3) {
  typedef typename SRof<T>::type SR;
  typedef typename SR::V V;
  hipLaunchKernelGGL(export3_kernel<SR>, dim3((L + 255) / 256, L), dim3(256), 0, s, L, (const V *)w.Sc.as<V>(),
                     (const V *)w.F.as<V>(), (const V *)nullptr, (const V *)w.Lc.as<V>(),
                     (const double *)w.C.as<double>(), 0, d_in3);
}

template <typename T>
void inside_import(hipStream_t s, InsideWork &w, int L, const double *d_in3) {
  typedef typename SRof<T>::type SR;
  typedef typename SR::V V;
  hipLaunchKernelGGL(import3_kernel<SR>, dim3((L + 255) / 256, L), dim3(256), 0, s, L, d_in3,
                     (const double *)w.C.as<double>(), 0, w.Sr.as<V>(), w.Sc.as<V>(), w.F.as<V>(), (V *)nullptr,
                     w.Lr.as<V>(), w.Lc.as<V>());
  std::vector<InstDesc<V>> d(1, inst_of<SR>(w));
  upload(w.inst1, d, s);
  hipLaunchKernelGGL(zlog_kernel<SR>, dim3(1), dim3(64), 0, s, (const InstDesc<V> *)w.inst1.as<InstDesc<V>>(), 1,
                     (const int *)nullptr);
}

template <typename T>
void outside_run(hipStream_t s, InsideWork &w, int L, bool full_L) {
  typedef typename SRof<T>::type SR;
  typedef typename SR::V V;
  Rules<SR> r = make_rules<SR>(1.0);
  const int TE = w.tile == 32 ? 32 : 16, NT = (L + TE - 1) / TE;
  const int band = (w.band >= 0 && w.band < L) ? w.band : -1;
  const int lband = full_L ? -1 : band;
  // launch tables: the bulk tile ranges of every launch (host, cached per shape)
  std::vector<int> total(NT);
  const int W = w.window;
  MESSI_REQUIRE(W == 1 || W == 2 || W == 4, "bulk window must be 1, 2 or 4");
  const int key[5] = {L, TE, band, lband, W};
  const bool fresh = std::equal(key, key + 5, w.orng_key);
  std::vector<OutRange> rg(fresh ? 0 : (size_t)NT * NT);
  for (int u = 0; u < NT; ++u) {
    const int du = NT - 1 - u;
    int cum = 0;
    for (int d = 0; d < du; ++d) {
      // offset m = du + 1 - d >= 2; a tile is visited when (m - 2) % W == 0, and
      // its tiles with a ready product in the window are those of offset m
      const int m = du + 1 - d;
      const bool visit = (m - 2) % W == 0;
      int rend = visit && out_row_ok(d, TE, lband) ? u : 0;
      int cbeg = rend, cend = rend;
      if (visit && out_col_ok(m, TE, band) && m <= NT - 1 - d) {
        cbeg = m;
        cend = NT - d;
        if (cbeg <= rend) {  // the two ranges meet: one range
          rend = cend;
          cbeg = cend;
        }
      }
      if (!fresh) rg[(size_t)u * NT + d] = OutRange{cum, rend, cbeg, cend};
      cum += rend + (cend - cbeg);
    }
    // bulk workgroups: WAVE_ITEMS items each in the scaled T = 16 form, one otherwise
    total[u] = (u + 1) + ((SR::scaled && TE == 16) ? (cum + WAVE_ITEMS - 1) / WAVE_ITEMS : cum);
  }
  if (!fresh) {
    upload(w.orng, rg, s);
    std::copy(key, key + 5, w.orng_key);
  }
  OutDesc od;
  od.L = L;
  od.NT = NT;
  od.band = band;
  od.lband = lband;
  od.iSr = w.Sr.p;
  od.iLc = w.Lc.p;
  od.ptv = w.ptv.p;
  od.oSr = w.oSr.p;
  od.oSc = w.oSc.p;
  od.oFr = w.oFr.p;
  od.oL = w.oL.p;
  od.Wr = w.oWr.p;
  od.Wc = w.oWc.p;
  od.accL = w.oaL.p;
  od.accS = w.oaS.p;
  od.WpA = w.oWpA.as<float>();
  od.WpB = w.oWpB.as<float>();
  od.iSpB = w.iSpB.as<float>();
  od.iLpA = w.iLpA.as<float>();
  od.WpAE = w.oWpAE.as<int>();
  od.WpBE = w.oWpBE.as<int>();
  od.iSpBE = w.iSpBE.as<int>();
  od.iLpAE = w.iLpAE.as<int>();
  hipLaunchKernelGGL(outside_tiled_init_kernel<SR>, dim3((L + 255) / 256, L), dim3(256), 0, s, L, w.oaL.as<V>(),
                     w.oaS.as<V>());
  if (SR::scaled && TE == 16) {
    MESSI_REQUIRE(w.oWpA.p && w.iSpB.p, "outside: product-ready buffers not allocated");
    hipLaunchKernelGGL(out_operands_kernel, dim3(NT, NT), dim3(256), 0, s, od);
  }
  const OutRange *dr = w.orng.as<OutRange>();
  for (int u = 0; u < NT; ++u) {
    auto go = [&](auto kern, int threads) { hipLaunchKernelGGL(kern, dim3(total[u]), dim3(threads), 0, s, od, dr, u, r); };
    if (TE == 32) {
      if (W == 4) go(outside_step_kernel<SR, 32, 4>, 32 * 32);
      else if (W == 2) go(outside_step_kernel<SR, 32, 2>, 32 * 32);
      else go(outside_step_kernel<SR, 32, 1>, 32 * 32);
    } else {
      if (W == 4) go(outside_step_kernel<SR, 16, 4>, 16 * 16);
      else if (W == 2) go(outside_step_kernel<SR, 16, 2>, 16 * 16);
      else go(outside_step_kernel<SR, 16, 1>, 16 * 16);
    }
  }
  MESSI_HIP(hipGetLastError());
}

template <typename T>
void outside_export(hipStream_t s, InsideWork &w, int L, double *d_out3) {
  typedef typename SRof<T>::type SR;
  typedef typename SR::V V;
  hipLaunchKernelGGL(export3_kernel<SR>, dim3((L + 255) / 256, L), dim3(256), 0, s, L,
                     (const V *)w.oSc.as<V>(), (const V *)w.oFr.as<V>(), (const V *)w.oL.as<V>(), (const V *)nullptr,
                     (const double *)w.C.as<double>(), 1, d_out3);
}

template <typename T>
void outside_import(hipStream_t s, InsideWork &w, int L, const double *d_out3) {
  typedef typename SRof<T>::type SR;
  typedef typename SR::V V;
  hipLaunchKernelGGL(import3_kernel<SR>, dim3((L + 255) / 256, L), dim3(256), 0, s, L, d_out3,
                     (const double *)w.C.as<double>(), 1, w.oSr.as<V>(), w.oSc.as<V>(), w.oFr.as<V>(),
                     nullptr, w.oL.as<V>(), (V *)nullptr);
}

template <typename T>
void bpp_run(hipStream_t s, InsideWork &w, int L, double *d_unpaired, double *d_pair) {
  typedef typename SRof<T>::type SR;
  typedef typename SR::V V;
  Rules<SR> r = make_rules<SR>(1.0);
  double *zz = w.red.as<double>();  // [0] log Z, [1] baseline-relative log Z (from zlog_kernel)
  double *res = zz + 16, *rmax = res + L, *rnorm = rmax + L;
  hipLaunchKernelGGL(bpp_logs_kernel<SR>, dim3((L + 255) / 256, L), dim3(256), 0, s, L, (const V *)w.oSr.as<V>(),
                     (const V *)w.oFr.as<V>(), (const V *)w.oL.as<V>(), (const V *)w.F.as<V>(),
                     (const V *)w.ptv.as<V>(), (const double *)w.uhat.as<double>(), r, (const double *)zz,
                     w.lp.as<double>(), res);
  hipLaunchKernelGGL(bpp_rownorm_kernel, dim3(L), dim3(256), 0, s, L, (const double *)w.lp.as<double>(),
                     (const double *)res, rmax, rnorm);
  hipLaunchKernelGGL(bpp_write_kernel, dim3((L + 255) / 256, L), dim3(256), 0, s, L,
                     (const double *)w.lp.as<double>(), (const double *)res, (const double *)rmax,
                     (const double *)rnorm, d_unpaired, d_pair);
  MESSI_HIP(hipGetLastError());
}

#define MESSI_INST(T)                                                                   \
  template void inside_alloc<T>(InsideWork &, int, bool);                               \
  template void inside_prepare<T>(hipStream_t, InsideWork &, int, double);              \
  template void inside_run<T>(hipStream_t, InsideWork &, int, double);                  \
  template int inside_batch_run<T>(hipStream_t, InsideWork *const *, int, InsideBatch &, double,     \
                                   const double *, hipEvent_t, hipEvent_t);                   \
  template double inside_logZ<T>(hipStream_t, InsideWork &, int);                       \
  template void inside_export<T>(hipStream_t, InsideWork &, int, double *);             \
  template void inside_import<T>(hipStream_t, InsideWork &, int, const double *);       \
  template void outside_run<T>(hipStream_t, InsideWork &, int, bool);                   \
  template void outside_export<T>(hipStream_t, InsideWork &, int, double *);            \
  template void outside_import<T>(hipStream_t, InsideWork &, int, const double *);      \
  template void bpp_run<T>(hipStream_t, InsideWork &, int, double *, double *);
MESSI_INST(double)
MESSI_INST(float)

}  // namespace messi
