// Maximum-expected-accuracy (posterior) decoding of a consensus structure from
// base-pair posteriors: the fill of computeposteriordecodingcuda
// (src/InsideCUDA.jl:183-215, kernel posteriordecoding src/cuda/inside.cu:286-321;
// the same recursion as iterativeposteriordecoding, src/RNATools.jl:6-42) and
// the traceback posteriordecodingtraceback (src/RNATools.jl:44-74).
//
// For i < j (E is zero on and below the diagonal):
//   U[i,k]  = alpha p(i,k) + E[i+1,k-1]                 (pair (i,k) closed over its interior)
//   e1      = single[i] + E[i+1,j]                       (i unpaired)
//   e2      = U[i,j]                                     (i pairs with j)
//   e3      = max_{i<k<j} U[i,k] + E[k+1,j], first k     (i pairs with k, then [k+1,j])
//   E[i,j]  = e1, replaced by e2, then e3, only when strictly greater;
//   S[i,j]  = -1, -2 or the 1-based k.
// FP64 rounds alpha p and each sum separately (the Julia loop); FP32 converts the
// inputs to float, reads the column-major posterior matrix at i*L + k (p(k,i)),
// and fuses alpha p + E into one fma, as nvcc's default contraction compiles
// the CUDA kernel.  e3 is a max-plus product of U (row i) and E (column j), not
// a sum of products, so it runs on the vector ALUs, not on MFMA.
//
// Schedule (tile edge T, NT tiles per side): launch t finalises the tiles
// (I, I+t).  The k in a tile K strictly between I and J form the split-point
// product U(I,K) x E(rows KT+1..KT+T, J); the product for K becomes ready when
// both operand tiles are final, after launch max(K-I, J-K), and runs in launch
// max(K-I, J-K)+1: in the finalising workgroup for K = I+1 and K = J-1; the rest
// in the tile's next bulk visit (a bulk workgroup per tile every MEA_W launches,
// folding the products that became ready since its previous visit), accumulated
// as (value, first k) in HBM.  The remaining k (inside tile I
// and tile J) are pulled by an in-LDS wavefront over the tile's cells, one
// column and one row term per cell per step.
#include <algorithm>

#include "prune.h"

// The FP64 path rounds alpha * p and each sum separately.
#pragma clang fp contract(off)

namespace messi {

namespace {

template <typename V> struct MeaDesc {
  int L, NT;
  const double *single;  // L
  const double *pair;    // L x L, column-major; p(i,k) = pair[i * psi + k * psk]
  int psi, psk;
  V alpha;
  V *E, *U;              // column-major L x L (cells i < j only)
  int *S;                // row-major L x L (cells i < j only)
  V *accV;               // bulk split-point maxima, row-major
  int *accK;             // their (0-based) first maximising k
};

__device__ __forceinline__ double pair_term(double alpha, double p, double e) { return alpha * p + e; }
__device__ __forceinline__ float pair_term(float alpha, float p, float e) { return fmaf(alpha, p, e); }

template <typename V> __device__ __forceinline__ V neg_inf();
template <> __device__ __forceinline__ double neg_inf<double>() { return -INFINITY; }
template <> __device__ __forceinline__ float neg_inf<float>() { return -INFINITY; }

// Candidate (v, k) replaces (bv, bk) when greater, or equal with a smaller k: the
// result of any fold order is the sequential loop's first maximum.
template <typename V> __device__ __forceinline__ void fold(V v, int k, V &bv, int &bk) {
  if (v > bv || (v == bv && k < bk)) {
    bv = v;
    bk = k;
  }
}

template <typename V, int T> struct MeaLDS {
  V Ud[T][T + 1];  // U of tile (I, I)
  V Ed[T][T + 1];  // E of tile (J, J)
  V cE[T][T + 1];  // E of the tile being finalised
  V cU[T][T + 1];  // U of the tile being finalised
  V cP[T][T + 1];  // p(i, j) of the tile being finalised
  V sg[T];         // single[i] of its rows
  V Ebot[T];       // E[(I+1)T, JT+b]
  V Ebot2[T];      // E[(I+1)T, JT+b-1]
  V Eleft[T];      // E[IT+a+1, JT-1]
  struct {
    V A[T][T + 1];  // U[IT+a, KT+c]
    V B[T][T + 1];  // E[KT+c+1, JT+b]
  } p[2];
};

// Bulk window: a bulk tile at distance d is visited in the launches t = d-1,
// d-1-W, ... (those with 2t - 2 >= d), each visit folding the products of the
// launches since the previous one, so its accumulator is read and written once
// per W launches.  The fold keeps the first maximum whatever the order, so the
// window does not change any result.  Bulk items of launch t: the tiles at
// distances t+1, t+1+W, ..., NT - d of them per distance d; mea_bulk_items(k)
// counts the items of the first k distances.
constexpr int MEA_W = 4;
__host__ __device__ __forceinline__ int mea_bulk_items(int NT, int t, int k) {
  return k * (NT - t - 1) - MEA_W * (k * (k - 1) / 2);
}

template <typename V>
__device__ __forceinline__ V load_E(const MeaDesc<V> &d, int r, int c) {
  return (r >= 0 && r < c && c < d.L) ? d.E[(size_t)c * d.L + r] : V(0);
}
template <typename V>
__device__ __forceinline__ V load_p(const MeaDesc<V> &d, int i, int k) {
  return (V)d.pair[(size_t)i * d.psi + (size_t)k * d.psk];
}

template <typename V, int T>
__global__ __launch_bounds__(T * T) void mea_step_kernel(MeaDesc<V> d, int t) {
  __shared__ MeaLDS<V, T> sm;
  const int tid = threadIdx.x;
  const int L = d.L, NT = d.NT;
  // ---- decode the work item: the NT - t finalising tiles are the first
  // workgroups in dispatch order (so the launch's critical path starts first);
  // the bulk items after them are renumbered so that each XCD takes a contiguous run
  const int nfin = NT - t;
  const int item = (int)blockIdx.x < nfin ? (int)blockIdx.x
                                           : nfin + xcd_chunk_id((int)blockIdx.x - nfin, (int)gridDim.x - nfin);
  int I, J;
  bool fin;
  if (item < NT - t) {
    fin = true;
    I = item;
    J = item + t;
  } else {
    fin = false;
    const int idx = item - (NT - t);
    const int nd = (min(2 * t - 2, NT - 1) - t - 1) / MEA_W + 1;  // distances with bulk items (>= 1 here)
    int lo = 0, hi = nd - 1;  // the distance: mea_bulk_items(lo) <= idx < mea_bulk_items(lo + 1)
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (mea_bulk_items(NT, t, mid) <= idx) lo = mid;
      else hi = mid - 1;
    }
    I = idx - mea_bulk_items(NT, t, lo);
    J = I + t + 1 + lo * MEA_W;
  }
  const int IT = I * T, JT = J * T;
  const int a = tid / T, b = tid % T;  // this thread's cell
  const int i = IT + a, j = JT + b;
  const bool inside_L = i < L && j < L;
  // ---- split-point products that became ready with launches t-W+1 .. t (bulk)
  // or t (finalising).  The products of a tile at distance d run in launches
  // smin = ceil(d/2)+1 .. d: the accumulator holds earlier ones once a bulk visit
  // came before (d >= 4 when finalising).
  const int dist = J - I, smin = (dist + 3) / 2;
  const bool prior = fin ? t >= 4 : t - MEA_W >= smin;
  V bv = neg_inf<V>();
  int bk = 0x7fffffff;
  if (prior && inside_L) {
    bv = d.accV[(size_t)i * L + j];
    bk = d.accK[(size_t)i * L + j];
  }
#pragma unroll 1
  for (int st = fin ? t : max(t - MEA_W + 1, smin); st <= t; ++st) {
    const int K1 = I + st - 1, K2 = J - st + 1;
    const bool do1 = st >= 2, do2 = st >= 2 && K2 != K1;
    if (!do1) continue;
    // coalesced loads: the fast index runs down a column
    const int r = tid % T, c = tid / T;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      if (q == 1 && !do2) continue;
      const int KT = (q ? K2 : K1) * T;
      const int ar = IT + r, ac = KT + c;
      sm.p[q].A[r][c] = (ar < ac && ac < L) ? d.U[(size_t)ac * L + ar] : neg_inf<V>();
      // E[JT, JT] (k = JT-1, b = 0) is on the diagonal: zero, never stored
      sm.p[q].B[r][c] = load_E(d, KT + r + 1, JT + c);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      if (q == 1 && !do2) continue;
      const int KT = (q ? K2 : K1) * T;
      V pv = neg_inf<V>();
      int pk = 0x7fffffff;
#pragma unroll
      for (int c = 0; c < T; ++c) {
        const V v = sm.p[q].A[a][c] + sm.p[q].B[c][b];
        if (v > pv) {
          pv = v;
          pk = KT + c;
        }
      }
      fold(pv, pk, bv, bk);
    }
    __syncthreads();  // the next launch's operands reuse the staging
  }
  if (!fin) {
    if (inside_L) {
      d.accV[(size_t)i * L + j] = bv;
      d.accK[(size_t)i * L + j] = bk;
    }
    return;
  }
  // ---- finalise tile (I, J): stage its inputs and borders
  {
    const int r = tid % T, c = tid / T;  // coalesced: r runs down a column
    const int ri = IT + r, cj = JT + c;
    sm.cP[r][c] = (ri < L && cj < L) ? load_p(d, ri, cj) : V(0);
    sm.cE[r][c] = V(0);
    sm.cU[r][c] = neg_inf<V>();
    if (t >= 1) {
      const int ci = IT + c;  // column of the diagonal tile (I, I)
      sm.Ud[r][c] = (ri < ci && ci < L) ? d.U[(size_t)ci * L + ri] : neg_inf<V>();
      sm.Ed[r][c] = load_E(d, JT + r, cj);
    }
  }
  if (tid < T) {
    const int q = tid;
    sm.sg[q] = (IT + q < L) ? (V)d.single[IT + q] : V(0);
    const int rb = (I + 1) * T;
    sm.Ebot[q] = t >= 1 ? load_E(d, rb, JT + q) : V(0);
    sm.Ebot2[q] = t >= 1 ? load_E(d, rb, JT + q - 1) : V(0);
    sm.Eleft[q] = t >= 1 ? load_E(d, IT + q + 1, JT - 1) : V(0);
  }
  __syncthreads();
  const bool mine = inside_L && (t > 0 || a < b);
  int sres = 0;
  if (t == 0) {
    const int own = b - a;
    for (int st = 1; st < T; ++st) {
      if (mine && st <= own) {
        if (own - st <= st - 1) {
          const int k1 = a + st - 1, k2 = b - st;
          if (k1 >= a + 1 && k1 <= b - 1) fold(sm.cU[a][k1] + sm.cE[k1 + 1][b], IT + k1, bv, bk);
          if (k2 >= a + 1 && k2 != k1) fold(sm.cU[a][k2] + sm.cE[k2 + 1][b], IT + k2, bv, bk);
        }
        if (st == own) {
          const V e1 = sm.sg[a] + sm.cE[a + 1][b];
          const V u = pair_term(d.alpha, sm.cP[a][b], sm.cE[a + 1][b - 1]);
          V val = e1;
          sres = -1;
          if (u > val) {
            val = u;
            sres = -2;
          }
          if (bv > val) {
            val = bv;
            sres = bk + 1;
          }
          sm.cE[a][b] = val;
          sm.cU[a][b] = u;
        }
      }
      __syncthreads();
    }
  } else {
    const int own = b + T - 1 - a;
    // split point k = IT+T-1: E[(I+1)T, j] lies in the tile below
    if (mine && a <= T - 2) fold(sm.Ud[a][T - 1] + sm.Ebot[b], IT + T - 1, bv, bk);
    // Per-cell constants in registers.  The cell below (a+1, b) is the column
    // producer of this cell's own step, so its E comes with that step's reads;
    // the diagonal neighbour (a+1, b-1) is final two steps earlier (own >= 2 for
    // it) and is read one step ahead from a fixed LDS position (its border value
    // for the last row / first column; primed before step 0).
    const V sgv = sm.sg[a], pv = sm.cP[a][b], ebot = sm.Ebot[b];
    const V *einsrc = (a + 1 < T && b >= 1) ? &sm.cE[a + 1][b - 1] : (a + 1 == T ? &sm.Ebot2[b] : &sm.Eleft[a]);
    const bool ezero = i + 1 > j - 1;
    V ein_next = *einsrc;
    auto step = [&](int st) {
      const V ein_f = ein_next;
      ein_next = *einsrc;
      const int ap = b + T - st;  // column producer (ap, b), final since step st-1
      const int bp = st - T + a;  // row producer (a, bp), final since step st-1
      const int apc = min(max(ap, 1), T - 1), bpc = min(max(bp, 0), T - 2);
      const V ecol = sm.cE[apc][b];
      const V tc = sm.Ud[a][apc - 1] + ecol, tr = sm.cU[a][bpc] + sm.Ed[bpc + 1][b];
      if (mine && st >= 1 && st <= own) {
        if (ap >= a + 2 && ap <= T - 1) fold(tc, IT + ap - 1, bv, bk);
        if (bp >= 0 && bp < b) fold(tr, JT + bp, bv, bk);
      }
      if (mine && st == own) {
        const V eb = a + 1 < T ? ecol : ebot;  // at this step apc = a + 1
        const V e1 = sgv + eb;
        const V u = pair_term(d.alpha, pv, ezero ? V(0) : ein_f);
        V val = e1;
        sres = -1;
        if (u > val) {
          val = u;
          sres = -2;
        }
        if (bv > val) {
          val = bv;
          sres = bk + 1;
        }
        sm.cE[a][b] = val;
        sm.cU[a][b] = u;
      }
      __syncthreads();
    };
#pragma unroll
    for (int st = 0; st <= 2 * T - 2; ++st) step(st);
  }
  // ---- write the tile back
  if (mine) d.S[(size_t)i * L + j] = sres;
  {
    const int r = tid % T, c = tid / T;
    const int ri = IT + r, cj = JT + c;
    if (ri < L && cj < L && ri < cj) {
      d.E[(size_t)cj * L + ri] = sm.cE[r][c];
      d.U[(size_t)cj * L + ri] = sm.cU[r][c];
    }
  }
}

// Traceback by one wavefront.  The stack of pending intervals (packed i | j << 16;
// they are disjoint and at least two long, so at most L/2 are pending) and the
// pairing (16-bit partners) are in LDS; the pairing goes to global memory at the
// end, so no global store is outstanding while the walk waits on its loads.  S
// is read through a 16 x 16 window (rows i..i+15, columns j-15..j) held four
// entries per lane, so runs of unpaired positions and stacked pairs cost one
// round of global loads per sixteen steps; a step reads its entry with
// v_readlane into a scalar register.  Every lane runs the same control flow.
__global__ __launch_bounds__(64) void mea_traceback_kernel(int L, const int *__restrict__ S, int *__restrict__ paired) {
  extern __shared__ unsigned smem[];
  constexpr int W = 16;
  const int lane = threadIdx.x;
  unsigned *stk = smem;                                         // L/2 + 1 entries
  unsigned short *pr = (unsigned short *)(smem + (L / 2 + 1));  // L entries
  for (int x = lane; x < L; x += 64) pr[x] = 0;
  __syncthreads();
  int top = 0;
  if (L >= 2) {
    stk[0] = 0u | ((unsigned)(L - 1) << 16);
    top = 1;
  }
  int wi = -1000, wj = -1000;
  int w0 = 0, w1 = 0, w2 = 0, w3 = 0;  // window entry lane + 64 q: row (lane + 64 q) / W
  while (top > 0) {
    --top;
    // every lane holds the same interval: keep the walk's state in scalar registers
    const unsigned e = (unsigned)__builtin_amdgcn_readfirstlane((int)stk[top]);
    int i = (int)(e & 0xffffu), j = (int)(e >> 16);
    while (i < j) {
      if (!(i >= wi && i <= wi + W - 1 && j <= wj && j >= wj - (W - 1))) {
        wi = i;
        wj = j;
        const int c = wj - (W - 1) + (lane % W);
        const int r0 = wi + lane / W;
        auto at = [&](int r) { return (r < c && c < L) ? S[(size_t)r * L + c] : 0; };
        w0 = at(r0);
        w1 = at(r0 + 4);
        w2 = at(r0 + 8);
        w3 = at(r0 + 12);
      }
      const int idx = (i - wi) * W + (j - (wj - (W - 1)));
      const int q = idx >> 6;
      const int s = __builtin_amdgcn_readlane(q == 0 ? w0 : q == 1 ? w1 : q == 2 ? w2 : w3, idx & 63);
      if (s == -1) {
        i += 1;
      } else if (s == -2) {
        if (lane == 0) {
          pr[i] = (unsigned short)(j + 1);
          pr[j] = (unsigned short)(i + 1);
        }
        i += 1;
        j -= 1;
      } else if (s > 0) {
        const int k = s - 1;
        if (lane == 0) {
          pr[i] = (unsigned short)(k + 1);
          pr[k] = (unsigned short)(i + 1);
        }
        if (k + 1 < j) stk[top++] = (unsigned)(k + 1) | ((unsigned)j << 16);
        i += 1;
        j = k - 1;
      } else {
        break;
      }
    }
  }
  __syncthreads();
  for (int x = lane; x < L; x += 64) paired[x] = pr[x];
}

template <typename V>
__global__ void mea_export_kernel(int L, const V *E, const int *S, double *Eout, int *Sout) {
  const size_t n = (size_t)L * L;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
    const int r = (int)(q % L), c = (int)(q / L);
    const bool up = r < c;
    if (Eout) Eout[q] = up ? (double)E[q] : 0.0;
    if (Sout) Sout[q] = up ? S[(size_t)r * L + c] : 0;
  }
}

}  // namespace

template <typename V>
void mea_run(hipStream_t s, MeaWork &w, int L, const double *d_single, const double *d_pair, double alpha,
             int *d_paired) {
  MESSI_REQUIRE(L >= 1 && L <= MEA_MAX_L, "MEA decoding supports 1 <= L <= 16383");
  constexpr int T = 16;
  const size_t LL = (size_t)L * L;
  w.E.ensure(sizeof(V) * LL);
  w.U.ensure(sizeof(V) * LL);
  w.S.ensure(sizeof(int) * LL);
  w.accV.ensure(sizeof(V) * LL);
  w.accK.ensure(sizeof(int) * LL);
  w.L = L;
  w.fp32 = sizeof(V) == 4;
  MeaDesc<V> d;
  d.L = L;
  d.NT = (L + T - 1) / T;
  d.single = d_single;
  d.pair = d_pair;
  if (sizeof(V) == 4) {  // the CUDA kernel's reading of the column-major matrix: p(k, i)
    d.psi = L;
    d.psk = 1;
  } else {
    d.psi = 1;
    d.psk = L;
  }
  d.alpha = (V)alpha;
  d.E = w.E.as<V>();
  d.U = w.U.as<V>();
  d.S = w.S.as<int>();
  d.accV = w.accV.as<V>();
  d.accK = w.accK.as<int>();
  const int NT = d.NT;
  for (int t = 0; t < NT; ++t) {
    const int dmax = std::min(2 * t - 2, NT - 1);
    const int nb = dmax >= t + 1 ? mea_bulk_items(NT, t, (dmax - t - 1) / MEA_W + 1) : 0;
    hipLaunchKernelGGL((mea_step_kernel<V, T>), dim3(NT - t + nb), dim3(T * T), 0, s, d, t);
  }
  MESSI_HIP(hipGetLastError());
  if (d_paired) mea_run_traceback(s, w, L, d_paired);
}

void mea_run_traceback(hipStream_t s, MeaWork &w, int L, int *d_paired) {
  MESSI_REQUIRE(w.L == L && L >= 1 && L <= MEA_MAX_L, "no MEA table of this length");
  const size_t lds = sizeof(unsigned) * (L / 2 + 1) + sizeof(unsigned short) * L;  // <= 64 KiB for L <= 16383
  hipLaunchKernelGGL(mea_traceback_kernel, dim3(1), dim3(64), lds, s, L, (const int *)w.S.as<int>(), d_paired);
  MESSI_HIP(hipGetLastError());
}

void mea_export(hipStream_t s, MeaWork &w, int L, double *d_E, int *d_S) {
  MESSI_REQUIRE(w.L == L, "no MEA table of this length");
  const size_t LL = (size_t)L * L;
  const unsigned blocks = (unsigned)std::min<size_t>((LL + 255) / 256, 8192);
  if (w.fp32)
    hipLaunchKernelGGL(mea_export_kernel<float>, dim3(blocks), dim3(256), 0, s, L, (const float *)w.E.as<float>(),
                       (const int *)w.S.as<int>(), d_E, d_S);
  else
    hipLaunchKernelGGL(mea_export_kernel<double>, dim3(blocks), dim3(256), 0, s, L,
                       (const double *)w.E.as<double>(), (const int *)w.S.as<int>(), d_E, d_S);
  MESSI_HIP(hipGetLastError());
}

template void mea_run<double>(hipStream_t, MeaWork &, int, const double *, const double *, double, int *);
template void mea_run<float>(hipStream_t, MeaWork &, int, const double *, const double *, double, int *);

}  // namespace messi
