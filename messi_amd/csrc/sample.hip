// Stochastic traceback of KH99 structures from an inside table:
// samplestructure / samplestructurehelper (src/InsideOutsideAlgorithm.jl:570-629),
// one wavefront per sample, many samples per launch.
//
// From (S, 0, L-1), an interval (symbol, x, y) with x < y chooses among the
// bifurcations h = x..y-1 of its type-3 rule (S->LS or F->LS: weight
// p L[x,h] S[h+1,y]) and, last, the pairing of its type-2 rule (p paired[x,y]
// F[x+1,y-1]), with probability proportional to weight^B; x == y is a terminal.
// The 64 lanes score 64 consecutive split points at a time from one row of L
// (row-major) and one column of S (column-major copy): a pass for the maximum,
// a pass for the total, and a chunked prefix-sum search for the first
// cumulative weight >= u * total.  The uniform u of an interval is keyed by
// (seed, sample, symbol, x, y) with the SplitMix64 finaliser (Steele, Lea and
// Flood 2014), so the draw does not depend on the visiting order.  In the
// scaled FP32 tables every candidate of an interval carries the same baseline
// factor exp(-sum of lambda over [x,y]), which cancels.
//
// The KH99 log prior of the sampled structure (calculateKH99prior,
// src/InsideOutsideAlgorithm.jl:86-144) is accumulated along the sampled
// derivation: the log probability of each rule used.
#include "prune.h"
#include "sr.h"

namespace messi {

namespace {

__host__ __device__ inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

__device__ inline double keyed_uniform(uint64_t seed, int sample, int sym, int x, int y) {
  const uint64_t key = mix64(seed + 0x9E3779B97F4A7C15ULL * (uint64_t)(sample + 1));
  const uint64_t z = mix64(key ^ (((uint64_t)(uint32_t)x << 32) | (uint64_t)(uint32_t)y) ^ ((uint64_t)sym << 62));
  return (double)((z >> 11) + 1) * 0x1.0p-53;  // (0, 1]
}

template <class SR> struct SampleArgs {
  typedef typename SR::V V;
  int L, n, first;
  uint64_t seed;
  double B;
  const V *Lr, *Sc, *F, *ptv;
  double lp[7];   // KH99 natural-log rule probabilities
  int *out;       // n x L, 1-based partners, 0 unpaired
  double *prior;  // n
};

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ __forceinline__ double wave_incl_scan(double v, int lane) {
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const double t = __shfl_up(v, o);
    if (lane >= o) v += t;
  }
  return v;
}

constexpr unsigned SYM_S = 0, SYM_F = 1, SYM_L = 2;

template <class SR>
__global__ __launch_bounds__(64) void sample_kernel(SampleArgs<SR> a) {
  extern __shared__ unsigned smem[];
  const int L = a.L, lane = threadIdx.x;
  const int sample = a.first + blockIdx.x;
  unsigned *stk = smem;                                         // L/2 + 1 pending L intervals
  unsigned short *pr = (unsigned short *)(smem + (L / 2 + 1));  // the pairing
  for (int q = lane; q < L; q += 64) pr[q] = 0;
  __syncthreads();
  const double B = a.B;
  double prior = 0.0;
  int top = 0;
  int sym = SYM_S, x = 0, y = L - 1;
  bool have = L >= 1;
  while (true) {
    if (!have) {
      if (top == 0) break;
      --top;
      const unsigned e = (unsigned)__builtin_amdgcn_readfirstlane((int)stk[top]);
      sym = (int)(e >> 28);
      x = (int)((e >> 14) & 0x3fffu);
      y = (int)(e & 0x3fffu);
    }
    have = false;
    if (x > y) continue;
    if (x == y) {  // terminal: S -> s or L -> s (F has none)
      prior += sym == SYM_S ? a.lp[1] : sym == SYM_L ? a.lp[5] : 0.0;
      continue;
    }
    const double t3 = sym == SYM_S ? a.lp[0] : a.lp[4];
    const double t2 = sym == SYM_S ? a.lp[2] : sym == SYM_F ? a.lp[3] : a.lp[6];
    const int nb = sym == SYM_L ? 0 : y - x;  // bifurcation candidates h = x .. y-1
    const typename SR::V *Lrow = a.Lr + (size_t)x * L + x;
    const typename SR::V *Scol = a.Sc + (size_t)y * L + x + 1;
    const double fin = x + 1 <= y - 1 ? SR::to_log(a.F[(size_t)(x + 1) * L + (y - 1)]) : -HUGE_VAL;
    const double lpair = ((t2 + SR::to_log(a.ptv[(size_t)x * L + y])) + fin) * B;
    // pass 1: the maximum
    double m = lpair;
    for (int base = 0; base < nb; base += 64) {
      const int k = base + lane;
      if (k < nb) m = fmax(m, ((t3 + SR::to_log(Lrow[k])) + SR::to_log(Scol[k])) * B);
    }
    m = wave_max(m);
    int choice = nb;  // index of the chosen candidate (nb = the pairing)
    if (m != -HUGE_VAL) {
      // pass 2: the total
      double part = 0.0;
      for (int base = 0; base < nb; base += 64) {
        const int k = base + lane;
        if (k < nb) part += exp(((t3 + SR::to_log(Lrow[k])) + SR::to_log(Scol[k])) * B - m);
      }
      const double wpair = exp(lpair - m);
      const double tot = wave_sum(part) + wpair;
      const double target = keyed_uniform(a.seed, sample, sym, x, y) * tot;
      // pass 3: the first cumulative weight >= target
      double run = 0.0;
      int lastpos = wpair > 0.0 ? nb : -1;
      bool found = false;
      for (int base = 0; base < nb; base += 64) {
        const int k = base + lane;
        const double w = k < nb ? exp(((t3 + SR::to_log(Lrow[k])) + SR::to_log(Scol[k])) * B - m) : 0.0;
        const double incl = wave_incl_scan(w, lane);
        const uint64_t pos = __ballot(w > 0.0);
        if (pos && lastpos < nb) lastpos = base + 63 - __clzll(pos);
        const uint64_t hit = __ballot(w > 0.0 && run + incl >= target);
        if (hit) {
          choice = base + __ffsll((unsigned long long)hit) - 1;
          found = true;
          break;
        }
        run += __shfl(incl, 63);
      }
      if (!found && !(wpair > 0.0 && run + wpair >= target)) choice = lastpos >= 0 ? lastpos : nb;
    }
    if (choice == nb) {  // pairing; then F over the interior
      prior += t2;
      if (lane == 0) {
        pr[x] = (unsigned short)(y + 1);
        pr[y] = (unsigned short)(x + 1);
      }
      sym = SYM_F;
      x = x + 1;
      y = y - 1;
      have = true;
    } else {             // bifurcation at h: L over [x, h], S over [h+1, y]
      const int h = x + choice;
      prior += t3;
      if (h == x) prior += a.lp[5];  // L -> s
      else stk[top++] = (SYM_L << 28) | ((unsigned)x << 14) | (unsigned)h;
      sym = SYM_S;
      x = h + 1;
      have = true;
    }
  }
  __syncthreads();
  int *o = a.out + (size_t)blockIdx.x * L;
  for (int q = lane; q < L; q += 64) o[q] = pr[q];
  if (lane == 0) a.prior[blockIdx.x] = prior;
}

// log(e^a + e^b) with the reference's -inf guards (src/CommonUtils.jl:145-155).
__device__ __forceinline__ double lse_states(double a, double b) {
  if (a == -HUGE_VAL) return b;
  if (b == -HUGE_VAL) return a;
  return a < b ? b + log1p(exp(a - b)) : a + log1p(exp(b - a));
}

// CommonUtils.sample (src/CommonUtils.jl:129-142) over w[0..n-1]: the first index
// whose running sum of w/total reaches u, else the last.
__device__ __forceinline__ int draw_index(const double *w, int n, double u) {
  double tot = 0.0;
  for (int i = 0; i < n; ++i) tot += w[i];
  double cum = 0.0;
  for (int i = 0; i < n; ++i) {
    cum += w[i] / tot;
    if (cum >= u) return i;
  }
  return n - 1;
}

constexpr int STATES_MAX_CATS = 8, STATES_MAX_COND = STATES_MAX_CATS * STATES_MAX_CATS;

// State sampling given a structure: the sampling part of computeuncachedlikelihood
// (src/FARCE.jl:1398-1486, called by samplestates, 2534-2538).  One thread per
// column.  An unpaired column draws its rate category from exp(sitec - lse)^B; the
// opener x of a pair (x, y) draws the category pair (c1, c2) from the pair
// conditionals (the components of (c1, c2) summed over the M lambda classes),
// normalised and raised to B, and writes both states.  The uniform of column x is
// keyed by (seed, 0, 3, x, x), as the structure sampler keys its intervals.
__global__ void state_sample_kernel(int L, int cats, int M, const double *__restrict__ sitec,
                                    const double *__restrict__ pairc, const double *__restrict__ logw16,
                                    const int *__restrict__ partner, double B, uint64_t seed,
                                    int *__restrict__ states, double *__restrict__ unpaired,
                                    double *__restrict__ pairedll) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= L) return;
  const int p = partner[x];
  double w[STATES_MAX_COND];
  if (p == 0) {
    double tot = -HUGE_VAL;
    for (int c = 0; c < cats; ++c) tot = c == 0 ? sitec[x] : lse_states(tot, sitec[(size_t)c * L + x]);
    for (int c = 0; c < cats; ++c) {
      const double pr = exp(sitec[(size_t)c * L + x] - tot);
      w[c] = B == 1.0 ? pr : pow(pr, B);
    }
    states[x] = draw_index(w, cats, keyed_uniform(seed, 0, 3, x, x)) + 1;
    unpaired[x] = tot;
    pairedll[x] = -HUGE_VAL;
  } else if (p - 1 > x) {
    const int nc = cats * cats;
    double tot = -HUGE_VAL;
    for (int i = 0; i < nc; ++i) {
      double ci = -HUGE_VAL;
      for (int m = 0; m < M; ++m) {
        const int k = i * M + m;
        const double v = logw16[k] + pairc[(size_t)k * L + x];
        tot = lse_states(tot, v);
        ci = lse_states(ci, v);
      }
      w[i] = ci;
    }
    for (int i = 0; i < nc; ++i) {
      const double pr = exp(w[i] - tot);
      w[i] = B == 1.0 ? pr : pow(pr, B);
    }
    const int r = draw_index(w, nc, keyed_uniform(seed, 0, 3, x, x));
    states[x] = r / cats + 1;
    states[p - 1] = r % cats + 1;
    unpaired[x] = 0.0;
    pairedll[x] = tot;
  } else {
    unpaired[x] = 0.0;  // the closing column of a pair: its state comes from the opener
    pairedll[x] = -HUGE_VAL;
  }
}

}  // namespace

void launch_state_sample(hipStream_t s, int L, int cats, int M, const double *sitec, const double *pairc,
                         const double *logw16, const int *partner, double B, uint64_t seed, int *states,
                         double *unpaired, double *pairedll) {
  MESSI_REQUIRE(cats >= 1 && cats <= STATES_MAX_CATS, "state sampling supports 1 .. 8 site categories");
  MESSI_REQUIRE(M >= 1, "need at least one lambda class");
  hipLaunchKernelGGL(state_sample_kernel, dim3((L + 255) / 256), dim3(256), 0, s, L, cats, M, sitec, pairc, logw16,
                     partner, B, seed, states, unpaired, pairedll);
  MESSI_HIP(hipGetLastError());
}

// Natural-log doubles of n scaled FP32 cells (the conversion the FP32 sampler
// would otherwise repeat for every candidate of every interval).
__global__ void xf_to_log_kernel(const XFv *__restrict__ src, double *__restrict__ dst, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = XF::to_log(src[i]);
}

template <class SR>
void sample_launch(hipStream_t s, const typename SR::V *Lr, const typename SR::V *Sc, const typename SR::V *F,
                   const typename SR::V *ptv, int L, int n, uint64_t seed, int first, double B, int *d_out,
                   double *d_prior) {
  SampleArgs<SR> a;
  a.L = L;
  a.n = n;
  a.first = first;
  a.seed = seed;
  a.B = B;
  a.Lr = Lr;
  a.Sc = Sc;
  a.F = F;
  a.ptv = ptv;
  kh99_logprobs(a.lp);
  a.out = d_out;
  a.prior = d_prior;
  const size_t lds = sizeof(unsigned) * (L / 2 + 1) + sizeof(unsigned short) * L;  // <= 64 KiB for L <= 16383
  hipLaunchKernelGGL(sample_kernel<SR>, dim3(n), dim3(64), lds, s, a);
  MESSI_HIP(hipGetLastError());
}

// The FP32 tables are converted to natural-log doubles once (XF::to_log, the
// conversion the FP32 kernel applied per candidate, so the draws are unchanged)
// and sampled by the FP64 kernel.
template <typename T>
void sample_run(hipStream_t s, InsideWork &w, SampleWork &sw, int L, int n, uint64_t seed, int first, double B,
                int *d_out, double *d_prior) {
  MESSI_REQUIRE(L >= 1 && L <= SAMPLE_MAX_L, "structure sampling supports 1 <= L <= 16383");
  MESSI_REQUIRE(n >= 0, "negative sample count");
  if (n == 0) return;
  if constexpr (sizeof(T) == 8) {
    sample_launch<LogSR>(s, w.Lr.as<double>(), w.Sc.as<double>(), w.F.as<double>(), w.ptv.as<double>(), L, n, seed,
                         first, B, d_out, d_prior);
  } else {
    const size_t LL = (size_t)L * L;
    sw.logs.ensure(sizeof(double) * 4 * LL);
    double *lg = sw.logs.as<double>();
    const XFv *src[4] = {w.Lr.as<XFv>(), w.Sc.as<XFv>(), w.F.as<XFv>(), w.ptv.as<XFv>()};
    const int blocks = (int)std::min<size_t>((LL + 255) / 256, 8192);
    for (int k = 0; k < 4; ++k)
      hipLaunchKernelGGL(xf_to_log_kernel, dim3(blocks), dim3(256), 0, s, src[k], lg + k * LL, LL);
    MESSI_HIP(hipGetLastError());
    sample_launch<LogSR>(s, lg, lg + LL, lg + 2 * LL, lg + 3 * LL, L, n, seed, first, B, d_out, d_prior);
  }
}

template void sample_run<double>(hipStream_t, InsideWork &, SampleWork &, int, int, uint64_t, int, double, int *,
                                 double *);
template void sample_run<float>(hipStream_t, InsideWork &, SampleWork &, int, int, uint64_t, int, double, int *,
                                double *);

}  // namespace messi
