/*
 * FP64 CPU oracle for the MESSI likelihood hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates the reference's serial CPU
 * algorithms so that the HIP path can be checked against them.  Only the test
 * suite, the smoke check and the CPU-baseline leg of the benchmark load it;
 * the product library never links or calls it.
 *
 * Restated functions (reference file:line):
 *   oracle_site_loglik   -- felsensteinstack, 4 states (src/FARCE.jl:2182-2254) as driven by
 *                           computeunpairedlikelihoodscat (src/FARCE.jl:742-773)
 *   oracle_pair_loglik   -- felsensteinstack, 16 states (src/FARCE.jl:2182-2254) as driven by
 *                           felsensteinall (src/FARCE.jl:1992-2046); leaf state vector of a
 *                           column pair is data[x] (x) data[y] with index h1*4+h2
 *                           (src/FARCE.jl:370-377)
 *   oracle_inside        -- computeinsideparallel/computeblock (src/InsideOutsideAlgorithm.jl:648-749)
 *   oracle_outside       -- computeoutside (src/InsideOutsideAlgorithm.jl:447-483)
 *   oracle_bpp           -- computebasepairprobs (src/InsideOutsideAlgorithm.jl:486-533)
 *   oracle_mea           -- iterativeposteriordecoding (src/RNATools.jl:6-42), FP64; the kernel
 *                           posteriordecoding (src/cuda/inside.cu:286-321), FP32
 *   oracle_mea_traceback -- posteriordecodingtraceback (src/RNATools.jl:44-74)
 *   oracle_sample_structure -- samplestructure / samplestructurehelper
 *                           (src/InsideOutsideAlgorithm.jl:570-629), keyed uniforms
 *   oracle_kh99_prior    -- calculateKH99prior (src/InsideOutsideAlgorithm.jl:86-144)
 *   oracle_sample_states -- the state sampling of computeuncachedlikelihood
 *                           (src/FARCE.jl:1398-1486) called by samplestates
 *                           (src/FARCE.jl:2534-2538), keyed uniforms
 *   lse                  -- CommonUtils.logsumexp (src/CommonUtils.jl:145-155)
 *
 * Conventions: trees are flat pre-order arrays with the root at index 0,
 * children[2*v], children[2*v+1] (-1 for leaves), seqindex[v] the alignment
 * row of leaf v.  Alignment cells are 4-bit state masks codes[s*ncol+c].
 * Transition matrices are column-major per node (Julia layout):
 * P[node*dim*dim + a + dim*d] = P(a -> d).  L x L matrices are column-major,
 * 3 x L x L tables use Julia's (3,L,L) layout: (s,i,j) at s + 3*(i + L*j).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* KH99 rule probabilities, src/InsideOutsideAlgorithm.jl:40-46. */
#define P_S_LS 0.868534
#define P_S_s 0.117609877998
#define P_S_dFd 0.013856122002
#define P_F_dFd 0.787640
#define P_F_LS 0.21236
#define P_L_s 0.894603
#define P_L_dFd 0.105397

/* Threads of the OpenMP loops above (the per-cell arithmetic, and so every
 * result, does not depend on it); n <= 0 keeps the runtime's default. */
void oracle_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

int oracle_get_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

static double lse(double a, double b) {
  if (b == -INFINITY) return a;
  if (a == -INFINITY) return b;
  if (a < b) return b + log1p(exp(a - b));
  return a + log1p(exp(b - a));
}

/* Post-order pruning of one column (dim 4) or column pair (dim 16).
 * part: nnodes*dim scratch, logm: nnodes scratch.  Rescaling follows
 * src/FARCE.jl:2225-2238: divide by the maximum only when it is below 1e-20. */
static double prune_one(int nnodes, const int *children, const int *seqindex, int ncol,
                        const uint8_t *codes, int x, int y, int dim, const double *P,
                        const double *pi, double *part, double *logm) {
  for (int v = nnodes - 1; v >= 0; --v) {   /* reverse pre-order: children first */
    double *pv = part + (size_t)v * dim;
    int l = children[2 * v], r = children[2 * v + 1];
    if (l < 0) {
      int s = seqindex[v];
      if (dim == 4) {
        uint8_t c = codes[(size_t)s * ncol + x];
        for (int a = 0; a < 4; ++a) pv[a] = (double)((c >> a) & 1);
      } else {
        uint8_t cx = codes[(size_t)s * ncol + x], cy = codes[(size_t)s * ncol + y];
        for (int h1 = 0; h1 < 4; ++h1)
          for (int h2 = 0; h2 < 4; ++h2)
            pv[h1 * 4 + h2] = (double)((cx >> h1) & 1) * (double)((cy >> h2) & 1);
      }
      logm[v] = 0.0;
      continue;
    }
    const double *Pl = P + (size_t)l * dim * dim, *Pr = P + (size_t)r * dim * dim;
    const double *vl = part + (size_t)l * dim, *vr = part + (size_t)r * dim;
    double m = 0.0;
    for (int a = 0; a < dim; ++a) {
      double sl = 0.0, sr = 0.0;
      for (int d = 0; d < dim; ++d) {
        sl += Pl[a + dim * d] * vl[d];
        sr += Pr[a + dim * d] * vr[d];
      }
      pv[a] = sl * sr;
      if (pv[a] > m) m = pv[a];
    }
    if (m < 1e-20) {
      for (int a = 0; a < dim; ++a) pv[a] /= m;
      logm[v] = log(m) + logm[l] + logm[r];
    } else {
      logm[v] = logm[l] + logm[r];
    }
  }
  double tot = 0.0;
  for (int a = 0; a < dim; ++a) tot += pi[a] * part[a];
  return log(tot) + logm[0];
}

void oracle_site_loglik(int nnodes, const int *children, const int *seqindex, int nseq, int ncol,
                        const uint8_t *codes, const double *P4, const double *pi4, double *out) {
  (void)nseq;
#pragma omp parallel
  {
    double *part = (double *)malloc(sizeof(double) * 4 * nnodes);
    double *logm = (double *)malloc(sizeof(double) * nnodes);
#pragma omp for schedule(static)
    for (int c = 0; c < ncol; ++c)
      out[c] = prune_one(nnodes, children, seqindex, ncol, codes, c, c, 4, P4, pi4, part, logm);
    free(part);
    free(logm);
  }
}

/* One mixture component of the pair log-likelihood matrix.  out is L x L
 * column-major; entries for x<y<=x+maxbp (and mask[x+L*y] != 0 when a mask is
 * given) are written symmetrically, everything else is left untouched. */
void oracle_pair_loglik(int nnodes, const int *children, const int *seqindex, int nseq, int ncol,
                        const uint8_t *codes, const double *P16, const double *pi16, int maxbp,
                        const uint8_t *mask, double *out) {
  (void)nseq;
#pragma omp parallel
  {
    double *part = (double *)malloc(sizeof(double) * 16 * nnodes);
    double *logm = (double *)malloc(sizeof(double) * nnodes);
#pragma omp for schedule(dynamic, 4)
    for (int x = 0; x < ncol; ++x) {
      int yend = x + maxbp < ncol - 1 ? x + maxbp : ncol - 1;
      for (int y = x + 1; y <= yend; ++y) {
        if (mask && !mask[x + (size_t)ncol * y]) continue;
        double v = prune_one(nnodes, children, seqindex, ncol, codes, x, y, 16, P16, pi16, part, logm);
        out[x + (size_t)ncol * y] = v;
        out[y + (size_t)ncol * x] = v;
      }
    }
    free(part);
    free(logm);
  }
}

/* Accumulate a component into a running mixture: acc = lse(acc, logw + comp),
 * elementwise over n values (src/FARCE.jl:1780-1792, src/FARCE.jl:1207-1218). */
void oracle_mix_accumulate(size_t n, double *acc, const double *comp, double logw, int first) {
  for (size_t i = 0; i < n; ++i) {
    if (first) acc[i] = logw + comp[i];
    else acc[i] = lse(acc[i], logw + comp[i]);
  }
}

#define IDX3(s, i, j) ((size_t)(s) + 3 * ((size_t)(i) + (size_t)L * (size_t)(j)))

/* KH99 inside, src/InsideOutsideAlgorithm.jl:648-749.  S=0, F=1, L=2. */
void oracle_inside(int L, const double *u, const double *paired, double B, double *in) {
  const double lp[7] = {log(P_S_LS), log(P_S_s), log(P_S_dFd), log(P_F_dFd),
                        log(P_F_LS), log(P_L_s), log(P_L_dFd)};
  size_t n = (size_t)3 * L * L;
  for (size_t k = 0; k < n; ++k) in[k] = -INFINITY;
  for (int i = 0; i < L; ++i) {
    in[IDX3(0, i, i)] = (u[i] + lp[1]) * B;
    in[IDX3(2, i, i)] = (u[i] + lp[5]) * B;
  }
  /* type-3 rules: S->LS (lp[0]) and F->LS (lp[4]); type-2 rules: S->dFd (lp[2]),
   * F->dFd (lp[3]), L->dFd (lp[6]), in the reference's order. */
  const int t3left[2] = {0, 1};
  const double t3lp[2] = {lp[0], lp[4]};
  const int t2left[3] = {0, 1, 2};
  const double t2lp[3] = {lp[2], lp[3], lp[6]};
  for (int d = 1; d < L; ++d) {
    /* the cells of one span depend only on shorter spans */
#pragma omp parallel for schedule(static) if (L - d > 64)
    for (int j = 0; j < L - d; ++j) {
      int y = j + d;
      for (int r = 0; r < 2; ++r) {
        double tmp = -INFINITY;
        for (int h = j; h <= y - 1; ++h) tmp = lse(tmp, in[IDX3(2, j, h)] + in[IDX3(0, h + 1, y)]);
        size_t c = IDX3(t3left[r], j, y);
        in[c] = lse(in[c], t3lp[r] * B + tmp);
      }
      for (int r = 0; r < 3; ++r) {
        size_t c = IDX3(t2left[r], j, y);
        double f = (j + 1 < L && y - 1 >= 0) ? in[IDX3(1, j + 1, y - 1)] : -INFINITY;
        in[c] = lse(in[c], (t2lp[r] + paired[j + (size_t)L * y]) * B + f);
      }
    }
  }
}

/* KH99 outside, src/InsideOutsideAlgorithm.jl:447-483. */
void oracle_outside(int L, const double *in, const double *u, const double *paired, double *out) {
  (void)u;
  const double lp[7] = {log(P_S_LS), log(P_S_s), log(P_S_dFd), log(P_F_dFd),
                        log(P_F_LS), log(P_L_s), log(P_L_dFd)};
  size_t n = (size_t)3 * L * L;
  for (size_t k = 0; k < n; ++k) out[k] = -INFINITY;
  out[IDX3(0, 0, L - 1)] = 0.0;
  /* type-3 rules (left, right1, right2, logp): S->LS, F->LS. */
  const int t3left[2] = {0, 1};
  const double t3lp[2] = {lp[0], lp[4]};
  const int t2left[3] = {0, 1, 2};
  const double t2lp[3] = {lp[2], lp[3], lp[6]};
  for (int c = 2; c <= L; ++c) {
    int b = L - c;
    /* the cells of one span depend only on longer spans */
#pragma omp parallel for schedule(static) if (L - b > 64)
    for (int j = 0; j <= L - b - 1; ++j) {
      for (int r = 0; r < 2; ++r) {
        int A = t3left[r];
        double tmp = -INFINITY;
        for (int k = j + b + 1; k <= L - 1; ++k)
          tmp = lse(tmp, out[IDX3(A, j, k)] + in[IDX3(0, j + b + 1, k)]);
        size_t cl = IDX3(2, j, j + b);
        out[cl] = lse(out[cl], tmp + t3lp[r]);
        tmp = -INFINITY;
        for (int k = 0; k <= j - 1; ++k)
          tmp = lse(tmp, out[IDX3(A, k, j + b)] + in[IDX3(2, k, j - 1)]);
        size_t cs = IDX3(0, j, j + b);
        out[cs] = lse(out[cs], tmp + t3lp[r]);
      }
      if (j >= 1 && j + b + 1 < L) {
        for (int r = 0; r < 3; ++r) {
          size_t cf = IDX3(1, j, j + b);
          out[cf] = lse(out[cf], t2lp[r] + paired[(j - 1) + (size_t)L * (j + b + 1)] +
                                     out[IDX3(t2left[r], j - 1, j + b + 1)]);
        }
      }
    }
  }
}

/* Base-pair posteriors, src/InsideOutsideAlgorithm.jl:486-533.  res (L) and
 * pair (L x L column-major) receive the row-normalised probabilities. */
void oracle_bpp(int L, const double *in, const double *out, const double *u, const double *paired,
                double *res, double *pair) {
  const double lp[7] = {log(P_S_LS), log(P_S_s), log(P_S_dFd), log(P_F_dFd),
                        log(P_F_LS), log(P_L_s), log(P_L_dFd)};
  double Z = in[IDX3(0, 0, L - 1)];
  for (int j = 0; j < L; ++j) {
    double tmp = -INFINITY;
    tmp = lse(tmp, out[IDX3(0, j, j)] + lp[1]);
    tmp = lse(tmp, out[IDX3(2, j, j)] + lp[5]);
    res[j] = tmp + u[j] - Z;
  }
  for (size_t k = 0; k < (size_t)L * L; ++k) pair[k] = -INFINITY;
  const int t2left[3] = {0, 1, 2};
  const double t2lp[3] = {lp[2], lp[3], lp[6]};
#pragma omp parallel for schedule(dynamic, 8)
  for (int j = 0; j < L; ++j) {
    for (int k = j + 1; k < L; ++k) {
      double tmp = -INFINITY;
      for (int r = 0; r < 3; ++r) {
        double inres = (j + 1 <= k - 1) ? in[IDX3(1, j + 1, k - 1)] : -INFINITY;
        double outres = out[IDX3(t2left[r], j, k)];
        tmp = lse(tmp, outres + t2lp[r] + inres);
      }
      double v = tmp + paired[j + (size_t)L * k] - Z;
      pair[j + (size_t)L * k] = v;
      pair[k + (size_t)L * j] = v;
    }
  }
  double m = -INFINITY;
  for (int j = 0; j < L; ++j) if (res[j] > m) m = res[j];
  for (size_t k = 0; k < (size_t)L * L; ++k) if (pair[k] > m) m = pair[k];
  for (int j = 0; j < L; ++j) res[j] = exp(res[j] - m);
  for (size_t k = 0; k < (size_t)L * L; ++k) pair[k] = exp(pair[k] - m);
#pragma omp parallel for schedule(static)
  for (int x = 0; x < L; ++x) {
    double norm = res[x];
    for (int y = 0; y < L; ++y) norm += pair[x + (size_t)L * y];
    res[x] /= norm;
    for (int y = 0; y < L; ++y) pair[x + (size_t)L * y] /= norm;
  }
}

/* Maximum-expected-accuracy (posterior) decoding, fill step.
 *
 * prec 0 (FP64) follows iterativeposteriordecoding (src/RNATools.jl:6-42): the
 * pair weight of (i, k) is pairprobs[i, k] (column-major element i + L*k), and
 * v = (alpha * p + E[i+1,k-1]) + E[k+1,j] with every operation rounded.
 * prec 1 (FP32) follows the kernel posteriordecoding (src/cuda/inside.cu:286-321)
 * as driven by computeposteriordecodingcuda (src/InsideCUDA.jl:183-215): inputs
 * converted to float, the device reads element i*L + k of the column-major matrix
 * (pairprobs[k, i]), and nvcc's default contraction fuses alpha * p + E into one
 * fused multiply-add, then adds E[k+1,j].
 * Both: E[i,i] = 0 (the diagonal is never set), e1 = single[i] + E[i+1,j],
 * e2 = the k = j pair term with E[i+1,j-1], e3 = the first maximal k in i < k < j
 * (strict >), E = e1, replaced by e2 and then e3 only when strictly greater
 * (the Julia loop's findmax agrees for inputs without NaN or negative zero; the
 * CUDA kernel starts its maximum at -1e10 rather than -inf, which only matters
 * for terms below -1e10, never for probabilities).
 * S[i,j] (column-major int) = -1 (e1), -2 (e2) or the 1-based k of e3.
 * E (L x L, column-major) receives the table as double. */
static double mea_pair_term(int prec, double alpha, const double *pair, int L, int i, int k, double e) {
  if (prec == 0) {
    double ap = alpha * pair[i + (size_t)L * k];
    return ap + e;
  }
  float p = (float)pair[k + (size_t)L * i];
  return (double)fmaf((float)alpha, p, (float)e);
}

void oracle_mea(int L, const double *single, const double *pair, double alpha, int prec, double *E, int *S) {
  for (size_t q = 0; q < (size_t)L * L; ++q) {
    E[q] = 0.0;
    S[q] = 0;
  }
#define EM(r, c) ((r) > (c) ? 0.0 : E[(r) + (size_t)L * (c)])
  for (int d = 1; d < L; ++d) {
    for (int i = 0; i + d < L; ++i) {
      int j = i + d;
      double val, best = -INFINITY;
      int bestk = -1, s;
      if (prec == 0) {
        double e1 = single[i] + EM(i + 1, j);
        double e2 = mea_pair_term(0, alpha, pair, L, i, j, EM(i + 1, j - 1));
        for (int k = i + 1; k <= j - 1; ++k) {
          double v = mea_pair_term(0, alpha, pair, L, i, k, EM(i + 1, k - 1)) + EM(k + 1, j);
          if (v > best) {
            best = v;
            bestk = k;
          }
        }
        val = e1;
        s = -1;
        if (e2 > val) {
          val = e2;
          s = -2;
        }
        if (best > val) {
          val = best;
          s = bestk + 1;
        }
      } else {
        float e1 = (float)single[i] + (float)EM(i + 1, j);
        float e2 = (float)mea_pair_term(1, alpha, pair, L, i, j, EM(i + 1, j - 1));
        float bf = -INFINITY;
        for (int k = i + 1; k <= j - 1; ++k) {
          float v = (float)mea_pair_term(1, alpha, pair, L, i, k, EM(i + 1, k - 1)) + (float)EM(k + 1, j);
          if (v > bf) {
            bf = v;
            bestk = k;
          }
        }
        float vf = e1;
        s = -1;
        if (e2 > vf) {
          vf = e2;
          s = -2;
        }
        if (bf > vf) {
          vf = bf;
          s = bestk + 1;
        }
        val = vf;
      }
      E[i + (size_t)L * j] = val;
      S[i + (size_t)L * j] = s;
    }
  }
#undef EM
}

/* Traceback of an S table into a pairing, posteriordecodingtraceback
 * (src/RNATools.jl:44-74): paired[i] = 1-based partner, 0 when unpaired. */
void oracle_mea_traceback(int L, const int *S, int *paired) {
  for (int i = 0; i < L; ++i) paired[i] = 0;
  if (L < 1) return;
  int *st = (int *)malloc(sizeof(int) * 2 * ((size_t)L + 1));
  int top = 0;
  st[0] = 0;
  st[1] = L - 1;
  top = 1;
  while (top > 0) {
    --top;
    int i = st[2 * top], j = st[2 * top + 1];
    int k = S[i + (size_t)L * j];
    if (k == -1) {
      st[2 * top] = i + 1;
      st[2 * top + 1] = j;
      ++top;
    } else if (k == -2) {
      paired[i] = j + 1;
      paired[j] = i + 1;
      if (i + 1 < j - 1) {
        st[2 * top] = i + 1;
        st[2 * top + 1] = j - 1;
        ++top;
      }
    } else if (k > 0) {
      int kk = k - 1;
      if (i < kk) {
        paired[i] = kk + 1;
        paired[kk] = i + 1;
      }
      if (i + 1 <= kk - 1) {
        st[2 * top] = i + 1;
        st[2 * top + 1] = kk - 1;
        ++top;
      }
      if (kk + 1 <= j) {
        st[2 * top] = kk + 1;
        st[2 * top + 1] = j;
        ++top;
      }
    }
  }
  free(st);
}

/* Keyed uniform in (0, 1] for the stochastic traceback: one draw per interval
 * (symbol, x, y) of a sample, so the result does not depend on the order in
 * which intervals are visited.  The mixer is the SplitMix64 finaliser (Steele,
 * Lea and Flood, "Fast splittable pseudorandom number generators", 2014).  The
 * reference draws from Julia's MersenneTwister instead; its stream is not
 * reproduced. */
static uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
double oracle_sample_uniform(uint64_t seed, int sample, int sym, int x, int y) {
  uint64_t key = mix64(seed + 0x9E3779B97F4A7C15ULL * (uint64_t)(sample + 1));
  uint64_t z = mix64(key ^ (((uint64_t)(uint32_t)x << 32) | (uint64_t)(uint32_t)y) ^ ((uint64_t)sym << 62));
  return (double)((z >> 11) + 1) * 0x1.0p-53;
}

/* Stochastic traceback of one structure, samplestructure / samplestructurehelper
 * (src/InsideOutsideAlgorithm.jl:570-629) from (S, 0, L-1): for an interval
 * (symbol, x, y), x < y, the candidates are the bifurcations h = x..y-1 of the
 * type-3 rule of the symbol (S->LS, F->LS: logprob + L[x,h] + S[h+1,y]) and then
 * the pairing of its type-2 rule (logprob + paired[x,y] + F[x+1,y-1]), all times B;
 * a candidate is drawn with probability exp(l - logsumexp(l)) by the first
 * normalised cumulative sum >= u (CommonUtils.sample, src/CommonUtils.jl:129-142).
 * in: 3 x L x L inside table (S=0, F=1, L=2); paired: column-major pair log-likelihoods.
 * partner receives 1-based partners (0 unpaired); the return value is the
 * structure's KH99 log prior (calculateKH99prior of the result). */
double oracle_kh99_prior(int L, const int *partner);

double oracle_sample_structure(int L, const double *in, const double *paired, double B, uint64_t seed, int sample,
                               int *partner) {
  const double lp[7] = {log(P_S_LS), log(P_S_s), log(P_S_dFd), log(P_F_dFd),
                        log(P_F_LS), log(P_L_s), log(P_L_dFd)};
  for (int i = 0; i < L; ++i) partner[i] = 0;
  if (L < 1) return 0.0;
  int *st = (int *)malloc(sizeof(int) * 3 * (2 * (size_t)L + 2));
  double *ll = (double *)malloc(sizeof(double) * ((size_t)L + 1));
  int top = 0;
  st[0] = 0;
  st[1] = 0;
  st[2] = L - 1;
  top = 1;
  while (top > 0) {
    --top;
    int sym = st[3 * top], x = st[3 * top + 1], y = st[3 * top + 2];
    if (x >= y) continue;
    int n = 0;
    double t3 = sym == 0 ? lp[0] : lp[4];
    if (sym != 2)
      for (int h = x; h <= y - 1; ++h) ll[n++] = (t3 + in[IDX3(2, x, h)] + in[IDX3(0, h + 1, y)]) * B;
    double t2 = sym == 0 ? lp[2] : sym == 1 ? lp[3] : lp[6];
    double f = (x + 1 <= y - 1) ? in[IDX3(1, x + 1, y - 1)] : -INFINITY;
    ll[n++] = ((t2 + paired[x + (size_t)L * y]) + f) * B;
    double s = -INFINITY;
    for (int k = 0; k < n; ++k) s = lse(s, ll[k]);
    double tot = 0.0;
    for (int k = 0; k < n; ++k) {
      ll[k] = exp(ll[k] - s);
      tot += ll[k];
    }
    double u = oracle_sample_uniform(seed, sample, sym, x, y);
    double cum = 0.0;
    int r = n - 1;
    for (int k = 0; k < n; ++k) {
      cum += ll[k] / tot;
      if (cum >= u) {
        r = k;
        break;
      }
    }
    if (r == n - 1) {  /* the pairing */
      partner[x] = y + 1;
      partner[y] = x + 1;
      st[3 * top] = 1;
      st[3 * top + 1] = x + 1;
      st[3 * top + 2] = y - 1;
      ++top;
    } else {           /* bifurcation at h = x + r: L (x, h), then S (h+1, y) popped first */
      int h = x + r;
      st[3 * top] = 2;
      st[3 * top + 1] = x;
      st[3 * top + 2] = h;
      ++top;
      st[3 * top] = 0;
      st[3 * top + 1] = h + 1;
      st[3 * top + 2] = y;
      ++top;
    }
  }
  free(st);
  free(ll);
  return oracle_kh99_prior(L, partner);
}

/* calculateKH99prior (src/InsideOutsideAlgorithm.jl:86-144): log probability of a
 * structure's KH99 derivation with the composite rules S -> sS | dFdS | s | dFd,
 * F -> dFd | sS | dFdS (S -> LS, F -> LS times L -> s or L -> dFd). */
double oracle_kh99_prior(int L, const int *partner) {
  if (L < 1) return 0.0;
  const double r_S_sS = log(P_S_LS * P_L_s), r_S_dFdS = log(P_S_LS * P_L_dFd), r_S_s = log(P_S_s),
               r_S_dFd = log(P_S_dFd), r_F_dFd = log(P_F_dFd), r_F_sS = log(P_F_LS * P_L_s),
               r_F_dFdS = log(P_F_LS * P_L_dFd);
  int *st = (int *)malloc(sizeof(int) * 3 * (2 * (size_t)L + 2));
  int top = 1;
  st[0] = 0;
  st[1] = 0;
  st[2] = L - 1;
  double ll = 0.0;
  while (top > 0) {
    --top;
    int sym = st[3 * top], s = st[3 * top + 1], e = st[3 * top + 2];
    if (s < 0 || s >= L) continue;
    int p = partner[s] - 1;  /* 0-based partner, -1 unpaired */
    if (partner[s] == 0) {
      if (s == e) {
        if (sym == 0) ll += r_S_s;          /* F has no single-position rule */
      } else if (s < e) {
        ll += sym == 0 ? r_S_sS : r_F_sS;
        st[3 * top] = 0;
        st[3 * top + 1] = s + 1;
        st[3 * top + 2] = e;
        ++top;
      }
    } else if (p > s) {
      if (p == e) {
        ll += sym == 0 ? r_S_dFd : r_F_dFd;
        st[3 * top] = 1;
        st[3 * top + 1] = s + 1;
        st[3 * top + 2] = p - 1;
        ++top;
      } else {
        ll += sym == 0 ? r_S_dFdS : r_F_dFdS;
        st[3 * top] = 1;
        st[3 * top + 1] = s + 1;
        st[3 * top + 2] = p - 1;
        ++top;
        st[3 * top] = 0;
        st[3 * top + 1] = p + 1;
        st[3 * top + 2] = e;
        ++top;
      }
    }
  }
  free(st);
  return ll;
}

/* CommonUtils.sample (src/CommonUtils.jl:129-142): the first index whose running
 * sum of w/total reaches u, else the last. */
static int draw_index(const double *w, int n, double u) {
  double tot = 0.0, cum = 0.0;
  for (int i = 0; i < n; ++i) tot += w[i];
  for (int i = 0; i < n; ++i) {
    cum += w[i] / tot;
    if (cum >= u) return i;
  }
  return n - 1;
}

/* State sampling given a structure, src/FARCE.jl:1398-1486.  sitec (K4 x L,
 * row k = site category k) holds log w_k + the column's log-likelihood under
 * category k (unpairedconditionals); pairc (K16 x L) the log-likelihood of the
 * pair opened at column x under pair component k, components ordered (c1, c2,
 * then M lambda classes) with log weights logw16 (fullconditionals = logw16 +
 * pairc).  An unpaired column x draws c with probability proportional to
 * exp(sitec - lse)^B; the opener x of (x, y) draws (c1, c2) from the lambda-summed
 * pair conditionals, normalised and raised to B, and sets states[x] = c1+1,
 * states[y] = c2+1.  unpaired[x] is the lse over categories (0 for paired
 * columns), pairedll[x] the lse over all pair components at openers (-inf
 * elsewhere).  The uniform of column x is oracle_sample_uniform(seed, 0, 3, x, x)
 * (the reference draws from its rng in column order). */
void oracle_sample_states(int L, int cats, int M, const double *sitec, const double *pairc, const double *logw16,
                          const int *partner, double B, uint64_t seed, int *states, double *unpaired,
                          double *pairedll) {
  const int nc = cats * cats;
  double *w = (double *)malloc(sizeof(double) * (size_t)(nc > cats ? nc : cats));
  for (int x = 0; x < L; ++x) {
    const int p = partner[x];
    if (p == 0) {
      double tot = -INFINITY;
      for (int c = 0; c < cats; ++c) tot = c == 0 ? sitec[x] : lse(tot, sitec[(size_t)c * L + x]);
      for (int c = 0; c < cats; ++c) {
        const double pr = exp(sitec[(size_t)c * L + x] - tot);
        w[c] = B == 1.0 ? pr : pow(pr, B);
      }
      states[x] = draw_index(w, cats, oracle_sample_uniform(seed, 0, 3, x, x)) + 1;
      unpaired[x] = tot;
      pairedll[x] = -INFINITY;
    } else if (p - 1 > x) {
      double tot = -INFINITY;
      for (int i = 0; i < nc; ++i) {
        double ci = -INFINITY;
        for (int m = 0; m < M; ++m) {
          const int k = i * M + m;
          const double v = logw16[k] + pairc[(size_t)k * L + x];
          tot = lse(tot, v);
          ci = lse(ci, v);
        }
        w[i] = ci;
      }
      for (int i = 0; i < nc; ++i) {
        const double pr = exp(w[i] - tot);
        w[i] = B == 1.0 ? pr : pow(pr, B);
      }
      const int r = draw_index(w, nc, oracle_sample_uniform(seed, 0, 3, x, x));
      states[x] = r / cats + 1;
      states[p - 1] = r % cats + 1;
      unpaired[x] = 0.0;
      pairedll[x] = tot;
    } else {
      unpaired[x] = 0.0;
      pairedll[x] = -INFINITY;
    }
  }
  free(w);
}
