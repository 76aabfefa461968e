/*
 * libmessi -- MI355X (gfx950) likelihood hot path of MESSI behind a C ABI.
 *
 * One context per host thread; one HIP stream per context.  Every call is
 * synchronous from the caller's view, reads and writes caller-owned host
 * arrays only during the call, and returns 0 on success or a non-zero code
 * with a message available from messi_last_error().
 *
 * Layouts are Julia-native (column-major):
 *   alignment  data[s + nseq*c + nseq*ncol*a]  (nseq, ncol, 4), 0/1 state indicators
 *   matrices   P[k][node][a + dim*d] = P(a -> d)  (dim, dim, nnodes, K)
 *   L x L      M[i + L*j]
 *   3 x L x L  T[s + 3*(i + L*j)], s = 0 (S), 1 (F), 2 (L) -- the KH99 nonterminals
 * Trees are pre-order with the root at index 0 (the reference's nodeindex - 1);
 * children[2*v], children[2*v+1] are -1 for leaves; seqindex[v] is the
 * alignment row (0-based) of leaf v and -1 for internal nodes.
 *
 * prec: MESSI_FP64 is the plain path (matches the reference CPU semantics);
 * MESSI_FP32 keeps pruning and the inside/outside bulk in FP32 with power-of-two
 * scaling per pattern and baseline-relative base-2 log tables.
 */
#ifndef MESSI_H
#define MESSI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MESSI_FP64 0
#define MESSI_FP32 1

typedef struct messi_ctx messi_ctx;
typedef struct messi_dataset messi_dataset;
typedef struct messi_batch messi_batch;
typedef struct messi_mlp messi_mlp;

/* Context and error handling.  Replaces the per-call CuContext/CuModuleFile setup
 * of src/FelsensteinCUDA.jl:240-253 and src/InsideCUDA.jl:96-104. */
int messi_ctx_create(int device, messi_ctx **out);
void messi_ctx_destroy(messi_ctx *ctx);
const char *messi_last_error(const messi_ctx *ctx);
/* Record per-phase device times (HIP events) for messi_last_timings. */
int messi_set_profiling(messi_ctx *ctx, int on);
/* Tile edge (16 or 32) of the inside schedule used by later calls on this context. */
int messi_set_tile(messi_ctx *ctx, int tile);
/* Bulk window (1, 2 or 4 launches) of the inside schedule used by later calls on
 * this context: each bulk accumulator tile is read and written once per window,
 * adding the split-point products of every launch in it.  Results do not depend
 * on it.  Default 4.  No counterpart in the reference, whose inside has no tiles. */
int messi_set_bulk_window(messi_ctx *ctx, int window);
/* Whether later inside / outside passes on this context skip the split-point
 * terms that a pairing band (pair scores -inf beyond a distance) makes -inf
 * (default 1).  Results do not depend on it; 0 runs the full-span recursions,
 * as the reference always does (src/InsideOutsideAlgorithm.jl:447-483,648-749). */
int messi_set_band_skip(messi_ctx *ctx, int on);
/* Whether later inside passes on this context run as one launch (1) instead of
 * one launch per tile-diagonal (0, default): resident workgroups take the same
 * work items from a ticket counter and wait on per-tile flags, so each tile waits
 * only for the tiles it reads and members of a batch proceed independently.
 * Results are bit-identical either way.  A wait that does not end within a few
 * seconds fails the call.  The allocation-free entry point always uses the
 * launch sequence.  No counterpart in the reference, whose inside has no tiles
 * (the schedule it replaces launches one kernel per anti-diagonal,
 * src/InsideCUDA.jl:87-181). */
int messi_set_single_launch(messi_ctx *ctx, int on);
/* Whether datasets created later on this context compute some 16-state pair
 * nodes inside their parent's level launch instead of storing them (1, default):
 * a node whose entries mostly serve one entry of its parent is then never
 * written to or read back from device memory.  Results are bit-identical either
 * way.  No counterpart in the reference, which stores every node's partials
 * (felsensteinstack, src/FARCE.jl:2182-2254). */
int messi_set_pair_fusion(messi_ctx *ctx, int on);
/* Checks the MFMA operand layouts the kernels rely on against a CPU product. */
int messi_selftest(messi_ctx *ctx);
/* Page-locked host memory on the context's device (hipHostMalloc), for callers
 * that receive large results repeatedly (an L x L posterior matrix is copied into
 * it at the link's full rate instead of through pageable staging); freed by
 * messi_host_free, which does not use the context (it may be destroyed
 * already).  No counterpart in the reference. */
int messi_host_alloc(messi_ctx *ctx, size_t bytes, void **out);
int messi_host_free(messi_ctx *ctx, void *p);

/* Dataset handle: one-hot alignment, rooted binary tree, base-pair distance
 * bound and optional pair mask (mask[x + ncol*y] != 0 allows the pair, NULL
 * allows all).  Builds the per-node pattern dedup (src/Dataset.jl:180-196) and
 * the per-node pair index lists (getindexdict, src/FelsensteinCUDA.jl:77-142)
 * once, replacing the reference's global caches (src/FelsensteinCUDA.jl:206-238). */
int messi_dataset_create(messi_ctx *ctx, int nseq, int ncol, const double *data, int nnodes,
                         const int *children, const int *seqindex, int maxbasepairdistance,
                         const uint8_t *mask, messi_dataset **out);
void messi_dataset_destroy(messi_dataset *ds);
/* Work counts: stats[0] = sum over internal nodes of distinct pair patterns
 * (sum U_v), stats[1] = root pair entries, stats[2] = in-range column pairs,
 * stats[3] = sum over internal nodes of distinct column patterns, stats[4] =
 * tree levels, stats[5] = device bytes held by the dataset. */
int messi_dataset_stats(const messi_dataset *ds, int64_t *stats);

/* Site log-likelihoods over K site-rate components:
 * u[c] = lse_k(logw[k] + log sum_a pi4[k][a] * root_a(c)).
 * Replaces computeunpairedlikelihoods (src/FARCE.jl:1198-1220), whose per-column
 * work is felsensteinstack (src/FARCE.jl:2182-2254). */
int messi_site_loglik(messi_dataset *ds, int K, const double *P4, const double *pi4,
                      const double *logw, int prec, double *u_out);

/* Pair log-likelihood matrix, mixed over K components with log weights.
 * out (ncol x ncol) is symmetric; the diagonal, pairs beyond the distance bound
 * and masked pairs are -inf.  percomp (K x ncol x ncol) receives the
 * per-component matrices when not NULL.
 * Replaces felsensteincuda (src/FelsensteinCUDA.jl:214-516). */
int messi_pair_loglik(messi_dataset *ds, int K, const double *P16, const double *pi16,
                      const double *logw, int prec, double *out, double *percomp);

/* KH99 inside.  Z receives log Z = S[0, L-1]; when zonly == 0, in3 (3 x L x L)
 * receives the inside table (-inf below the diagonal).
 * Replaces computeinsideKH99 (src/InsideCUDA.jl:23-53) and its CPU semantics
 * computeinsideparallel (src/InsideOutsideAlgorithm.jl:683-749). */
int messi_inside(messi_ctx *ctx, int L, const double *u, const double *paired, double B, int zonly,
                 int prec, double *Z, double *in3);

/* messi_inside without any device allocation.  No counterpart in the reference:
 * this variant is for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, results and errors as messi_inside, for 1 <= L <= MESSI_NOALLOC_MAX_L
 * (larger L is rejected with an error).  Every buffer of the call is a fixed
 * region of a static device array of about 25 MiB that each GPU the library
 * runs on holds until the process ends; all calls on one GPU share it, so calls
 * that overlap on one GPU (from any context or thread) race. */
#define MESSI_NOALLOC_MAX_L 512
int messi_inside_noalloc(messi_ctx *ctx, int L, const double *u, const double *paired, double B, int zonly,
                         int prec, double *Z, double *in3);

/* KH99 outside from an inside table.  Replaces computeoutsideKH99
 * (src/InsideCUDA.jl:55-61) and computeoutside (src/InsideOutsideAlgorithm.jl:447-483). */
int messi_outside(messi_ctx *ctx, int L, const double *in3, const double *u, const double *paired,
                  int prec, double *out3);

/* Row-normalised base-pair posteriors.  Replaces computebasepairprobs
 * (src/InsideOutsideAlgorithm.jl:486-533). */
int messi_bpp(messi_ctx *ctx, int L, const double *in3, const double *out3, const double *u,
              const double *paired, int prec, double *unpaired, double *pairprob);

/* One full-likelihood evaluation on the device with no intermediate host
 * transfers: site log-likelihoods, pair matrix, inside log Z.  Replaces
 * computestructureintegratedlikelihood without its prior term
 * (src/FARCE.jl:2322-2345). */
int messi_eval(messi_dataset *ds, int K4, const double *P4, const double *pi4, const double *logw4,
               int K16, const double *P16, const double *pi16, const double *logw16, double B,
               int prec, double *logZ);

/* Evaluation plus outside and base-pair posteriors (the processmax path,
 * src/MESSI.jl:1097-1103).  unpaired (L) and pairprob (L x L) may be NULL. */
int messi_posterior(messi_dataset *ds, int K4, const double *P4, const double *pi4,
                    const double *logw4, int K16, const double *P16, const double *pi16,
                    const double *logw16, int prec, double *logZ, double *unpaired,
                    double *pairprob);

/* Maximum-expected-accuracy consensus structure from base-pair posteriors
 * (getPosteriorDecodingConsensusStructure, src/RNATools.jl:76-85): the fill of
 * computeposteriordecodingcuda (src/InsideCUDA.jl:183-215, kernel
 * src/cuda/inside.cu:286-321) and posteriordecodingtraceback (src/RNATools.jl:44-74).
 * unpaired (L) and pairprob (L x L, column-major) are as messi_bpp returns them.
 * prec MESSI_FP64 follows iterativeposteriordecoding (src/RNATools.jl:6-42) with
 * its arguments in the order of its signature; MESSI_FP32 follows the CUDA kernel
 * (float inputs, pairprob read at i*L + k, alpha*p + E fused).  paired (L)
 * receives the 1-based partner of each position, 0 when unpaired; emat (double)
 * and smat (-1, -2 or the 1-based split point) receive the L x L tables,
 * column-major, zero on and below the diagonal, or are NULL.  1 <= L <= 16383. */
int messi_mea(messi_ctx *ctx, int L, const double *unpaired, const double *pairprob, double alpha, int prec,
              int *paired, double *emat, int *smat);

/* Milliseconds of the last MEA call with profiling on: host upload, fill,
 * traceback (the upload is 0 after messi_posterior_consensus).  No counterpart in
 * the reference. */
int messi_mea_timings(messi_ctx *ctx, double *t);

/* Evaluation, outside, base-pair posteriors and MEA consensus on the device with
 * only log Z and the pairing returned (the processmax path, src/MESSI.jl:1097-1115).
 * prec selects the likelihood path, mea_prec the decoding. */
int messi_posterior_consensus(messi_dataset *ds, int K4, const double *P4, const double *pi4,
                              const double *logw4, int K16, const double *P16, const double *pi16,
                              const double *logw16, int prec, int mea_prec, double alpha, double *logZ,
                              int *paired);

/* Stochastic tracebacks: nsamples KH99 structures drawn from an inside table, as
 * samplestructure (src/InsideOutsideAlgorithm.jl:623-627) draws one from (S, 1, L).
 * in3 (3 x L x L), paired (L x L) and unpaired (L) as for messi_outside; prec
 * selects the table the draws read (FP64 log values or the scaled FP32 copy).
 * Interval (symbol, x, y) of sample k uses the uniform keyed by (seed, k, symbol,
 * x, y) (SplitMix64 finaliser), not the reference's MersenneTwister stream.
 * structures (nsamples x L, sample-major) receives 1-based partners, 0 when
 * unpaired; logprior (nsamples, may be NULL) the KH99 log prior of each structure
 * (calculateKH99prior, src/InsideOutsideAlgorithm.jl:86-144).  1 <= L <= 16383. */
int messi_sample_structures(messi_ctx *ctx, int L, const double *in3, const double *paired, const double *unpaired,
                            int prec, int nsamples, uint64_t seed, double B, int *structures, double *logprior);

/* Evaluation and sampling without the inside table leaving the device (the MCMC
 * step's structure draw, src/MCMC.jl:1736-1746, for many samples at once). */
int messi_sample_posterior(messi_dataset *ds, int K4, const double *P4, const double *pi4, const double *logw4,
                           int K16, const double *P16, const double *pi16, const double *logw16, int prec,
                           int nsamples, uint64_t seed, double B, double *logZ, int *structures,
                           double *logprior);

/* Fixed-structure log-likelihood: the sum of the site log-likelihoods of the
 * unpaired columns and of the pair log-likelihoods of the pairs of the structure
 * paired (L, 1-based partners, 0 unpaired); parts (2, may be NULL) receives the
 * two sums.  Replaces felsensteincudapaired (src/FelsensteinCUDA.jl:721-912) with
 * the unpaired sum of computemodellikelihood's GPU branch (src/FARCE.jl:2479-2487);
 * the reference's leftnumpairs offset (src/FelsensteinCUDA.jl:858-859) is not
 * carried over.  Every pair must be in the dataset's pair list (distance bound and
 * mask): a dataset created with the structure's pairs as its mask prunes only those. */
int messi_structure_loglik(messi_dataset *ds, int K4, const double *P4, const double *pi4, const double *logw4,
                           int K16, const double *P16, const double *pi16, const double *logw16,
                           const int *paired, int prec, double *loglik, double *parts);
/* State sampling given a structure: samplestates (src/FARCE.jl:2534-2538), i.e.
 * the sampling branch of computeuncachedlikelihood (src/FARCE.jl:1336-1486).  The
 * site components are the cats rate categories (K4 == cats); the pair components
 * are ordered category 1, category 2, then K16 / cats^2 lambda classes, as
 * felsensteincuda orders them.  Every unpaired column draws a category with
 * probability proportional to (w_c L_c / sum)^B; the opener of each pair draws the
 * category pair from the pair conditionals (summed over lambda classes),
 * normalised and raised to B.  Outputs (L each): states (1-based categories),
 * unpaired (the column's mixture log-likelihood, 0 for paired columns), pairedll
 * (the pair's mixture log-likelihood at its opener, -inf elsewhere).  Optional:
 * siteconds (K4 x L: log w_c + log L_c) and paircomps (K16 x L: each pair
 * component's log-likelihood at the openers, -inf elsewhere).  The reference's
 * rng becomes seed: the draw of column x uses a uniform keyed by (seed, x), as
 * messi_sample_structures keys its intervals.  Every pair must be in the
 * dataset's pair list (distance bound and mask); 1 <= cats <= 8. */
int messi_sample_states(messi_dataset *ds, int K4, const double *P4, const double *pi4, const double *logw4, int K16,
                        const double *P16, const double *pi16, const double *logw16, int cats, const int *paired,
                        double B, uint64_t seed, int prec, int *states, double *unpaired, double *pairedll,
                        double *siteconds, double *paircomps);

/* Posterior weights of the site-rate categories of every column:
 * felsensteinposteriorsiterateconditionalscuda (src/FelsensteinCUDA.jl:518-717).
 * out (cats x L, column-major: out[c + cats*x]) receives, for category c and
 * column x, the log-sum-exp of log w_c + log unpaired[x] + log L_c(x) (the site
 * likelihood of category c) and, for every pair (x, y) of the dataset's pair list
 * and every pair component k = (c1, c2, lambda class), log pairprob[min + L*max]
 * + logw16[k] + log L_k(x, y), counted for c1 at the pair's first column and for
 * c2 at its second.  unpaired (L) and pairprob (L x L, column-major) are the
 * posteriors messi_bpp returns; the components are ordered as for
 * messi_sample_states (K4 == cats, K16 == cats^2 * lambda classes).  The reference
 * computes this with FP32 pair likelihoods; both precisions are offered here.
 * Its only caller in the reference sits in a disabled branch of processmax
 * (src/MESSI.jl:1066-1068). */
int messi_site_rate_posteriors(messi_dataset *ds, int K4, const double *P4, const double *pi4, const double *logw4,
                               int K16, const double *P16, const double *pi16, const double *logw16, int cats,
                               const double *unpaired, const double *pairprob, int prec, double *out);

/* ---- Input formats (host only, no device context) ----
 * Alignment handle from a FASTA file and the first line of a Newick file
 * (getalignmentfromfastaandnewick, src/AlignmentIO.jl:18-48; the Dataset
 * constructor, src/Dataset.jl:126-199): state masks, gap frequency and observed
 * base frequencies (case-sensitive ACGTU, src/Dataset.jl:135-152), the tree
 * parsed (src/Tree.jl:114-171), rooted at its first child (roottree,
 * src/Tree.jl:45-55) and numbered in pre-order (annotatetree, src/Tree.jl:190-199).
 * Failures return non-zero with the message in messi_io_error(). */
typedef struct messi_alignment messi_alignment;
const char *messi_io_error(void);
int messi_alignment_read(const char *fasta_path, const char *newick_path, messi_alignment **out);
void messi_alignment_destroy(messi_alignment *a);
int messi_alignment_info(const messi_alignment *a, int *nseq, int *ncol, int *nnodes);
/* data: (nseq, ncol, 4) 0/1, column-major; codes: nseq x ncol state masks,
 * row-major; gapfrequency: ncol; obsfreqs: 4.  Any may be NULL. */
int messi_alignment_data(const messi_alignment *a, double *data, uint8_t *codes, double *gapfrequency,
                         double *obsfreqs);
/* children (2 per node, -1 for leaves), seqindex (alignment row of a leaf, -1
 * for internal nodes), branch lengths (root 0), pre-order with the root at 0. */
int messi_alignment_tree(const messi_alignment *a, int *children, int *seqindex, double *branchlengths);
/* FASTA name of alignment row `row` (NUL-terminated, truncated to cap; 3 if truncated). */
int messi_alignment_name(const messi_alignment *a, int row, char *buf, int cap);
/* The dataset of an alignment handle, as messi_dataset_create builds it. */
int messi_dataset_from_alignment(messi_ctx *ctx, const messi_alignment *a, int maxbasepairdistance,
                                 const uint8_t *mask, messi_dataset **out);
/* readctfile (src/RNATools.jl:165-193): *L receives the length from the header;
 * with paired non-NULL (cap >= L) the 1-based partners, and with sequence
 * non-NULL (cap + 1 bytes) the sequence column. */
int messi_read_ct(const char *path, int cap, int *L, int *paired, char *sequence);
/* parsedbn (src/RNATools.jl:109-140): brackets ( < A B C D E close with ) > a b c d e;
 * a closing bracket pairs only with the innermost open bracket of its kind. */
int messi_parse_dbn(const char *dbn, int cap, int *L, int *paired);

/* Batch of members, each a dataset (the same dataset may appear several times,
 * e.g. one member per MCMC chain or proposal).  messi_batch_eval evaluates all
 * members together: each tree level and each tile-diagonal of the inside
 * schedule is one launch for every member.  Inputs are member-major: member m's
 * transition matrices are K x nnodes_m x dim x dim (column-major per matrix) and
 * follow member m-1's; pi is n x K x dim, logw is n x K; logZ receives n values.
 * No counterpart in the reference (it evaluates one parameter vector per call,
 * src/FARCE.jl:2322-2345); the per-member result equals messi_eval's. */
int messi_batch_create(messi_ctx *ctx, int n, messi_dataset *const *members, messi_batch **out);
void messi_batch_destroy(messi_batch *b);
int messi_batch_eval(messi_batch *b, int K4, const double *P4, const double *pi4, const double *logw4,
                     int K16, const double *P16, const double *pi16, const double *logw16, double B,
                     int prec, double *logZ);
int messi_batch_timings(const messi_batch *b, double *t);
/* Device time of the inside step kernels alone in the last evaluation made with
 * profiling on (messi_set_profiling): ms[0], launches[0] for the tile-16
 * members, ms[1], launches[1] for the tile-32 ones (0 where none), measured with
 * HIP events on the context's stream around the launches.  No counterpart in
 * the reference. */
int messi_batch_step_timings(messi_batch *b, double *ms, int *launches);
/* messi_batch_eval in two steps: messi_batch_load copies the mixture
 * components (same arguments and layouts as messi_batch_eval) into device
 * memory, converted to the precision prec; messi_batch_eval_loaded evaluates
 * every member from them, as often as wanted, with no host-to-device transfer
 * besides its launch parameters.  An evaluation in the other precision needs a
 * new load.  No counterpart in the reference. */
int messi_batch_load(messi_batch *b, int K4, const double *P4, const double *pi4, const double *logw4,
                     int K16, const double *P16, const double *pi16, const double *logw16, int prec);
int messi_batch_eval_loaded(messi_batch *b, double B, int prec, double *logZ);
/* messi_batch_load with the transition matrices built on the device from
 * eigendecompositions of the rate matrices, as gettransitionmatriceslist does on the
 * host for every evaluation (src/FARCE.jl:410-433, absmat 389-408): for node v >= 1
 * of member m with branch length t, P = V diag(exp(w t)) Vi, entries <= 0 set to
 * 1e-50, the identity when t == 0.  Member-major inputs: member m's V and Vi are
 * K x dim x dim (column-major per matrix: V[a + dim*c] is entry (a, c), the columns
 * of V the eigenvectors, Vi their inverse), w is K x dim; branchlengths holds each
 * member's nnodes_m values (pre-order, root first), members concatenated.  pi and
 * logw as for messi_batch_eval.  The result is what messi_batch_load of those
 * matrices would hold, for messi_batch_eval_loaded. */
int messi_batch_load_eigen(messi_batch *b, int K4, const double *V4, const double *w4, const double *Vi4,
                           const double *pi4, const double *logw4, int K16, const double *V16, const double *w16,
                           const double *Vi16, const double *pi16, const double *logw16,
                           const double *branchlengths, int prec);
/* Inside tile edge per member (n values, each 16, 32, or 0 for the context's
 * setting; NULL resets all to 0).  Members of one edge share one launch
 * sequence.  Set by the autotuner (messi_amd/autotune.py).  No counterpart in
 * the reference. */
int messi_batch_set_tiles(messi_batch *b, const int *tiles);
/* Member groups of an evaluation (1 .. 4; 0, the default, picks 2 groups for 4
 * or more members, else 1).  Each group runs its pair phase and then its inside
 * on a stream of its own, the pair phase of a group after that of the group
 * before, so the bandwidth-bound pruning of one group overlaps the
 * latency-bound inside launches of the others.  Results do not depend on the
 * grouping.  With profiling on, an evaluation runs as one group.  No
 * counterpart in the reference. */
int messi_batch_set_groups(messi_batch *b, int groups);
/* Bulk window of this batch's inside schedule (1, 2 or 4; 0, the default, takes
 * the context's, messi_set_bulk_window).  Results do not depend on it.  Set by
 * the autotuner.  No counterpart in the reference. */
int messi_batch_set_window(messi_batch *b, int window);

/* Learned MCMC proposal network: din inputs -> tanh(dh) -> tanh(dh) -> 2*din
 * outputs (mean shift mu, then log-scale ell, of a diagonal Gaussian proposal
 * z' ~ N(z + mu(z), diag exp(2 ell(z)))).  Products are bf16 MFMA with FP32
 * accumulation; weights and Adam state are FP32.  Training minimises the mean
 * negative log density  sum_d ell_d + ((t_d - z_d - mu_d) exp(-ell_d))^2 / 2
 * over rows (z, t).  Row-major host arrays: x and t are n x din, out n x 2*din.
 * The flat parameter vector (messi_mlp_num_params floats) holds W1 (D0 x dh),
 * b1, W2 (dh x dh), b2, W3 (dh x D3), b3 with D0 = din and D3 = 2*din rounded
 * up to multiples of 16 (padding entries are zero and never updated).
 * No counterpart in the reference: its MCMC proposals are random walks and
 * adaptive Metropolis (src/MCMC.jl:393-449, 1306-1633); this network replaces
 * those moves with an MH-corrected learned proposal. */
int messi_mlp_create(messi_ctx *ctx, int din, int dh, uint64_t seed, messi_mlp **out);
void messi_mlp_destroy(messi_mlp *m);
int messi_mlp_num_params(const messi_mlp *m, int64_t *n);
int messi_mlp_get_params(messi_mlp *m, float *out);
/* Replaces the weights and restarts the optimiser (Adam moments and step count zeroed). */
int messi_mlp_set_params(messi_mlp *m, const float *in);
int messi_mlp_forward(messi_mlp *m, int n, const float *x, float *out);
/* Loss and gradient (flat, same layout as the parameters; grad may be NULL). */
int messi_mlp_loss_grad(messi_mlp *m, int n, const float *x, const float *t, float *loss, float *grad);
/* One AdamW step (beta1 0.9, beta2 0.999, eps 1e-8, decoupled weight decay
 * w -= lr * weight_decay * w) on the loss over the n rows. */
int messi_mlp_train_step(messi_mlp *m, int n, const float *x, const float *t, float lr, float weight_decay,
                         float *loss);

/* Device time in milliseconds of the phases of the last messi_eval /
 * messi_posterior call on this dataset: t[0] site, t[1] pair, t[2] inside,
 * t[3] outside, t[4] bpp (HIP events on the context's stream). */
int messi_last_timings(const messi_dataset *ds, double *t);

#ifdef __cplusplus
}
#endif

#endif
