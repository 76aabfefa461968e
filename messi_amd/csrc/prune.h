// Declarations shared by the pruning kernels and the dataset/ABI layer.
#pragma once

#include "common.h"

namespace messi {

#define PAIR_NB 4   // 16-entry blocks per wavefront item of the pair kernel
#define PAIR_VNB 2  // the same for nodes whose first child is computed in registers

// Per internal node descriptor of a level launch (one batch member's node).
// Offsets are in entries of the batch arenas.
struct NodeDesc {
  int64_t U;                // number of entries of this node
  int64_t idx_l, idx_r;     // offsets of the entry -> child-entry maps in the index arena
  int64_t part;             // offset of this node's entries in the partial arena
  int64_t cpart_l, cpart_r; // child's partial offset, or leaf-code offset for leaf children
  int64_t out;              // root only: offset of the member's root entries in the root arrays
  int leaf_l, leaf_r;       // child is a leaf
  int node_l, node_r;       // global transition-matrix index of each child (member base + node)
  int member;               // batch member
  int root;                 // this node is its tree's root
  // A node whose first child c is not stored but computed in registers from c's
  // own two children (a, b) for each of the node's entries: idx_l, cpart_l and
  // leaf_l then describe a, idx_v2, cpart_v2 and leaf_v2 describe b (the maps
  // take the node's entries straight to a's and b's entries), node_l is c's
  // transition matrix and node_va, node_vb are a's and b's.
  int64_t idx_v2 = 0, cpart_v2 = 0;
  int leaf_v2 = 0, node_va = 0, node_vb = 0;
  int virt = 0;             // the first child is computed in registers
};

// Several mixture components in one launch (grid row c = component c): each
// component's pointers advance by these strides (in elements of their arrays);
// the default is one component.
struct CompBatch {
  int K = 1;
  int64_t part = 0, expo = 0, root = 0, P = 0, pi = 0, logw = 0;
};

// P / pi / logw point at one mixture component for all members: the matrix of global
// node g is P + 256 g (16 g for 4 states), member m uses pi + 16 m (4 m) and logw[m].
template <typename T, int NB, bool VL = false>
void launch_pair_level(hipStream_t s, const int2 *items, int nitems, const NodeDesc *nodes,
                       const int *idx, const uint16_t *leafc, T *part, const T *P,
                       const double *pi16, const double *logw, int first, double *rootacc,
                       double *rootcomp, const CompBatch &cb = CompBatch());
template <typename T>
void launch_site_level(hipStream_t s, const int2 *items, int nitems, const NodeDesc *nodes,
                       const int *idx, const uint8_t *leafc, T *part, int *expo, const T *P4,
                       const double *pi4, const double *logw, int first, double *rootacc,
                       const CompBatch &cb = CompBatch());

// Expansion of root entries to each member's outputs.
struct ExpandDesc {
  int L, maxbp, n0;
  const int *pid0;
  const int *T0;
  const uint8_t *mask;
  int64_t rootoff;   // offset of the member's root entries
  double *u;         // site output (L)
  double *paired;    // pair output (L x L)
};
void launch_site_expand(hipStream_t s, int n, int Lmax, const ExpandDesc *d, const double *acc);
void launch_pair_expand(hipStream_t s, int n, int Lmax, const ExpandDesc *d, const double *acc,
                        double *const *outs);
// Fixed-structure sums of one member (descriptor d, pair root values acc):
// out[0] unpaired site log-likelihoods, out[1] pair log-likelihoods, out[2] missing pairs.
void launch_structure_sum(hipStream_t s, const ExpandDesc *d, const double *acc, const int *partner, double *out);
// State sampling given a structure (one member, descriptors ds / dp of its site and
// pair root entries, partner 1-based):
// out[x] = site root value of column x (one site component's pass);
// (K components: acc + c * root_stride into out + c * L.)
void launch_site_gather(hipStream_t s, int L, int K, int64_t root_stride, const ExpandDesc *ds, const double *acc,
                        double *out);
// out[x] = pair root value of the pair opened at x (-inf elsewhere); pairs the
// member's pair list lacks are counted in *missing (not cleared here).
// Posterior weights of the site-rate categories of every column
// (felsensteinposteriorsiterateconditionalscuda, src/FelsensteinCUDA.jl:518-717):
// out[c + cats * x] = log-sum-exp of the column's site term sitec[c L + x] +
// log single[x] and, over every pair (x, y) of the pair list and every pair
// component k = (c1 * cats + c2) * M + m, log pair[min + L max] + logw16[k] +
// L_k(x, y), counted for category c1 at the pair's first column and c2 at its
// second.  paircomp holds each component's root entries (stride root_stride).
void launch_site_rate_posterior(hipStream_t s, int L, int cats, int K16, int64_t root_stride,
                                const ExpandDesc *dp, const double *paircomp, const double *logw16,
                                const double *sitec, const double *single, const double *pairprob, double *out);
void launch_pair_gather(hipStream_t s, int L, int K, int64_t root_stride, const ExpandDesc *dp, const double *acc,
                        const int *partner, double *out, int *missing);
// One categorical draw per unpaired column and per base pair (keyed uniforms, as
// the structure sampler): sitec (K4 x L) the site conditionals log w_c + ll_c,
// pairc (K16 x L) the pair component log-likelihoods at the openers, components
// ordered (category 1, category 2, M lambda classes), logw16 their log weights.
void launch_state_sample(hipStream_t s, int L, int cats, int M, const double *sitec, const double *pairc,
                         const double *logw16, const int *partner, double B, uint64_t seed, int *states,
                         double *unpaired, double *pairedll);
template <typename T> bool mfma_probe(hipStream_t s);
bool mfma_bf16_probe(hipStream_t s);

template <typename T> void inside_alloc(InsideWork &w, int L, bool outside);
template <typename T> void inside_prepare(hipStream_t s, InsideWork &w, int L, double B);
template <typename T> void inside_run(hipStream_t s, InsideWork &w, int L, double B);
// With rootacc the pair term of each member is gathered from its pair root
// entries (its work's exp descriptor) instead of being read from its L x L pair
// matrix.  All members take the tile edge of ws[0].  steps_begin / steps_end,
// when given, are recorded around the step kernels; returns their launch count.
template <typename T>
int inside_batch_run(hipStream_t s, InsideWork *const *ws, int n, InsideBatch &ib, double B,
                     const double *rootacc = nullptr, hipEvent_t steps_begin = nullptr,
                     hipEvent_t steps_end = nullptr);
template <typename T> double inside_logZ(hipStream_t s, InsideWork &w, int L);
template <typename T> void inside_export(hipStream_t s, InsideWork &w, int L, double *d_in3);
template <typename T> void inside_import(hipStream_t s, InsideWork &w, int L, const double *d_in3);
// full_L: also the L outside table beyond the pairing band (the bpp needs it only within)
template <typename T> void outside_run(hipStream_t s, InsideWork &w, int L, bool full_L);
template <typename T> void outside_export(hipStream_t s, InsideWork &w, int L, double *d_out3);
template <typename T> void outside_import(hipStream_t s, InsideWork &w, int L, const double *d_out3);
template <typename T> void bpp_run(hipStream_t s, InsideWork &w, int L, double *d_unpaired, double *d_pair);
// MEA decoding of (single, pair) on the device; d_paired (optional) receives the
// traceback (1-based partners, 0 unpaired).  V = double: FP64 semantics, float: FP32.
template <typename V>
void mea_run(hipStream_t s, MeaWork &w, int L, const double *d_single, const double *d_pair, double alpha,
             int *d_paired);
// Traceback of the last mea_run's S table into d_paired.
void mea_run_traceback(hipStream_t s, MeaWork &w, int L, int *d_paired);
// Stochastic tracebacks (one wavefront each) of samples first .. first+n-1 from w's
// inside tables: d_out (n x L) receives 1-based partners, d_prior (n) the KH99 log priors.
constexpr int SAMPLE_MAX_L = 16383;
template <typename T>
void sample_run(hipStream_t s, InsideWork &w, SampleWork &sw, int L, int n, uint64_t seed, int first, double B,
                int *d_out, double *d_prior);
// E (as double) and S of the last mea_run, column-major, zero on and below the diagonal.
void mea_export(hipStream_t s, MeaWork &w, int L, double *d_E, int *d_S);

}  // namespace messi
