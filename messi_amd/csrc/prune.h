// Declarations shared by the pruning kernels and the dataset/ABI layer.
#pragma once

#include "common.h"

namespace messi {

#define PAIR_NB 4  // 16-entry blocks per wavefront item of the pair kernel

// Per internal node descriptor of a level launch.  Offsets are in entries.
struct NodeDesc {
  int64_t U;                // number of entries of this node
  int64_t idx_l, idx_r;     // offsets of the entry -> child-entry maps in the index arena
  int64_t part;             // offset of this node's entries in the partial arena
  int64_t cpart_l, cpart_r; // child's partial offset, or leaf-code offset for leaf children
  int leaf_l, leaf_r;       // child is a leaf
  int node_l, node_r;       // child node ids (index of their transition matrix)
};

template <typename T, int NB>
void launch_pair_level(hipStream_t s, const int2 *items, int nitems, bool root, const NodeDesc *nodes,
                       const int *idx, const uint16_t *leafc, T *part, int *expo, const T *P,
                       const double *pi16, double logw, int first, double *rootacc,
                       double *rootcomp);
template <typename T>
void launch_site_level(hipStream_t s, const int2 *items, int nitems, bool root, const NodeDesc *nodes,
                       const int *idx, const uint8_t *leafc, T *part, int *expo, const T *P4,
                       const double *pi4, double logw, int first, double *rootacc);
void launch_site_expand(hipStream_t s, int L, const int *pid0, const double *acc, double *u);
void launch_pair_expand(hipStream_t s, int L, int maxbp, int n0, const int *pid0, const int *T0,
                        const uint8_t *mask, const double *acc, double *out);
template <typename T> bool mfma_probe(hipStream_t s);

template <typename T> void inside_alloc(InsideWork &w, int L, bool outside);
template <typename T> void inside_prepare(hipStream_t s, InsideWork &w, int L, double B);
template <typename T> void inside_run(hipStream_t s, InsideWork &w, int L, double B);
template <typename T> double inside_logZ(hipStream_t s, InsideWork &w, int L);
template <typename T> void inside_export(hipStream_t s, InsideWork &w, int L, double *d_in3);
template <typename T> void inside_import(hipStream_t s, InsideWork &w, int L, const double *d_in3);
template <typename T> void outside_run(hipStream_t s, InsideWork &w, int L);
template <typename T> void outside_export(hipStream_t s, InsideWork &w, int L, double *d_out3);
template <typename T> void outside_import(hipStream_t s, InsideWork &w, int L, const double *d_out3);
template <typename T> void bpp_run(hipStream_t s, InsideWork &w, int L, double *d_unpaired, double *d_pair);

}  // namespace messi
