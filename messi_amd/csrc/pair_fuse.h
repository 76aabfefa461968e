// Which internal nodes of the 16-state pair pruning are computed in their
// parent's level launch instead of stored, shared by the dataset build (api.hip)
// and a host-side checker (tests/pair_fuse_check.cpp).  Plain C++.
//
// A non-root internal node c with parent p can skip its own launch: p's launch
// then gathers c's two children through composed maps and runs c's products and
// scaling in registers for every entry of p.  Counting the bytes of an FP32
// launch (64 per stored entry, 2 per leaf code, 4 per index), storing c costs
// its store, its own children's reads and indices, and one read of c per entry
// of p; computing it costs the reads and indices of c's children once per entry
// of p.  The computed form also repeats c's products and scaling for every entry
// of p that uses the same entry of c, which the byte count leaves out, so c is a
// candidate only when p has fewer than kMaxReuse entries per entry of c (at
// 50 x 1500 a bound of 1.4 beat no bound by 1.5 % per FP32 batch-16 step).
//
// The chosen set maximises the total saving under two rules: no computed node
// has a computed parent or child (the computing launch gathers stored entries
// only), and no node has two computed children (one kernel form).
#pragma once

#include <algorithm>
#include <cstdint>
#include <vector>

namespace messi {
namespace fuse {

constexpr double kMaxReuse = 1.4;

// Bytes of one child entry read: a stored entry or a leaf's two codes.
inline double read_bytes(bool leaf) { return leaf ? 2.0 : 64.0; }

// Saving of computing a node with Uc entries under a parent with Up entries,
// whose children read ca and cb bytes per entry; negative or zero when the node
// is not a candidate (it then stays stored).
inline double saving(int64_t Uc, int64_t Up, double ca, double cb) {
  if (!((double)Up < kMaxReuse * (double)Uc)) return 0.0;
  const double keep = (double)Uc * (64.0 + ca + cb + 8.0) + (double)Up * (64.0 + 4.0);
  const double comp = (double)Up * (ca + cb + 8.0);
  return keep - comp;
}

// ch: children of node v at ch[2v], ch[2v+1] (-1 for a leaf), every child after
// its parent, root 0; U: entries per node.  Returns 1 for the computed nodes.
inline std::vector<char> pick(int n, const int *ch, const int64_t *U) {
  std::vector<char> virt(n, 0);
  auto leaf = [&](int v) { return ch[2 * v] < 0; };
  std::vector<int> parent(n, -1);
  for (int v = 0; v < n; ++v)
    if (!leaf(v)) parent[ch[2 * v]] = parent[ch[2 * v + 1]] = v;
  // best saving below v with v stored (fs) and with v computed (fv, -1: not allowed)
  std::vector<double> fs(n, 0.0), fv(n, -1.0);
  for (int v = n - 1; v >= 0; --v) {
    if (leaf(v)) continue;
    const int l = ch[2 * v], r = ch[2 * v + 1];
    double base = 0.0, best = 0.0;
    for (int c : {l, r})
      if (!leaf(c)) {
        base += fs[c];
        if (fv[c] >= 0.0) best = std::max(best, fv[c] - fs[c]);
      }
    fs[v] = base + best;
    if (v != 0) {
      const double g = saving(U[v], U[parent[v]], read_bytes(leaf(l)), read_bytes(leaf(r)));
      if (g > 0.0) fv[v] = g + base;  // its children stay stored
    }
  }
  for (int v = 0; v < n; ++v) {  // pre-order: a node's status is known before its children's
    if (leaf(v) || virt[v]) continue;
    int pickc = -1;
    double best = 0.0;
    for (int c : {ch[2 * v], ch[2 * v + 1]})
      if (!leaf(c) && fv[c] >= 0.0 && fv[c] - fs[c] > best) {
        best = fv[c] - fs[c];
        pickc = c;
      }
    if (pickc >= 0) virt[pickc] = 1;
  }
  return virt;
}

}  // namespace fuse
}  // namespace messi
