// Host check of the choice of pair nodes computed in their parent's launch
// (messi_amd/csrc/pair_fuse.h, the function the dataset build calls).  On random
// binary trees in pre-order with random entry counts it checks that the choice
//   * holds only candidates (internal, not the root, positive saving);
//   * has no computed node with a computed parent or child;
//   * gives no node two computed children;
//   * reaches the largest total saving over every set that obeys these rules,
//     found by enumerating all subsets on trees of up to 14 candidates.
// Prints "ok" and exits 0, or prints the first violation and exits 1.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

#include "../messi_amd/csrc/pair_fuse.h"

using namespace messi;

struct Tree {
  std::vector<int> ch;
  std::vector<int64_t> U;
  int n = 0;
};

// A random tree with `leaves` leaves, nodes numbered in pre-order; entry counts
// grow towards the root by random factors, so parent/child ratios straddle the
// reuse bound.
static int build(Tree &t, int leaves, std::mt19937_64 &rng) {
  const int v = t.n++;
  t.ch.push_back(-1);
  t.ch.push_back(-1);
  t.U.push_back(0);
  if (leaves == 1) {
    t.U[v] = 1 + (int64_t)(rng() % 64);
    return v;
  }
  const int k = 1 + (int)(rng() % (uint64_t)(leaves - 1));
  const int l = build(t, k, rng);
  const int r = build(t, leaves - k, rng);
  t.ch[2 * v] = l;
  t.ch[2 * v + 1] = r;
  std::uniform_real_distribution<double> f(0.9, 2.2);
  const double base = (double)std::max(t.U[l], t.U[r]);
  t.U[v] = (int64_t)std::llround(base * f(rng)) + 1;
  return v;
}

static bool is_leaf(const Tree &t, int v) { return t.ch[2 * v] < 0; }

static double node_saving(const Tree &t, const std::vector<int> &parent, int c) {
  const int a = t.ch[2 * c], b = t.ch[2 * c + 1];
  return fuse::saving(t.U[c], t.U[parent[c]], fuse::read_bytes(is_leaf(t, a)), fuse::read_bytes(is_leaf(t, b)));
}

int main() {
  std::mt19937_64 rng(12345);
  int trees = 0, enumerated = 0;
  for (int iter = 0; iter < 3000; ++iter) {
    Tree t;
    const int leaves = 2 + (int)(rng() % (iter < 2000 ? 16 : 200));
    build(t, leaves, rng);
    const int n = t.n;
    std::vector<int> parent(n, -1);
    for (int v = 0; v < n; ++v)
      if (!is_leaf(t, v)) parent[t.ch[2 * v]] = parent[t.ch[2 * v + 1]] = v;
    const std::vector<char> virt = fuse::pick(n, t.ch.data(), t.U.data());
    double got = 0.0;
    for (int v = 0; v < n; ++v) {
      if (!virt[v]) continue;
      if (v == 0 || is_leaf(t, v)) return std::printf("violation: root or leaf chosen (tree %d)\n", iter), 1;
      const double g = node_saving(t, parent, v);
      if (!(g > 0.0)) return std::printf("violation: non-candidate chosen (tree %d, node %d)\n", iter, v), 1;
      if (virt[parent[v]]) return std::printf("violation: computed parent (tree %d, node %d)\n", iter, v), 1;
      got += g;
    }
    for (int v = 0; v < n; ++v)
      if (!is_leaf(t, v) && virt[t.ch[2 * v]] && virt[t.ch[2 * v + 1]])
        return std::printf("violation: two computed children (tree %d, node %d)\n", iter, v), 1;
    ++trees;
    // exhaustive optimum over the candidates
    std::vector<int> cand;
    for (int v = 1; v < n; ++v)
      if (!is_leaf(t, v) && node_saving(t, parent, v) > 0.0) cand.push_back(v);
    if (cand.size() > 14) continue;
    ++enumerated;
    double best = 0.0;
    for (uint32_t mask = 0; mask < (1u << cand.size()); ++mask) {
      std::vector<char> in(n, 0);
      double sum = 0.0;
      for (size_t i = 0; i < cand.size(); ++i)
        if (mask >> i & 1u) {
          in[cand[i]] = 1;
          sum += node_saving(t, parent, cand[i]);
        }
      bool okset = true;
      for (int v = 0; v < n && okset; ++v) {
        if (in[v] && parent[v] >= 0 && in[parent[v]]) okset = false;
        if (!is_leaf(t, v) && in[t.ch[2 * v]] && in[t.ch[2 * v + 1]]) okset = false;
      }
      if (okset) best = std::max(best, sum);
    }
    if (std::fabs(best - got) > 1e-6 * std::max(1.0, best))
      return std::printf("violation: saving %.3f below the optimum %.3f (tree %d)\n", got, best, iter), 1;
  }
  std::printf("ok %d trees (%d enumerated)\n", trees, enumerated);
  return 0;
}
