// Host check of the single-launch (persistent) inside schedule
// (messi_amd/csrc/inside_sched.h, the functions the kernels use).  For batches
// of instances of assorted tile counts, bands and bulk windows it builds the
// ticket table, decodes every ticket and checks:
//   * every finalisation (launch t, tile (I, I+t)) and every bulk item of every
//     launch appears exactly once;
//   * a finalisation comes after both neighbouring finalisations it waits for
//     ((I, J-1) and (I+1, J)) and after the tile's latest bulk visit, which is
//     the one last_visit names (no visit when it names none);
//   * a bulk visit comes after every finalisation of launch t-1 (the diagonal
//     counter it waits for) and after the tile's previous visit, again the one
//     last_visit names;
//   * every visit of a tile comes before its finalisation.
// So each item waits only for items of smaller tickets: with tickets handed out
// in order, the oldest unfinished item can always run.  Prints "ok" and exits 0,
// or prints the first violation and exits 1.
#include <cstdio>
#include <map>
#include <tuple>
#include <vector>

#include "../messi_amd/csrc/inside_sched.h"

using namespace messi::sched;

static int fail(const char *what, int m, int t, int I, int J) {
  std::printf("violation: %s (instance %d, launch %d, tile %d,%d)\n", what, m, t, I, J);
  return 1;
}

static int check(const std::vector<int> &NT, const std::vector<int> &KB, int W) {
  const int n = (int)NT.size();
  int NP = 0;
  for (int x : NT) NP = x > NP ? x : NP;
  NP += 1;
  std::vector<int> tab(table_size(n, NP));
  build_table(n, NT.data(), KB.data(), W, NP, tab.data());
  const int total = tab[NP];
  const int BP = bulk_per(W);
  // ticket of each finalisation and of each visit, per instance
  std::vector<std::map<std::pair<int, int>, int>> fin(n);
  std::vector<std::map<std::tuple<int, int, int>, int>> vis(n);  // (I, J, t) -> ticket
  std::vector<std::map<std::pair<int, int>, std::vector<int>>> vlist(n);  // (I, J) -> launches
  for (int k = 0; k < total; ++k) {
    const Ticket tk = decode_ticket(tab.data(), n, NP, k);
    const int m = tk.m, t = tk.t;
    if (m < 0 || m >= n || t < 0 || t >= NT[m]) return fail("ticket decodes out of range", m, t, -1, -1);
    if (tk.kind == 0) {
      const int I = tk.item, J = I + t;
      if (I < 0 || J >= NT[m]) return fail("finalisation tile out of range", m, t, I, J);
      if (!fin[m].emplace(std::make_pair(I, J), k).second) return fail("finalisation twice", m, t, I, J);
    } else {
      const int crit = crit_groups(NT[m], t, W, KB[m]);
      const int g = tk.item + (tk.kind == 2 ? crit : 0);
      if (tk.kind == 1 && g >= crit) return fail("critical group out of range", m, t, g, -1);
      if (g >= bulk_groups(NT[m], t, W, KB[m])) return fail("bulk group out of range", m, t, g, -1);
      const int nd = bulk_dists(NT[m], t, W, KB[m]), cnt = bulk_count(NT[m], t, W, KB[m]);
      for (int u = 0; u < BP; ++u) {
        const int idx = BP * g + u;
        if (idx >= cnt) continue;
        int I = 0, J = 0;
        bulk_tile(NT[m], t, W, nd, idx, I, J);
        if (I < 0 || J >= NT[m] || J - I <= t) return fail("bulk tile out of range", m, t, I, J);
        if (tk.kind == 1 && W > 1 && J - I != t + 1) return fail("critical item not at distance t+1", m, t, I, J);
        if (tk.kind == 2 && J - I == t + 1) return fail("distance t+1 item deferred", m, t, I, J);
        if (!vis[m].emplace(std::make_tuple(I, J, t), k).second) return fail("visit twice", m, t, I, J);
        vlist[m][std::make_pair(I, J)].push_back(t);
      }
    }
  }
  long nfin = 0, nvis = 0;
  for (int m = 0; m < n; ++m) {
    for (int t = 0; t < NT[m]; ++t) {
      nfin += NT[m] - t;
      nvis += bulk_count(NT[m], t, W, KB[m]);
    }
  }
  long gfin = 0, gvis = 0;
  for (int m = 0; m < n; ++m) {
    gfin += (long)fin[m].size();
    gvis += (long)vis[m].size();
  }
  if (gfin != nfin || gvis != nvis) {
    std::printf("violation: %ld of %ld finalisations, %ld of %ld visits\n", gfin, nfin, gvis, nvis);
    return 1;
  }
  for (int m = 0; m < n; ++m) {
    auto latest_before = [&](int I, int J, int t) {
      int best = -1;
      auto it = vlist[m].find(std::make_pair(I, J));
      if (it != vlist[m].end())
        for (int v : it->second)
          if (v < t && v > best) best = v;
      return best;
    };
    for (auto &f : fin[m]) {
      const int I = f.first.first, J = f.first.second, t = J - I, k = f.second;
      if (t >= 1) {
        if (fin[m].at(std::make_pair(I, J - 1)) >= k) return fail("left neighbour after", m, t, I, J);
        if (fin[m].at(std::make_pair(I + 1, J)) >= k) return fail("lower neighbour after", m, t, I, J);
      }
      const int lv = last_visit(t, t, W, KB[m], true);
      if (lv != latest_before(I, J, t + 1)) return fail("last_visit (finalisation) disagrees", m, t, I, J);
      if (lv >= 0 && vis[m].at(std::make_tuple(I, J, lv)) >= k) return fail("last visit after finalisation", m, t, I, J);
    }
    for (auto &v : vis[m]) {
      const int I = std::get<0>(v.first), J = std::get<1>(v.first), t = std::get<2>(v.first), k = v.second;
      for (int i = 0; i + t - 1 < NT[m]; ++i)
        if (t >= 1 && fin[m].at(std::make_pair(i, i + t - 1)) >= k)
          return fail("visit before the previous diagonal", m, t, I, J);
      const int pv = last_visit(t, J - I, W, KB[m], false);
      if (pv != latest_before(I, J, t)) return fail("last_visit (bulk) disagrees", m, t, I, J);
      if (pv >= 0 && vis[m].at(std::make_tuple(I, J, pv)) >= k) return fail("previous visit after", m, t, I, J);
      if (fin[m].at(std::make_pair(I, J)) <= k) return fail("visit after finalisation", m, t, I, J);
    }
  }
  return 0;
}

int main() {
  const std::vector<std::vector<int>> sizes = {{1}, {2}, {3}, {7}, {12}, {31}, {94}, {94, 94, 94},
                                               {5, 38, 12, 1, 20}, {17, 3, 40, 40, 9, 2, 25}};
  int cases = 0;
  for (int W : {1, 2, 4})
    for (const auto &NT : sizes)
      for (int band : {-1, 0, 1, 2, 3, 5, 8}) {
        std::vector<int> KB(NT.size());
        for (size_t m = 0; m < NT.size(); ++m) KB[m] = band < 0 ? NT[m] : (band < NT[m] ? band : NT[m]);
        if (check(NT, KB, W)) {
          std::printf("case W=%d band=%d n=%zu\n", W, band, NT.size());
          return 1;
        }
        ++cases;
      }
  std::printf("ok %d cases\n", cases);
  return 0;
}
