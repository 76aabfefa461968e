// Work schedule of the tiled inside pass, shared by the kernels (inside.hip), the
// host code that sizes their launches and a host-side checker
// (tests/inside_sched_check.cpp).  Plain C++: constexpr functions, which HIP
// compiles for host and device alike.
//
// Launch t finalises the tiles (I, I+t) and runs the bulk visits of launch t: a
// tile at distance d is visited in the launches t = d-1, d-1-W, ... (those with
// 2t - 2 >= d), each visit adding the split-point products that became ready
// since the previous one.  Bulk items of launch t are the tiles at distances
// t+1, t+1+W, ... (NT - d of them per distance d), numbered by distance first.
//
// The single-launch (persistent) form runs the same items from one ticket
// counter, in phases: phase p holds the finalisations of launch p, then the
// bulk items of launch p at distance p+1 (the "critical" ones: launch p+1
// finalises those tiles), then the other bulk items of launch p-1 (with a
// window W >= 2 their tiles are next needed in launch p-1+W >= p+1).  Every
// item depends only on items of earlier tickets, so the oldest unfinished item
// can always proceed whatever the number of resident workgroups.
#pragma once

namespace messi {
namespace sched {

// Bulk items per bulk workgroup (their loads overlap): two with a window of one
// launch, one with a wider window.
constexpr int BULK_PER = 2;
constexpr int bulk_per(int W) { return W == 1 ? BULK_PER : 1; }

// Bulk items of the first k distances of launch t.
constexpr int bulk_items(int NT, int t, int W, int k) { return k * (NT - t - 1) - W * (k * (k - 1) / 2); }

// Number of bulk distances of launch t (tile offset band KB: beyond distance
// t-1+KB only a product with s-1 <= KB could be in the band, and such a product
// is ready only for distances <= 2KB < t-1+KB).
constexpr int bulk_dists(int NT, int t, int W, int KB) {
  int dmax = 2 * t - 2 < NT - 1 ? 2 * t - 2 : NT - 1;
  if (t - 1 > KB && t - 1 + KB < dmax) dmax = t - 1 + KB;
  return dmax >= t + 1 ? (dmax - t - 1) / W + 1 : 0;
}

constexpr int bulk_count(int NT, int t, int W, int KB) { return bulk_items(NT, t, W, bulk_dists(NT, t, W, KB)); }

// Tile (I, J) of bulk item idx < bulk_count of launch t (nd = bulk_dists).
constexpr void bulk_tile(int NT, int t, int W, int nd, int idx, int &I, int &J) {
  int lo = 0, hi = nd - 1;  // the distance: bulk_items(lo) <= idx < bulk_items(lo + 1)
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (bulk_items(NT, t, W, mid) <= idx) lo = mid;
    else hi = mid - 1;
  }
  I = idx - bulk_items(NT, t, W, lo);
  J = I + t + 1 + lo * W;
}

// The latest bulk visit of a tile at distance d before launch t (fin: before
// its finalisation in launch d), or -1.  The visits are t' = last, last-W, ...
// >= (d+3)/2, each scheduled unless t'-1 > KB and d > t'-1+KB.  Every
// scheduled visit writes the tile's accumulator, so the accumulator is read
// only when there was one and needs no clearing.
constexpr int last_visit(int t, int d, int W, int KB, bool fin) {
  const int smin = (d + 3) / 2, last = fin ? d - 1 : t - W;
  if (last < smin) return -1;
  if (last - 1 <= KB || d <= last - 1 + KB) return last;
  // only visits with t' - 1 <= KB are scheduled below this one
  const int tp = last - W * ((last - (KB + 1) + W - 1) / W);
  return tp >= smin ? tp : -1;
}

constexpr bool had_visit(int t, int d, int W, int KB, bool fin) { return last_visit(t, d, W, KB, fin) >= 0; }

// Bulk workgroups of launch t (items grouped bulk_per(W) at a time).
constexpr int bulk_groups(int NT, int t, int W, int KB) {
  return (bulk_count(NT, t, W, KB) + bulk_per(W) - 1) / bulk_per(W);
}

// Critical bulk workgroups of launch t (its first ones): those at distance t+1
// with a window W >= 2, all of them with W = 1.
constexpr int crit_groups(int NT, int t, int W, int KB) {
  return W == 1 ? bulk_groups(NT, t, W, KB) : (bulk_dists(NT, t, W, KB) >= 1 ? NT - t - 1 : 0);
}

// ---------------------------------------------------------------------------
// Ticket table of the persistent form, for n instances of NT[m] tiles per side
// and band offsets KB[m]:
//   tab[0 .. NP]           first ticket of each phase (NP = max NT + 1 phases)
//   tab[NP+1 + p*3*(n+1)]  per phase p, three prefix arrays of n+1 entries:
//                          finalisations of launch p, critical bulk groups of
//                          launch p, other bulk groups of launch p-1.
// Returns the number of ints (table_size).
constexpr int table_size(int n, int NP) { return NP + 1 + NP * 3 * (n + 1); }

inline int build_table(int n, const int *NT, const int *KB, int W, int NP, int *tab) {
  int tk = 0;
  for (int p = 0; p < NP; ++p) {
    tab[p] = tk;
    int *seg = tab + NP + 1 + p * 3 * (n + 1);
    for (int s = 0; s < 3; ++s) {
      int *pre = seg + s * (n + 1);
      int c = 0;
      for (int m = 0; m < n; ++m) {
        pre[m] = c;
        const int t = s == 2 ? p - 1 : p;
        if (t < 0 || t >= NT[m]) continue;
        if (s == 0) c += NT[m] - t;
        else if (s == 1) c += crit_groups(NT[m], t, W, KB[m]);
        else c += bulk_groups(NT[m], t, W, KB[m]) - crit_groups(NT[m], t, W, KB[m]);
      }
      pre[n] = c;
      tk += c;
    }
  }
  tab[NP] = tk;
  return table_size(n, NP);
}

// Work item of ticket k < tab[NP]: kind 0 finalisation of launch t (item = I),
// kind 1 critical bulk workgroup of launch t (item = its index among the
// launch's bulk workgroups of instance m), kind 2 other bulk workgroup of
// launch t (item = its index past the launch's critical groups: add
// crit_groups for the index among all of them).
struct Ticket {
  int kind, m, t, item;
};

constexpr Ticket decode_ticket(const int *tab, int n, int NP, int k) {
  int lo = 0, hi = NP - 1;  // phase: tab[lo] <= k < tab[lo + 1]
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tab[mid] <= k) lo = mid;
    else hi = mid - 1;
  }
  const int p = lo;
  const int *seg = tab + NP + 1 + p * 3 * (n + 1);
  int x = k - tab[p], s = 0;
  while (s < 2 && x >= seg[s * (n + 1) + n]) {
    x -= seg[s * (n + 1) + n];
    ++s;
  }
  const int *pre = seg + s * (n + 1);
  int a = 0, b = n - 1;  // instance: pre[a] <= x < pre[a + 1]
  while (a < b) {
    const int mid = (a + b + 1) >> 1;
    if (pre[mid] <= x) a = mid;
    else b = mid - 1;
  }
  return Ticket{s, a, s == 2 ? p - 1 : p, x - pre[a]};
}

}  // namespace sched
}  // namespace messi
