// C-ABI of libmessi: context, dataset (pattern dedup and pair index lists built
// natively at creation), batches of (dataset, parameter set) members, and the
// likelihood entry points.  Single-dataset calls run as a batch of one.
#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>

#include "common.h"
#include "prune.h"
#include "pair_fuse.h"

using namespace messi;

namespace {

// Open-addressing map int64 -> int for pattern and pair-pattern dedup.
struct HashI64 {
  std::vector<int64_t> k;
  std::vector<int> v;
  size_t mask;
  explicit HashI64(size_t n) {
    size_t c = 16;
    while (c < 2 * n + 16) c <<= 1;
    k.assign(c, -1);
    v.assign(c, 0);
    mask = c - 1;
  }
  static inline uint64_t mix(uint64_t x) {
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
  }
  int get_or_insert(int64_t key, int nv, bool &ins) {
    size_t h = mix((uint64_t)key) & mask;
    for (;;) {
      if (k[h] == -1) {
        k[h] = key;
        v[h] = nv;
        ins = true;
        return nv;
      }
      if (k[h] == key) {
        ins = false;
        return v[h];
      }
      h = (h + 1) & mask;
    }
  }
};

struct Events {
  hipEvent_t e[6] = {};
  bool made = false;
  void make() {
    if (made) return;
    for (auto &x : e) MESSI_HIP(hipEventCreate(&x));
    made = true;
  }
  ~Events() {
    if (made)
      for (auto &x : e) (void)hipEventDestroy(x);
  }
};

// Ordering events of a pipelined batch evaluation: per member group the end
// of its pair phase and its end.
struct PipeEvents {
  static constexpr int N = messi_ctx::NPIPE + 1;
  hipEvent_t pair[N] = {}, done[N] = {};
  bool made = false;
  void make() {
    if (made) return;
    for (int g = 0; g < N; ++g) {
      MESSI_HIP(hipEventCreateWithFlags(&pair[g], hipEventDisableTiming));
      MESSI_HIP(hipEventCreateWithFlags(&done[g], hipEventDisableTiming));
    }
    made = true;
  }
  ~PipeEvents() {
    if (!made) return;
    for (int g = 0; g < N; ++g) {
      (void)hipEventDestroy(pair[g]);
      (void)hipEventDestroy(done[g]);
    }
  }
};

}  // namespace

struct messi_dataset {
  messi_ctx *ctx = nullptr;
  int nseq = 0, L = 0, nn = 0, maxbp = 0, H = 0;
  std::vector<int> ch, seq, height, npat;
  std::vector<std::vector<int>> pid, rep;
  // Node descriptors with dataset-local offsets (member 0, local node ids) and the
  // level work items (index h-1 holds height h).
  std::vector<NodeDesc> site_nodes, pair_nodes;
  std::vector<std::vector<int2>> site_items, pair_items;
  std::vector<std::vector<int2>> pair_vitems;  // pair items of nodes whose first child is computed in registers
  int64_t site_entries = 0, pair_entries = 0, U_root = 0, sumU = 0, sumN = 0, inrange = 0;
  int64_t site_idx_n = 0, pair_idx_n = 0, site_leaf_n = 0, pair_leaf_n = 0;
  DevBuf d_site_idx, d_site_leaf, d_pair_idx, d_pair_leaf;
  DevBuf d_pid0, d_T0, d_mask;
  DevBuf d_partner, d_sums;     // fixed-structure likelihood: the structure and its sums
  DevBuf d_sitec, d_pairc, d_sout, d_miss;  // state sampling: conditionals, outputs, missing pairs
  bool has_mask = false;
  messi_batch *self = nullptr;  // batch of one, created on first use
  ~messi_dataset();
};

// Pinned host staging buffer (grow-only).
struct HostBuf {
  void *p = nullptr;
  size_t bytes = 0;
  HostBuf() = default;
  HostBuf(const HostBuf &) = delete;
  HostBuf &operator=(const HostBuf &) = delete;
  ~HostBuf() {
    if (p) (void)hipHostFree(p);
  }
  void ensure(size_t n) {
    if (n <= bytes) return;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
    MESSI_HIP(hipHostMalloc(&p, n < 256 ? 256 : n, hipHostMallocDefault));
    bytes = n < 256 ? 256 : n;
  }
  template <typename T> T *as() const { return static_cast<T *>(p); }
};

// Mixture components resident on the device, component-major: the matrix of
// component k for global node g is P + (k * nn_total + g) * dim2, member m's pi
// and log weight are pi + (k * n + m) * pidim and logw[k * n + m].
struct Components {
  DevBuf P, pi, logw;
  HostBuf stage;
  int K = 0;
  int prec = -1;  // element type of P (MESSI_FP64 / MESSI_FP32), -1 when nothing is loaded
};

struct messi_batch {
  messi_ctx *ctx = nullptr;
  int n = 0, Hmax = 0;
  std::vector<messi_dataset *> mem;
  bool owns_arrays = false;
  DevBuf c_site_idx, c_site_leaf, c_pair_idx, c_pair_leaf;  // concatenation when datasets differ
  const int *site_idx = nullptr, *pair_idx = nullptr;
  const uint8_t *site_leaf = nullptr;
  const uint16_t *pair_leaf = nullptr;
  DevBuf site_nodes, pair_nodes, site_items, pair_items, pair_vitems;
  std::vector<int> site_lv, pair_lv, pair_vlv;
  std::vector<int> pair_lvm;   // first pair item of member m at level h: [(h - 1) * (n + 1) + m]
  std::vector<int> pair_vlvm;  // the same for the items of nodes with a child computed in registers
  std::vector<int64_t> node_base, site_root, pair_root;
  int64_t nn_total = 0, site_entries = 0, pair_entries = 0, site_root_total = 0, pair_root_total = 0;
  int Lmax = 0;
  DevBuf site_part, site_expo, site_acc, pair_part, pair_acc, pair_comp;  // pair entries carry their exponents
  Components c4, c16;  // loaded mixture components (site, pair)
  DevBuf dexp_site, dexp_pair, douts;
  std::vector<ExpandDesc> hexp_site, hexp_pair;  // what dexp_site / dexp_pair hold
  std::vector<std::unique_ptr<InsideWork>> w64, w32;
  InsideBatch ib, ib32;     // launch tables of the tile-16 and tile-32 groups
  int groups = 0;           // member groups of a pipelined evaluation (0: automatic)
  int window = 0;           // bulk window of the inside schedule (0: the context's)
  InsideBatch pib[messi_ctx::NPIPE + 1], pib32[messi_ctx::NPIPE + 1];  // their launch tables
  PipeEvents pev;
  std::vector<int> tiles;   // per-member inside tile edge (0: the context's)
  DevBuf zall;              // gathered log Z of all members (2 per member)
  DevBuf eig, gnode;        // messi_batch_load_eigen: eigendecompositions + branch lengths; (member, node) per global node
  HostBuf eig_stage;        // their pinned staging
  Events ev;
  Events sev;                       // profiling: begin / end of the tile-16 and tile-32 step kernels
  int steps_launches[2] = {0, 0};   // step kernel launches of the last profiled evaluation, per tile edge
  double t[5] = {0, 0, 0, 0, 0};
};

messi_dataset::~messi_dataset() { delete self; }

namespace {

template <typename F>
int guard(messi_ctx *ctx, F &&f) {
  try {
    f();
    return 0;
  } catch (const std::exception &e) {
    if (ctx) ctx->err = e.what();
    return 1;
  } catch (...) {
    if (ctx) ctx->err = "unknown error";
    return 1;
  }
}

void build_dataset(messi_dataset *ds, const double *data, const int *children, const int *seqindex,
                   const uint8_t *mask) {
  const int N = ds->nseq, L = ds->L, n = ds->nn;
  hipStream_t s = ds->ctx->stream;
  MESSI_REQUIRE(N >= 1 && L >= 1, "empty alignment");
  MESSI_REQUIRE(n >= 3, "tree must have at least two leaves");
  // 0/1 state indicators -> 4-bit masks.
  std::vector<uint8_t> codes((size_t)N * L);
  for (int sq = 0; sq < N; ++sq)
    for (int c = 0; c < L; ++c) {
      unsigned m = 0;
      for (int a = 0; a < 4; ++a) {
        double x = data[sq + (size_t)N * c + (size_t)N * L * a];
        MESSI_REQUIRE(x == 0.0 || x == 1.0, "alignment data must be 0/1 state indicators");
        if (x == 1.0) m |= 1u << a;
      }
      MESSI_REQUIRE(m != 0, "alignment cell with no allowed state");
      codes[(size_t)sq * L + c] = (uint8_t)m;
    }
  // Tree: pre-order, root 0, children after their parent.
  ds->ch.assign(children, children + 2 * n);
  ds->seq.assign(seqindex, seqindex + n);
  std::vector<int> nparent(n, 0);
  for (int v = 0; v < n; ++v) {
    int l = ds->ch[2 * v], r = ds->ch[2 * v + 1];
    if (l < 0) {
      MESSI_REQUIRE(r < 0, "leaf with one child");
      MESSI_REQUIRE(ds->seq[v] >= 0 && ds->seq[v] < N, "leaf without a valid sequence index");
    } else {
      MESSI_REQUIRE(l > v && r > v && l < n && r < n && l != r,
                    "children must follow their parent in pre-order");
      nparent[l]++;
      nparent[r]++;
    }
  }
  MESSI_REQUIRE(ds->ch[0] >= 0, "root must be internal");
  for (int v = 1; v < n; ++v) MESSI_REQUIRE(nparent[v] == 1, "every non-root node needs one parent");
  // Per-node column patterns, children before parents.
  ds->pid.assign(n, {});
  ds->rep.assign(n, {});
  ds->npat.assign(n, 0);
  ds->height.assign(n, 0);
  for (int v = n - 1; v >= 0; --v) {
    std::vector<int> &p = ds->pid[v];
    p.resize(L);
    int l = ds->ch[2 * v], r = ds->ch[2 * v + 1];
    int np = 0;
    if (l < 0) {
      int tab[16];
      for (int &x : tab) x = -1;
      for (int c = 0; c < L; ++c) {
        int k = codes[(size_t)ds->seq[v] * L + c];
        if (tab[k] < 0) {
          tab[k] = np++;
          ds->rep[v].push_back(c);
        }
        p[c] = tab[k];
      }
    } else {
      HashI64 h(L);
      const std::vector<int> &pl = ds->pid[l], &pr = ds->pid[r];
      for (int c = 0; c < L; ++c) {
        bool ins;
        int id = h.get_or_insert((int64_t)pl[c] * (L + 1) + pr[c], np, ins);
        if (ins) {
          np++;
          ds->rep[v].push_back(c);
        }
        p[c] = id;
      }
      ds->height[v] = 1 + std::max(ds->height[l], ds->height[r]);
    }
    ds->npat[v] = np;
  }
  ds->H = ds->height[0];
  const int H = ds->H;
  auto isleaf = [&](int v) { return ds->ch[2 * v] < 0; };

  // ---- site (4-state) structures ----
  {
    std::vector<NodeDesc> nodes(n);
    std::vector<int> idx;
    std::vector<uint8_t> leaf;
    std::vector<int64_t> off(n, 0);
    int64_t part = 0;
    for (int v = 0; v < n; ++v) {
      if (isleaf(v)) {
        off[v] = (int64_t)leaf.size();
        for (int c : ds->rep[v]) leaf.push_back(codes[(size_t)ds->seq[v] * L + c]);
      } else if (v != 0) {
        off[v] = part;
        part += ds->npat[v];
      }
      if (!isleaf(v)) ds->sumN += ds->npat[v];
    }
    ds->site_entries = part;
    for (int v = 0; v < n; ++v) {
      if (isleaf(v)) continue;
      int l = ds->ch[2 * v], r = ds->ch[2 * v + 1];
      NodeDesc d{};
      d.U = ds->npat[v];
      d.idx_l = (int64_t)idx.size();
      for (int c : ds->rep[v]) idx.push_back(ds->pid[l][c]);
      d.idx_r = (int64_t)idx.size();
      for (int c : ds->rep[v]) idx.push_back(ds->pid[r][c]);
      d.part = off[v];
      d.cpart_l = off[l];
      d.cpart_r = off[r];
      d.leaf_l = isleaf(l);
      d.leaf_r = isleaf(r);
      d.node_l = l;
      d.node_r = r;
      d.root = v == 0;
      nodes[v] = d;
    }
    ds->site_items.assign(H, {});
    for (int v = 0; v < n; ++v)
      if (!isleaf(v))
        for (int e = 0; e < ds->npat[v]; e += 256) ds->site_items[ds->height[v] - 1].push_back(make_int2(v, e));
    ds->site_nodes = nodes;
    ds->site_idx_n = (int64_t)idx.size();
    ds->site_leaf_n = (int64_t)leaf.size();
    upload(ds->d_site_idx, idx, s);
    upload(ds->d_site_leaf, leaf, s);
  }

  // ---- pair (16-state) structures: getindexdict semantics ----
  {
    const int n0 = ds->npat[0];
    MESSI_REQUIRE((int64_t)n0 * n0 < ((int64_t)1 << 31), "too many distinct root column patterns");
    std::vector<int> T0((size_t)n0 * n0, -1);
    std::vector<std::vector<int>> ex(n), ey(n);
    std::vector<std::vector<int>> il(n), ir(n);
    std::vector<int64_t> U(n, 0);
    const std::vector<int> &p0 = ds->pid[0];
    const int D = ds->maxbp;
    for (int x = 0; x < L; ++x) {
      int yend = (int)std::min<int64_t>((int64_t)x + D, L - 1);
      for (int y = x + 1; y <= yend; ++y) {
        if (mask && !mask[(size_t)x + (size_t)L * y]) continue;
        ds->inrange++;
        int &slot = T0[(size_t)p0[x] * n0 + p0[y]];
        if (slot < 0) {
          slot = (int)U[0]++;
          ex[0].push_back(x);
          ey[0].push_back(y);
        }
      }
    }
    MESSI_REQUIRE(U[0] < ((int64_t)1 << 31), "too many root pair entries");
    for (int v = 0; v < n; ++v) {
      if (isleaf(v)) continue;
      for (int side = 0; side < 2; ++side) {
        int c = ds->ch[2 * v + side];
        std::vector<int> &ix = side ? ir[v] : il[v];
        ix.resize(U[v]);
        HashI64 h((size_t)U[v]);
        const std::vector<int> &pc = ds->pid[c];
        const int64_t nc = ds->npat[c];
        for (int64_t e = 0; e < U[v]; ++e) {
          int x = ex[v][e], y = ey[v][e];
          bool ins;
          int id = h.get_or_insert((int64_t)pc[x] * nc + pc[y], (int)U[c], ins);
          if (ins) {
            U[c]++;
            ex[c].push_back(x);
            ey[c].push_back(y);
          }
          ix[e] = id;
        }
      }
      if (v != 0) {
        std::vector<int>().swap(ex[v]);
        std::vector<int>().swap(ey[v]);
      }
    }
    // nodes computed in their parent's launch instead of stored (pair_fuse.h)
    const std::vector<char> virt =
        ds->ctx->pair_fuse ? fuse::pick(n, ds->ch.data(), U.data()) : std::vector<char>(n, 0);
    std::vector<int64_t> off(n, 0);
    std::vector<uint16_t> leaf;
    int64_t part = 0;
    for (int v = 0; v < n; ++v) {
      if (isleaf(v)) {
        off[v] = (int64_t)leaf.size();
        for (int64_t e = 0; e < U[v]; ++e) {
          unsigned cx = codes[(size_t)ds->seq[v] * L + ex[v][e]];
          unsigned cy = codes[(size_t)ds->seq[v] * L + ey[v][e]];
          leaf.push_back((uint16_t)(cx | (cy << 8)));
        }
      } else {
        ds->sumU += U[v];
        if (v != 0 && !virt[v]) {
          off[v] = part;
          part += U[v];
        }
      }
    }
    ds->pair_entries = part;
    ds->U_root = U[0];
    // the pair kernel addresses a node's entries with 32-bit byte offsets (16 values
    // of up to 8 bytes per stored entry; one 8-byte output per root entry)
    for (int v = 0; v < n; ++v)
      if (!isleaf(v))
        MESSI_REQUIRE((uint64_t)U[v] * (v == 0 ? sizeof(double) : 16 * sizeof(double)) <= 0xFFFFFFFFull,
                      "a tree node has too many distinct column-pattern pairs for 32-bit entry offsets");
    std::vector<NodeDesc> nodes(n);
    std::vector<int> idx;
    idx.reserve((size_t)2 * ds->sumU);
    // A node's children in pre-order; a computed child goes first (swapping the
    // two sides changes no result: the sides' products are multiplied, their
    // exponents added).  A computed child's own maps stay until its parent,
    // which comes before it, has composed them.
    for (int v = 0; v < n; ++v) {
      if (isleaf(v) || virt[v]) continue;
      int l = ds->ch[2 * v], r = ds->ch[2 * v + 1];
      const bool sw = virt[r] != 0;
      if (sw) std::swap(l, r);
      const std::vector<int> &ml = sw ? ir[v] : il[v], &mr = sw ? il[v] : ir[v];
      NodeDesc d{};
      d.U = U[v];
      if (virt[l]) {
        const int a = ds->ch[2 * l], bb = ds->ch[2 * l + 1];
        d.virt = 1;
        d.idx_l = (int64_t)idx.size();
        for (int64_t e = 0; e < U[v]; ++e) idx.push_back(il[l][ml[e]]);
        d.idx_v2 = (int64_t)idx.size();
        for (int64_t e = 0; e < U[v]; ++e) idx.push_back(ir[l][ml[e]]);
        d.cpart_l = off[a];
        d.leaf_l = isleaf(a);
        d.cpart_v2 = off[bb];
        d.leaf_v2 = isleaf(bb);
        d.node_va = a;
        d.node_vb = bb;
      } else {
        d.idx_l = (int64_t)idx.size();
        idx.insert(idx.end(), ml.begin(), ml.end());
        d.cpart_l = off[l];
        d.leaf_l = isleaf(l);
      }
      d.idx_r = (int64_t)idx.size();
      idx.insert(idx.end(), mr.begin(), mr.end());
      d.part = off[v];
      d.cpart_r = off[r];
      d.leaf_r = isleaf(r);
      d.node_l = l;
      d.node_r = r;
      d.root = v == 0;
      nodes[v] = d;
      std::vector<int>().swap(il[v]);
      std::vector<int>().swap(ir[v]);
    }
    const int step = 16 * PAIR_NB, vstep = 16 * PAIR_VNB;
    ds->pair_items.assign(H, {});
    ds->pair_vitems.assign(H, {});
    for (int v = 0; v < n; ++v)
      if (!isleaf(v) && !virt[v]) {
        std::vector<int2> &list = nodes[v].virt ? ds->pair_vitems[ds->height[v] - 1] : ds->pair_items[ds->height[v] - 1];
        const int st = nodes[v].virt ? vstep : step;
        for (int64_t e = 0; e < U[v]; e += st) list.push_back(make_int2(v, (int)e));
      }
    ds->pair_nodes = nodes;
    ds->pair_idx_n = (int64_t)idx.size();
    ds->pair_leaf_n = (int64_t)leaf.size();
    upload(ds->d_pair_idx, idx, s);
    upload(ds->d_pair_leaf, leaf, s);
    upload(ds->d_T0, T0, s);
  }
  upload(ds->d_pid0, ds->pid[0], s);
  if (mask) {
    std::vector<uint8_t> m(mask, mask + (size_t)L * L);
    upload(ds->d_mask, m, s);
    ds->has_mask = true;
  }
  MESSI_HIP(hipStreamSynchronize(s));
}

// Combine the members' descriptors into one launch structure.
void build_batch(messi_batch *b) {
  hipStream_t s = b->ctx->stream;
  const int n = b->n;
  std::map<messi_dataset *, int> distinct;
  std::vector<messi_dataset *> dlist;
  for (auto *d : b->mem)
    if (!distinct.count(d)) {
      distinct[d] = (int)dlist.size();
      dlist.push_back(d);
    }
  std::vector<int64_t> sidx_base(dlist.size(), 0), pidx_base(dlist.size(), 0), sleaf_base(dlist.size(), 0),
      pleaf_base(dlist.size(), 0);
  if (dlist.size() == 1) {
    b->site_idx = dlist[0]->d_site_idx.as<int>();
    b->pair_idx = dlist[0]->d_pair_idx.as<int>();
    b->site_leaf = dlist[0]->d_site_leaf.as<uint8_t>();
    b->pair_leaf = dlist[0]->d_pair_leaf.as<uint16_t>();
  } else {
    int64_t a = 0, c = 0, e = 0, f = 0;
    for (size_t i = 0; i < dlist.size(); ++i) {
      sidx_base[i] = a;
      pidx_base[i] = c;
      sleaf_base[i] = e;
      pleaf_base[i] = f;
      a += dlist[i]->site_idx_n;
      c += dlist[i]->pair_idx_n;
      e += dlist[i]->site_leaf_n;
      f += dlist[i]->pair_leaf_n;
    }
    b->c_site_idx.ensure(std::max<int64_t>(1, a) * sizeof(int));
    b->c_pair_idx.ensure(std::max<int64_t>(1, c) * sizeof(int));
    b->c_site_leaf.ensure(std::max<int64_t>(1, e));
    b->c_pair_leaf.ensure(std::max<int64_t>(1, f) * sizeof(uint16_t));
    for (size_t i = 0; i < dlist.size(); ++i) {
      messi_dataset *d = dlist[i];
      if (d->site_idx_n)
        MESSI_HIP(hipMemcpyAsync(b->c_site_idx.as<int>() + sidx_base[i], d->d_site_idx.p,
                                 d->site_idx_n * sizeof(int), hipMemcpyDeviceToDevice, s));
      if (d->pair_idx_n)
        MESSI_HIP(hipMemcpyAsync(b->c_pair_idx.as<int>() + pidx_base[i], d->d_pair_idx.p,
                                 d->pair_idx_n * sizeof(int), hipMemcpyDeviceToDevice, s));
      if (d->site_leaf_n)
        MESSI_HIP(hipMemcpyAsync(b->c_site_leaf.as<uint8_t>() + sleaf_base[i], d->d_site_leaf.p, d->site_leaf_n,
                                 hipMemcpyDeviceToDevice, s));
      if (d->pair_leaf_n)
        MESSI_HIP(hipMemcpyAsync(b->c_pair_leaf.as<uint16_t>() + pleaf_base[i], d->d_pair_leaf.p,
                                 d->pair_leaf_n * sizeof(uint16_t), hipMemcpyDeviceToDevice, s));
    }
    b->site_idx = b->c_site_idx.as<int>();
    b->pair_idx = b->c_pair_idx.as<int>();
    b->site_leaf = b->c_site_leaf.as<uint8_t>();
    b->pair_leaf = b->c_pair_leaf.as<uint16_t>();
    b->owns_arrays = true;
  }
  b->node_base.assign(n, 0);
  b->site_root.assign(n, 0);
  b->pair_root.assign(n, 0);
  std::vector<int64_t> spart(n), ppart(n);
  int64_t nb = 0, sp = 0, pp = 0, sr = 0, pr = 0;
  b->Hmax = 0;
  b->Lmax = 0;
  for (int m = 0; m < n; ++m) {
    messi_dataset *d = b->mem[m];
    b->node_base[m] = nb;
    spart[m] = sp;
    ppart[m] = pp;
    b->site_root[m] = sr;
    b->pair_root[m] = pr;
    nb += d->nn;
    sp += d->site_entries;
    pp += d->pair_entries;
    sr += d->npat[0];
    pr += d->U_root;
    b->Hmax = std::max(b->Hmax, d->H);
    b->Lmax = std::max(b->Lmax, d->L);
  }
  b->nn_total = nb;
  b->site_entries = sp;
  b->pair_entries = pp;
  b->site_root_total = sr;
  b->pair_root_total = pr;
  std::vector<NodeDesc> sn(nb), pn(nb);
  std::vector<int2> si, pi, vi;
  b->site_lv.assign(b->Hmax + 1, 0);
  b->pair_lv.assign(b->Hmax + 1, 0);
  b->pair_vlv.assign(b->Hmax + 1, 0);
  b->pair_lvm.assign((size_t)b->Hmax * (n + 1), 0);
  b->pair_vlvm.assign((size_t)b->Hmax * (n + 1), 0);
  for (int m = 0; m < n; ++m) {
    messi_dataset *d = b->mem[m];
    const int di = distinct[d];
    for (int v = 0; v < d->nn; ++v) {
      if (d->ch[2 * v] < 0) continue;
      for (int kind = 0; kind < 2; ++kind) {
        NodeDesc nd = kind ? d->pair_nodes[v] : d->site_nodes[v];
        const int64_t ib = kind ? pidx_base[di] : sidx_base[di];
        const int64_t lb = kind ? pleaf_base[di] : sleaf_base[di];
        const int64_t pb = kind ? ppart[m] : spart[m];
        nd.idx_l += ib;
        nd.idx_r += ib;
        nd.part += pb;
        nd.cpart_l += nd.leaf_l ? lb : pb;
        nd.cpart_r += nd.leaf_r ? lb : pb;
        nd.node_l += (int)b->node_base[m];
        nd.node_r += (int)b->node_base[m];
        if (nd.virt) {
          nd.idx_v2 += ib;
          nd.cpart_v2 += nd.leaf_v2 ? lb : pb;
          nd.node_va += (int)b->node_base[m];
          nd.node_vb += (int)b->node_base[m];
        }
        nd.member = m;
        nd.out = kind ? b->pair_root[m] : b->site_root[m];
        (kind ? pn : sn)[b->node_base[m] + v] = nd;
      }
    }
  }
  for (int h = 1; h <= b->Hmax; ++h) {
    b->site_lv[h - 1] = (int)si.size();
    b->pair_lv[h - 1] = (int)pi.size();
    b->pair_vlv[h - 1] = (int)vi.size();
    for (int m = 0; m < n; ++m) {
      b->pair_lvm[(size_t)(h - 1) * (n + 1) + m] = (int)pi.size();
      b->pair_vlvm[(size_t)(h - 1) * (n + 1) + m] = (int)vi.size();
      messi_dataset *d = b->mem[m];
      if (h > d->H) continue;
      for (int2 it : d->site_items[h - 1]) si.push_back(make_int2(it.x + (int)b->node_base[m], it.y));
      for (int2 it : d->pair_items[h - 1]) pi.push_back(make_int2(it.x + (int)b->node_base[m], it.y));
      for (int2 it : d->pair_vitems[h - 1]) vi.push_back(make_int2(it.x + (int)b->node_base[m], it.y));
    }
    b->pair_lvm[(size_t)(h - 1) * (n + 1) + n] = (int)pi.size();
    b->pair_vlvm[(size_t)(h - 1) * (n + 1) + n] = (int)vi.size();
  }
  b->site_lv[b->Hmax] = (int)si.size();
  b->pair_lv[b->Hmax] = (int)pi.size();
  b->pair_vlv[b->Hmax] = (int)vi.size();
  upload(b->site_nodes, sn, s);
  upload(b->pair_nodes, pn, s);
  upload(b->site_items, si, s);
  upload(b->pair_items, pi, s);
  upload(b->pair_vitems, vi, s);
  b->w64.clear();
  b->w32.clear();
  for (int m = 0; m < n; ++m) {
    b->w64.emplace_back(new InsideWork());
    b->w32.emplace_back(new InsideWork());
  }
  MESSI_HIP(hipStreamSynchronize(s));
}

messi_batch *self_batch(messi_dataset *ds) {
  if (!ds->self) {
    auto *b = new messi_batch();
    b->ctx = ds->ctx;
    b->n = 1;
    b->mem.push_back(ds);
    try {
      build_batch(b);
    } catch (...) {
      delete b;
      throw;
    }
    ds->self = b;
  }
  return ds->self;
}

// Member-major input (member m: K x nn_m x dim^2) -> component-major device arena.
template <typename T>
void load_components(messi_batch *b, Components &c, int K, int dim2, const double *P, const double *pi,
                     int pidim, const double *logw, int prec) {
  hipStream_t s = b->ctx->stream;
  const int n = b->n;
  const size_t nP = (size_t)K * b->nn_total * dim2, nq = (size_t)K * n * pidim, nw = (size_t)K * n;
  const size_t bP = nP * sizeof(T), bq = nq * sizeof(double), bw = nw * sizeof(double);
  MESSI_HIP(hipStreamSynchronize(s));  // the staging buffer may still feed an earlier copy
  c.stage.ensure(bP + bq + bw);
  T *Ph = c.stage.as<T>();
  double *pih = reinterpret_cast<double *>(c.stage.as<char>() + bP);
  double *lwh = pih + nq;
  size_t src = 0;
  for (int m = 0; m < n; ++m) {
    const int nn = b->mem[m]->nn;
    for (int k = 0; k < K; ++k) {
      T *dst = Ph + ((size_t)k * b->nn_total + b->node_base[m]) * dim2;
      for (size_t i = 0; i < (size_t)nn * dim2; ++i) dst[i] = (T)P[src + i];
      src += (size_t)nn * dim2;
      for (int a = 0; a < pidim; ++a) pih[((size_t)k * n + m) * pidim + a] = pi[((size_t)m * K + k) * pidim + a];
      lwh[(size_t)k * n + m] = logw[(size_t)m * K + k];
    }
  }
  c.P.ensure(bP);
  c.pi.ensure(bq);
  c.logw.ensure(bw);
  MESSI_HIP(hipMemcpyAsync(c.P.p, Ph, bP, hipMemcpyHostToDevice, s));
  MESSI_HIP(hipMemcpyAsync(c.pi.p, pih, bq, hipMemcpyHostToDevice, s));
  MESSI_HIP(hipMemcpyAsync(c.logw.p, lwh, bw, hipMemcpyHostToDevice, s));
  c.K = K;
  c.prec = prec;
}

// Transition matrices of every member, component and node from eigendecompositions,
// gettransitionmatriceslist with absmat (src/FARCE.jl:389-433): for a node v >= 1 of
// member m with branch length t, P(a -> d) = sum_c V(a, c) exp(w_c t) Vi(c, d), each
// entry <= 0 replaced by 1e-50, and the identity when t == 0; the root's slot (v = 0,
// no branch) is zero, as the reference's first list entry.  One thread per entry of
// the component-major arena (component k, global node g, column-major entry a + dim d);
// V and Vi are column-major per (member, component), w per (member, component).
template <typename T>
__global__ void eigen_transition_kernel(int dim, int K, int nn_total, const int2 *__restrict__ gnode,
                                        const double *__restrict__ t, const double *__restrict__ V,
                                        const double *__restrict__ w, const double *__restrict__ Vi, T *P) {
  const size_t d2 = (size_t)dim * dim;
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)K * nn_total * d2) return;
  const int e = (int)(idx % d2);
  const size_t kg = idx / d2;
  const int g = (int)(kg % nn_total), k = (int)(kg / nn_total);
  const int a = e % dim, d = e / dim;
  const int2 mv = gnode[g];
  const size_t mk = (size_t)mv.x * K + k;
  const double tv = t[g];
  double v;
  if (mv.y == 0) {
    v = 0.0;
  } else if (tv == 0.0) {
    v = a == d ? 1.0 : 0.0;
  } else {
    const double *Vm = V + mk * d2, *wm = w + mk * dim, *Vim = Vi + mk * d2;
    v = 0.0;
    for (int c = 0; c < dim; ++c) v += Vm[a + dim * c] * exp(wm[c] * tv) * Vim[c + dim * d];
    if (v <= 0.0) v = 1e-50;
  }
  P[idx] = (T)v;
}

// Member-major eigendecompositions (member m: K x (dim^2 V, dim w, dim^2 Vi)),
// branch lengths in global node order, pi and log weights as load_components
// takes them -> the resident component arenas of both models.
template <typename T>
void load_eigen_impl(messi_batch *b, int K4, const double *V4, const double *w4, const double *Vi4,
                     const double *pi4, const double *logw4, int K16, const double *V16, const double *w16,
                     const double *Vi16, const double *pi16, const double *logw16, const double *t) {
  MESSI_REQUIRE(K4 >= 1 && K16 >= 1, "need at least one site and one pair component");
  hipStream_t s = b->ctx->stream;
  const int n = b->n, prec = sizeof(T) == 4 ? MESSI_FP32 : MESSI_FP64;
  const int64_t nt = b->nn_total;
  if (b->gnode.bytes < sizeof(int2) * (size_t)nt) {
    std::vector<int2> gn((size_t)nt);
    for (int m = 0; m < n; ++m)
      for (int v = 0; v < b->mem[m]->nn; ++v) gn[(size_t)b->node_base[m] + v] = make_int2(m, v);
    upload(b->gnode, gn, s);
  }
  // staging: V4 w4 Vi4 V16 w16 Vi16 t, then each model's pi and log weights (component-major)
  const size_t n4 = (size_t)n * K4, n16 = (size_t)n * K16;
  const size_t sz[7] = {n4 * 16, n4 * 4, n4 * 16, n16 * 256, n16 * 16, n16 * 256, (size_t)nt};
  const double *src[7] = {V4, w4, Vi4, V16, w16, Vi16, t};
  size_t tot = 0;
  for (size_t z : sz) tot += z;
  MESSI_HIP(hipStreamSynchronize(s));  // the staging buffer may still feed an earlier copy
  b->eig_stage.ensure(sizeof(double) * (tot + n4 * 5 + n16 * 17));
  double *h = b->eig_stage.as<double>();
  size_t off[7], o = 0;
  for (int i = 0; i < 7; ++i) {
    off[i] = o;
    std::memcpy(h + o, src[i], sizeof(double) * sz[i]);
    o += sz[i];
  }
  b->eig.ensure(sizeof(double) * tot);
  MESSI_HIP(hipMemcpyAsync(b->eig.p, h, sizeof(double) * tot, hipMemcpyHostToDevice, s));
  const double *de = b->eig.as<double>();
  auto comps = [&](Components &c, int K, int dim, size_t oV, size_t ow, size_t oVi, const double *pi,
                   const double *logw, double *hq) {
    const size_t dim2 = (size_t)dim * dim;
    double *hw = hq + (size_t)K * n * dim;
    for (int m = 0; m < n; ++m)
      for (int k = 0; k < K; ++k) {
        for (int q = 0; q < dim; ++q) hq[((size_t)k * n + m) * dim + q] = pi[((size_t)m * K + k) * dim + q];
        hw[(size_t)k * n + m] = logw[(size_t)m * K + k];
      }
    c.P.ensure(sizeof(T) * (size_t)K * nt * dim2);
    c.pi.ensure(sizeof(double) * (size_t)K * n * dim);
    c.logw.ensure(sizeof(double) * (size_t)K * n);
    MESSI_HIP(hipMemcpyAsync(c.pi.p, hq, sizeof(double) * (size_t)K * n * dim, hipMemcpyHostToDevice, s));
    MESSI_HIP(hipMemcpyAsync(c.logw.p, hw, sizeof(double) * (size_t)K * n, hipMemcpyHostToDevice, s));
    const size_t total = (size_t)K * nt * dim2;
    hipLaunchKernelGGL(eigen_transition_kernel<T>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, dim,
                       K, (int)nt, b->gnode.as<int2>(), de + off[6], de + oV, de + ow, de + oVi, c.P.as<T>());
    c.K = K;
    c.prec = prec;
  };
  comps(b->c4, K4, 4, off[0], off[1], off[2], pi4, logw4, h + tot);
  comps(b->c16, K16, 16, off[3], off[4], off[5], pi16, logw16, h + tot + n4 * 5);
  MESSI_HIP(hipGetLastError());
}

void require_loaded(const Components &c, int prec, const char *what) {
  MESSI_REQUIRE(c.prec >= 0, std::string("no ") + what + " components loaded");
  MESSI_REQUIRE(c.prec == prec, std::string(what) + " components were loaded for the other precision");
}

InsideWork &member_work(messi_batch *b, int m, int prec) {
  return prec == MESSI_FP32 ? *b->w32[m] : *b->w64[m];
}

// Expansion descriptors of the site (pairs = false) or pair root entries; the
// device copy is refreshed only when a buffer it points to has moved.
const ExpandDesc *upload_expand(messi_batch *b, int prec, bool pairs) {
  std::vector<ExpandDesc> e(b->n);
  for (int m = 0; m < b->n; ++m) {
    messi_dataset *d = b->mem[m];
    InsideWork &w = member_work(b, m, prec);
    e[m].L = d->L;
    e[m].maxbp = d->maxbp;
    e[m].n0 = d->npat[0];
    e[m].pid0 = d->d_pid0.as<int>();
    e[m].T0 = d->d_T0.as<int>();
    e[m].mask = d->has_mask ? d->d_mask.as<uint8_t>() : nullptr;
    e[m].rootoff = pairs ? b->pair_root[m] : b->site_root[m];
    e[m].u = w.u.as<double>();
    e[m].paired = w.paired.as<double>();
  }
  std::vector<ExpandDesc> &h = pairs ? b->hexp_pair : b->hexp_site;
  DevBuf &d = pairs ? b->dexp_pair : b->dexp_site;
  if (h.size() != e.size() || std::memcmp(h.data(), e.data(), e.size() * sizeof(ExpandDesc)) != 0) {
    upload(d, e, b->ctx->stream);
    h = e;
  }
  return d.as<ExpandDesc>();
}

template <typename T>
void run_site(messi_batch *b, int prec) {
  hipStream_t s = b->ctx->stream;
  require_loaded(b->c4, prec, "site");
  const int K = b->c4.K;
  b->site_part.ensure(std::max<int64_t>(1, b->site_entries) * 4 * sizeof(T));
  b->site_expo.ensure(std::max<int64_t>(1, b->site_entries) * sizeof(int));
  b->site_acc.ensure(std::max<int64_t>(1, b->site_root_total) * sizeof(double));
  for (int m = 0; m < b->n; ++m) member_work(b, m, prec).u.ensure(sizeof(double) * b->mem[m]->L);
  for (int k = 0; k < K; ++k)
    for (int h = 1; h <= b->Hmax; ++h) {
      int lo = b->site_lv[h - 1], hi = b->site_lv[h];
      launch_site_level<T>(s, b->site_items.as<int2>() + lo, hi - lo, b->site_nodes.as<NodeDesc>(), b->site_idx,
                           b->site_leaf, b->site_part.as<T>(), b->site_expo.as<int>(),
                           b->c4.P.as<T>() + (size_t)k * b->nn_total * 16,
                           b->c4.pi.as<double>() + (size_t)k * b->n * 4, b->c4.logw.as<double>() + (size_t)k * b->n,
                           k == 0, b->site_acc.as<double>());
    }
  const ExpandDesc *dexp = upload_expand(b, prec, false);
  launch_site_expand(s, b->n, b->Lmax, dexp, b->site_acc.as<double>());
  MESSI_HIP(hipGetLastError());
}

// The level launches of pair component k for members m0 .. m1-1 on stream s
// (the items of a level are ordered by member); comp keeps the component's own
// root values in b->pair_comp besides their sum over components in b->pair_acc.
template <typename T>
void pair_levels(messi_batch *b, hipStream_t s, int k, int m0, int m1, bool comp = false) {
  const int n = b->n;
  for (int h = 1; h <= b->Hmax; ++h) {
    const size_t r = (size_t)(h - 1) * (n + 1);
    const int lo = b->pair_lvm[r + m0], hi = b->pair_lvm[r + m1];
    const int vlo = b->pair_vlvm[r + m0], vhi = b->pair_vlvm[r + m1];
    const T *P = b->c16.P.as<T>() + (size_t)k * b->nn_total * 256;
    const double *pi = b->c16.pi.as<double>() + (size_t)k * b->n * 16;
    const double *logw = b->c16.logw.as<double>() + (size_t)k * b->n;
    double *pc = comp ? b->pair_comp.as<double>() : nullptr;
    launch_pair_level<T, PAIR_NB>(s, b->pair_items.as<int2>() + lo, hi - lo, b->pair_nodes.as<NodeDesc>(),
                                  b->pair_idx, b->pair_leaf, b->pair_part.as<T>(), P, pi, logw, k == 0,
                                  b->pair_acc.as<double>(), pc);
    launch_pair_level<T, PAIR_VNB, true>(s, b->pair_vitems.as<int2>() + vlo, vhi - vlo,
                                         b->pair_nodes.as<NodeDesc>(), b->pair_idx, b->pair_leaf,
                                         b->pair_part.as<T>(), P, pi, logw, k == 0, b->pair_acc.as<double>(), pc);
  }
}

template <typename T>
// expand: write each member's L x L pair matrix (the evaluation path instead
// builds the inside pair term straight from the root entries).  pair_only:
// only size the buffers and upload the expansion descriptors (the launches are
// made per member group by the pipelined evaluation).
void run_pair(messi_batch *b, int prec, double *h_percomp, bool expand, bool pair_only = false) {
  hipStream_t s = b->ctx->stream;
  require_loaded(b->c16, prec, "pair");
  const int K = b->c16.K;
  b->pair_part.ensure(std::max<int64_t>(1, b->pair_entries) * 16 * sizeof(T));
  b->pair_acc.ensure(std::max<int64_t>(1, b->pair_root_total) * sizeof(double));
  b->pair_comp.ensure(std::max<int64_t>(1, b->pair_root_total) * sizeof(double));
  if (expand || h_percomp)
    for (int m = 0; m < b->n; ++m)
      member_work(b, m, prec).paired.ensure(sizeof(double) * b->mem[m]->L * b->mem[m]->L);
  const ExpandDesc *dexp = upload_expand(b, prec, true);
  if (pair_only) return;
  for (int k = 0; k < K; ++k) {
    pair_levels<T>(b, s, k, 0, b->n, h_percomp != nullptr);
    if (h_percomp) {
      MESSI_REQUIRE(b->n == 1, "per-component output needs a single member");
      const int L = b->mem[0]->L;
      InsideWork &w = member_work(b, 0, prec);
      w.out3.ensure(sizeof(double) * L * L);
      double *o = w.out3.as<double>();
      std::vector<double *> outs(1, o);
      upload(b->douts, outs, s);
      launch_pair_expand(s, 1, L, dexp, b->pair_comp.as<double>(), b->douts.as<double *>());
      MESSI_HIP(hipMemcpyAsync(h_percomp + (size_t)k * L * L, o, sizeof(double) * L * L, hipMemcpyDeviceToHost, s));
      MESSI_HIP(hipStreamSynchronize(s));
    }
  }
  if (expand) launch_pair_expand(s, b->n, b->Lmax, dexp, b->pair_acc.as<double>(), nullptr);
  MESSI_HIP(hipGetLastError());
}

InsideWork &work(messi_ctx *ctx, int prec) {
  MESSI_REQUIRE(prec == MESSI_FP32 || prec == MESSI_FP64, "prec must be MESSI_FP64 or MESSI_FP32");
  return prec == MESSI_FP32 ? ctx->w32 : ctx->w64;
}

void check_prec(int prec) {
  MESSI_REQUIRE(prec == MESSI_FP32 || prec == MESSI_FP64, "prec must be MESSI_FP64 or MESSI_FP32");
}

void record(messi_batch *b, int i) {
  if (b->ctx->profiling) MESSI_HIP(hipEventRecord(b->ev.e[i], b->ctx->stream));
}

void collect(messi_batch *b, int upto) {
  if (!b->ctx->profiling) return;
  MESSI_HIP(hipEventSynchronize(b->ev.e[upto]));
  for (int i = 0; i < upto; ++i) {
    float ms = 0;
    MESSI_HIP(hipEventElapsedTime(&ms, b->ev.e[i], b->ev.e[i + 1]));
    b->t[i] = ms;
  }
  for (int i = upto; i < 5; ++i) b->t[i] = 0;
}

// Site + pair + inside for every member; logZ[m] on return.
template <typename T>
void load_impl(messi_batch *b, int K4, const double *P4, const double *pi4, const double *logw4, int K16,
               const double *P16, const double *pi16, const double *logw16) {
  MESSI_REQUIRE(K4 >= 1 && K16 >= 1, "need at least one site and one pair component");
  const int prec = sizeof(T) == 4 ? MESSI_FP32 : MESSI_FP64;
  load_components<T>(b, b->c4, K4, 16, P4, pi4, 4, logw4, prec);
  load_components<T>(b, b->c16, K16, 256, P16, pi16, 16, logw16, prec);
}

// Site + pair + inside for every member from the loaded components; logZ[m] on return.
template <typename T>
void eval_impl(messi_batch *b, double B, bool keep_outside, double *logZ) {
  messi_ctx *ctx = b->ctx;
  hipStream_t s = ctx->stream;
  const int prec = sizeof(T) == 4 ? MESSI_FP32 : MESSI_FP64;
  std::vector<InsideWork *> ws(b->n);
  b->zall.ensure(sizeof(double) * 2 * b->n);
  for (int m = 0; m < b->n; ++m) {
    ws[m] = &member_work(b, m, prec);
    ws[m]->tile = (m < (int)b->tiles.size() && b->tiles[m]) ? b->tiles[m] : ctx->tile;
    ws[m]->window = b->window ? b->window : ctx->window;
    ws[m]->persist = ctx->persist;
    ws[m]->zout = b->zall.as<double>() + 2 * m;
    // pair scores beyond the distance bound are -inf (pair_expand writes them so)
    const int D = b->mem[m]->maxbp, Lm = b->mem[m]->L;
    ws[m]->band = (ctx->band_skip && D < Lm - 1) ? D : -1;
    inside_alloc<T>(*ws[m], b->mem[m]->L, keep_outside);
  }
  if (ctx->profiling) {
    b->ev.make();
    b->sev.make();
    b->steps_launches[0] = b->steps_launches[1] = 0;
  }
  // Inside launch sequences of members m0 .. m1-1 on stream st, one per tile edge.
  auto insides = [&](hipStream_t st, int m0, int m1, InsideBatch &i16, InsideBatch &i32) {
    for (int TE : {16, 32}) {
      std::vector<InsideWork *> g;
      for (int m = m0; m < m1; ++m)
        if (ws[m]->tile == TE) {
          ws[m]->exp = b->dexp_pair.as<ExpandDesc>() + m;
          ws[m]->dset = b->mem[m];
          g.push_back(ws[m]);
        }
      if (g.empty()) continue;
      // profiling (one stream): events around each tile edge's step kernels
      const int k = TE == 32 ? 2 : 0;
      const bool timed = ctx->profiling;
      const int nl = inside_batch_run<T>(st, g.data(), (int)g.size(), TE == 32 ? i32 : i16, B,
                                         b->pair_acc.as<double>(), timed ? b->sev.e[k] : nullptr,
                                         timed ? b->sev.e[k + 1] : nullptr);
      if (timed) b->steps_launches[k / 2] = nl;
    }
  };
  // Member groups: with G > 1 each group runs its pair phase and then its
  // inside on a stream of its own, the pair phase of group g after that of
  // group g-1, so the (bandwidth-bound) pruning of one group overlaps the
  // (latency-bound) inside launches of the groups before it.  Profiling runs
  // the phases one after another on the context's stream.
  int G = b->groups > 0 ? b->groups : (b->n >= 4 ? 2 : 1);
  G = std::min(std::min(G, messi_ctx::NPIPE + 1), b->n);
  if (ctx->profiling) G = 1;
  record(b, 0);
  run_site<T>(b, prec);
  record(b, 1);
  if (G == 1) {
    run_pair<T>(b, prec, nullptr, false);
    record(b, 2);
    insides(s, 0, b->n, b->ib, b->ib32);
  } else {
    run_pair<T>(b, prec, nullptr, false, true);
    b->pev.make();
    // group 0 runs on the context's stream, group g >= 1 on pipe[g - 1]: at
    // most 4 streams, one per hardware queue
    for (int g = 0; g < G; ++g) {
      const int m0 = g * b->n / G, m1 = (g + 1) * b->n / G;
      hipStream_t sg = g == 0 ? s : ctx->pipe[g - 1];
      if (g > 0) MESSI_HIP(hipStreamWaitEvent(sg, b->pev.pair[g - 1], 0));
      for (int k = 0; k < b->c16.K; ++k) pair_levels<T>(b, sg, k, m0, m1);
      MESSI_HIP(hipEventRecord(b->pev.pair[g], sg));
      insides(sg, m0, m1, b->pib[g], b->pib32[g]);
      if (g > 0) MESSI_HIP(hipEventRecord(b->pev.done[g], sg));
    }
    for (int g = 1; g < G; ++g) MESSI_HIP(hipStreamWaitEvent(s, b->pev.done[g], 0));
  }
  for (int m = 0; m < b->n; ++m) {
    ws[m]->exp = nullptr;
    ws[m]->dset = nullptr;
    ws[m]->zout = nullptr;
  }
  record(b, 3);
  std::vector<double> z(2 * b->n);
  MESSI_HIP(hipMemcpyAsync(z.data(), b->zall.p, sizeof(double) * 2 * b->n, hipMemcpyDeviceToHost, s));
  MESSI_HIP(hipStreamSynchronize(s));
  for (int m = 0; m < b->n; ++m) {
    logZ[m] = z[2 * m];
    MESSI_REQUIRE(!(ctx->persist && std::isnan(logZ[m])),
                  "single-launch inside gave up waiting for a dependency (log Z not computed)");
  }
  if (!keep_outside) collect(b, 3);
}

// Every loaded component of the batch's single member in the same launches
// (grid row = component), each with its own partials and root entries, one
// launch per tree level and kind: sitec (K4 x L) receives log w_c + log L_c(x)
// per site component, b->pair_comp each pair component's root entries with a
// stride of the returned count (root entries of the member); *dpair the pair
// expansion descriptor.
template <typename T>
int64_t component_roots(messi_batch *b, int prec, double *sitec, const ExpandDesc **dpair) {
  hipStream_t s = b->ctx->stream;
  const int K4 = b->c4.K, K16 = b->c16.K, L = b->mem[0]->L;
  const int64_t se = std::max<int64_t>(1, b->site_entries), sr = std::max<int64_t>(1, b->site_root_total);
  const int64_t pe = std::max<int64_t>(1, b->pair_entries), pr = std::max<int64_t>(1, b->pair_root_total);
  b->site_part.ensure(sizeof(T) * 4 * se * K4);
  b->site_expo.ensure(sizeof(int) * se * K4);
  b->site_acc.ensure(sizeof(double) * sr * K4);
  b->pair_part.ensure(sizeof(T) * 16 * pe * K16);
  b->pair_acc.ensure(sizeof(double) * pr * K16);
  b->pair_comp.ensure(sizeof(double) * pr * K16);
  const ExpandDesc *dsite = upload_expand(b, prec, false);
  CompBatch cs;
  cs.K = K4;
  cs.part = 4 * se;
  cs.expo = se;
  cs.root = sr;
  cs.P = b->nn_total * 16;
  cs.pi = (int64_t)b->n * 4;
  cs.logw = b->n;
  for (int h = 1; h <= b->Hmax; ++h) {
    const int lo = b->site_lv[h - 1], hi = b->site_lv[h];
    launch_site_level<T>(s, b->site_items.as<int2>() + lo, hi - lo, b->site_nodes.as<NodeDesc>(), b->site_idx,
                         b->site_leaf, b->site_part.as<T>(), b->site_expo.as<int>(), b->c4.P.as<T>(),
                         b->c4.pi.as<double>(), b->c4.logw.as<double>(), 1, b->site_acc.as<double>(), cs);
  }
  launch_site_gather(s, L, K4, sr, dsite, b->site_acc.as<double>(), sitec);
  *dpair = upload_expand(b, prec, true);
  CompBatch cp;
  cp.K = K16;
  cp.part = 16 * pe;
  cp.root = pr;
  cp.P = b->nn_total * 256;
  cp.pi = (int64_t)b->n * 16;
  cp.logw = b->n;
  for (int h = 1; h <= b->Hmax; ++h) {
    const int lo = b->pair_lv[h - 1], hi = b->pair_lv[h];
    const int vlo = b->pair_vlv[h - 1], vhi = b->pair_vlv[h];
    launch_pair_level<T, PAIR_NB>(s, b->pair_items.as<int2>() + lo, hi - lo, b->pair_nodes.as<NodeDesc>(),
                                  b->pair_idx, b->pair_leaf, b->pair_part.as<T>(),
                                  b->c16.P.as<T>(), b->c16.pi.as<double>(), b->c16.logw.as<double>(), 1,
                                  b->pair_acc.as<double>(), b->pair_comp.as<double>(), cp);
    launch_pair_level<T, PAIR_VNB, true>(s, b->pair_vitems.as<int2>() + vlo, vhi - vlo,
                                         b->pair_nodes.as<NodeDesc>(), b->pair_idx, b->pair_leaf,
                                         b->pair_part.as<T>(), b->c16.P.as<T>(), b->c16.pi.as<double>(),
                                         b->c16.logw.as<double>(), 1, b->pair_acc.as<double>(),
                                         b->pair_comp.as<double>(), cp);
  }
  return pr;
}

}  // namespace

namespace {

// Events of one MEA call (profiling on): [0] start, [1] after the upload, [2] after
// the fill, [3] after the traceback.
struct MeaEvents {
  hipEvent_t e[4] = {nullptr, nullptr, nullptr, nullptr};
  double ms[3] = {0, 0, 0};
};
MeaEvents &mea_events(messi_ctx *ctx) {
  static MeaEvents ev[64];
  MeaEvents &m = ev[ctx->device & 63];
  if (!m.e[0])
    for (auto &e : m.e) MESSI_HIP(hipEventCreate(&e));
  return m;
}

template <typename V>
void mea_timed(messi_ctx *ctx, int L, const double *d_single, const double *d_pair, double alpha, int *d_paired) {
  hipStream_t s = ctx->stream;
  if (ctx->profiling) {
    MeaEvents &m = mea_events(ctx);
    MESSI_HIP(hipEventRecord(m.e[1], s));
    mea_run<V>(s, ctx->mea, L, d_single, d_pair, alpha, nullptr);
    MESSI_HIP(hipEventRecord(m.e[2], s));
    mea_run_traceback(s, ctx->mea, L, d_paired);
    MESSI_HIP(hipEventRecord(m.e[3], s));
  } else {
    mea_run<V>(s, ctx->mea, L, d_single, d_pair, alpha, d_paired);
  }
}

void mea_collect(messi_ctx *ctx, bool upload) {
  if (!ctx->profiling) return;
  MeaEvents &m = mea_events(ctx);
  MESSI_HIP(hipEventSynchronize(m.e[3]));
  float a = 0.f, b = 0.f, c = 0.f;
  if (upload) MESSI_HIP(hipEventElapsedTime(&a, m.e[0], m.e[1]));
  MESSI_HIP(hipEventElapsedTime(&b, m.e[1], m.e[2]));
  MESSI_HIP(hipEventElapsedTime(&c, m.e[2], m.e[3]));
  m.ms[0] = upload ? a : 0.0;
  m.ms[1] = b;
  m.ms[2] = c;
}

}  // namespace

namespace {

// ---------------------------------------------------------------------------
// Static scratch of messi_inside_noalloc.  Every buffer of that call is a fixed
// region of this array, sized for the tables of L = MESSI_NOALLOC_MAX_L (both
// precisions use 8-byte cells), so the call allocates no device memory.  Each
// device holds its own copy; its address is taken once per device, with the
// device current, when the first context on it is created (hipGetSymbolAddress),
// and only that address is used on the host side.  All calls on one device share
// the array: calls that overlap on a device race.  Nothing in it has to start at
// zero (every cell the inside pass reads was written earlier in the same call),
// and the launch tables are uploaded on every call, since another context may
// have overwritten them.
constexpr size_t na_round(size_t n) { return (n + 255) & ~(size_t)255; }
constexpr size_t NA_L = MESSI_NOALLOC_MAX_L;
constexpr size_t NA_TABLE = na_round(NA_L * NA_L * 8);                  // one L x L table of 8-byte cells
constexpr size_t NA_N16 = (NA_L + 15) / 16, NA_N32 = (NA_L + 31) / 32;
constexpr size_t NA_CELLS = std::max(NA_N16 * NA_N16 * 256, NA_N32 * NA_N32 * 1024);  // tile-major cells
constexpr size_t NA_ROWS = std::max(NA_N16 * NA_N16 * 16, NA_N32 * NA_N32 * 32);      // tile rows
constexpr size_t NA_VEC = na_round((NA_L + 1) * 8);
constexpr size_t NA_RED = na_round((3 * NA_L + 16) * 8);
constexpr size_t NA_DESC = 4096;                                          // descriptors, launch tables
constexpr size_t NOALLOC_BYTES = 7 * NA_TABLE + na_round(NA_CELLS * 8) + 2 * na_round(NA_CELLS * 4) +
                                 2 * na_round(NA_ROWS * 4) + 4 * NA_VEC + NA_RED + 3 * NA_TABLE + 3 * NA_DESC;

}  // namespace

static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

namespace {

std::mutex g_scratch_mu;
std::vector<unsigned char *> g_scratch;  // per device: the address of scratch on it (null until taken)
std::vector<std::string> g_scratch_err;  // per device: why it could not be taken

// Called with the device current.  A failure is kept for messi_inside_noalloc to
// report; it does not stop the context from being created.
void resolve_scratch(int device) {
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if ((int)g_scratch.size() <= device) {
    g_scratch.resize(device + 1, nullptr);
    g_scratch_err.resize(device + 1);
  }
  if (g_scratch[device]) return;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    g_scratch_err[device] = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    return;
  }
  g_scratch[device] = static_cast<unsigned char *>(p);
}

unsigned char *scratch_of(int device) {
  std::lock_guard<std::mutex> g(g_scratch_mu);
  MESSI_REQUIRE(device < (int)g_scratch.size(), "static scratch not resolved on this device");
  MESSI_REQUIRE(g_scratch[device] != nullptr, "static scratch unavailable: " + g_scratch_err[device]);
  return g_scratch[device];
}

// Every buffer of an inside call, borrowed from the scratch at fixed offsets.
void bind_noalloc(InsideWork &w, unsigned char *base) {
  size_t off = 0;
  auto take = [&](DevBuf &b, size_t n) {
    b.borrow(base + off, n);
    off += n;
  };
  for (DevBuf *b : {&w.Lr, &w.Lc, &w.Sr, &w.Sc, &w.F, &w.ptv, &w.paired}) take(*b, NA_TABLE);
  take(w.acc, na_round(NA_CELLS * 8));
  take(w.Lp, na_round(NA_CELLS * 4));
  take(w.Sp, na_round(NA_CELLS * 4));
  take(w.LpE, na_round(NA_ROWS * 4));
  take(w.SpE, na_round(NA_ROWS * 4));
  for (DevBuf *b : {&w.lam, &w.C, &w.uhat, &w.u}) take(*b, NA_VEC);
  take(w.red, NA_RED);
  take(w.out3, 3 * NA_TABLE);
  take(w.inst1, NA_DESC);
  take(w.ib1.inst, NA_DESC);
  take(w.ib1.prefix, NA_DESC);
  if (off > NOALLOC_BYTES) throw std::logic_error("static scratch layout exceeds its size");
  w.ib1.h_inst.clear();
  w.ib1.h_prefix.clear();
}

}  // namespace

extern "C" {

int messi_ctx_create(int device, messi_ctx **out) {
  if (!out) return 2;
  *out = nullptr;
  auto *ctx = new messi_ctx();
  int rc = guard(ctx, [&] {
    int n = 0;
    MESSI_HIP(hipGetDeviceCount(&n));
    MESSI_REQUIRE(device >= 0 && device < n, "no such HIP device");
    MESSI_HIP(hipSetDevice(device));
    ctx->device = device;
    resolve_scratch(device);
    MESSI_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    for (auto &p : ctx->pipe) MESSI_HIP(hipStreamCreateWithFlags(&p, hipStreamNonBlocking));
  });
  *out = ctx;
  return rc;
}

void messi_ctx_destroy(messi_ctx *ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (auto p : ctx->pipe)
    if (p) (void)hipStreamSynchronize(p);
  ctx->w64.~InsideWork();
  new (&ctx->w64) InsideWork();
  ctx->w32.~InsideWork();
  new (&ctx->w32) InsideWork();
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  for (auto p : ctx->pipe)
    if (p) (void)hipStreamDestroy(p);
  delete ctx;
}

const char *messi_last_error(const messi_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int messi_host_alloc(messi_ctx *ctx, size_t bytes, void **out) {
  if (!ctx || !out) return 2;
  *out = nullptr;
  return guard(ctx, [&] {
    MESSI_REQUIRE(bytes > 0, "empty host allocation");
    MESSI_HIP(hipSetDevice(ctx->device));
    MESSI_HIP(hipHostMalloc(out, bytes, hipHostMallocDefault));
  });
}

// Does not touch the context (it may be destroyed already when a caller frees).
int messi_host_free(messi_ctx *ctx, void *p) {
  (void)ctx;
  return p && hipHostFree(p) != hipSuccess ? 1 : 0;
}

int messi_set_profiling(messi_ctx *ctx, int on) {
  if (!ctx) return 2;
  ctx->profiling = on != 0;
  return 0;
}

int messi_set_tile(messi_ctx *ctx, int tile) {
  if (!ctx) return 2;
  return guard(ctx, [&] {
    MESSI_REQUIRE(tile == 16 || tile == 32, "tile must be 16 or 32");
    ctx->tile = tile;
    ctx->w64.tile = tile;
    ctx->w32.tile = tile;
  });
}

int messi_set_bulk_window(messi_ctx *ctx, int window) {
  if (!ctx) return 2;
  return guard(ctx, [&] {
    MESSI_REQUIRE(window == 1 || window == 2 || window == 4, "bulk window must be 1, 2 or 4");
    ctx->window = window;
    ctx->w64.window = window;
    ctx->w32.window = window;
  });
}

int messi_set_band_skip(messi_ctx *ctx, int on) {
  if (!ctx) return 2;
  ctx->band_skip = on != 0;
  return 0;
}

int messi_set_single_launch(messi_ctx *ctx, int on) {
  if (!ctx) return 2;
  ctx->persist = on != 0;
  return 0;
}

int messi_set_pair_fusion(messi_ctx *ctx, int on) {
  if (!ctx) return 2;
  ctx->pair_fuse = on != 0;
  return 0;
}

int messi_selftest(messi_ctx *ctx) {
  if (!ctx) return 2;
  return guard(ctx, [&] {
    MESSI_HIP(hipSetDevice(ctx->device));
    MESSI_REQUIRE(mfma_probe<double>(ctx->stream), "FP64 MFMA 16x16x4 layout check failed");
    MESSI_REQUIRE(mfma_probe<float>(ctx->stream), "FP32 MFMA 16x16x4 layout check failed");
    MESSI_REQUIRE(mfma_bf16_probe(ctx->stream), "bf16 MFMA 16x16x16 layout check failed");
  });
}

int messi_dataset_create(messi_ctx *ctx, int nseq, int ncol, const double *data, int nnodes,
                         const int *children, const int *seqindex, int maxbasepairdistance,
                         const uint8_t *mask, messi_dataset **out) {
  if (!ctx || !out) return 2;
  *out = nullptr;
  auto ds = std::make_unique<messi_dataset>();
  int rc = guard(ctx, [&] {
    MESSI_REQUIRE(data && children && seqindex, "null input array");
    MESSI_HIP(hipSetDevice(ctx->device));
    ds->ctx = ctx;
    ds->nseq = nseq;
    ds->L = ncol;
    ds->nn = nnodes;
    ds->maxbp = maxbasepairdistance < 0 ? 0 : maxbasepairdistance;
    build_dataset(ds.get(), data, children, seqindex, mask);
  });
  if (rc == 0) *out = ds.release();
  return rc;
}

void messi_dataset_destroy(messi_dataset *ds) {
  if (!ds) return;
  (void)hipSetDevice(ds->ctx->device);
  (void)hipStreamSynchronize(ds->ctx->stream);
  delete ds;
}

int messi_dataset_stats(const messi_dataset *ds, int64_t *stats) {
  if (!ds || !stats) return 2;
  stats[0] = ds->sumU;
  stats[1] = ds->U_root;
  stats[2] = ds->inrange;
  stats[3] = ds->sumN;
  stats[4] = ds->H;
  stats[5] = (int64_t)(ds->d_site_idx.bytes + ds->d_site_leaf.bytes + ds->d_pair_idx.bytes +
                       ds->d_pair_leaf.bytes + ds->d_pid0.bytes + ds->d_T0.bytes + ds->d_mask.bytes);
  return 0;
}

int messi_site_loglik(messi_dataset *ds, int K, const double *P4, const double *pi4, const double *logw, int prec,
                      double *u_out) {
  if (!ds) return 2;
  messi_ctx *ctx = ds->ctx;
  return guard(ctx, [&] {
    MESSI_REQUIRE(P4 && pi4 && logw && u_out, "null argument");
    check_prec(prec);
    MESSI_HIP(hipSetDevice(ctx->device));
    MESSI_REQUIRE(K >= 1, "need at least one site component");
    messi_batch *b = self_batch(ds);
    if (prec == MESSI_FP32) {
      load_components<float>(b, b->c4, K, 16, P4, pi4, 4, logw, prec);
      run_site<float>(b, prec);
    } else {
      load_components<double>(b, b->c4, K, 16, P4, pi4, 4, logw, prec);
      run_site<double>(b, prec);
    }
    InsideWork &w = member_work(b, 0, prec);
    MESSI_HIP(hipMemcpyAsync(u_out, w.u.p, sizeof(double) * ds->L, hipMemcpyDeviceToHost, ctx->stream));
    MESSI_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int messi_pair_loglik(messi_dataset *ds, int K, const double *P16, const double *pi16, const double *logw, int prec,
                      double *out, double *percomp) {
  if (!ds) return 2;
  messi_ctx *ctx = ds->ctx;
  return guard(ctx, [&] {
    MESSI_REQUIRE(P16 && pi16 && logw && out, "null argument");
    check_prec(prec);
    MESSI_HIP(hipSetDevice(ctx->device));
    MESSI_REQUIRE(K >= 1, "need at least one pair component");
    messi_batch *b = self_batch(ds);
    if (prec == MESSI_FP32) {
      load_components<float>(b, b->c16, K, 256, P16, pi16, 16, logw, prec);
      run_pair<float>(b, prec, percomp, true);
    } else {
      load_components<double>(b, b->c16, K, 256, P16, pi16, 16, logw, prec);
      run_pair<double>(b, prec, percomp, true);
    }
    InsideWork &w = member_work(b, 0, prec);
    MESSI_HIP(hipMemcpyAsync(out, w.paired.p, sizeof(double) * ds->L * ds->L, hipMemcpyDeviceToHost,
                             ctx->stream));
    MESSI_HIP(hipStreamSynchronize(ctx->stream));
  });
}

// Largest y - j (j < y) with a finite pair score in a column-major L x L matrix
// (0 when there is none), or -1 when it reaches the full span.  The inside and
// outside passes skip the terms that a smaller band makes -inf.
static int pairing_band(int L, const double *paired) {
  int band = 0;
  for (int y = 1; y < L; ++y) {
    const double *col = paired + (size_t)y * L;
    for (int j = 0; j < y - band; ++j)
      if (col[j] != -HUGE_VAL) {
        band = y - j;
        break;
      }
  }
  return band >= L - 1 ? -1 : band;
}

// The inside call of messi_inside and messi_inside_noalloc on the work w.
static void inside_call(messi_ctx *ctx, InsideWork &w, int L, const double *u, const double *paired, double B,
                        int zonly, int prec, double *Z, double *in3) {
  MESSI_REQUIRE(L >= 1 && u && paired && Z, "bad argument");
  MESSI_REQUIRE(B > 0.0, "B must be positive");
  MESSI_REQUIRE(zonly || in3, "in3 required when zonly == 0");
  MESSI_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  w.tile = ctx->tile;
  w.window = ctx->window;
  w.persist = ctx->persist && &w != &ctx->wna;  // the allocation-free call keeps the launch sequence
  w.band = ctx->band_skip ? pairing_band(L, paired) : -1;
  const size_t LL = (size_t)L * L;
  auto go = [&](auto tag) {
    using T = decltype(tag);
    inside_alloc<T>(w, L, false);
    MESSI_HIP(hipMemcpyAsync(w.u.p, u, sizeof(double) * L, hipMemcpyHostToDevice, s));
    MESSI_HIP(hipMemcpyAsync(w.paired.p, paired, sizeof(double) * LL, hipMemcpyHostToDevice, s));
    inside_run<T>(s, w, L, B);
    *Z = inside_logZ<T>(s, w, L);
    MESSI_REQUIRE(!(w.persist && std::isnan(*Z)),
                  "single-launch inside gave up waiting for a dependency (log Z not computed)");
    if (!zonly) {
      w.out3.ensure(sizeof(double) * 3 * LL);
      inside_export<T>(s, w, L, w.out3.as<double>());
      MESSI_HIP(hipMemcpyAsync(in3, w.out3.p, sizeof(double) * 3 * LL, hipMemcpyDeviceToHost, s));
    }
    MESSI_HIP(hipStreamSynchronize(s));
  };
  if (prec == MESSI_FP32) go(float{});
  else go(double{});
}

int messi_inside(messi_ctx *ctx, int L, const double *u, const double *paired, double B, int zonly, int prec,
                 double *Z, double *in3) {
  if (!ctx) return 2;
  return guard(ctx, [&] { inside_call(ctx, work(ctx, prec), L, u, paired, B, zonly, prec, Z, in3); });
}

int messi_inside_noalloc(messi_ctx *ctx, int L, const double *u, const double *paired, double B, int zonly,
                         int prec, double *Z, double *in3) {
  if (!ctx) return 2;
  return guard(ctx, [&] {
    MESSI_REQUIRE(L >= 1 && L <= MESSI_NOALLOC_MAX_L,
                  "messi_inside_noalloc supports 1 <= L <= " + std::to_string(MESSI_NOALLOC_MAX_L) +
                      " (its static scratch holds the tables of that length)");
    check_prec(prec);
    MESSI_HIP(hipSetDevice(ctx->device));
    bind_noalloc(ctx->wna, scratch_of(ctx->device));
    inside_call(ctx, ctx->wna, L, u, paired, B, zonly, prec, Z, in3);
  });
}

int messi_outside(messi_ctx *ctx, int L, const double *in3, const double *u, const double *paired, int prec,
                  double *out3) {
  if (!ctx) return 2;
  return guard(ctx, [&] {
    MESSI_REQUIRE(L >= 1 && in3 && u && paired && out3, "bad argument");
    MESSI_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    InsideWork &w = work(ctx, prec);
    w.band = ctx->band_skip ? pairing_band(L, paired) : -1;
    const size_t LL = (size_t)L * L;
    auto go = [&](auto tag) {
      using T = decltype(tag);
      inside_alloc<T>(w, L, true);
      w.out3.ensure(sizeof(double) * 3 * LL);
      MESSI_HIP(hipMemcpyAsync(w.u.p, u, sizeof(double) * L, hipMemcpyHostToDevice, s));
      MESSI_HIP(hipMemcpyAsync(w.paired.p, paired, sizeof(double) * LL, hipMemcpyHostToDevice, s));
      MESSI_HIP(hipMemcpyAsync(w.out3.p, in3, sizeof(double) * 3 * LL, hipMemcpyHostToDevice, s));
      inside_prepare<T>(s, w, L, 1.0);
      inside_import<T>(s, w, L, w.out3.as<double>());
      outside_run<T>(s, w, L, true);
      outside_export<T>(s, w, L, w.out3.as<double>());
      MESSI_HIP(hipMemcpyAsync(out3, w.out3.p, sizeof(double) * 3 * LL, hipMemcpyDeviceToHost, s));
      MESSI_HIP(hipStreamSynchronize(s));
    };
    if (prec == MESSI_FP32) go(float{});
    else go(double{});
  });
}

int messi_bpp(messi_ctx *ctx, int L, const double *in3, const double *out3, const double *u, const double *paired,
              int prec, double *unpaired, double *pairprob) {
  if (!ctx) return 2;
  return guard(ctx, [&] {
    MESSI_REQUIRE(L >= 1 && in3 && out3 && u && paired && pairprob, "bad argument");
    MESSI_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    InsideWork &w = work(ctx, prec);
    const size_t LL = (size_t)L * L;
    auto go = [&](auto tag) {
      using T = decltype(tag);
      inside_alloc<T>(w, L, true);
      w.out3.ensure(sizeof(double) * 3 * LL);
      DevBuf du, dp;
      du.ensure(sizeof(double) * L);
      dp.ensure(sizeof(double) * LL);
      MESSI_HIP(hipMemcpyAsync(w.u.p, u, sizeof(double) * L, hipMemcpyHostToDevice, s));
      MESSI_HIP(hipMemcpyAsync(w.paired.p, paired, sizeof(double) * LL, hipMemcpyHostToDevice, s));
      inside_prepare<T>(s, w, L, 1.0);
      MESSI_HIP(hipMemcpyAsync(w.out3.p, in3, sizeof(double) * 3 * LL, hipMemcpyHostToDevice, s));
      inside_import<T>(s, w, L, w.out3.as<double>());
      MESSI_HIP(hipMemcpyAsync(w.out3.p, out3, sizeof(double) * 3 * LL, hipMemcpyHostToDevice, s));
      outside_import<T>(s, w, L, w.out3.as<double>());
      bpp_run<T>(s, w, L, du.as<double>(), dp.as<double>());
      if (unpaired) MESSI_HIP(hipMemcpyAsync(unpaired, du.p, sizeof(double) * L, hipMemcpyDeviceToHost, s));
      MESSI_HIP(hipMemcpyAsync(pairprob, dp.p, sizeof(double) * LL, hipMemcpyDeviceToHost, s));
      MESSI_HIP(hipStreamSynchronize(s));
    };
    if (prec == MESSI_FP32) go(float{});
    else go(double{});
  });
}

int messi_eval(messi_dataset *ds, int K4, const double *P4, const double *pi4, const double *logw4, int K16,
               const double *P16, const double *pi16, const double *logw16, double B, int prec, double *logZ) {
  if (!ds) return 2;
  messi_ctx *ctx = ds->ctx;
  return guard(ctx, [&] {
    MESSI_REQUIRE(P4 && pi4 && logw4 && P16 && pi16 && logw16 && logZ, "null argument");
    MESSI_REQUIRE(B > 0.0, "B must be positive");
    check_prec(prec);
    MESSI_HIP(hipSetDevice(ctx->device));
    messi_batch *b = self_batch(ds);
    if (prec == MESSI_FP32) {
      load_impl<float>(b, K4, P4, pi4, logw4, K16, P16, pi16, logw16);
      eval_impl<float>(b, B, false, logZ);
    } else {
      load_impl<double>(b, K4, P4, pi4, logw4, K16, P16, pi16, logw16);
      eval_impl<double>(b, B, false, logZ);
    }
  });
}

int messi_posterior(messi_dataset *ds, int K4, const double *P4, const double *pi4, const double *logw4, int K16,
                    const double *P16, const double *pi16, const double *logw16, int prec, double *logZ,
                    double *unpaired, double *pairprob) {
  if (!ds) return 2;
  messi_ctx *ctx = ds->ctx;
  return guard(ctx, [&] {
    MESSI_REQUIRE(P4 && pi4 && logw4 && P16 && pi16 && logw16 && logZ, "null argument");
    check_prec(prec);
    MESSI_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    const int L = ds->L;
    const size_t LL = (size_t)L * L;
    messi_batch *b = self_batch(ds);
    auto go = [&](auto tag) {
      using T = decltype(tag);
      InsideWork &w = member_work(b, 0, sizeof(T) == 4 ? MESSI_FP32 : MESSI_FP64);
      load_impl<T>(b, K4, P4, pi4, logw4, K16, P16, pi16, logw16);
      eval_impl<T>(b, 1.0, true, logZ);
      outside_run<T>(s, w, L, false);
      record(b, 4);
      w.out3.ensure(sizeof(double) * (LL + L));
      double *dpair = w.out3.as<double>();
      double *dunp = dpair + LL;
      bpp_run<T>(s, w, L, dunp, dpair);
      record(b, 5);
      if (unpaired) MESSI_HIP(hipMemcpyAsync(unpaired, dunp, sizeof(double) * L, hipMemcpyDeviceToHost, s));
      if (pairprob) MESSI_HIP(hipMemcpyAsync(pairprob, dpair, sizeof(double) * LL, hipMemcpyDeviceToHost, s));
      MESSI_HIP(hipStreamSynchronize(s));
      collect(b, 5);
    };
    if (prec == MESSI_FP32) go(float{});
    else go(double{});
  });
}

int messi_mea(messi_ctx *ctx, int L, const double *unpaired, const double *pairprob, double alpha, int prec,
              int *paired, double *emat, int *smat) {
  if (!ctx) return 2;
  return guard(ctx, [&] {
    MESSI_REQUIRE(L >= 1 && L <= MEA_MAX_L, "MEA decoding supports 1 <= L <= 16383");
    MESSI_REQUIRE(unpaired && pairprob && paired, "null argument");
    MESSI_REQUIRE(std::isfinite(alpha), "alpha must be finite");
    check_prec(prec);
    MESSI_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    MeaWork &w = ctx->mea;
    const size_t LL = (size_t)L * L;
    w.single.ensure(sizeof(double) * L);
    w.pair.ensure(sizeof(double) * LL);
    w.paired.ensure(sizeof(int) * L);
    if (ctx->profiling) MESSI_HIP(hipEventRecord(mea_events(ctx).e[0], s));
    MESSI_HIP(hipMemcpyAsync(w.single.p, unpaired, sizeof(double) * L, hipMemcpyHostToDevice, s));
    MESSI_HIP(hipMemcpyAsync(w.pair.p, pairprob, sizeof(double) * LL, hipMemcpyHostToDevice, s));
    if (prec == MESSI_FP32)
      mea_timed<float>(ctx, L, w.single.as<double>(), w.pair.as<double>(), alpha, w.paired.as<int>());
    else
      mea_timed<double>(ctx, L, w.single.as<double>(), w.pair.as<double>(), alpha, w.paired.as<int>());
    MESSI_HIP(hipMemcpyAsync(paired, w.paired.p, sizeof(int) * L, hipMemcpyDeviceToHost, s));
    if (emat || smat) {
      DevBuf de, ds;
      if (emat) de.ensure(sizeof(double) * LL);
      if (smat) ds.ensure(sizeof(int) * LL);
      mea_export(s, w, L, emat ? de.as<double>() : nullptr, smat ? ds.as<int>() : nullptr);
      if (emat) MESSI_HIP(hipMemcpyAsync(emat, de.p, sizeof(double) * LL, hipMemcpyDeviceToHost, s));
      if (smat) MESSI_HIP(hipMemcpyAsync(smat, ds.p, sizeof(int) * LL, hipMemcpyDeviceToHost, s));
      MESSI_HIP(hipStreamSynchronize(s));
    }
    MESSI_HIP(hipStreamSynchronize(s));
    mea_collect(ctx, true);
  });
}

int messi_mea_timings(messi_ctx *ctx, double *t) {
  if (!ctx || !t) return 2;
  const MeaEvents &m = mea_events(ctx);
  for (int q = 0; q < 3; ++q) t[q] = m.ms[q];
  return 0;
}

int messi_posterior_consensus(messi_dataset *ds, int K4, const double *P4, const double *pi4, const double *logw4,
                              int K16, const double *P16, const double *pi16, const double *logw16, int prec,
                              int mea_prec, double alpha, double *logZ, int *paired) {
  if (!ds) return 2;
  messi_ctx *ctx = ds->ctx;
  return guard(ctx, [&] {
    MESSI_REQUIRE(P4 && pi4 && logw4 && P16 && pi16 && logw16 && logZ && paired, "null argument");
    MESSI_REQUIRE(std::isfinite(alpha), "alpha must be finite");
    check_prec(prec);
    check_prec(mea_prec);
    const int L = ds->L;
    MESSI_REQUIRE(L <= MEA_MAX_L, "MEA decoding supports L <= 16383");
    MESSI_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    const size_t LL = (size_t)L * L;
    messi_batch *b = self_batch(ds);
    auto go = [&](auto tag) {
      using T = decltype(tag);
      InsideWork &w = member_work(b, 0, sizeof(T) == 4 ? MESSI_FP32 : MESSI_FP64);
      load_impl<T>(b, K4, P4, pi4, logw4, K16, P16, pi16, logw16);
      eval_impl<T>(b, 1.0, true, logZ);
      outside_run<T>(s, w, L, false);
      record(b, 4);
      w.out3.ensure(sizeof(double) * (LL + L));
      double *dpair = w.out3.as<double>();
      double *dunp = dpair + LL;
      bpp_run<T>(s, w, L, dunp, dpair);
      record(b, 5);
      collect(b, 5);
      ctx->mea.paired.ensure(sizeof(int) * L);
      if (ctx->profiling) MESSI_HIP(hipEventRecord(mea_events(ctx).e[0], s));
      if (mea_prec == MESSI_FP32) mea_timed<float>(ctx, L, dunp, dpair, alpha, ctx->mea.paired.as<int>());
      else mea_timed<double>(ctx, L, dunp, dpair, alpha, ctx->mea.paired.as<int>());
      MESSI_HIP(hipMemcpyAsync(paired, ctx->mea.paired.p, sizeof(int) * L, hipMemcpyDeviceToHost, s));
      MESSI_HIP(hipStreamSynchronize(s));
      mea_collect(ctx, false);
    };
    if (prec == MESSI_FP32) go(float{});
    else go(double{});
  });
}

int messi_structure_loglik(messi_dataset *ds, int K4, const double *P4, const double *pi4, const double *logw4,
                           int K16, const double *P16, const double *pi16, const double *logw16,
                           const int *paired, int prec, double *loglik, double *parts) {
  if (!ds) return 2;
  messi_ctx *ctx = ds->ctx;
  return guard(ctx, [&] {
    MESSI_REQUIRE(P4 && pi4 && logw4 && P16 && pi16 && logw16 && paired && loglik, "null argument");
    check_prec(prec);
    const int L = ds->L;
    for (int x = 0; x < L; ++x) {
      const int p = paired[x];
      MESSI_REQUIRE(p >= 0 && p <= L && p != x + 1, "paired must hold 1-based partners (0 unpaired)");
      MESSI_REQUIRE(p == 0 || paired[p - 1] == x + 1, "paired is not symmetric");
    }
    MESSI_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    messi_batch *b = self_batch(ds);
    ds->d_partner.ensure(sizeof(int) * L);
    ds->d_sums.ensure(sizeof(double) * 4);
    MESSI_HIP(hipMemcpyAsync(ds->d_partner.p, paired, sizeof(int) * L, hipMemcpyHostToDevice, s));
    auto go = [&](auto tag) {
      using T = decltype(tag);
      load_impl<T>(b, K4, P4, pi4, logw4, K16, P16, pi16, logw16);
      run_site<T>(b, prec);
      run_pair<T>(b, prec, nullptr, false);
      const ExpandDesc *dexp = upload_expand(b, prec, true);
      launch_structure_sum(s, dexp, b->pair_acc.as<double>(), ds->d_partner.as<int>(), ds->d_sums.as<double>());
      MESSI_HIP(hipGetLastError());
    };
    if (prec == MESSI_FP32) go(float{});
    else go(double{});
    double h[3];
    MESSI_HIP(hipMemcpyAsync(h, ds->d_sums.p, sizeof(double) * 3, hipMemcpyDeviceToHost, s));
    MESSI_HIP(hipStreamSynchronize(s));
    MESSI_REQUIRE(h[2] == 0.0, "the dataset's pair list lacks a pair of the structure (distance bound or mask)");
    *loglik = h[0] + h[1];
    if (parts) {
      parts[0] = h[0];
      parts[1] = h[1];
    }
  });
}

int messi_sample_states(messi_dataset *ds, int K4, const double *P4, const double *pi4, const double *logw4, int K16,
                        const double *P16, const double *pi16, const double *logw16, int cats, const int *paired,
                        double B, uint64_t seed, int prec, int *states, double *unpaired, double *pairedll,
                        double *siteconds, double *paircomps) {
  if (!ds) return 2;
  messi_ctx *ctx = ds->ctx;
  return guard(ctx, [&] {
    MESSI_REQUIRE(P4 && pi4 && logw4 && P16 && pi16 && logw16 && paired && states && unpaired && pairedll,
                  "null argument");
    check_prec(prec);
    MESSI_REQUIRE(cats >= 1 && cats <= 8, "cats (site categories) must be 1 .. 8");
    MESSI_REQUIRE(K4 == cats, "K4 must equal the number of site categories");
    MESSI_REQUIRE(K16 >= cats * cats && K16 % (cats * cats) == 0,
                  "K16 must be cats^2 times the number of lambda classes");
    MESSI_REQUIRE(B > 0.0 && std::isfinite(B), "B must be positive");
    const int L = ds->L;
    for (int x = 0; x < L; ++x) {
      const int p = paired[x];
      MESSI_REQUIRE(p >= 0 && p <= L && p != x + 1, "paired must hold 1-based partners (0 unpaired)");
      MESSI_REQUIRE(p == 0 || paired[p - 1] == x + 1, "paired is not symmetric");
    }
    MESSI_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    messi_batch *b = self_batch(ds);
    ds->d_partner.ensure(sizeof(int) * L);
    ds->d_sitec.ensure(sizeof(double) * (size_t)K4 * L);
    ds->d_pairc.ensure(sizeof(double) * (size_t)K16 * L);
    ds->d_sout.ensure((sizeof(int) + 2 * sizeof(double)) * (size_t)L);
    ds->d_miss.ensure(sizeof(int));
    MESSI_HIP(hipMemcpyAsync(ds->d_partner.p, paired, sizeof(int) * L, hipMemcpyHostToDevice, s));
    MESSI_HIP(hipMemsetAsync(ds->d_miss.p, 0, sizeof(int), s));
    double *sitec = ds->d_sitec.as<double>(), *pairc = ds->d_pairc.as<double>();
    double *d_unp = ds->d_sout.as<double>(), *d_pll = d_unp + L;
    int *d_states = reinterpret_cast<int *>(d_pll + L);
    auto go = [&](auto tag) {
      using T = decltype(tag);
      load_impl<T>(b, K4, P4, pi4, logw4, K16, P16, pi16, logw16);
      const ExpandDesc *dpair = nullptr;
      const int64_t pr = component_roots<T>(b, prec, sitec, &dpair);
      launch_pair_gather(s, L, K16, pr, dpair, b->pair_comp.as<double>(), ds->d_partner.as<int>(), pairc,
                         ds->d_miss.as<int>());
      MESSI_HIP(hipGetLastError());
    };
    if (prec == MESSI_FP32) go(float{});
    else go(double{});
    launch_state_sample(s, L, cats, K16 / (cats * cats), sitec, pairc, b->c16.logw.as<double>(),
                        ds->d_partner.as<int>(), B, seed, d_states, d_unp, d_pll);
    int miss = 0;
    MESSI_HIP(hipMemcpyAsync(&miss, ds->d_miss.p, sizeof(int), hipMemcpyDeviceToHost, s));
    MESSI_HIP(hipMemcpyAsync(states, d_states, sizeof(int) * L, hipMemcpyDeviceToHost, s));
    MESSI_HIP(hipMemcpyAsync(unpaired, d_unp, sizeof(double) * L, hipMemcpyDeviceToHost, s));
    MESSI_HIP(hipMemcpyAsync(pairedll, d_pll, sizeof(double) * L, hipMemcpyDeviceToHost, s));
    if (siteconds)
      MESSI_HIP(hipMemcpyAsync(siteconds, sitec, sizeof(double) * (size_t)K4 * L, hipMemcpyDeviceToHost, s));
    if (paircomps)
      MESSI_HIP(hipMemcpyAsync(paircomps, pairc, sizeof(double) * (size_t)K16 * L, hipMemcpyDeviceToHost, s));
    MESSI_HIP(hipStreamSynchronize(s));
    MESSI_REQUIRE(miss == 0, "the dataset's pair list lacks a pair of the structure (distance bound or mask)");
  });
}

int messi_site_rate_posteriors(messi_dataset *ds, int K4, const double *P4, const double *pi4, const double *logw4,
                               int K16, const double *P16, const double *pi16, const double *logw16, int cats,
                               const double *unpaired, const double *pairprob, int prec, double *out) {
  if (!ds) return 2;
  messi_ctx *ctx = ds->ctx;
  return guard(ctx, [&] {
    MESSI_REQUIRE(P4 && pi4 && logw4 && P16 && pi16 && logw16 && unpaired && pairprob && out, "null argument");
    check_prec(prec);
    MESSI_REQUIRE(cats >= 1 && cats <= 8, "cats (site categories) must be 1 .. 8");
    MESSI_REQUIRE(K4 == cats, "K4 must equal the number of site categories");
    MESSI_REQUIRE(K16 >= cats * cats && K16 % (cats * cats) == 0,
                  "K16 must be cats^2 times the number of lambda classes");
    const int L = ds->L;
    MESSI_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    messi_batch *b = self_batch(ds);
    ds->d_sitec.ensure(sizeof(double) * (size_t)K4 * L);
    ds->d_sout.ensure(sizeof(double) * ((size_t)L * L + (size_t)L + (size_t)cats * L));
    double *sitec = ds->d_sitec.as<double>();
    double *d_pp = ds->d_sout.as<double>(), *d_single = d_pp + (size_t)L * L, *d_out = d_single + L;
    MESSI_HIP(hipMemcpyAsync(d_pp, pairprob, sizeof(double) * L * L, hipMemcpyHostToDevice, s));
    MESSI_HIP(hipMemcpyAsync(d_single, unpaired, sizeof(double) * L, hipMemcpyHostToDevice, s));
    auto go = [&](auto tag) {
      using T = decltype(tag);
      load_impl<T>(b, K4, P4, pi4, logw4, K16, P16, pi16, logw16);
      const ExpandDesc *dpair = nullptr;
      const int64_t pr = component_roots<T>(b, prec, sitec, &dpair);
      launch_site_rate_posterior(s, L, cats, K16, pr, dpair, b->pair_comp.as<double>(), b->c16.logw.as<double>(),
                                 sitec, d_single, d_pp, d_out);
      MESSI_HIP(hipGetLastError());
    };
    if (prec == MESSI_FP32) go(float{});
    else go(double{});
    MESSI_HIP(hipMemcpyAsync(out, d_out, sizeof(double) * (size_t)cats * L, hipMemcpyDeviceToHost, s));
    MESSI_HIP(hipStreamSynchronize(s));
  });
}

int messi_sample_structures(messi_ctx *ctx, int L, const double *in3, const double *paired, const double *unpaired,
                            int prec, int nsamples, uint64_t seed, double B, int *structures, double *logprior) {
  if (!ctx) return 2;
  return guard(ctx, [&] {
    MESSI_REQUIRE(L >= 1 && L <= SAMPLE_MAX_L, "structure sampling supports 1 <= L <= 16383");
    MESSI_REQUIRE(in3 && paired && unpaired && structures && nsamples >= 0, "bad argument");
    MESSI_REQUIRE(std::isfinite(B), "B must be finite");
    check_prec(prec);
    MESSI_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    InsideWork &w = work(ctx, prec);
    const size_t LL = (size_t)L * L;
    SampleWork &sw = ctx->samp;
    sw.out.ensure(sizeof(int) * (size_t)nsamples * L);
    sw.prior.ensure(sizeof(double) * (size_t)nsamples);
    auto go = [&](auto tag) {
      using T = decltype(tag);
      inside_alloc<T>(w, L, false);
      w.out3.ensure(sizeof(double) * 3 * LL);
      MESSI_HIP(hipMemcpyAsync(w.u.p, unpaired, sizeof(double) * L, hipMemcpyHostToDevice, s));
      MESSI_HIP(hipMemcpyAsync(w.paired.p, paired, sizeof(double) * LL, hipMemcpyHostToDevice, s));
      inside_prepare<T>(s, w, L, 1.0);
      MESSI_HIP(hipMemcpyAsync(w.out3.p, in3, sizeof(double) * 3 * LL, hipMemcpyHostToDevice, s));
      inside_import<T>(s, w, L, w.out3.as<double>());
      sample_run<T>(s, w, sw, L, nsamples, seed, 0, B, sw.out.as<int>(), sw.prior.as<double>());
    };
    if (prec == MESSI_FP32) go(float{});
    else go(double{});
    if (nsamples > 0) {
      MESSI_HIP(hipMemcpyAsync(structures, sw.out.p, sizeof(int) * (size_t)nsamples * L, hipMemcpyDeviceToHost, s));
      if (logprior)
        MESSI_HIP(hipMemcpyAsync(logprior, sw.prior.p, sizeof(double) * nsamples, hipMemcpyDeviceToHost, s));
    }
    MESSI_HIP(hipStreamSynchronize(s));
  });
}

int messi_sample_posterior(messi_dataset *ds, int K4, const double *P4, const double *pi4, const double *logw4,
                           int K16, const double *P16, const double *pi16, const double *logw16, int prec,
                           int nsamples, uint64_t seed, double B, double *logZ, int *structures, double *logprior) {
  if (!ds) return 2;
  messi_ctx *ctx = ds->ctx;
  return guard(ctx, [&] {
    MESSI_REQUIRE(P4 && pi4 && logw4 && P16 && pi16 && logw16 && logZ && structures && nsamples >= 0,
                  "bad argument");
    MESSI_REQUIRE(std::isfinite(B), "B must be finite");
    check_prec(prec);
    const int L = ds->L;
    MESSI_REQUIRE(L <= SAMPLE_MAX_L, "structure sampling supports L <= 16383");
    MESSI_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    messi_batch *b = self_batch(ds);
    SampleWork &sw = ctx->samp;
    sw.out.ensure(sizeof(int) * (size_t)nsamples * L);
    sw.prior.ensure(sizeof(double) * (size_t)nsamples);
    auto go = [&](auto tag) {
      using T = decltype(tag);
      InsideWork &w = member_work(b, 0, sizeof(T) == 4 ? MESSI_FP32 : MESSI_FP64);
      load_impl<T>(b, K4, P4, pi4, logw4, K16, P16, pi16, logw16);
      eval_impl<T>(b, 1.0, true, logZ);
      sample_run<T>(s, w, sw, L, nsamples, seed, 0, B, sw.out.as<int>(), sw.prior.as<double>());
    };
    if (prec == MESSI_FP32) go(float{});
    else go(double{});
    if (nsamples > 0) {
      MESSI_HIP(hipMemcpyAsync(structures, sw.out.p, sizeof(int) * (size_t)nsamples * L, hipMemcpyDeviceToHost, s));
      if (logprior)
        MESSI_HIP(hipMemcpyAsync(logprior, sw.prior.p, sizeof(double) * nsamples, hipMemcpyDeviceToHost, s));
    }
    MESSI_HIP(hipStreamSynchronize(s));
  });
}

int messi_last_timings(const messi_dataset *ds, double *t) {
  if (!ds || !t) return 2;
  for (int i = 0; i < 5; ++i) t[i] = ds->self ? ds->self->t[i] : 0.0;
  return 0;
}

int messi_batch_create(messi_ctx *ctx, int n, messi_dataset *const *members, messi_batch **out) {
  if (!ctx || !out) return 2;
  *out = nullptr;
  auto b = std::make_unique<messi_batch>();
  int rc = guard(ctx, [&] {
    MESSI_REQUIRE(n >= 1 && members, "a batch needs at least one member");
    MESSI_HIP(hipSetDevice(ctx->device));
    b->ctx = ctx;
    b->n = n;
    for (int m = 0; m < n; ++m) {
      MESSI_REQUIRE(members[m] && members[m]->ctx == ctx, "batch members must belong to the context");
      b->mem.push_back(members[m]);
    }
    build_batch(b.get());
  });
  if (rc == 0) *out = b.release();
  return rc;
}

void messi_batch_destroy(messi_batch *b) {
  if (!b) return;
  (void)hipSetDevice(b->ctx->device);
  (void)hipStreamSynchronize(b->ctx->stream);
  for (auto p : b->ctx->pipe)
    if (p) (void)hipStreamSynchronize(p);
  delete b;
}

int messi_batch_eval(messi_batch *b, int K4, const double *P4, const double *pi4, const double *logw4, int K16,
                     const double *P16, const double *pi16, const double *logw16, double B, int prec,
                     double *logZ) {
  if (!b) return 2;
  messi_ctx *ctx = b->ctx;
  return guard(ctx, [&] {
    MESSI_REQUIRE(P4 && pi4 && logw4 && P16 && pi16 && logw16 && logZ, "null argument");
    MESSI_REQUIRE(B > 0.0, "B must be positive");
    check_prec(prec);
    MESSI_HIP(hipSetDevice(ctx->device));
    if (prec == MESSI_FP32) {
      load_impl<float>(b, K4, P4, pi4, logw4, K16, P16, pi16, logw16);
      eval_impl<float>(b, B, false, logZ);
    } else {
      load_impl<double>(b, K4, P4, pi4, logw4, K16, P16, pi16, logw16);
      eval_impl<double>(b, B, false, logZ);
    }
  });
}

int messi_batch_load(messi_batch *b, int K4, const double *P4, const double *pi4, const double *logw4, int K16,
                     const double *P16, const double *pi16, const double *logw16, int prec) {
  if (!b) return 2;
  messi_ctx *ctx = b->ctx;
  return guard(ctx, [&] {
    MESSI_REQUIRE(P4 && pi4 && logw4 && P16 && pi16 && logw16, "null argument");
    check_prec(prec);
    MESSI_HIP(hipSetDevice(ctx->device));
    if (prec == MESSI_FP32) load_impl<float>(b, K4, P4, pi4, logw4, K16, P16, pi16, logw16);
    else load_impl<double>(b, K4, P4, pi4, logw4, K16, P16, pi16, logw16);
    MESSI_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int messi_batch_load_eigen(messi_batch *b, int K4, const double *V4, const double *w4, const double *Vi4,
                           const double *pi4, const double *logw4, int K16, const double *V16, const double *w16,
                           const double *Vi16, const double *pi16, const double *logw16,
                           const double *branchlengths, int prec) {
  if (!b) return 2;
  messi_ctx *ctx = b->ctx;
  return guard(ctx, [&] {
    MESSI_REQUIRE(V4 && w4 && Vi4 && pi4 && logw4 && V16 && w16 && Vi16 && pi16 && logw16 && branchlengths,
                  "null argument");
    check_prec(prec);
    MESSI_HIP(hipSetDevice(ctx->device));
    if (prec == MESSI_FP32)
      load_eigen_impl<float>(b, K4, V4, w4, Vi4, pi4, logw4, K16, V16, w16, Vi16, pi16, logw16, branchlengths);
    else
      load_eigen_impl<double>(b, K4, V4, w4, Vi4, pi4, logw4, K16, V16, w16, Vi16, pi16, logw16, branchlengths);
    MESSI_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int messi_batch_eval_loaded(messi_batch *b, double B, int prec, double *logZ) {
  if (!b) return 2;
  messi_ctx *ctx = b->ctx;
  return guard(ctx, [&] {
    MESSI_REQUIRE(logZ, "null argument");
    MESSI_REQUIRE(B > 0.0, "B must be positive");
    check_prec(prec);
    MESSI_HIP(hipSetDevice(ctx->device));
    if (prec == MESSI_FP32) eval_impl<float>(b, B, false, logZ);
    else eval_impl<double>(b, B, false, logZ);
  });
}

int messi_batch_set_tiles(messi_batch *b, const int *tiles) {
  if (!b) return 2;
  return guard(b->ctx, [&] {
    if (!tiles) {
      b->tiles.clear();
      return;
    }
    for (int m = 0; m < b->n; ++m)
      MESSI_REQUIRE(tiles[m] == 0 || tiles[m] == 16 || tiles[m] == 32, "tile must be 0, 16 or 32");
    b->tiles.assign(tiles, tiles + b->n);
  });
}

int messi_batch_set_groups(messi_batch *b, int groups) {
  if (!b) return 2;
  return guard(b->ctx, [&] {
    MESSI_REQUIRE(groups >= 0 && groups <= messi_ctx::NPIPE + 1, "groups must be 0 (automatic) .. 4");
    b->groups = groups;
  });
}

int messi_batch_set_window(messi_batch *b, int window) {
  if (!b) return 2;
  return guard(b->ctx, [&] {
    MESSI_REQUIRE(window == 0 || window == 1 || window == 2 || window == 4, "bulk window must be 0, 1, 2 or 4");
    b->window = window;
  });
}

int messi_batch_timings(const messi_batch *b, double *t) {
  if (!b || !t) return 2;
  for (int i = 0; i < 5; ++i) t[i] = b->t[i];
  return 0;
}

int messi_batch_step_timings(messi_batch *b, double *ms, int *launches) {
  if (!b || !ms || !launches) return 2;
  return guard(b->ctx, [&] {
    MESSI_REQUIRE(b->sev.made, "no profiled evaluation on this batch");
    for (int k = 0; k < 2; ++k) {
      float v = 0.0f;
      if (b->steps_launches[k] > 0) {
        MESSI_HIP(hipEventSynchronize(b->sev.e[2 * k + 1]));
        MESSI_HIP(hipEventElapsedTime(&v, b->sev.e[2 * k], b->sev.e[2 * k + 1]));
      }
      ms[k] = v;
      launches[k] = b->steps_launches[k];
    }
  });
}

}  // extern "C"
