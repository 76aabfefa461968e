// C-ABI of libmessi: context, dataset (pattern dedup and pair index lists built
// natively at creation), and the likelihood entry points.
#include <algorithm>
#include <cstring>
#include <memory>

#include "common.h"
#include "prune.h"

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

}  // namespace

struct messi_dataset {
  messi_ctx *ctx = nullptr;
  int nseq = 0, L = 0, nn = 0, maxbp = 0, H = 0;
  std::vector<int> ch, seq, height, npat;
  std::vector<std::vector<int>> pid, rep;
  std::vector<int> site_lv, pair_lv;  // item offsets per level (H+1 entries)
  int64_t site_entries = 0, pair_entries = 0, U_root = 0, sumU = 0, sumN = 0, inrange = 0;
  DevBuf d_site_nodes, d_site_idx, d_site_leaf, d_site_items;
  DevBuf d_pair_nodes, d_pair_idx, d_pair_leaf, d_pair_items;
  DevBuf d_pid0, d_T0, d_mask;
  bool has_mask = false;
  DevBuf site_part, site_expo, site_acc;
  DevBuf pair_part, pair_expo, pair_acc, pair_comp;
  DevBuf dP, dpi, dtmp;
  Events ev;
  double t[5] = {0, 0, 0, 0, 0};
  size_t dev_bytes() const {
    return d_site_nodes.bytes + d_site_idx.bytes + d_site_leaf.bytes + d_site_items.bytes +
           d_pair_nodes.bytes + d_pair_idx.bytes + d_pair_leaf.bytes + d_pair_items.bytes +
           d_pid0.bytes + d_T0.bytes + d_mask.bytes + site_part.bytes + site_expo.bytes +
           site_acc.bytes + pair_part.bytes + pair_expo.bytes + pair_acc.bytes + pair_comp.bytes;
  }
};

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
      nodes[v] = d;
    }
    std::vector<int2> items;
    ds->site_lv.assign(H + 1, 0);
    for (int h = 1; h <= H; ++h) {
      ds->site_lv[h - 1] = (int)items.size();
      for (int v = 0; v < n; ++v)
        if (!isleaf(v) && ds->height[v] == h)
          for (int e = 0; e < ds->npat[v]; e += 256) items.push_back(make_int2(v, e));
    }
    ds->site_lv[H] = (int)items.size();
    upload(ds->d_site_nodes, nodes, s);
    upload(ds->d_site_idx, idx, s);
    upload(ds->d_site_leaf, leaf, s);
    upload(ds->d_site_items, items, s);
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
        // Only leaves keep their representative pairs (for their state codes).
        std::vector<int>().swap(ex[v]);
        std::vector<int>().swap(ey[v]);
      }
    }
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
        if (v != 0) {
          off[v] = part;
          part += U[v];
        }
      }
    }
    ds->pair_entries = part;
    ds->U_root = U[0];
    std::vector<NodeDesc> nodes(n);
    std::vector<int> idx;
    idx.reserve((size_t)2 * ds->sumU);
    for (int v = 0; v < n; ++v) {
      if (isleaf(v)) continue;
      int l = ds->ch[2 * v], r = ds->ch[2 * v + 1];
      NodeDesc d{};
      d.U = U[v];
      d.idx_l = (int64_t)idx.size();
      idx.insert(idx.end(), il[v].begin(), il[v].end());
      d.idx_r = (int64_t)idx.size();
      idx.insert(idx.end(), ir[v].begin(), ir[v].end());
      d.part = off[v];
      d.cpart_l = off[l];
      d.cpart_r = off[r];
      d.leaf_l = isleaf(l);
      d.leaf_r = isleaf(r);
      d.node_l = l;
      d.node_r = r;
      nodes[v] = d;
      std::vector<int>().swap(il[v]);
      std::vector<int>().swap(ir[v]);
    }
    const int step = 16 * PAIR_NB;
    std::vector<int2> items;
    ds->pair_lv.assign(H + 1, 0);
    for (int h = 1; h <= H; ++h) {
      ds->pair_lv[h - 1] = (int)items.size();
      for (int v = 0; v < n; ++v)
        if (!isleaf(v) && ds->height[v] == h)
          for (int64_t e = 0; e < U[v]; e += step) items.push_back(make_int2(v, (int)e));
    }
    ds->pair_lv[H] = (int)items.size();
    upload(ds->d_pair_nodes, nodes, s);
    upload(ds->d_pair_idx, idx, s);
    upload(ds->d_pair_leaf, leaf, s);
    upload(ds->d_pair_items, items, s);
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

template <typename T>
std::vector<T> convert_P(const double *P, size_t n) {
  std::vector<T> out(n);
  for (size_t i = 0; i < n; ++i) out[i] = (T)P[i];
  return out;
}

template <typename T>
void run_site(messi_dataset *ds, int K, const double *P4, const double *pi4, const double *logw,
              double *d_u) {
  hipStream_t s = ds->ctx->stream;
  const int nn = ds->nn;
  MESSI_REQUIRE(K >= 1, "need at least one site component");
  std::vector<T> Ph = convert_P<T>(P4, (size_t)K * nn * 16);
  std::vector<double> pih(pi4, pi4 + 4 * K);
  upload(ds->dP, Ph, s);
  upload(ds->dpi, pih, s);
  ds->site_part.ensure(std::max<int64_t>(1, ds->site_entries) * 4 * sizeof(T));
  ds->site_expo.ensure(std::max<int64_t>(1, ds->site_entries) * sizeof(int));
  ds->site_acc.ensure(std::max(1, ds->npat[0]) * sizeof(double));
  for (int k = 0; k < K; ++k) {
    for (int h = 1; h <= ds->H; ++h) {
      int b = ds->site_lv[h - 1], e = ds->site_lv[h];
      launch_site_level<T>(s, ds->d_site_items.as<int2>() + b, e - b, h == ds->H,
                           ds->d_site_nodes.as<NodeDesc>(), ds->d_site_idx.as<int>(),
                           ds->d_site_leaf.as<uint8_t>(), ds->site_part.as<T>(),
                           ds->site_expo.as<int>(), ds->dP.as<T>() + (size_t)k * nn * 16,
                           ds->dpi.as<double>() + 4 * k, logw[k], k == 0, ds->site_acc.as<double>());
    }
  }
  launch_site_expand(s, ds->L, ds->d_pid0.as<int>(), ds->site_acc.as<double>(), d_u);
  MESSI_HIP(hipGetLastError());
}

template <typename T>
void run_pair(messi_dataset *ds, int K, const double *P16, const double *pi16, const double *logw,
              double *d_out, double *h_percomp) {
  hipStream_t s = ds->ctx->stream;
  const int nn = ds->nn, L = ds->L;
  MESSI_REQUIRE(K >= 1, "need at least one pair component");
  std::vector<T> Ph = convert_P<T>(P16, (size_t)K * nn * 256);
  std::vector<double> pih(pi16, pi16 + 16 * K);
  upload(ds->dP, Ph, s);
  upload(ds->dpi, pih, s);
  ds->pair_part.ensure(std::max<int64_t>(1, ds->pair_entries) * 16 * sizeof(T));
  ds->pair_expo.ensure(std::max<int64_t>(1, ds->pair_entries) * sizeof(int));
  ds->pair_acc.ensure(std::max<int64_t>(1, ds->U_root) * sizeof(double));
  ds->pair_comp.ensure(std::max<int64_t>(1, ds->U_root) * sizeof(double));
  for (int k = 0; k < K; ++k) {
    for (int h = 1; h <= ds->H; ++h) {
      int b = ds->pair_lv[h - 1], e = ds->pair_lv[h];
      launch_pair_level<T, PAIR_NB>(s, ds->d_pair_items.as<int2>() + b, e - b, h == ds->H,
                                    ds->d_pair_nodes.as<NodeDesc>(), ds->d_pair_idx.as<int>(),
                                    ds->d_pair_leaf.as<uint16_t>(), ds->pair_part.as<T>(),
                                    ds->pair_expo.as<int>(), ds->dP.as<T>() + (size_t)k * nn * 256,
                                    ds->dpi.as<double>() + 16 * k, logw[k], k == 0,
                                    ds->pair_acc.as<double>(), ds->pair_comp.as<double>());
    }
    if (h_percomp) {
      ds->dtmp.ensure((size_t)L * L * sizeof(double));
      launch_pair_expand(s, L, ds->maxbp, ds->npat[0], ds->d_pid0.as<int>(), ds->d_T0.as<int>(),
                         ds->has_mask ? ds->d_mask.as<uint8_t>() : nullptr,
                         ds->pair_comp.as<double>(), ds->dtmp.as<double>());
      MESSI_HIP(hipMemcpyAsync(h_percomp + (size_t)k * L * L, ds->dtmp.p, (size_t)L * L * sizeof(double),
                               hipMemcpyDeviceToHost, s));
    }
  }
  launch_pair_expand(s, L, ds->maxbp, ds->npat[0], ds->d_pid0.as<int>(), ds->d_T0.as<int>(),
                     ds->has_mask ? ds->d_mask.as<uint8_t>() : nullptr, ds->pair_acc.as<double>(),
                     d_out);
  MESSI_HIP(hipGetLastError());
}

InsideWork &work(messi_ctx *ctx, int prec) {
  MESSI_REQUIRE(prec == MESSI_FP64 || prec == MESSI_FP32, "prec must be MESSI_FP64 or MESSI_FP32");
  return prec == MESSI_FP32 ? ctx->w32 : ctx->w64;
}

void record(messi_dataset *ds, int i) {
  if (ds->ctx->profiling) MESSI_HIP(hipEventRecord(ds->ev.e[i], ds->ctx->stream));
}

void collect(messi_dataset *ds, int upto) {
  if (!ds->ctx->profiling) return;
  MESSI_HIP(hipEventSynchronize(ds->ev.e[upto]));
  for (int i = 0; i < upto; ++i) {
    float ms = 0;
    MESSI_HIP(hipEventElapsedTime(&ms, ds->ev.e[i], ds->ev.e[i + 1]));
    ds->t[i] = ms;
  }
  for (int i = upto; i < 5; ++i) ds->t[i] = 0;
}

template <typename T>
double eval_impl(messi_dataset *ds, int K4, const double *P4, const double *pi4, const double *logw4,
                 int K16, const double *P16, const double *pi16, const double *logw16, double B,
                 bool keep_outside) {
  messi_ctx *ctx = ds->ctx;
  InsideWork &w = work(ctx, sizeof(T) == 4 ? MESSI_FP32 : MESSI_FP64);
  const int L = ds->L;
  inside_alloc<T>(w, L, keep_outside);
  if (ctx->profiling) ds->ev.make();
  record(ds, 0);
  run_site<T>(ds, K4, P4, pi4, logw4, w.u.as<double>());
  record(ds, 1);
  run_pair<T>(ds, K16, P16, pi16, logw16, w.paired.as<double>(), nullptr);
  record(ds, 2);
  inside_prepare<T>(ctx->stream, w, L, B);
  inside_run<T>(ctx->stream, w, L, B);
  record(ds, 3);
  double z = inside_logZ<T>(ctx->stream, w, L);
  if (!keep_outside) collect(ds, 3);
  return z;
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
    MESSI_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  });
  if (rc) {
    // Leave the context to report its error.
    *out = ctx;
    return rc;
  }
  *out = ctx;
  return 0;
}

void messi_ctx_destroy(messi_ctx *ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  ctx->w64.~InsideWork();
  new (&ctx->w64) InsideWork();
  ctx->w32.~InsideWork();
  new (&ctx->w32) InsideWork();
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char *messi_last_error(const messi_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int messi_set_profiling(messi_ctx *ctx, int on) {
  if (!ctx) return 2;
  ctx->profiling = on != 0;
  return 0;
}

int messi_selftest(messi_ctx *ctx) {
  if (!ctx) return 2;
  return guard(ctx, [&] {
    MESSI_HIP(hipSetDevice(ctx->device));
    MESSI_REQUIRE(mfma_probe<double>(ctx->stream), "FP64 MFMA 16x16x4 layout check failed");
    MESSI_REQUIRE(mfma_probe<float>(ctx->stream), "FP32 MFMA 16x16x4 layout check failed");
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
  stats[5] = (int64_t)ds->dev_bytes();
  return 0;
}

int messi_site_loglik(messi_dataset *ds, int K, const double *P4, const double *pi4,
                      const double *logw, int prec, double *u_out) {
  if (!ds) return 2;
  messi_ctx *ctx = ds->ctx;
  return guard(ctx, [&] {
    MESSI_REQUIRE(P4 && pi4 && logw && u_out, "null argument");
    MESSI_HIP(hipSetDevice(ctx->device));
    InsideWork &w = work(ctx, prec);
    w.u.ensure(sizeof(double) * ds->L);
    if (prec == MESSI_FP32) run_site<float>(ds, K, P4, pi4, logw, w.u.as<double>());
    else run_site<double>(ds, K, P4, pi4, logw, w.u.as<double>());
    MESSI_HIP(hipMemcpyAsync(u_out, w.u.p, sizeof(double) * ds->L, hipMemcpyDeviceToHost, ctx->stream));
    MESSI_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int messi_pair_loglik(messi_dataset *ds, int K, const double *P16, const double *pi16,
                      const double *logw, int prec, double *out, double *percomp) {
  if (!ds) return 2;
  messi_ctx *ctx = ds->ctx;
  return guard(ctx, [&] {
    MESSI_REQUIRE(P16 && pi16 && logw && out, "null argument");
    MESSI_HIP(hipSetDevice(ctx->device));
    InsideWork &w = work(ctx, prec);
    const size_t LL = (size_t)ds->L * ds->L;
    w.paired.ensure(sizeof(double) * LL);
    if (prec == MESSI_FP32) run_pair<float>(ds, K, P16, pi16, logw, w.paired.as<double>(), percomp);
    else run_pair<double>(ds, K, P16, pi16, logw, w.paired.as<double>(), percomp);
    MESSI_HIP(hipMemcpyAsync(out, w.paired.p, sizeof(double) * LL, hipMemcpyDeviceToHost, ctx->stream));
    MESSI_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int messi_inside(messi_ctx *ctx, int L, const double *u, const double *paired, double B, int zonly,
                 int prec, double *Z, double *in3) {
  if (!ctx) return 2;
  return guard(ctx, [&] {
    MESSI_REQUIRE(L >= 1 && u && paired && Z, "bad argument");
    MESSI_REQUIRE(B > 0.0, "B must be positive");
    MESSI_REQUIRE(zonly || in3, "in3 required when zonly == 0");
    MESSI_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    InsideWork &w = work(ctx, prec);
    const size_t LL = (size_t)L * L;
    auto go = [&](auto tag) {
      using T = decltype(tag);
      inside_alloc<T>(w, L, false);
      MESSI_HIP(hipMemcpyAsync(w.u.p, u, sizeof(double) * L, hipMemcpyHostToDevice, s));
      MESSI_HIP(hipMemcpyAsync(w.paired.p, paired, sizeof(double) * LL, hipMemcpyHostToDevice, s));
      inside_prepare<T>(s, w, L, B);
      inside_run<T>(s, w, L, B);
      *Z = inside_logZ<T>(s, w, L);
      if (!zonly) {
        w.out3.ensure(sizeof(double) * 3 * LL);
        inside_export<T>(s, w, L, w.out3.as<double>());
        MESSI_HIP(hipMemcpyAsync(in3, w.out3.p, sizeof(double) * 3 * LL, hipMemcpyDeviceToHost, s));
      }
      MESSI_HIP(hipStreamSynchronize(s));
    };
    if (prec == MESSI_FP32) go(float{});
    else go(double{});
  });
}

int messi_outside(messi_ctx *ctx, int L, const double *in3, const double *u, const double *paired,
                  int prec, double *out3) {
  if (!ctx) return 2;
  return guard(ctx, [&] {
    MESSI_REQUIRE(L >= 1 && in3 && u && paired && out3, "bad argument");
    MESSI_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    InsideWork &w = work(ctx, prec);
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
      outside_run<T>(s, w, L);
      outside_export<T>(s, w, L, w.out3.as<double>());
      MESSI_HIP(hipMemcpyAsync(out3, w.out3.p, sizeof(double) * 3 * LL, hipMemcpyDeviceToHost, s));
      MESSI_HIP(hipStreamSynchronize(s));
    };
    if (prec == MESSI_FP32) go(float{});
    else go(double{});
  });
}

int messi_bpp(messi_ctx *ctx, int L, const double *in3, const double *out3, const double *u,
              const double *paired, int prec, double *unpaired, double *pairprob) {
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
      if (unpaired)
        MESSI_HIP(hipMemcpyAsync(unpaired, du.p, sizeof(double) * L, hipMemcpyDeviceToHost, s));
      MESSI_HIP(hipMemcpyAsync(pairprob, dp.p, sizeof(double) * LL, hipMemcpyDeviceToHost, s));
      MESSI_HIP(hipStreamSynchronize(s));
    };
    if (prec == MESSI_FP32) go(float{});
    else go(double{});
  });
}

int messi_eval(messi_dataset *ds, int K4, const double *P4, const double *pi4, const double *logw4,
               int K16, const double *P16, const double *pi16, const double *logw16, double B, int prec,
               double *logZ) {
  if (!ds) return 2;
  messi_ctx *ctx = ds->ctx;
  return guard(ctx, [&] {
    MESSI_REQUIRE(P4 && pi4 && logw4 && P16 && pi16 && logw16 && logZ, "null argument");
    MESSI_REQUIRE(B > 0.0, "B must be positive");
    MESSI_HIP(hipSetDevice(ctx->device));
    (void)work(ctx, prec);
    if (prec == MESSI_FP32)
      *logZ = eval_impl<float>(ds, K4, P4, pi4, logw4, K16, P16, pi16, logw16, B, false);
    else
      *logZ = eval_impl<double>(ds, K4, P4, pi4, logw4, K16, P16, pi16, logw16, B, false);
  });
}

int messi_posterior(messi_dataset *ds, int K4, const double *P4, const double *pi4,
                    const double *logw4, int K16, const double *P16, const double *pi16,
                    const double *logw16, int prec, double *logZ, double *unpaired,
                    double *pairprob) {
  if (!ds) return 2;
  messi_ctx *ctx = ds->ctx;
  return guard(ctx, [&] {
    MESSI_REQUIRE(P4 && pi4 && logw4 && P16 && pi16 && logw16 && logZ, "null argument");
    MESSI_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    const int L = ds->L;
    const size_t LL = (size_t)L * L;
    auto go = [&](auto tag) {
      using T = decltype(tag);
      InsideWork &w = work(ctx, sizeof(T) == 4 ? MESSI_FP32 : MESSI_FP64);
      *logZ = eval_impl<T>(ds, K4, P4, pi4, logw4, K16, P16, pi16, logw16, 1.0, true);
      outside_run<T>(s, w, L);
      record(ds, 4);
      ds->dtmp.ensure(sizeof(double) * (LL + L));
      double *dpair = ds->dtmp.as<double>();
      double *dunp = dpair + LL;
      bpp_run<T>(s, w, L, dunp, dpair);
      record(ds, 5);
      if (unpaired)
        MESSI_HIP(hipMemcpyAsync(unpaired, dunp, sizeof(double) * L, hipMemcpyDeviceToHost, s));
      if (pairprob)
        MESSI_HIP(hipMemcpyAsync(pairprob, dpair, sizeof(double) * LL, hipMemcpyDeviceToHost, s));
      MESSI_HIP(hipStreamSynchronize(s));
      collect(ds, 5);
    };
    if (prec == MESSI_FP32) go(float{});
    else go(double{});
  });
}

int messi_last_timings(const messi_dataset *ds, double *t) {
  if (!ds || !t) return 2;
  for (int i = 0; i < 5; ++i) t[i] = ds->t[i];
  return 0;
}

}  // extern "C"
