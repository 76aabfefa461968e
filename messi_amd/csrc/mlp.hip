// Small MLP for the learned MCMC proposal: training and inference on gfx950.
//
// Network (all dimensions padded to multiples of 16 with zeros):
//   x (din) -> tanh(x W1 + b1) (H) -> tanh(. W2 + b2) (H) -> . W3 + b3 (2*din)
// The output holds a mean shift mu (first din) and a log-scale ell (next din) of a
// diagonal Gaussian proposal z' ~ N(z + mu(z), diag exp(2 ell(z))).  Training
// minimises the mean negative log density of target vectors t given inputs z:
//   nll = sum_d ell_d + 0.5 ((t_d - z_d - mu_d) / exp(ell_d))^2.
// Weights are kept in FP32 (Adam state too) with bf16 copies feeding the
// products; every product runs on V_MFMA_F32_16X16X16_BF16 with FP32
// accumulation.  Activations entering a product are rounded to bf16, the rest of
// the arithmetic is FP32.
#include <algorithm>
#include <cstring>
#include <memory>
#include <random>

#include "common.h"
#include "prune.h"

namespace messi {

typedef uint16_t bf16_t;

__device__ __forceinline__ bf16_t f2bf(float f) {
  // round to nearest even
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (bf16_t)(u >> 16);  // inf / nan
  u += 0x7fffu + ((u >> 16) & 1u);
  return (bf16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(bf16_t b) { return __uint_as_float(((uint32_t)b) << 16); }

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// C[M x N] (fp32, row-major, ldc) = op(A)[M x K] * op(B)[K x N] (+ bias[N]) with
// bf16 operands.  op(A)(i,k) = A[i*sa0 + k*sa1]; op(B)(k,j) = B[k*sb0 + j*sb1].
// M, N, K multiples of 16.  One wavefront per 16x16 output tile.
// A operand of the 16x16x16 bf16 MFMA: lane l holds A[l%16][4*(l/16) + v],
// B operand: B[4*(l/16) + v][l%16], result as the f32 16x16x4 variant
// (row 4*(l/16) + r, column l%16); checked by the self-test.
__global__ __launch_bounds__(256) void gemm_bf16_kernel(int M, int N, int K, const bf16_t *__restrict__ A,
                                                        int sa0, int sa1, const bf16_t *__restrict__ B, int sb0,
                                                        int sb1, float *__restrict__ C, int ldc,
                                                        const float *__restrict__ bias) {
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  const int tn = N / 16, tiles = (M / 16) * tn;
  if (wave >= tiles) return;
  const int i0 = (wave / tn) * 16, j0 = (wave % tn) * 16;
  const int r = lane & 15, g = lane >> 4;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += 16) {
    bf16x4 a, b;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int k = k0 + 4 * g + v;
      a[v] = (short)A[(size_t)(i0 + r) * sa0 + (size_t)k * sa1];
      b[v] = (short)B[(size_t)k * sb0 + (size_t)(j0 + r) * sb1];
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int i = i0 + 4 * g + q, j = j0 + r;
    C[(size_t)i * ldc + j] = acc[q] + (bias ? bias[j] : 0.0f);
  }
}

// y = tanh(x) in place on n values, plus a bf16 copy.
__global__ void tanh_kernel(int n, float *x, bf16_t *xb) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float y = tanhf(x[i]);
  x[i] = y;
  xb[i] = f2bf(y);
}

__global__ void to_bf16_kernel(int n, const float *x, bf16_t *xb) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) xb[i] = f2bf(x[i]);
}

// Input rows (n x din, fp32) -> padded bf16 (Mp x D0), zero padding.
__global__ void pack_input_kernel(int n, int din, int Mp, int D0, const float *x, bf16_t *xb) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Mp * D0) return;
  int row = i / D0, c = i % D0;
  xb[i] = (row < n && c < din) ? f2bf(x[(size_t)row * din + c]) : (bf16_t)0;
}

// Loss and its gradient with respect to the output layer (mean over n rows).
// O: Mp x D3 (mu | ell | padding); x: n x din inputs; t: n x din targets.
__global__ void nll_grad_kernel(int n, int din, int D3, const float *O, const float *x, const float *t,
                                float *dO, bf16_t *dOb, float *rowloss) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;  // row
  if (i >= n) return;
  float l = 0.0f;
  const float inv = 1.0f / (float)n;
  for (int d = 0; d < din; ++d) {
    float mu = O[(size_t)i * D3 + d], ell = O[(size_t)i * D3 + din + d];
    float s = expf(-ell);
    float rr = (t[(size_t)i * din + d] - x[(size_t)i * din + d] - mu) * s;
    l += ell + 0.5f * rr * rr;
    float gmu = -rr * s * inv, gell = (1.0f - rr * rr) * inv;
    dO[(size_t)i * D3 + d] = gmu;
    dO[(size_t)i * D3 + din + d] = gell;
    dOb[(size_t)i * D3 + d] = f2bf(gmu);
    dOb[(size_t)i * D3 + din + d] = f2bf(gell);
  }
  rowloss[i] = l;
}

// dA = (dH from a product) * (1 - H^2), in place, plus a bf16 copy.
__global__ void tanh_back_kernel(int n, float *dA, const float *H, bf16_t *dAb) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float h = H[i];
  float v = dA[i] * (1.0f - h * h);
  dA[i] = v;
  dAb[i] = f2bf(v);
}

// Column sums of an M x N fp32 matrix (bias gradients); one thread per column.
__global__ void colsum_kernel(int M, int N, const float *X, float *out) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  float s = 0.0f;
  for (int i = 0; i < M; ++i) s += X[(size_t)i * N + j];
  out[j] = s;
}

// Adam with decoupled weight decay (AdamW) on the flat parameter vector; mask
// zeroes updates of padding entries.
__global__ void adam_kernel(int n, float *w, const float *g, float *m, float *v, const uint8_t *mask, float lr,
                            float b1, float b2, float eps, float c1, float c2, float wd) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !mask[i]) return;
  float gi = g[i];
  float mi = b1 * m[i] + (1.0f - b1) * gi;
  float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  float wi = w[i];
  w[i] = wi - lr * ((mi / c1) / (sqrtf(vi / c2) + eps) + wd * wi);
}

// Self-test of the bf16 MFMA layout: D = A (16x16) * B (16x16).
__global__ void mfma_bf16_probe_kernel(const bf16_t *A, const bf16_t *B, float *D) {
  int l = threadIdx.x, r = l & 15, g = l >> 4;
  bf16x4 a, b;
  for (int v = 0; v < 4; ++v) {
    a[v] = (short)A[r * 16 + 4 * g + v];
    b[v] = (short)B[(4 * g + v) * 16 + r];
  }
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  for (int q = 0; q < 4; ++q) D[(4 * g + q) * 16 + r] = acc[q];
}

static uint16_t host_bf16(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
static float host_bf2f(uint16_t b) {
  uint32_t u = ((uint32_t)b) << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

bool mfma_bf16_probe(hipStream_t s) {
  std::vector<uint16_t> A(256), B(256);
  std::vector<float> D(256), R(256, 0.0f);
  for (int i = 0; i < 256; ++i) {
    A[i] = host_bf16((float)(1 + (i * 7) % 13));
    B[i] = host_bf16((float)(1 + (i * 5) % 11));
  }
  for (int i = 0; i < 16; ++i)
    for (int j = 0; j < 16; ++j)
      for (int k = 0; k < 16; ++k) R[i * 16 + j] += host_bf2f(A[i * 16 + k]) * host_bf2f(B[k * 16 + j]);
  DevBuf dA, dB, dD;
  upload(dA, A, s);
  upload(dB, B, s);
  dD.ensure(256 * sizeof(float));
  hipLaunchKernelGGL(mfma_bf16_probe_kernel, dim3(1), dim3(64), 0, s, dA.as<bf16_t>(), dB.as<bf16_t>(),
                     dD.as<float>());
  MESSI_HIP(hipGetLastError());
  MESSI_HIP(hipMemcpyAsync(D.data(), dD.p, 256 * sizeof(float), hipMemcpyDeviceToHost, s));
  MESSI_HIP(hipStreamSynchronize(s));
  for (int i = 0; i < 256; ++i)
    if (D[i] != R[i]) return false;
  return true;
}

}  // namespace messi

using namespace messi;

struct messi_mlp {
  messi_ctx *ctx = nullptr;
  int din = 0, dh = 0, D0 = 0, D3 = 0;
  // flat fp32 parameters: W1 (D0 x dh), b1 (dh), W2 (dh x dh), b2 (dh), W3 (dh x D3), b3 (D3)
  size_t oW1 = 0, ob1 = 0, oW2 = 0, ob2 = 0, oW3 = 0, ob3 = 0, np = 0;
  DevBuf w, g, m, v, mask, wb;  // wb: bf16 copies of W1, W2, W3 at the same offsets
  int step = 0;
  int Mcap = 0;
  DevBuf xb, A1, H1b, A2, H2b, O, dO, dOb, dH2, dH2b, dH1, dH1b, rowloss, xin, tin;
};

namespace {

int ceil16(int x) { return (x + 15) / 16 * 16; }

void gemm(hipStream_t s, int M, int N, int K, const bf16_t *A, int sa0, int sa1, const bf16_t *B, int sb0, int sb1,
          float *C, int ldc, const float *bias) {
  int tiles = (M / 16) * (N / 16);
  hipLaunchKernelGGL(gemm_bf16_kernel, dim3((tiles + 3) / 4), dim3(256), 0, s, M, N, K, A, sa0, sa1, B, sb0, sb1, C,
                     ldc, bias);
}

void refresh_bf16(messi_mlp *p) {
  hipStream_t s = p->ctx->stream;
  hipLaunchKernelGGL(to_bf16_kernel, dim3((unsigned)((p->np + 255) / 256)), dim3(256), 0, s, (int)p->np,
                     p->w.as<float>(), p->wb.as<bf16_t>());
}

void ensure_rows(messi_mlp *p, int Mp) {
  if (Mp <= p->Mcap) return;
  p->Mcap = Mp;
  p->xb.ensure((size_t)Mp * p->D0 * 2);
  p->A1.ensure((size_t)Mp * p->dh * 4);
  p->H1b.ensure((size_t)Mp * p->dh * 2);
  p->A2.ensure((size_t)Mp * p->dh * 4);
  p->H2b.ensure((size_t)Mp * p->dh * 2);
  p->O.ensure((size_t)Mp * p->D3 * 4);
  p->dO.ensure((size_t)Mp * p->D3 * 4);
  p->dOb.ensure((size_t)Mp * p->D3 * 2);
  p->dH2.ensure((size_t)Mp * p->dh * 4);
  p->dH2b.ensure((size_t)Mp * p->dh * 2);
  p->dH1.ensure((size_t)Mp * p->dh * 4);
  p->dH1b.ensure((size_t)Mp * p->dh * 2);
  p->rowloss.ensure((size_t)Mp * 4);
}

// Forward pass of n rows already in p->xin (device, n x din).
void forward(messi_mlp *p, int n) {
  hipStream_t s = p->ctx->stream;
  const int Mp = ceil16(n), D0 = p->D0, H = p->dh, D3 = p->D3;
  const bf16_t *wb = p->wb.as<bf16_t>();
  const float *w = p->w.as<float>();
  hipLaunchKernelGGL(pack_input_kernel, dim3((Mp * D0 + 255) / 256), dim3(256), 0, s, n, p->din, Mp, D0,
                     p->xin.as<float>(), p->xb.as<bf16_t>());
  gemm(s, Mp, H, D0, p->xb.as<bf16_t>(), D0, 1, wb + p->oW1, H, 1, p->A1.as<float>(), H, w + p->ob1);
  hipLaunchKernelGGL(tanh_kernel, dim3((Mp * H + 255) / 256), dim3(256), 0, s, Mp * H, p->A1.as<float>(),
                     p->H1b.as<bf16_t>());
  gemm(s, Mp, H, H, p->H1b.as<bf16_t>(), H, 1, wb + p->oW2, H, 1, p->A2.as<float>(), H, w + p->ob2);
  hipLaunchKernelGGL(tanh_kernel, dim3((Mp * H + 255) / 256), dim3(256), 0, s, Mp * H, p->A2.as<float>(),
                     p->H2b.as<bf16_t>());
  gemm(s, Mp, D3, H, p->H2b.as<bf16_t>(), H, 1, wb + p->oW3, D3, 1, p->O.as<float>(), D3, w + p->ob3);
}

// Loss, gradients into p->g (flat, fp32) for n rows with targets in p->tin.
float loss_grad(messi_mlp *p, int n) {
  hipStream_t s = p->ctx->stream;
  const int Mp = ceil16(n), D0 = p->D0, H = p->dh, D3 = p->D3;
  const bf16_t *wb = p->wb.as<bf16_t>();
  float *g = p->g.as<float>();
  forward(p, n);
  MESSI_HIP(hipMemsetAsync(p->dO.p, 0, (size_t)Mp * D3 * 4, s));
  MESSI_HIP(hipMemsetAsync(p->dOb.p, 0, (size_t)Mp * D3 * 2, s));
  hipLaunchKernelGGL(nll_grad_kernel, dim3((n + 255) / 256), dim3(256), 0, s, n, p->din, D3, p->O.as<float>(),
                     p->xin.as<float>(), p->tin.as<float>(), p->dO.as<float>(), p->dOb.as<bf16_t>(),
                     p->rowloss.as<float>());
  // layer 3: dW3 = H2b^T dOb, db3 = colsum dO, dH2 = dOb W3^T
  gemm(s, H, D3, Mp, p->H2b.as<bf16_t>(), 1, H, p->dOb.as<bf16_t>(), D3, 1, g + p->oW3, D3, nullptr);
  hipLaunchKernelGGL(colsum_kernel, dim3((D3 + 63) / 64), dim3(64), 0, s, Mp, D3, p->dO.as<float>(), g + p->ob3);
  gemm(s, Mp, H, D3, p->dOb.as<bf16_t>(), D3, 1, wb + p->oW3, 1, D3, p->dH2.as<float>(), H, nullptr);
  // A2 holds tanh(layer-2 pre-activation) after the forward pass.
  hipLaunchKernelGGL(tanh_back_kernel, dim3((Mp * H + 255) / 256), dim3(256), 0, s, Mp * H, p->dH2.as<float>(),
                     p->A2.as<float>(), p->dH2b.as<bf16_t>());
  // layer 2
  gemm(s, H, H, Mp, p->H1b.as<bf16_t>(), 1, H, p->dH2b.as<bf16_t>(), H, 1, g + p->oW2, H, nullptr);
  hipLaunchKernelGGL(colsum_kernel, dim3((H + 63) / 64), dim3(64), 0, s, Mp, H, p->dH2.as<float>(), g + p->ob2);
  gemm(s, Mp, H, H, p->dH2b.as<bf16_t>(), H, 1, wb + p->oW2, 1, H, p->dH1.as<float>(), H, nullptr);
  hipLaunchKernelGGL(tanh_back_kernel, dim3((Mp * H + 255) / 256), dim3(256), 0, s, Mp * H, p->dH1.as<float>(),
                     p->A1.as<float>(), p->dH1b.as<bf16_t>());
  // layer 1
  gemm(s, D0, H, Mp, p->xb.as<bf16_t>(), 1, D0, p->dH1b.as<bf16_t>(), H, 1, g + p->oW1, H, nullptr);
  hipLaunchKernelGGL(colsum_kernel, dim3((H + 63) / 64), dim3(64), 0, s, Mp, H, p->dH1.as<float>(), g + p->ob1);
  MESSI_HIP(hipGetLastError());
  std::vector<float> rl(n);
  MESSI_HIP(hipMemcpyAsync(rl.data(), p->rowloss.p, (size_t)n * 4, hipMemcpyDeviceToHost, s));
  MESSI_HIP(hipStreamSynchronize(s));
  double tot = 0.0;
  for (float x : rl) tot += x;
  return (float)(tot / n);
}

template <typename F>
int guard_mlp(messi_ctx *ctx, F &&f) {
  try {
    f();
    return 0;
  } catch (const std::exception &e) {
    if (ctx) ctx->err = e.what();
    return 1;
  }
}

void put_rows(messi_mlp *p, DevBuf &dst, int n, const float *h) {
  dst.ensure((size_t)std::max(1, n) * p->din * 4);
  MESSI_HIP(hipMemcpyAsync(dst.p, h, (size_t)n * p->din * 4, hipMemcpyHostToDevice, p->ctx->stream));
}

}  // namespace

extern "C" {

int messi_mlp_create(messi_ctx *ctx, int din, int dh, uint64_t seed, messi_mlp **out) {
  if (!ctx || !out) return 2;
  *out = nullptr;
  auto p = std::make_unique<messi_mlp>();
  int rc = guard_mlp(ctx, [&] {
    MESSI_REQUIRE(din >= 1 && din <= 64 && dh >= 16 && dh % 16 == 0 && dh <= 1024, "bad MLP dimensions");
    MESSI_HIP(hipSetDevice(ctx->device));
    p->ctx = ctx;
    p->din = din;
    p->dh = dh;
    p->D0 = ceil16(din);
    p->D3 = ceil16(2 * din);
    const int D0 = p->D0, H = dh, D3 = p->D3;
    p->oW1 = 0;
    p->ob1 = p->oW1 + (size_t)D0 * H;
    p->oW2 = p->ob1 + H;
    p->ob2 = p->oW2 + (size_t)H * H;
    p->oW3 = p->ob2 + H;
    p->ob3 = p->oW3 + (size_t)H * D3;
    p->np = p->ob3 + D3;
    std::vector<float> w(p->np, 0.0f);
    std::vector<uint8_t> mask(p->np, 0);
    std::mt19937_64 rng(seed);
    auto fill = [&](size_t off, int rows, int cols, int vr, int vc, double scale) {
      std::normal_distribution<double> nd(0.0, scale);
      for (int i = 0; i < rows; ++i)
        for (int j = 0; j < cols; ++j)
          if (i < vr && j < vc) {
            w[off + (size_t)i * cols + j] = (float)nd(rng);
            mask[off + (size_t)i * cols + j] = 1;
          }
    };
    fill(p->oW1, D0, H, din, H, 1.0 / std::sqrt((double)din));
    fill(p->oW2, H, H, H, H, 1.0 / std::sqrt((double)H));
    fill(p->oW3, H, D3, H, 2 * din, 0.01 / std::sqrt((double)H));
    for (int j = 0; j < H; ++j) mask[p->ob1 + j] = mask[p->ob2 + j] = 1;
    for (int j = 0; j < 2 * din; ++j) mask[p->ob3 + j] = 1;
    upload(p->w, w, ctx->stream);
    upload(p->mask, mask, ctx->stream);
    std::vector<float> zero(p->np, 0.0f);
    upload(p->g, zero, ctx->stream);
    upload(p->m, zero, ctx->stream);
    upload(p->v, zero, ctx->stream);
    p->wb.ensure(p->np * 2);
    refresh_bf16(p.get());
    MESSI_HIP(hipStreamSynchronize(ctx->stream));
  });
  if (rc == 0) *out = p.release();
  return rc;
}

void messi_mlp_destroy(messi_mlp *p) {
  if (!p) return;
  (void)hipStreamSynchronize(p->ctx->stream);
  delete p;
}

int messi_mlp_num_params(const messi_mlp *p, int64_t *n) {
  if (!p || !n) return 2;
  *n = (int64_t)p->np;
  return 0;
}

int messi_mlp_get_params(messi_mlp *p, float *out) {
  if (!p || !out) return 2;
  return guard_mlp(p->ctx, [&] {
    MESSI_HIP(hipMemcpyAsync(out, p->w.p, p->np * 4, hipMemcpyDeviceToHost, p->ctx->stream));
    MESSI_HIP(hipStreamSynchronize(p->ctx->stream));
  });
}

int messi_mlp_set_params(messi_mlp *p, const float *in) {
  if (!p || !in) return 2;
  return guard_mlp(p->ctx, [&] {
    MESSI_HIP(hipMemcpyAsync(p->w.p, in, p->np * 4, hipMemcpyHostToDevice, p->ctx->stream));
    MESSI_HIP(hipMemsetAsync(p->m.p, 0, p->np * 4, p->ctx->stream));
    MESSI_HIP(hipMemsetAsync(p->v.p, 0, p->np * 4, p->ctx->stream));
    p->step = 0;
    refresh_bf16(p);
    MESSI_HIP(hipStreamSynchronize(p->ctx->stream));
  });
}

int messi_mlp_forward(messi_mlp *p, int n, const float *x, float *out) {
  if (!p || !x || !out) return 2;
  return guard_mlp(p->ctx, [&] {
    MESSI_REQUIRE(n >= 1, "need at least one row");
    MESSI_HIP(hipSetDevice(p->ctx->device));
    ensure_rows(p, ceil16(n));
    put_rows(p, p->xin, n, x);
    forward(p, n);
    // copy the 2*din useful output columns of each row
    std::vector<float> O((size_t)ceil16(n) * p->D3);
    MESSI_HIP(hipMemcpyAsync(O.data(), p->O.p, O.size() * 4, hipMemcpyDeviceToHost, p->ctx->stream));
    MESSI_HIP(hipStreamSynchronize(p->ctx->stream));
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < 2 * p->din; ++j) out[(size_t)i * 2 * p->din + j] = O[(size_t)i * p->D3 + j];
  });
}

int messi_mlp_loss_grad(messi_mlp *p, int n, const float *x, const float *t, float *loss, float *grad) {
  if (!p || !x || !t || !loss) return 2;
  return guard_mlp(p->ctx, [&] {
    MESSI_REQUIRE(n >= 1, "need at least one row");
    MESSI_HIP(hipSetDevice(p->ctx->device));
    ensure_rows(p, ceil16(n));
    put_rows(p, p->xin, n, x);
    put_rows(p, p->tin, n, t);
    *loss = loss_grad(p, n);
    if (grad) {
      MESSI_HIP(hipMemcpyAsync(grad, p->g.p, p->np * 4, hipMemcpyDeviceToHost, p->ctx->stream));
      MESSI_HIP(hipStreamSynchronize(p->ctx->stream));
    }
  });
}

int messi_mlp_train_step(messi_mlp *p, int n, const float *x, const float *t, float lr, float weight_decay,
                         float *loss) {
  if (!p || !x || !t || !loss) return 2;
  return guard_mlp(p->ctx, [&] {
    MESSI_REQUIRE(n >= 1, "need at least one row");
    MESSI_HIP(hipSetDevice(p->ctx->device));
    ensure_rows(p, ceil16(n));
    put_rows(p, p->xin, n, x);
    put_rows(p, p->tin, n, t);
    *loss = loss_grad(p, n);
    p->step += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - std::pow(b1, (float)p->step), c2 = 1.0f - std::pow(b2, (float)p->step);
    hipLaunchKernelGGL(adam_kernel, dim3((unsigned)((p->np + 255) / 256)), dim3(256), 0, p->ctx->stream, (int)p->np,
                       p->w.as<float>(), p->g.as<float>(), p->m.as<float>(), p->v.as<float>(),
                       p->mask.as<uint8_t>(), lr, b1, b2, 1e-8f, c1, c2, weight_decay);
    refresh_bf16(p);
    MESSI_HIP(hipGetLastError());
    MESSI_HIP(hipStreamSynchronize(p->ctx->stream));
  });
}

}  // extern "C"
