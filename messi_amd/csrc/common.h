// Shared host/device definitions of libmessi (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/messi.h"

#define MESSI_HIP(expr)                                                                      \
  do {                                                                                       \
    hipError_t _e = (expr);                                                                  \
    if (_e != hipSuccess)                                                                    \
      throw std::runtime_error(std::string(#expr) + ": " + hipGetErrorString(_e));           \
  } while (0)

#define MESSI_REQUIRE(cond, msg)                                                             \
  do {                                                                                       \
    if (!(cond)) throw std::invalid_argument(msg);                                           \
  } while (0)

namespace messi {

// KH99 rule log-probabilities (src/InsideOutsideAlgorithm.jl:40-46), natural log.
// Index: 0 S->LS, 1 S->s, 2 S->dFd, 3 F->dFd, 4 F->LS, 5 L->s, 6 L->dFd.
inline void kh99_logprobs(double lp[7]) {
  lp[0] = std::log(0.868534);
  lp[1] = std::log(0.117609877998);
  lp[2] = std::log(0.013856122002);
  lp[3] = std::log(0.787640);
  lp[4] = std::log(0.21236);
  lp[5] = std::log(0.894603);
  lp[6] = std::log(0.105397);
}

// A pointer read from device memory (a descriptor field) is generic to the
// compiler, so its loads and stores become flat instructions, which also count
// against the LDS / scalar-memory counter and make every later wait on that
// counter wait for them too.  gp() restates such a pointer as a global one
// (the address space exists only in the device pass of the compiler).
#if defined(__HIP_DEVICE_COMPILE__)
#define MESSI_GLOBAL __attribute__((address_space(1)))
#else
#define MESSI_GLOBAL
#endif
template <class X> __device__ __forceinline__ MESSI_GLOBAL X *gp(X *p) { return (MESSI_GLOBAL X *)p; }

// Workgroups are dispatched to the 8 XCDs of an MI355X round-robin by workgroup
// id, and each XCD has its own L2.  This permutation gives XCD k a contiguous
// chunk of the logical work items (in their natural order), so items that share
// operands run on the same XCD and close in time.  (On another XCD count it is
// still a permutation.)
__device__ __forceinline__ int xcd_chunk_id(int x, int n) {
  constexpr int NX = 8;
  const int k = x % NX, q = n / NX, r = n % NX;
  return k * q + (k < r ? k : r) + x / NX;
}

// Device buffer that only grows.  A borrowed buffer is a fixed region of memory
// the buffer does not own (a slice of a preallocated scratch array): it never
// allocates or frees, and a request beyond its capacity is an error.
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  bool borrowed = false;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p && !borrowed) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    borrowed = false;
  }
  void borrow(void *q, size_t cap) {
    release();
    p = q;
    bytes = cap;
    borrowed = true;
  }
  void ensure(size_t n) {
    if (n <= bytes) return;
    if (borrowed) throw std::length_error("scratch region too small for this call");
    release();
    MESSI_HIP(hipMalloc(&p, n < 256 ? 256 : n));
    bytes = n < 256 ? 256 : n;
  }
  template <typename T> T *as() const { return static_cast<T *>(p); }
};

template <typename T>
inline void upload(DevBuf &b, const std::vector<T> &v, hipStream_t s) {
  b.ensure(v.size() * sizeof(T));
  if (!v.empty())
    MESSI_HIP(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
}

// Device tables of a batch of inside instances.
struct InsideBatch {
  DevBuf inst;    // instance descriptors
  DevBuf prefix;  // per launch t: first workgroup of each instance (n+1 ints)
  DevBuf ptab;    // single-launch form: ticket table
  DevBuf pflags;  // single-launch form: ticket counter, give-up flag, then every instance's flags
  DevBuf groups;  // runs of consecutive members on one dataset (first member, count)
  std::vector<unsigned char> h_inst, h_prefix, h_ptab, h_groups;  // what the tables hold (uploads only on change)
};

// Scratch of the inside/outside/bpp phase, per precision.
struct InsideWork {
  int L = 0;
  int tile = 16;                  // tile edge of the inside schedule (16 or 32)
  int window = 4;                 // bulk window of the inside schedule (1, 2 or 4 launches)
  bool persist = false;           // the inside pass as one launch (single-launch form)
  int band = -1;                  // paired[j,y] is -inf for y - j > band (-1: no such bound)
  const void *exp = nullptr;      // pair-root expansion descriptor of this member (device), or null
  const void *dset = nullptr;     // identity of the member's dataset while exp is set (host), or null
  double *zout = nullptr;         // where log Z (absolute, baseline-relative) is also written (device)
  DevBuf Lr, Lc, Sr, Sc, F;       // inside tables (row- and column-major copies)
  DevBuf acc;                     // bulk split-point accumulator of the tiled inside
  DevBuf Lp, LpE, Sp, SpE;        // scaled path: product-ready operand copies (see inside.hip)
  DevBuf oSr, oSc, oFr, oL;       // outside tables
  DevBuf oWr, oWc, oaL, oaS;      // outside: parent weight W (row/column-major), accumulators
  // scaled path, tile 16: product-ready operands of the outside's bulk products
  // (W as A and B operand, inside S and L shifted by one, see inside.hip)
  DevBuf oWpA, oWpAE, oWpB, oWpBE, iSpB, iSpBE, iLpA, iLpAE;
  DevBuf orng;                    // outside launch tables
  int orng_key[5] = {0, 0, 0, 0, 0};  // (L, tile, band, lband, window) the launch tables were built for
  DevBuf ptv;                     // scaled pair term per cell
  DevBuf lam, C;                  // per-position baseline and its prefix sums (double)
  DevBuf uhat;                    // u - lambda (double)
  DevBuf paired;                  // double L x L (row-major == column-major: symmetric)
  DevBuf u;                       // double L
  DevBuf lp;                      // bpp log table
  DevBuf red;                     // reductions
  DevBuf out3;                    // double 3 x L x L staging
  DevBuf inst1;                   // this work's own instance descriptor (batch of one)
  InsideBatch ib1;                // launch tables when run as a batch of one
};

// Tables of MEA decoding (L x L, cells i < j; E and U column-major, S and the
// split-point accumulators row-major).
constexpr int MEA_MAX_L = 16383;  // the traceback stack (L/2 + 1 words) and pairing (L halfwords) fit 64 KiB of LDS
struct MeaWork {
  int L = 0;
  bool fp32 = false;
  DevBuf E, U, S, accV, accK;
  DevBuf single, pair, paired;  // staging of host inputs and the result
};

// Outputs of structure sampling.
struct SampleWork {
  DevBuf out, prior;
  DevBuf logs;  // FP32 tables: L, S, F and the pair term as natural-log doubles (4 L^2)
};

}  // namespace messi

struct messi_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  static constexpr int NPIPE = 3;
  hipStream_t pipe[NPIPE] = {};  // streams of member groups 1.. of a pipelined batch evaluation
  std::string err;
  bool profiling = false;
  int tile = 16;  // tile edge of the inside schedule
  int window = 4;  // bulk window of the inside schedule
  bool band_skip = true;  // skip the split-point terms a pairing band makes -inf
  bool persist = false;   // inside pass as one launch (messi_set_single_launch)
  bool pair_fuse = true;  // datasets created later compute some pair nodes in their parent's launch
  messi::InsideWork w64, w32;
  messi::InsideWork wna;  // messi_inside_noalloc: every buffer borrowed from the device's static scratch
  messi::MeaWork mea;
  messi::SampleWork samp;
};
