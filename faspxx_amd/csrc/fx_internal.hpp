// fx_internal.hpp -- shared types, error handling and device building blocks
// of the gfx950 iterative-solve library.
//
// Compiled with -ffp-contract=off: every fused multiply-add below is written
// out with fma(), so each value is rounded exactly where the reference's
// g++ build (FMA contraction on) rounds it.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <vector>

#include "faspxx_hip.h"

namespace fx {

constexpr int kThreads = 256;   // threads per workgroup (4 wave64)
constexpr int kWaves = kThreads / 64;
constexpr int kXcds = 8;        // MI355X: 8 XCDs, workgroups dispatched round-robin
constexpr int kMaxRedBlocks = 2048;  // grid cap of streaming / reduction kernels
constexpr double kCloseZero = 1e-20;  // include/Faspxx.hxx:34
constexpr double kSmallDbl = -1e60;   // include/Faspxx.hxx:33
constexpr double kChkRatio = 0.95;    // include/Faspxx.hxx:43
constexpr uint32_t kMaxStag = 20;     // include/Faspxx.hxx:44

void set_error(const char* fmt, ...);
void clear_error();

}  // namespace fx

#define FX_HIP(call)                                                                 \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess) {                                                          \
      fx::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, \
                    __LINE__);                                                       \
      return FX_ERROR_DEVICE;                                                        \
    }                                                                                \
  } while (0)

#define FX_CHECK(cond, code, ...)  \
  do {                             \
    if (!(cond)) {                 \
      fx::set_error(__VA_ARGS__);  \
      return (code);               \
    }                              \
  } while (0)

#define FX_TRY(call)          \
  do {                        \
    int rc_ = (call);         \
    if (rc_ != 0) return rc_; \
  } while (0)

// ---------------------------------------------------------------------------
// Handles
// ---------------------------------------------------------------------------

// Reduction workspace: partial sums of one reduction kernel are written to
// `partials` ([slot][block]); the last workgroup to finish (ticket on
// `counter`) reduces them in a fixed order and resets the counter, so the
// result is deterministic and graph replays need no memset.
struct fx_red_ws {
  double* partials = nullptr;  // kMaxRedSlots * kMaxRedBlocks
  unsigned* counter = nullptr;
};

struct fx_ctx_s {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  fx_red_ws ws;
  double* dscal = nullptr;  // device scalar slots for standalone reductions
  double* hpin = nullptr;   // pinned host mirror of dscal
  int sms = 256;
};

struct fx_vec_s {
  fx_ctx ctx = nullptr;
  uint32_t n = 0;
  int dtype = FX_F64;
  void* d = nullptr;
  bool owned = false;
};

// Row-block table of the SpMV kernels: block b covers rows
// [blk[b], blk[b+1]).  A block either holds at most kBlockRows rows whose
// entries fit the LDS stage (<= kBlockNnz), or exactly one row longer than
// that ("long row", reduced by the whole workgroup).
struct fx_csr_s {
  fx_ctx ctx = nullptr;
  uint32_t nrow = 0, mcol = 0, nnz = 0;
  bool has_values = false;
  uint32_t* rp = nullptr;
  uint32_t* ci = nullptr;
  uint32_t* dp = nullptr;
  double* va = nullptr;
  float* va32 = nullptr;
  uint32_t* blk = nullptr;
  uint32_t nblk = 0;
  uint32_t max_row = 0;
};

struct fx_pc_s {
  fx_ctx ctx = nullptr;
  int kind = 0;  // 0 identity, 1 jacobi
  fx_csr A = nullptr;
  double weight = 1.0;
  uint32_t iters = 1;
  double* dinv = nullptr;   // f64 weight * (1/diag)
  float* dinv32 = nullptr;  // f32 copy for mixed precision (lazy)
};

namespace fx {

// Row-block geometry of the SpMV kernels.
constexpr uint32_t kBlockRows = kThreads;
constexpr uint32_t kBlockNnz = 2048;

int csr_spmv_dispatch(fx_ctx ctx, fx_csr A, const void* x, void* y, int dtype);
int ensure_dinv32(fx_pc pc);

// Host-side launch geometry for streaming kernels over n elements.
inline unsigned stream_blocks(uint32_t n, uint32_t per_thread = 4) {
  uint64_t b = ((uint64_t)n + (uint64_t)kThreads * per_thread - 1) / ((uint64_t)kThreads * per_thread);
  if (b < 1) b = 1;
  if (b > (uint64_t)kMaxRedBlocks) b = kMaxRedBlocks;
  return (unsigned)b;
}

// Grid for the row-block SpMV kernels: padded to a multiple of the XCD count
// so that xcd_block() can give each XCD one contiguous range of row blocks.
inline unsigned spmv_grid(uint32_t nblk) {
  unsigned g = nblk ? nblk : 1;
  return (g + kXcds - 1) / kXcds * kXcds;
}

// ---------------------------------------------------------------------------
// Device building blocks
// ---------------------------------------------------------------------------

// Workgroup id -> logical row block.  Hardware hands workgroup i to XCD
// i % 8; this maps XCD k to the contiguous logical range [k*G/8, (k+1)*G/8),
// so neighbouring row blocks (which share x entries) share an L2.
__device__ __forceinline__ unsigned xcd_block() {
  const unsigned g = gridDim.x, i = blockIdx.x;
  return (i % kXcds) * (g / kXcds) + i / kXcds;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}
// Sum within aligned groups of `width` lanes (width a power of two <= 64).
__device__ __forceinline__ double group_sum(double v, int width) {
  for (int off = width >> 1; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Reduce K per-thread values over the workgroup.  Slot k is a max-reduction
// when bit k of MaxMask is set, else a sum.  The result is valid in every
// thread.  Fixed order: wave butterflies, then waves 0..3 in sequence.
template <int K, unsigned MaxMask = 0u>
__device__ __forceinline__ void block_reduce(double (&v)[K]) {
  __shared__ double red[K][kWaves];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = ((MaxMask >> k) & 1u) ? wave_max(v[k]) : wave_sum(v[k]);
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) red[k][wid] = v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double a = red[k][0];
#pragma unroll
    for (int w = 1; w < kWaves; ++w) a = ((MaxMask >> k) & 1u) ? fmax(a, red[k][w]) : a + red[k][w];
    v[k] = a;
  }
  __syncthreads();
}

// Two-stage grid reduction.  Every workgroup of the grid calls this once with
// its per-thread values.  Returns true in all threads of the last workgroup
// to arrive, with v holding the grid-wide results; false elsewhere.  The
// partials of workgroup b for slot k live at partials[k * gridDim.x + b].
template <int K, unsigned MaxMask = 0u>
__device__ __forceinline__ bool grid_reduce(double (&v)[K], double* partials, unsigned* counter) {
  __shared__ int is_last;
  block_reduce<K, MaxMask>(v);
  const unsigned G = gridDim.x;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) partials[k * G + blockIdx.x] = v[k];
    __threadfence();
    const unsigned ticket = atomicAdd(counter, 1u);
    is_last = (ticket == G - 1);
  }
  __syncthreads();
  if (!is_last) return false;
  __threadfence();
  double acc[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double a = 0.0;
    for (unsigned b = threadIdx.x; b < G; b += kThreads) {
      const double p = partials[k * G + b];
      a = ((MaxMask >> k) & 1u) ? fmax(a, p) : a + p;
    }
    acc[k] = a;
  }
  block_reduce<K, MaxMask>(acc);
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = acc[k];
  if (threadIdx.x == 0) *counter = 0u;
  return true;
}

}  // namespace fx
