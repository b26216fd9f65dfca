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
#include <type_traits>
#include <utility>
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

// A process-wide increasing id, given to each matrix and preconditioner
// handle when it is made.  A solver remembers the ids of what it was set up
// with: a handle freed and made again may come back at the same address,
// and its device arrays, sizes and captured kernel arguments are then new.
uint64_t next_handle_gen();

// The device address of fx_pc_apply_noalloc's scratch array on `device`
// (which must be current), taken once per device (fx_noalloc.hip).
int noalloc_init(int device, double** out);

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

// Reduction workspace: per-workgroup partials of a reduction kernel
// ([slot][workgroup]), combined by a one-workgroup finisher kernel.
struct fx_red_ws {
  double* partials = nullptr;  // 16 slots x 65536 workgroups
};

struct fx_ctx_s {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  fx_red_ws ws;
  double* dscal = nullptr;  // device scalar slots for standalone reductions
  double* hpin = nullptr;   // pinned host mirror of dscal
  hipEvent_t tev[2] = {nullptr, nullptr};  // dispatch-attached timing events (fx_spmv_timed)
  int sms = 256;
  double* scratch = nullptr;  // this device's fx_pc_apply_noalloc scratch (fx_noalloc.hip)
};

struct fx_vec_s {
  fx_ctx ctx = nullptr;
  uint32_t n = 0;
  int dtype = FX_F64;
  void* d = nullptr;
  bool owned = false;
};

// Row-block table of the SpMV kernels: nblk + 1 pairs (first row, first
// entry), so block b covers rows [blk[2b], blk[2b+2]) and entries
// [blk[2b+1], blk[2b+3]) -- one scalar load gives a block its whole range.
// A block either holds at most kBlockRows rows whose entries fit the LDS
// stage (<= kBlockNnz), or exactly one row longer than that ("long row",
// reduced by the whole workgroup).
struct fx_csr_s {
  fx_ctx ctx = nullptr;
  uint64_t gen = fx::next_handle_gen();
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
  // Column deltas read by the SpMV kernels in place of ci: cd[k] = ci[k] - r0
  // with r0 the first row of entry k's row block.  Present (non-null) only
  // when every block's columns lie within [r0 - 32768, r0 + 32767], as for
  // banded and stencil matrices; 2 bytes per entry instead of 4.
  int16_t* cd = nullptr;
  // Row pointers relative to the row's block's first entry, 2 bytes each
  // (a block holds at most kBlockNnz entries or one row): what the SpMV
  // kernels read in place of rp.
  uint16_t* rpb = nullptr;
  // Row patterns, in place of cd / ci in the SpMV kernels, for matrices made
  // from host arrays whose rows repeat at most 255 column-offset patterns of
  // at most kPatMaxLen entries (stencil matrices): row i's entry j is in
  // column i + pdict[pid[i] * plen + j].  1 byte per row instead of 2 per
  // entry.
  uint8_t* pid = nullptr;
  int32_t* pdict = nullptr;
  uint32_t plen = 0, npat = 0;
  uint32_t pat_rows = 256;  // rows per row block of a row-pattern matrix (256 or 512)
  // set on a transpose: entry p of this matrix is entry src[p] of its source
  uint32_t* src = nullptr;
  // Sliced rows, for the f64 kernels of a matrix made from host arrays
  // whose columns are too scattered for 16-bit deltas or row patterns
  // (irregular matrices; the f32 kernels keep the row blocks above, which
  // measured faster for them).  Row blocks of kSellRows rows ("windows",
  // sblk / nwin in the format of blk / nblk).  In each window the rows of
  // at most kSellMax entries are sorted by length (longest first) into
  // slices of 64 rows stored column-major: slot soff[s] + j * 64 + lane
  // holds entry j of the slice's lane-th row, so a wave reads a slice with
  // whole-line loads and no LDS staging.  The window's longer rows are
  // listed in lrow (longest first) and read from the CSR arrays.
  bool sell = false;
  uint32_t nwin = 0;
  uint32_t* sblk = nullptr;
  uint32_t nslice = 0, sslots = 0;  // slices; padded slots (soff[nslice])
  uint32_t* swin = nullptr;   // nwin + 1: first slice of each window
  uint32_t* soff = nullptr;   // nslice + 1: first slot of each slice
  uint32_t* slane = nullptr;  // nslice * 64: (row - window's first row) << 16 | length,
                              // 0xFFFFFFFF for a lane without a row
  uint32_t* scol = nullptr;   // sslots: columns (padding: 0)
  uint32_t* smap = nullptr;   // sslots: the CSR entry in each slot (padding: UINT32_MAX)
  double* sval = nullptr;     // sslots: f64 values (padding: 0)
  uint32_t* lwin = nullptr;   // nwin + 1: first long row of each window in lrow
  uint32_t* lrow = nullptr;   // rows of more than kSellMax entries
  // Rows sorted by length, for matrices made from host arrays whose columns
  // take neither row patterns nor 16-bit deltas (irregular matrices), in
  // both precisions: a copy of the matrix with its rows in order of
  // decreasing length (ties in row order), cut into row blocks as above
  // (oblk / onblk / orpb / ocol / oval / oval32 in the formats of blk / nblk
  // / rpb / ci / va / va32).  Sorted row s is row omap[s]; each row keeps
  // its entries in their order, so every row is summed as before and only
  // the assignment of rows to workgroups changes.
  bool sorted = false;
  uint32_t onblk = 0;
  uint32_t* oblk = nullptr;
  uint16_t* orpb = nullptr;
  uint32_t* orp = nullptr;   // n + 1: row pointers of the sorted copy
  uint32_t* omap = nullptr;  // n: the matrix row of each sorted row
  uint32_t* ocol = nullptr;
  double* oval = nullptr;
  float* oval32 = nullptr;
};

struct fx_pc_s {
  fx_ctx ctx = nullptr;
  uint64_t gen = fx::next_handle_gen();
  int kind = 0;  // 0 identity, 1 jacobi, 2 symmetric multicolor Gauss-Seidel
  fx_mcgs mcgs = nullptr;  // kind 2
  fx_csr A = nullptr;
  double weight = 1.0;
  uint32_t iters = 1;
  double* dinv = nullptr;   // f64 weight * (1/diag)
  float* dinv32 = nullptr;  // f32 copy for mixed precision (lazy)
};

namespace fx {

// Row-block geometry of the SpMV kernels (LDS stage of kBlockNnz entries).
// 1792 entries = 256 rows of 7: with 16-bit columns an f64 block takes
// 19 KB of LDS, so eight blocks fit in a CU's 160 KB.
#ifndef FX_BLOCK_NNZ
#define FX_BLOCK_NNZ 1792
#endif
static_assert(FX_BLOCK_NNZ % 256 == 0, "a whole number of staged entries per thread");
#ifndef FX_BLOCK_ROWS
#define FX_BLOCK_ROWS 256
#endif
constexpr uint32_t kBlockRows = FX_BLOCK_ROWS;  // a multiple of kThreads
static_assert(kBlockRows % kThreads == 0, "row blocks hold whole passes of the workgroup");
constexpr uint32_t kBlockNnz = FX_BLOCK_NNZ;
constexpr uint32_t kPatMaxLen = 32;  // longest row the row-pattern format takes
constexpr uint32_t kPatMax = 255;    // most patterns it takes (1-byte ids)
constexpr uint32_t kPatDict = 512;   // largest dictionary (patterns x plen offsets) it takes
// Row blocks of a row-pattern matrix: 256 rows of up to 7 entries on
// average, as the general blocks; or, for the refinement's f32 copy of the
// matrix (pat_rows == 512), 512 rows (two per thread).
constexpr uint32_t kPatRows = 256;
constexpr uint32_t kPatNnz = 7 * kPatRows;
constexpr uint32_t kPat2Rows = 512;
constexpr uint32_t kPat2Nnz = 7 * kPat2Rows;
// Sliced rows (irregular matrices): windows of kSellRows rows, four slices
// of 64 rows at most; rows of at most kSellMax entries go into slices.
constexpr uint32_t kSellRows = kThreads;
#ifndef FX_SELL_MAX
#define FX_SELL_MAX 32
#endif
constexpr uint32_t kSellMax = FX_SELL_MAX;
constexpr uint32_t kSellNoRow = 0xFFFFFFFFu;

int csr_spmv_dispatch(fx_ctx ctx, fx_csr A, const void* x, void* y, int dtype,
                      hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
int ensure_dinv32(fx_pc pc);
// Slice values of a sliced-row matrix from CSR-ordered values (fx_csr.hip).
int sell_fill64(fx_csr A, const double* va, double* dst);
// The sorted copy's values from the matrix's (f64: A->va into A->oval;
// f32: src, in the matrix's entry order, into dst).
int sorted_fill(fx_csr A);
int sorted_fill_f32(fx_csr A, const float* src, float* dst);

// Host-side launch geometry for streaming kernels over n elements.
inline unsigned stream_blocks(uint32_t n, uint32_t per_thread = 4) {
  uint64_t b = ((uint64_t)n + (uint64_t)kThreads * per_thread - 1) / ((uint64_t)kThreads * per_thread);
  if (b < 1) b = 1;
  if (b > (uint64_t)kMaxRedBlocks) b = kMaxRedBlocks;
  return (unsigned)b;
}

// Grid of the row-block SpMV kernels: one workgroup per row block, padded to
// a multiple of the XCD count so that xcd_block() can give each XCD one
// contiguous range of row blocks.
inline unsigned spmv_grid(uint32_t nblk, int /*cus*/) {
  unsigned g = nblk ? nblk : 1;
  return (g + kXcds - 1) / kXcds * kXcds;
}

// Row blocks an SpMV of matrix-value type MT runs on (one workgroup and one
// reduction partial each): the sliced-row windows for f64 when the matrix
// has them, else the row blocks.
template <class MT>
inline uint32_t spmv_units(const fx_csr_s* A) {
  if (A->sorted) return A->onblk;
  return (sizeof(MT) == 8 && A->sell) ? A->nwin : A->nblk;
}
template <class MT>
inline unsigned spmv_grid_of(const fx_csr_s* A, int cus = 0) {
  return spmv_grid(spmv_units<MT>(A), cus);
}

// ---------------------------------------------------------------------------
// Device building blocks
// ---------------------------------------------------------------------------

// 16-byte packs for the streaming kernels: 4 floats or 2 doubles per lane
// per load (dwordx4).  Vectors are hipMalloc'd, so pack c of a vector is
// 16-byte aligned.
template <class T>
struct alignas(16) Pack16 {
  static constexpr int W = 16 / (int)sizeof(T);
  T v[W];
};
template <class T>
__device__ __forceinline__ Pack16<T> ld16(const T* __restrict__ p, uint32_t c) {
  return reinterpret_cast<const Pack16<T>*>(p)[c];
}
template <class T>
__device__ __forceinline__ void st16(T* __restrict__ p, uint32_t c, const Pack16<T>& v) {
  reinterpret_cast<Pack16<T>*>(p)[c] = v;
}
// Element count of a vector allocation rounded up to whole 16-byte packs of
// any element type (a multiple of 4 elements, at least 4), so that pack
// loads never leave the allocation.
inline size_t pack_round(uint32_t n) { return ((size_t)n + 3) / 4 * 4 + (n ? 0 : 4); }

// Grid of the pack-streaming kernels: two packs per thread and pass.
inline unsigned pack_blocks(uint32_t n, int W) {
  return stream_blocks((n + W - 1) / W, 2);
}

// Workgroup id -> logical row block.  Hardware hands workgroup i to XCD
// i % 8; this maps XCD k to the contiguous logical range [k*G/8, (k+1)*G/8),
// so neighbouring row blocks (which share x entries) share an L2.
#ifndef FX_XCD_REMAP
#define FX_XCD_REMAP 1
#endif
__device__ __forceinline__ unsigned xcd_block() {
  const unsigned g = gridDim.x, i = blockIdx.x;
  if (!FX_XCD_REMAP) return i;
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

// Two-stage grid reduction without fences or atomics.  Stage 1: every
// workgroup of a producer kernel reduces its per-thread values and writes
// one partial per slot to partials[k * gridDim.x + blockIdx.x].  Stage 2: a
// one-workgroup finisher kernel (k_finish) launched right after it on the
// same stream reduces the G partials in a fixed order and runs the caller's
// finisher.  The kernel boundary makes the partials visible; the order of
// summation depends only on G, so results are deterministic.  (A per-workgroup
// device-scope fence would write back the whole L2 on gfx950, for every
// workgroup.)
template <int K, unsigned MaxMask = 0u>
__device__ __forceinline__ void block_partials(double (&v)[K], double* partials) {
  block_reduce<K, MaxMask>(v);
  if (threadIdx.x == 0) {
    const unsigned G = gridDim.x;
#pragma unroll
    for (int k = 0; k < K; ++k) partials[k * G + blockIdx.x] = v[k];
  }
}

// Optional per-thread set-up of kernel functors (e.g. loading a device scalar
// once): called when the functor has an init() member.
template <class T, class = void>
struct has_init : std::false_type {};
template <class T>
struct has_init<T, std::void_t<decltype(std::declval<T&>().init())>> : std::true_type {};
template <class T>
__device__ __forceinline__ void maybe_init(T& t) {
  if constexpr (has_init<T>::value) t.init();
}

// Optional workgroup hooks of SpMV epilogues, called by every thread of
// every workgroup that did not skip: block_start() after init(), and
// block_end(acc) with the thread's accumulators before the partial sums
// are written.  Both may synchronise the workgroup.
template <class T, class = void>
struct has_block_start : std::false_type {};
template <class T>
struct has_block_start<T, std::void_t<decltype(std::declval<T&>().block_start())>>
    : std::true_type {};
template <class T>
__device__ __forceinline__ void maybe_block_start(T& t) {
  if constexpr (has_block_start<T>::value) t.block_start();
}
template <class T, class A, class = void>
struct has_block_end : std::false_type {};
template <class T, class A>
struct has_block_end<T, A, std::void_t<decltype(std::declval<T&>().block_end(std::declval<A&>()))>>
    : std::true_type {};
template <class T, class A>
__device__ __forceinline__ void maybe_block_end(T& t, A& acc) {
  if constexpr (has_block_end<T, A>::value) t.block_end(acc);
}

// Finisher functor: bool skip() const; void fin(double (&sum)[K]) (all
// threads of the finisher workgroup call it; it usually acts in thread 0);
// optionally void begin(), called first, so that loads of solver state
// overlap with the partials' reads.
// 1024 threads.  Up to kFinThreads * kFinUnroll partials per slot every
// load of every slot is issued before the first wait (indices are clamped,
// the identity substituted after the load), so the whole read costs about
// one memory latency; larger G falls back to a strided loop.
constexpr int kFinThreads = 1024;
constexpr int kFinUnroll = 8;
template <class T, class = void>
struct has_begin : std::false_type {};
template <class T>
struct has_begin<T, std::void_t<decltype(std::declval<T&>().begin())>> : std::true_type {};

// Per-thread part of a partials reduction over NT threads for G <= PER * NT:
// PER loads per slot, all issued before the first use.
template <int K, unsigned MaxMask, int PER, int NT>
__device__ __forceinline__ void fin_load(const double* __restrict__ partials, unsigned G, int tid,
                                         double (&acc)[K]) {
  double v[K][PER];
#pragma unroll
  for (int k = 0; k < K; ++k)
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const unsigned b = tid + j * NT;
      v[k][j] = partials[k * G + (b < G ? b : G - 1)];
    }
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const bool mx = (MaxMask >> k) & 1u;
    const double ident = mx ? -INFINITY : 0.0;
    double a = ident;
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const double x = ((unsigned)(tid + j * NT) < G) ? v[k][j] : ident;
      a = mx ? fmax(a, x) : a + x;
    }
    acc[k] = a;
  }
}

// Per-thread sums of the G partials of each of K slots (layout
// partials[k * G + g]) over a workgroup of NT threads; the same order for
// every workgroup that calls it, so every caller gets the same bits.
template <int K, unsigned MaxMask, int NT>
__device__ __forceinline__ void partials_thread(const double* __restrict__ partials, unsigned G,
                                                int tid, double (&acc)[K]) {
  if (G <= (unsigned)NT) {
    fin_load<K, MaxMask, 1, NT>(partials, G, tid, acc);
  } else if (G <= 2u * NT) {
    fin_load<K, MaxMask, 2, NT>(partials, G, tid, acc);
  } else if (G <= 4u * NT) {
    fin_load<K, MaxMask, 4, NT>(partials, G, tid, acc);
  } else if (G <= 8u * NT) {
    fin_load<K, MaxMask, 8, NT>(partials, G, tid, acc);
  } else if (G <= 16u * NT) {
    fin_load<K, MaxMask, 16, NT>(partials, G, tid, acc);
  } else {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const bool mx = (MaxMask >> k) & 1u;
      const double ident = mx ? -INFINITY : 0.0;
      double a = ident;
      for (unsigned base = 0; base < G; base += NT * kFinUnroll) {
        double v[kFinUnroll];
#pragma unroll
        for (int j = 0; j < kFinUnroll; ++j) {
          const unsigned b = base + tid + j * NT;
          v[j] = b < G ? partials[k * G + b] : ident;
        }
#pragma unroll
        for (int j = 0; j < kFinUnroll; ++j) a = mx ? fmax(a, v[j]) : a + v[j];
      }
      acc[k] = a;
    }
  }
}

// Combine the per-thread values of partials_thread over the workgroup (wave
// butterflies, then the waves in order); the result is valid in every thread.
template <int K, unsigned MaxMask, int NT>
__device__ __forceinline__ void partials_combine(double (&acc)[K]) {
  __shared__ double red[K][NT / 64];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const bool mx = (MaxMask >> k) & 1u;
    const double a = mx ? wave_max(acc[k]) : wave_sum(acc[k]);
    if (lane == 0) red[k][wid] = a;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const bool mx = (MaxMask >> k) & 1u;
    double a = red[k][0];
#pragma unroll
    for (int w = 1; w < NT / 64; ++w) a = mx ? fmax(a, red[k][w]) : a + red[k][w];
    acc[k] = a;
  }
  __syncthreads();  // red may be reused by the next call
}

// block_partials for a workgroup of NT threads.
template <int K, unsigned MaxMask, int NT>
__device__ __forceinline__ void block_partials_n(double (&v)[K], double* partials) {
  partials_combine<K, MaxMask, NT>(v);
  if (threadIdx.x == 0) {
    const unsigned G = gridDim.x;
#pragma unroll
    for (int k = 0; k < K; ++k) partials[k * G + blockIdx.x] = v[k];
  }
}

template <int K, unsigned MaxMask, class Fin>
__global__ __launch_bounds__(kFinThreads) void k_finish(const double* __restrict__ partials,
                                                        unsigned G, Fin f0) {
  Fin f = f0;
  if constexpr (has_begin<Fin>::value) f.begin();
  double acc[K];
  partials_thread<K, MaxMask, kFinThreads>(partials, G, threadIdx.x, acc);
  // the stop test comes after the loads are issued (a stopped solver's
  // stale partials are read and dropped), so its latency is not added
  if (f.skip()) return;
  partials_combine<K, MaxMask, kFinThreads>(acc);
  f.fin(acc);
}

// Building a matrix on the device (fx_csr.hip): allocate rp / ci / dp / va
// for A's nrow and nnz; free a matrix; exclusive scan of cnt[0..m) into
// out[0..m] (out[m] = total; one workgroup); and, once rp / ci / va are
// filled, form diagPtr (first match, as FormDiagPtr) and the row-block
// tables.
int csr_storage(fx_csr A);
// Row blocks of kPat2Rows rows for a row-pattern matrix (its row pointers
// are read back from the device): the (first row, first entry) table, the
// 2-byte block-relative row pointers and the block count, in new device
// allocations owned by the caller.
int csr_pattern_blocks512(fx_csr A, uint32_t** blk, uint16_t** rpb, uint32_t* nblk);
void csr_release(fx_csr A);
int scan_u32(hipStream_t st, const uint32_t* cnt, uint32_t m, uint32_t* out);
int csr_finish_device(fx_csr A);

template <int K, unsigned MaxMask, class Fin>
inline int launch_finish(hipStream_t st, const double* partials, unsigned G, Fin f) {
  hipLaunchKernelGGL((k_finish<K, MaxMask, Fin>), dim3(1), dim3(kFinThreads), 0, st, partials, G,
                     f);
  FX_HIP(hipGetLastError());
  return 0;
}

// Finisher that stores the reduced slots; slot 0 optionally transformed
// (post0 = 1: sqrt, 2: negate).
template <int K>
struct FinStore {
  double* out;
  int post0;
  __device__ bool skip() const { return false; }
  __device__ void fin(double (&s)[K]) const {
    if (threadIdx.x != 0) return;
#pragma unroll
    for (int k = 0; k < K; ++k)
      out[k] = (k == 0 && post0 == 1) ? sqrt(s[k]) : (k == 0 && post0 == 2) ? -s[k] : s[k];
  }
};

}  // namespace fx
