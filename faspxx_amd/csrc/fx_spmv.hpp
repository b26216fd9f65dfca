// fx_spmv.hpp -- the row-blocked CSR SpMV kernel (MAT::Apply /
// MAT::Residual, src/MAT.cxx:373-476) with pluggable per-row epilogues and
// an optional fused grid reduction.
//
// One workgroup (256 threads) per row block:
//   * the block's row pointers are staged in LDS;
//   * its column indices and values are read once, fully coalesced, into LDS;
//     the columns come as 2-byte deltas to the block's first row when the
//     matrix has them (banded / stencil matrices), else as u32; a matrix
//     stored as row patterns (stencils) stages only its values and the small
//     pattern dictionary, and forms each column as row + offset (pat_block);
//   * rows are then reduced with x gathered straight from memory.  While the
//     block's rows average at most 32 entries each row is summed by one
//     thread, left to right, in the reference's order (w = v0*x0, then
//     w = fma(vk, xk, w)), so the result equals the reference's FMA-contracted
//     build bit for bit; for banded / stencil matrices a wave's k-th gathers
//     then hit consecutive x entries.  Blocks of longer rows give each row a
//     lane group of 2..64 lanes (butterfly sum), and a row of more than
//     kLongRow entries among short ones is summed by the whole workgroup;
//   * a row longer than the LDS stage gets a whole workgroup.
// The epilogue receives (row, value) and may store, update other vectors and
// add to per-thread reduction slots; the slots are then reduced over the grid
// and a one-workgroup finisher kernel runs the epilogue's fin().
#pragma once

#include <hip/hip_ext.h>

#include <type_traits>

#include "fx_internal.hpp"

namespace fx {

enum { ROW_AX = 0, ROW_B_MINUS_AX = 1 };

template <class MT, class VT>
struct SpmvIn {
  const uint32_t* __restrict__ rp;
  const uint16_t* __restrict__ rpb;  // row pointers relative to the block's first entry
  const uint32_t* __restrict__ ci;
  const int16_t* __restrict__ cd;  // column deltas to the block's first row, or null
  const MT* __restrict__ va;  // null: structure-only matrix (values are 1)
  const uint32_t* __restrict__ blk;  // (first row, first entry) of each block, nblk + 1 pairs
  uint32_t nblk;
  uint32_t nnz;  // >= 1 whenever a kernel runs
  const uint8_t* __restrict__ pid;     // row patterns: pattern of each row, or null
  const int32_t* __restrict__ pdict;   // offsets col - row, plen per pattern
  uint32_t plen;
  uint32_t dlen;  // npat * plen
  // sliced rows (see fx_csr_s): slices of each window, their slots, lanes,
  // columns and values; the windows' long rows
  const uint32_t* __restrict__ swin;
  const uint32_t* __restrict__ soff;
  const uint32_t* __restrict__ slane;
  const uint32_t* __restrict__ scol;
  const MT* __restrict__ sval;  // null: structure-only matrix
  const uint32_t* __restrict__ lwin;
  const uint32_t* __restrict__ lrow;
  // rows sorted by length (u32 columns): the matrix row of each block row,
  // or null when block rows are matrix rows
  const uint32_t* __restrict__ rmap;
};

// Column formats of the staged entries: u32 columns, 16-bit deltas to the
// block's first row, or row patterns (no per-entry column data) in blocks
// of 256 rows (CF_PAT) or 512 rows (CF_PAT2).
enum { CF_U32 = 0, CF_D16 = 1, CF_PAT = 2, CF_PAT2 = 3, CF_SELL = 4 };
constexpr bool is_pat(int cf) { return cf == CF_PAT || cf == CF_PAT2; }

template <class MT, class VT>
inline SpmvIn<MT, VT> spmv_in(fx_csr A);

// f64 kernels of a sliced-row matrix run on its windows; everything else
// on the row blocks (the f32 kernels get no slice arrays).
template <>
inline SpmvIn<double, double> spmv_in<double, double>(fx_csr A) {
  if (A->sorted)
    return {A->orp,   A->orpb, A->ocol, nullptr, A->has_values ? A->oval : nullptr,
            A->oblk,  A->onblk, A->nnz, nullptr, nullptr,
            0,        0,
            nullptr, nullptr, nullptr, nullptr, nullptr,
            nullptr, nullptr, A->omap};
  return {A->rp,   A->rpb,  A->ci,  A->cd,    A->has_values ? A->va : nullptr,
          A->sell ? A->sblk : A->blk,  A->sell ? A->nwin : A->nblk, A->nnz, A->pid,   A->pdict,
          A->plen, A->npat * A->plen,
          A->swin, A->soff, A->slane, A->scol, A->has_values ? A->sval : nullptr,
          A->lwin, A->lrow, nullptr};
}
template <>
inline SpmvIn<float, float> spmv_in<float, float>(fx_csr A) {
  if (A->sorted)
    return {A->orp,   A->orpb, A->ocol, nullptr, A->has_values ? A->oval32 : nullptr,
            A->oblk,  A->onblk, A->nnz, nullptr, nullptr,
            0,        0,
            nullptr, nullptr, nullptr, nullptr, nullptr,
            nullptr, nullptr, A->omap};
  return {A->rp,   A->rpb,  A->ci,  A->cd,    A->has_values ? A->va32 : nullptr,
          A->blk,  A->nblk, A->nnz, A->pid,   A->pdict,
          A->plen, A->npat * A->plen,
          nullptr, nullptr, nullptr, nullptr, nullptr,
          nullptr, nullptr, nullptr};
}

// Threads per row for a block of nr rows holding nz entries: one thread per
// row (the reference's summation order) while rows average at most
// kSeqRowLen entries; longer rows get the largest power-of-two lane group
// that both fits 256/nr and leaves each lane at least kSeqRowLen/2 entries,
// capped at one wave.
constexpr uint32_t kSeqRowLen = 32;
// A row longer than this in a block whose rows are summed one per thread is
// summed by the whole workgroup instead.
#ifndef FX_LONG_ROW
#define FX_LONG_ROW 128
#endif
constexpr uint32_t kLongRow = FX_LONG_ROW;
__device__ __forceinline__ int threads_per_row(uint32_t nr, uint32_t nz) {
  if (nr == 0 || nz <= kSeqRowLen * nr) return 1;
  const uint32_t by_rows = kThreads / nr;
  const uint32_t by_len = (2 * nz) / (kSeqRowLen * nr);
  const uint32_t q = by_rows < by_len ? by_rows : by_len;
  int t = 1;
  while ((uint32_t)(t * 2) <= q && t < 64) t *= 2;
  return t;
}

// Matrix-stream loads: ordinary (default) or non-temporal (FX_NT_MATRIX=1;
// the u32 column form is always non-temporal, see stream_load).
// Ordinary loads leave the matrix in the 256 MB Infinity Cache between the
// iterations of a solve when it fits (3D Poisson 128^3: 188 MB in f64):
// measured 12% less time per f64 CG iteration and 28 against 37 us for a
// warm SpMV, and the same time for an SpMV after a cache flush.
#ifndef FX_NT_MATRIX
#define FX_NT_MATRIX 0
#endif
// The plain SpMV's result: non-temporal stores (FX_NT_Y=1, default) or
// ordinary ones.
#ifndef FX_NT_Y
#define FX_NT_Y 1
#endif
// Row pointers: 2-byte block-relative (FX_RP16=1, default) or u32.
#ifndef FX_RP16
#define FX_RP16 1
#endif

// LDS image of a row block: row pointers relative to the block's first
// entry, then the staged (column, value) pairs.  CF_D16: the columns stay in
// LDS as the 16-bit deltas to the block's first row r0 that were loaded,
// 10 bytes per f64 entry instead of 12, so that more blocks fit in a CU.
// CF_PAT: values, and a copy of the pattern dictionary (columns come from
// the row's pattern).
template <class MT, int CF>
struct SpmvLds {
  using RpT = std::conditional_t<FX_RP16 != 0, uint16_t, uint32_t>;
  using CiT = std::conditional_t<CF == CF_D16, int16_t, uint32_t>;
  static constexpr uint32_t kRows = CF == CF_SELL   ? 1
                                    : CF == CF_PAT2 ? kPat2Rows
                                    : CF == CF_PAT  ? kPatRows
                                                    : kBlockRows;
  static constexpr uint32_t kNnz = CF == CF_SELL   ? 1
                                   : CF == CF_PAT2 ? kPat2Nnz
                                   : CF == CF_PAT  ? kPatNnz
                                                   : kBlockNnz;
  static constexpr uint32_t kCi = (is_pat(CF) || CF == CF_SELL) ? 1 : kBlockNnz;
  static constexpr uint32_t kDict = is_pat(CF) ? kPatDict : 1;
  RpT rp[kRows + 1];
  CiT ci[kCi];
  int32_t dict[kDict];
  MT va[kNnz];
  __device__ __forceinline__ uint32_t col(uint32_t k, uint32_t r0) const {
    if constexpr (CF == CF_D16) return (uint32_t)((int32_t)r0 + (int32_t)ci[k]);
    else return ci[k];
  }
};
// The u32 column form (scattered columns, as in the power-law matrix)
// always streams its matrix non-temporally, so that it does not evict the
// gathered x from L2: 214 against 221 us for that SpMV, 5.33 against 5.43 ms
// for its CG.
template <bool NT = FX_NT_MATRIX != 0, class T>
__device__ __forceinline__ T stream_load(const T* p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}

// Stage a row block into LDS: its nr + 1 row pointers (nr <= kBlockRows;
// stored relative to k0) and its nz entries [k0, k0 +
// nz).  The entry loads are unconditional (indices past the block's end are
// clamped to its last entry) and branch-free, and the row-pointer loads come
// after them, so all of a thread's loads are in flight before the first
// wait.  CF: the column format (CF_U32 or CF_D16); VA: the matrix has
// values.
template <int CF, bool VA, class MT, class VT>
__device__ __forceinline__ void stage_block(const SpmvIn<MT, VT>& in, SpmvLds<MT, CF>& lds,
                                            uint32_t r0, uint32_t k0, uint32_t nr, uint32_t nz,
                                            int tid) {
  constexpr int kPer = kBlockNnz / kThreads;
  // last entry index a load may touch: the block's last entry, or for a
  // block of empty rows the matrix's last entry (nothing is stored then)
  const uint32_t klast = nz ? k0 + nz - 1 : in.nnz - 1;
  constexpr bool kNT = FX_NT_MATRIX != 0 || CF == CF_U32;
  typename SpmvLds<MT, CF>::CiT cix[kPer];
  MT vv[kPer];
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const uint32_t k = min(k0 + (uint32_t)(tid + j * kThreads), klast);
    if constexpr (CF == CF_D16)
      cix[j] = stream_load<kNT>(&in.cd[k]);
    else if constexpr (CF == CF_U32)
      cix[j] = stream_load<kNT>(&in.ci[k]);
    if constexpr (VA) vv[j] = stream_load<kNT>(&in.va[k]);
  }
#if FX_RP16
  // 2-byte block-relative row pointers; the block's end is nz
  constexpr int kRp = kBlockRows / kThreads;  // row pointers per thread
  uint32_t rpv[kRp];
#pragma unroll
  for (int j = 0; j < kRp; ++j) {
    const uint32_t i = tid + j * kThreads;
    rpv[j] = in.rpb[r0 + (i < nr ? i : nr - 1)];
  }
#else
  constexpr int kRp = kBlockRows / kThreads + 1;  // row pointers per thread
  uint32_t rpv[kRp];
#pragma unroll
  for (int j = 0; j < kRp; ++j) {
    const uint32_t i = tid + j * kThreads;
    rpv[j] = in.rp[r0 + (i <= nr ? i : nr)] - k0;
  }
#endif
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const uint32_t k = tid + j * kThreads;
    if (k < nz) {
      lds.ci[k] = cix[j];
      lds.va[k] = VA ? vv[j] : (MT)1;
    }
  }
#pragma unroll
  for (int j = 0; j < kRp; ++j) {
    const uint32_t i = tid + j * kThreads;
#if FX_RP16
    if (i < nr) lds.rp[i] = rpv[j];
#else
    if (i <= nr) lds.rp[i] = rpv[j];
#endif
  }
#if FX_RP16
  if (tid == 0) lds.rp[nr] = nz;
#endif
}

// A row-pattern block (CF_PAT / CF_PAT2: at most 256 / 512 rows of at most
// kPatMaxLen entries, one / two rows per thread).  The loads go out in the
// order: the
// pattern dictionary (small, L2-resident), the row's pattern id and bounds,
// then the block's values (f32), or the values first (f64).  Each row is
// summed left to right as in the thread-per-row loop of k_spmv, with
// column r + dict[pattern * plen + j].
// f32 (early gathers): after a first barrier (the dictionary in LDS) each
// row's first 8 x gathers are issued while the values are still in flight;
// the dictionary stores and these gathers are branch-free (threads without
// a row gather x[0]; every thread stores its own dictionary slots, those
// past the dictionary's end with values never read), so that the waits for
// the values never wait for the gathers.  f64: values first and the
// gathers after the one barrier,
// which measured faster for the HBM-bound f64 SpMV (32.2 against 33.5 us
// on the 128^3 stencil) though slower inside CG (20.4 against 20.3 ms);
// the f32 order won every f32 measurement (IR 11.3 against 11.6 ms).
template <int CF, bool VA, int MODE, int NRED, class MT, class VT, class XF, class BF, class Epi>
__device__ __forceinline__ void pat_block(const SpmvIn<MT, VT>& in, SpmvLds<MT, CF>& lds,
                                          uint32_t r0, uint32_t k0, uint32_t nr, uint32_t nz,
                                          int tid, const XF& xf, const BF& bf, const Epi& epi,
                                          double (&acc)[NRED > 0 ? NRED : 1]) {
  constexpr int R = SpmvLds<MT, CF>::kRows / kThreads;  // rows per thread: tid + q * kThreads
  constexpr int kPer = SpmvLds<MT, CF>::kNnz / kThreads;
  constexpr int kDv = kPatDict / kThreads;
  constexpr bool kEarly = sizeof(MT) == 4;
  const uint32_t klast = nz ? k0 + nz - 1 : in.nnz - 1;
  bool has[R];
  uint32_t row[R], pidv[R], s[R], e[R];
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const uint32_t rq = tid + q * kThreads;
    has[q] = rq < nr;
    row[q] = has[q] ? rq : nr - 1;
    s[q] = e[q] = 0;
  }
  int32_t dv[kDv];
  MT vv[kPer];
  // f32: dictionary, pattern ids, the rows' bounds, then values
  // f64: values, the rows' starts (to LDS), pattern ids, dictionary (only
  //      the slots it has)
  if constexpr (kEarly) {
#pragma unroll
    for (int j = 0; j < kDv; ++j) {
      const uint32_t i = tid + j * kThreads;
      dv[j] = in.pdict[i < in.dlen ? i : in.dlen - 1];
    }
#pragma unroll
    for (int q = 0; q < R; ++q) {
      pidv[q] = in.pid[r0 + row[q]];
#if FX_RP16
      s[q] = in.rpb[r0 + row[q]];
      const uint32_t e_next = in.rpb[r0 + (row[q] + 1 < nr ? row[q] + 1 : row[q])];
      e[q] = row[q] + 1 < nr ? e_next : nz;
#else
      s[q] = in.rp[r0 + row[q]] - k0;
      e[q] = in.rp[r0 + row[q] + 1] - k0;
#endif
    }
  }
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const uint32_t k = min(k0 + (uint32_t)(tid + j * kThreads), klast);
    vv[j] = VA ? stream_load(&in.va[k]) : (MT)1;
  }
  uint32_t rstart[R];
  if constexpr (!kEarly) {
#pragma unroll
    for (int q = 0; q < R; ++q) {
#if FX_RP16
      rstart[q] = in.rpb[r0 + row[q]];
#else
      rstart[q] = in.rp[r0 + row[q]] - k0;
#endif
      pidv[q] = in.pid[r0 + row[q]];
    }
#pragma unroll
    for (int j = 0; j < kDv; ++j) {
      const uint32_t i = tid + j * kThreads;
      if (i < in.dlen) dv[j] = in.pdict[i];
    }
  }
  double xv[R][8];
  if constexpr (kEarly) {
#pragma unroll
    for (int j = 0; j < kDv; ++j) lds.dict[tid + j * kThreads] = dv[j];  // slots >= dlen: unread
    __syncthreads();
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const bool live = has[q] && e[q] > s[q];
      const int32_t* pd = lds.dict + (live ? pidv[q] * in.plen : 0u);
      const uint32_t lm1 = live ? e[q] - s[q] - 1 : 0u;
      const uint32_t ri = r0 + row[q];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const uint32_t c = (uint32_t)((int32_t)ri + pd[min((uint32_t)j, lm1)]);
        xv[q][j] = (double)xf(live ? c : 0u);
      }
    }
  } else {
#pragma unroll
    for (int j = 0; j < kDv; ++j) {
      const uint32_t i = tid + j * kThreads;
      if (i < in.dlen) lds.dict[i] = dv[j];
    }
#pragma unroll
    for (int q = 0; q < R; ++q)
      if (has[q]) lds.rp[row[q]] = rstart[q];
    if (tid == 0) lds.rp[nr] = nz;
  }
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const uint32_t k = tid + j * kThreads;
    if (k < nz) lds.va[k] = vv[j];
  }
  __syncthreads();
  if constexpr (!kEarly) {
    // every row's first chunk of gathers goes out before any sum
#pragma unroll
    for (int q = 0; q < R; ++q) {
      s[q] = lds.rp[row[q]];
      e[q] = lds.rp[row[q] + 1];
      const bool live = has[q] && e[q] > s[q];
      const int32_t* pd = lds.dict + (live ? pidv[q] * in.plen : 0u);
      const uint32_t lm1 = live ? e[q] - s[q] - 1 : 0u;
      const uint32_t ri = r0 + row[q];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const uint32_t c = (uint32_t)((int32_t)ri + pd[min((uint32_t)j, lm1)]);
        xv[q][j] = (double)xf(live ? c : 0u);
      }
    }
  }
#pragma unroll
  for (int q = 0; q < R; ++q) {
    if (!has[q]) continue;
    const uint32_t ri = r0 + row[q], sq = s[q], eq = e[q];
    const int32_t* pd = lds.dict + pidv[q] * in.plen;
    double t = (MODE == ROW_AX) ? 0.0 : (double)bf(ri);
    for (uint32_t base = sq; base < eq; base += 8) {
      double av[8], xw[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        if (base == sq) {
          xw[j] = xv[q][j];
        } else {
          const uint32_t k = min(base + j, eq - 1) - sq;
          xw[j] = (double)xf((uint32_t)((int32_t)ri + pd[k]));
        }
      }
#pragma unroll
      for (int j = 0; j < 8; ++j) av[j] = (double)lds.va[min(base + j, eq - 1)];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const uint32_t k = base + j;
        if (k < eq) {
          if (MODE == ROW_AX)
            t = (k == sq) ? av[j] * xw[j] : fma(av[j], xw[j], t);
          else
            t = fma(-av[j], xw[j], t);
        }
      }
    }
    epi.row(ri, t, acc);
  }
}

// Rows longer than this in a sliced-row window are summed by the whole
// workgroup rather than by one wave.
#ifndef FX_SELL_HUGE
#define FX_SELL_HUGE 256
#endif
constexpr uint32_t kSellHuge = FX_SELL_HUGE;
// Entries per lane per step of a slice (loads, then gathers, then fmas).
#ifndef FX_SELL_CHUNK
#define FX_SELL_CHUNK 8
#endif
constexpr int kSellChunk = FX_SELL_CHUNK;

// A window of a sliced-row matrix (CF_SELL): wave k sums slice k of the
// window (64 rows of at most kSellMax entries, sorted by length), each lane
// its own row left to right in the reference's order, reading the slice's
// columns and values with whole-line loads straight into registers (8
// entries' loads, then their 8 gathers, then the fma chain; slots past the
// lane's length are clamped to its last entry and their terms dropped).
// Then the window's long rows (more than kSellMax entries, longest first)
// go round the waves, a wave per row: 4 strides of 64 entries in flight,
// summed per lane in stride order, then across the wave; rows of more than
// kSellHuge entries get the whole workgroup (4 strides of 256).  The row sums go
// to LDS, and the epilogue then runs in row order, thread i on the
// window's row i, so that its loads and stores are whole lines.
template <int MODE, int NRED, class MT, class VT, class XF, class BF, class Epi>
__device__ __forceinline__ void sell_block(const SpmvIn<MT, VT>& in, uint32_t w, int tid,
                                           const XF& xf, const BF& bf, const Epi& epi,
                                           double (&acc)[NRED > 0 ? NRED : 1]) {
  __shared__ double rsum[kSellRows];
  const int wid = tid >> 6, lane = tid & 63;
  const uint32_t r0 = w * kSellRows;
  const uint32_t nr = min(kSellRows, in.blk[2 * w + 2] - r0);
  const uint32_t s = in.swin[w] + (uint32_t)wid;
  if (s < in.swin[w + 1]) {
    const uint32_t meta = stream_load<true>(&in.slane[s * 64u + (uint32_t)lane]);
    const uint32_t base = in.soff[s], width = (in.soff[s + 1] - base) >> 6;
    const bool has = meta != kSellNoRow;
    const uint32_t len = has ? (meta & 0xFFFFu) : 0u, lm1 = len ? len - 1 : 0u;
    const uint32_t row = r0 + (meta >> 16);
    double t = (MODE == ROW_AX || !has) ? 0.0 : (double)bf(row);
    const uint32_t* __restrict__ sc = in.scol + base + (uint32_t)lane;
    const MT* __restrict__ sv = in.sval ? in.sval + base + (uint32_t)lane : nullptr;
    // software pipelined: chunk k + 1's matrix loads go out before chunk k's
    // gathers are waited for
    uint32_t cj[kSellChunk];
    MT av[kSellChunk];
    if (width) {
#pragma unroll
      for (int u = 0; u < kSellChunk; ++u) {
        const uint32_t j = min((uint32_t)u, lm1);
        cj[u] = stream_load<true>(&sc[j * 64u]);
        av[u] = sv ? stream_load<true>(&sv[j * 64u]) : (MT)1;
      }
    }
    for (uint32_t j0 = 0; j0 < width; j0 += kSellChunk) {
      double xv[kSellChunk];
#pragma unroll
      for (int u = 0; u < kSellChunk; ++u) xv[u] = (double)xf(cj[u]);
      const MT* ac = av;
      MT anext[kSellChunk];
      uint32_t cnext[kSellChunk];
      if (j0 + kSellChunk < width) {
#pragma unroll
        for (int u = 0; u < kSellChunk; ++u) {
          const uint32_t j = min(j0 + kSellChunk + (uint32_t)u, lm1);
          cnext[u] = stream_load<true>(&sc[j * 64u]);
          anext[u] = sv ? stream_load<true>(&sv[j * 64u]) : (MT)1;
        }
      }
#pragma unroll
      for (int u = 0; u < kSellChunk; ++u) {
        const uint32_t j = j0 + (uint32_t)u;
        if (j < len) {
          const double a = (double)ac[u];
          if (MODE == ROW_AX)
            t = (j == 0) ? a * xv[u] : fma(a, xv[u], t);
          else
            t = fma(-a, xv[u], t);
        }
      }
      if (j0 + kSellChunk < width) {
#pragma unroll
        for (int u = 0; u < kSellChunk; ++u) {
          cj[u] = cnext[u];
          av[u] = anext[u];
        }
      }
    }
    if (has) rsum[row - r0] = t;
  }
  // the window's long rows, longest first: those of more than kSellHuge
  // entries are left to the whole workgroup below
  const uint32_t q0 = in.lwin[w], q1 = in.lwin[w + 1];
  uint32_t qh = q0;
  while (qh < q1 && in.rp[in.lrow[qh] + 1] - in.rp[in.lrow[qh]] > kSellHuge) ++qh;
  for (uint32_t q = qh + (uint32_t)wid; q < q1; q += kWaves) {
    const uint32_t r = in.lrow[q], k0 = in.rp[r], k1 = in.rp[r + 1];
    double t = 0.0;
    for (uint32_t kb = k0 + (uint32_t)lane; kb < k1; kb += 4 * 64) {
      double a[4], xv[4];
      uint32_t c[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const uint32_t k = min(kb + (uint32_t)j * 64u, k1 - 1);
        c[j] = stream_load<true>(&in.ci[k]);
        a[j] = in.va ? (double)stream_load<true>(&in.va[k]) : 1.0;
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) xv[j] = (double)xf(c[j]);
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (kb + (uint32_t)j * 64u < k1) t = fma(a[j], xv[j], t);
    }
    t = group_sum(t, 64);
    if (lane == 0) rsum[r - r0] = (MODE == ROW_AX) ? t : (double)bf(r) - t;
  }
  for (uint32_t q = q0; q < qh; ++q) {
    const uint32_t r = in.lrow[q], k0 = in.rp[r], k1 = in.rp[r + 1];
    double t = 0.0;
    for (uint32_t kb = k0 + (uint32_t)tid; kb < k1; kb += 4 * kThreads) {
      double a[4], xv[4];
      uint32_t c[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const uint32_t k = min(kb + (uint32_t)j * kThreads, k1 - 1);
        c[j] = stream_load<true>(&in.ci[k]);
        a[j] = in.va ? (double)stream_load<true>(&in.va[k]) : 1.0;
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) xv[j] = (double)xf(c[j]);
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (kb + (uint32_t)j * kThreads < k1) t = fma(a[j], xv[j], t);
    }
    double tt[1] = {t};
    block_reduce<1>(tt);
    if (tid == 0) rsum[r - r0] = (MODE == ROW_AX) ? tt[0] : (double)bf(r) - tt[0];
  }
  __syncthreads();
  if ((uint32_t)tid < nr) epi.row(r0 + (uint32_t)tid, rsum[tid], acc);
}

// An epilogue may take the row's position in the matrix's stored row order
// as well (row_at(pos, row, v, acc)), for a matrix whose rows are stored
// sorted by length: vectors kept in that order are then read in whole lines.
template <class T, class A, class = void>
struct has_row_at : std::false_type {};
template <class T, class A>
struct has_row_at<T, A, std::void_t<decltype(std::declval<const T&>().row_at(
                             0u, 0u, 0.0, std::declval<A&>()))>> : std::true_type {};
template <class Epi, class A>
__device__ __forceinline__ void emit_row(const Epi& epi, uint32_t pos, uint32_t row, double v,
                                         A& acc) {
  if constexpr (has_row_at<Epi, A>::value) epi.row_at(pos, row, v, acc);
  else epi.row(row, v, acc);
}

// Generic kernel.  X is a functor giving the gathered vector entry x_j
// (a plain load, or a value formed on the fly); B gives b_i for
// ROW_B_MINUS_AX.  Epi must provide:
//   bool skip() const                      uniform early exit (solver stop flag)
//   void row(uint32_t r, double v, double (&acc)[NRED])
//   void fin(double (&sum)[NRED])          run by k_finish after the kernel
template <class MT, class VT, int MODE, int NRED, unsigned MAXMASK, class XF, class BF, class Epi,
          int CF>
__global__ __launch_bounds__(kThreads) void k_spmv(SpmvIn<MT, VT> in, XF xf0, BF bf, Epi epi0,
                                                   double* partials) {
  if (epi0.skip()) return;
  XF xf = xf0;
  Epi epi = epi0;
  maybe_init(xf);
  maybe_init(epi);
  maybe_block_start(epi);
  __shared__ SpmvLds<MT, CF> lds;
  double acc[NRED > 0 ? NRED : 1];
#pragma unroll
  for (int k = 0; k < (NRED > 0 ? NRED : 1); ++k) acc[k] = 0.0;
  const unsigned bid = xcd_block();
  const int tid = threadIdx.x;
  if constexpr (CF == CF_SELL) {
    if (bid < in.nblk) sell_block<MODE, NRED>(in, bid, tid, xf, bf, epi, acc);
  } else if (bid < in.nblk) {
    const uint32_t r0 = in.blk[2 * bid], k0 = in.blk[2 * bid + 1];
    const uint32_t r1 = in.blk[2 * bid + 2], k1 = in.blk[2 * bid + 3];
    const uint32_t nr = r1 - r0, nz = k1 - k0;
    // the matrix row of block row r (rows sorted by length: u32 columns only)
    auto orow = [&](uint32_t r) -> uint32_t {
      if constexpr (CF == CF_U32) return in.rmap ? in.rmap[r] : r;
      else return r;
    };
    if (!is_pat(CF) && nz > kBlockNnz) {  // (row-pattern rows are short)
      // a single long row: strided partial sums (4 strides' loads in
      // flight at a time, summed in stride order), then a workgroup reduction
      double s = 0.0;
      for (uint32_t kb = k0 + tid; kb < k1; kb += 4 * kThreads) {
        double a[4], xv[4];
        uint32_t c[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const uint32_t k = min(kb + j * kThreads, k1 - 1);
          a[j] = in.va ? (double)in.va[k] : 1.0;
          if constexpr (CF == CF_D16)
            c[j] = (uint32_t)((int32_t)r0 + (int32_t)in.cd[k]);
          else
            c[j] = in.ci[k];
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) xv[j] = (double)xf(c[j]);
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if (kb + j * kThreads < k1) s = fma(a[j], xv[j], s);
      }
      double t[1] = {s};
      block_reduce<1>(t);
      if (tid == 0) {
        const uint32_t orr = orow(r0);
        const double v = (MODE == ROW_AX) ? t[0] : (double)bf(orr) - t[0];
        emit_row(epi, r0, orr, v, acc);
      }
    } else if constexpr (is_pat(CF)) {
      in.va ? pat_block<CF, true, MODE, NRED>(in, lds, r0, k0, nr, nz, tid, xf, bf, epi, acc)
            : pat_block<CF, false, MODE, NRED>(in, lds, r0, k0, nr, nz, tid, xf, bf, epi, acc);
    } else {
      in.va ? stage_block<CF, true>(in, lds, r0, k0, nr, nz, tid)
            : stage_block<CF, false>(in, lds, r0, k0, nr, nz, tid);
      __syncthreads();
      const int tpr = threads_per_row(nr, nz);
      if (tpr == 1) {
        // A row longer than kLongRow in a block of short rows is left to the
        // whole workgroup afterwards (in row order: a ballot prefix gives
        // each its slot), so that it does not hold the block's other
        // threads back; it is summed in strides like a lane-group row.
        uint32_t nlong = 0;
        if constexpr (kBlockRows == kThreads) {
          __shared__ uint16_t long_row[kBlockRows];
          __shared__ uint32_t wave_long[kWaves];
          const bool is_long = (uint32_t)tid < nr && lds.rp[tid + 1] - lds.rp[tid] > kLongRow;
          const uint64_t m = __ballot(is_long);
          const int lane = tid & 63, wid = tid >> 6;
          if (lane == 0) wave_long[wid] = (uint32_t)__popcll(m);
          __syncthreads();
          uint32_t before = 0;
#pragma unroll
          for (int w = 0; w < kWaves; ++w) {
            before += w < wid ? wave_long[w] : 0u;
            nlong += wave_long[w];
          }
          if (is_long) long_row[before + __popcll(m & ((1ull << lane) - 1ull))] = tid;
          if (nlong) {
            __syncthreads();
            for (uint32_t q = 0; q < nlong; ++q) {
              const uint32_t rr = long_row[q], s = lds.rp[rr], e = lds.rp[rr + 1];
              double t = 0.0;
              for (uint32_t kb = s + tid; kb < e; kb += 4 * kThreads) {
                double a[4], xv[4];
                uint32_t c[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                  const uint32_t k = min(kb + j * kThreads, e - 1);
                  a[j] = (double)lds.va[k];
                  c[j] = lds.col(k, r0);
                }
#pragma unroll
                for (int j = 0; j < 4; ++j) xv[j] = (double)xf(c[j]);
#pragma unroll
                for (int j = 0; j < 4; ++j)
                  if (kb + j * kThreads < e) t = fma(a[j], xv[j], t);
              }
              double tt[1] = {t};
              block_reduce<1>(tt);
              if (tid == 0) {
                const uint32_t orr = orow(r0 + rr);
                const double v = (MODE == ROW_AX) ? tt[0] : (double)bf(orr) - tt[0];
                emit_row(epi, r0 + rr, orr, v, acc);
              }
            }
          }
        }
        for (uint32_t row = tid; row < nr; row += kThreads) {
          const uint32_t s = lds.rp[row], e = lds.rp[row + 1];
          if (nlong && e - s > kLongRow) continue;  // summed above
          const uint32_t orr = orow(r0 + row);
          double t = (MODE == ROW_AX) ? 0.0 : (double)bf(orr);
          // chunks of 8: the chunk's LDS reads, then its gathers, are all
          // issued before its fma chain (indices past the row's end are
          // clamped to its last entry and their terms dropped)
          for (uint32_t base = s; base < e; base += 8) {
            uint32_t cj[8];
            double av[8], xv[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
              const uint32_t k = min(base + j, e - 1);
              cj[j] = lds.col(k, r0);
              av[j] = (double)lds.va[k];
            }
#pragma unroll
            for (int j = 0; j < 8; ++j) xv[j] = (double)xf(cj[j]);
#pragma unroll
            for (int j = 0; j < 8; ++j) {
              const uint32_t k = base + j;
              if (k < e) {
                if (MODE == ROW_AX)
                  t = (k == s) ? av[j] * xv[j] : fma(av[j], xv[j], t);
                else
                  t = fma(-av[j], xv[j], t);
              }
            }
          }
          emit_row(epi, r0 + row, orr, t, acc);
        }
      } else {
        const int grp = tid / tpr, lane = tid % tpr, ngrp = kThreads / tpr;
        for (uint32_t rr = grp; rr < nr; rr += ngrp) {
          const uint32_t s = lds.rp[rr], e = lds.rp[rr + 1];
          double t = 0.0;
          // 4 strides' gathers in flight at a time, summed in stride order
          for (uint32_t kb = s + lane; kb < e; kb += 4 * tpr) {
            double a[4], xv[4];
            uint32_t c[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
              const uint32_t k = min(kb + j * tpr, e - 1);
              a[j] = (double)lds.va[k];
              c[j] = lds.col(k, r0);
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) xv[j] = (double)xf(c[j]);
#pragma unroll
            for (int j = 0; j < 4; ++j)
              if (kb + j * tpr < e) t = fma(a[j], xv[j], t);
          }
          t = group_sum(t, tpr);
          if (lane == 0) {
            const uint32_t orr = orow(r0 + rr);
            const double v = (MODE == ROW_AX) ? t : (double)bf(orr) - t;
            emit_row(epi, r0 + rr, orr, v, acc);
          }
        }
      }
    }
  }
  maybe_block_end(epi, acc);
  if constexpr (NRED > 0) block_partials<NRED, MAXMASK>(acc, partials);
}

// Functors for the common cases.
#ifndef FX_NT_GATHER
#define FX_NT_GATHER 0
#endif
template <class VT>
struct XLoad {
  const VT* __restrict__ x;
  __device__ __forceinline__ VT operator()(uint32_t j) const {
    if constexpr (FX_NT_GATHER != 0) return __builtin_nontemporal_load(&x[j]);
    else return x[j];
  }
};
struct BNone {
  __device__ __forceinline__ double operator()(uint32_t) const { return 0.0; }
};
template <class VT>
struct BLoad {
  const VT* __restrict__ b;
  __device__ __forceinline__ double operator()(uint32_t i) const { return (double)b[i]; }
};

// y = v (store) -- MAT::Apply
template <class VT>
struct EpiStore {
  VT* __restrict__ y;
  __device__ bool skip() const { return false; }
  __device__ __forceinline__ void row(uint32_t r, double v, double (&)[1]) const {
#if FX_NT_Y
    __builtin_nontemporal_store((VT)v, &y[r]);
#else
    y[r] = (VT)v;
#endif
  }
  __device__ void fin(double (&)[1]) const {}
};

// With events e0 / e1 the kernel is launched with them attached to its
// dispatch (hipExtLaunchKernelGGL), so that they time the kernel alone.
// The kernel is instantiated for each column format; A->pid / A->cd pick one.
template <class MT, class VT, int MODE, int NRED, unsigned MAXMASK, class XF, class BF, class Epi,
          int CF>
inline void launch_spmv_k(fx_csr A, hipStream_t st, XF xf, BF bf, Epi epi, double* partials,
                          hipEvent_t e0, hipEvent_t e1, unsigned g) {
  if (e0)
    hipExtLaunchKernelGGL((k_spmv<MT, VT, MODE, NRED, MAXMASK, XF, BF, Epi, CF>), dim3(g),
                          dim3(kThreads), 0, st, e0, e1, 0, spmv_in<MT, VT>(A), xf, bf, epi,
                          partials);
  else
    hipLaunchKernelGGL((k_spmv<MT, VT, MODE, NRED, MAXMASK, XF, BF, Epi, CF>), dim3(g),
                       dim3(kThreads), 0, st, spmv_in<MT, VT>(A), xf, bf, epi, partials);
}

template <class MT, class VT, int MODE, int NRED, unsigned MAXMASK, class XF, class BF, class Epi>
inline int launch_spmv(fx_csr A, hipStream_t st, XF xf, BF bf, Epi epi, double* partials,
                       hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr, bool finisher = true) {
  const unsigned g = spmv_grid_of<MT>(A, A->ctx->sms);
  if (A->pid && sizeof(MT) == 4 && A->pat_rows == kPat2Rows)  // f32 copies only
    launch_spmv_k<MT, VT, MODE, NRED, MAXMASK, XF, BF, Epi,
                  sizeof(MT) == 4 ? CF_PAT2 : CF_PAT>(A, st, xf, bf, epi, partials, e0, e1, g);
  else if (A->pid)
    launch_spmv_k<MT, VT, MODE, NRED, MAXMASK, XF, BF, Epi, CF_PAT>(A, st, xf, bf, epi, partials,
                                                                    e0, e1, g);
  else if (A->sell && sizeof(MT) == 8)
    launch_spmv_k<MT, VT, MODE, NRED, MAXMASK, XF, BF, Epi, CF_SELL>(A, st, xf, bf, epi,
                                                                     partials, e0, e1, g);
  else if (A->cd)
    launch_spmv_k<MT, VT, MODE, NRED, MAXMASK, XF, BF, Epi, CF_D16>(A, st, xf, bf, epi, partials,
                                                                    e0, e1, g);
  else
    launch_spmv_k<MT, VT, MODE, NRED, MAXMASK, XF, BF, Epi, CF_U32>(A, st, xf, bf, epi, partials,
                                                                    e0, e1, g);
  FX_HIP(hipGetLastError());
  if constexpr (NRED > 0) {
    if (finisher) return launch_finish<NRED, MAXMASK>(st, partials, g, epi);
  }
  return 0;
}

}  // namespace fx
