// fx_spmv.hpp -- the row-blocked CSR SpMV kernel (MAT::Apply /
// MAT::Residual, src/MAT.cxx:373-476) with pluggable per-row epilogues and
// an optional fused grid reduction.
//
// One workgroup (256 threads) per row block:
//   * the block's row pointers are staged in LDS;
//   * its column indices and values are read once, fully coalesced (non-
//     temporal: they are streamed, x is what should stay in cache), into LDS;
//   * rows are then reduced with x gathered straight from memory.  While the
//     block's rows average at most 32 entries each row is summed by one
//     thread, left to right, in the reference's order (w = v0*x0, then
//     w = fma(vk, xk, w)), so the result equals the reference's FMA-contracted
//     build bit for bit; for banded / stencil matrices a wave's k-th gathers
//     then hit consecutive x entries.  Blocks of longer rows give each row a
//     lane group of 2..64 lanes (butterfly sum);
//   * a row longer than the LDS stage gets a whole workgroup.
// The epilogue receives (row, value) and may store, update other vectors and
// add to per-thread reduction slots; the slots are then reduced over the grid
// and a one-workgroup finisher kernel runs the epilogue's fin().
#pragma once

#include "fx_internal.hpp"

namespace fx {

enum { ROW_AX = 0, ROW_B_MINUS_AX = 1 };

template <class MT, class VT>
struct SpmvIn {
  const uint32_t* __restrict__ rp;
  const uint32_t* __restrict__ ci;
  const MT* __restrict__ va;  // null: structure-only matrix (values are 1)
  const uint32_t* __restrict__ blk;
  uint32_t nblk;
};

template <class MT, class VT>
inline SpmvIn<MT, VT> spmv_in(fx_csr A);

template <>
inline SpmvIn<double, double> spmv_in<double, double>(fx_csr A) {
  return {A->rp, A->ci, A->has_values ? A->va : nullptr, A->blk, A->nblk};
}
template <>
inline SpmvIn<float, float> spmv_in<float, float>(fx_csr A) {
  return {A->rp, A->ci, A->has_values ? A->va32 : nullptr, A->blk, A->nblk};
}

// LDS footprint: row pointers + staged (column, value) pairs.
template <class MT, class VT>
struct SpmvLds {
  uint32_t rp[kBlockRows + 1];
  uint32_t ci[kBlockNnz];
  MT va[kBlockNnz];
};

// Threads per row for a block of nr rows holding nz entries: one thread per
// row (the reference's summation order) while rows average at most
// kSeqRowLen entries; longer rows get the largest power-of-two lane group
// that both fits 256/nr and leaves each lane at least kSeqRowLen/2 entries,
// capped at one wave.
constexpr uint32_t kSeqRowLen = 32;
__device__ __forceinline__ int threads_per_row(uint32_t nr, uint32_t nz) {
  if (nr == 0 || nz <= kSeqRowLen * nr) return 1;
  const uint32_t by_rows = kThreads / nr;
  const uint32_t by_len = (2 * nz) / (kSeqRowLen * nr);
  const uint32_t q = by_rows < by_len ? by_rows : by_len;
  int t = 1;
  while ((uint32_t)(t * 2) <= q && t < 64) t *= 2;
  return t;
}

// Generic kernel.  X is a functor giving the gathered vector entry x_j
// (a plain load, or a value formed on the fly); B gives b_i for
// ROW_B_MINUS_AX.  Epi must provide:
//   bool skip() const                      uniform early exit (solver stop flag)
//   void row(uint32_t r, double v, double (&acc)[NRED])
//   void fin(double (&sum)[NRED])          run by k_finish after the kernel
template <class MT, class VT, int MODE, int NRED, unsigned MAXMASK, class XF, class BF, class Epi>
__global__ __launch_bounds__(kThreads) void k_spmv(SpmvIn<MT, VT> in, XF xf0, BF bf, Epi epi0,
                                                   double* partials) {
  if (epi0.skip()) return;
  XF xf = xf0;
  Epi epi = epi0;
  maybe_init(xf);
  maybe_init(epi);
  __shared__ SpmvLds<MT, VT> lds;
  double acc[NRED > 0 ? NRED : 1];
#pragma unroll
  for (int k = 0; k < (NRED > 0 ? NRED : 1); ++k) acc[k] = 0.0;
  const unsigned bid = xcd_block();
  const int tid = threadIdx.x;
  if (bid < in.nblk) {
    const uint32_t r0 = in.blk[bid], r1 = in.blk[bid + 1];
    const uint32_t k0 = in.rp[r0], k1 = in.rp[r1];
    const uint32_t nr = r1 - r0, nz = k1 - k0;
    if (nz > kBlockNnz) {
      // a single long row: strided partial sums, then a workgroup reduction
      double s = 0.0;
      for (uint32_t k = k0 + tid; k < k1; k += kThreads) {
        const double a = in.va ? (double)in.va[k] : 1.0;
        s = fma(a, (double)xf(in.ci[k]), s);
      }
      double t[1] = {s};
      block_reduce<1>(t);
      if (tid == 0) {
        const double v = (MODE == ROW_AX) ? t[0] : (double)bf(r0) - t[0];
        epi.row(r0, v, acc);
      }
    } else {
      for (uint32_t i = tid; i <= nr; i += kThreads) lds.rp[i] = in.rp[r0 + i] - k0;
      // Stage: all of a thread's column / value loads are issued together.
      constexpr int kPer = kBlockNnz / kThreads;
      uint32_t cix[kPer];
      MT vv[kPer];
#pragma unroll
      for (int j = 0; j < kPer; ++j) {
        const uint32_t k = tid + j * kThreads;
        cix[j] = k < nz ? __builtin_nontemporal_load(&in.ci[k0 + k]) : 0u;
        vv[j] = (k < nz && in.va) ? __builtin_nontemporal_load(&in.va[k0 + k]) : (MT)1;
      }
#pragma unroll
      for (int j = 0; j < kPer; ++j) {
        const uint32_t k = tid + j * kThreads;
        if (k < nz) {
          lds.ci[k] = cix[j];
          lds.va[k] = vv[j];
        }
      }
      __syncthreads();
      const int tpr = threads_per_row(nr, nz);
      if (tpr == 1) {
        if ((uint32_t)tid < nr) {
          const uint32_t s = lds.rp[tid], e = lds.rp[tid + 1];
          double t = (MODE == ROW_AX) ? 0.0 : (double)bf(r0 + tid);
          // chunks of 8: the chunk's gathers are issued before its fma chain
          for (uint32_t base = s; base < e; base += 8) {
            double av[8], xv[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
              const uint32_t k = base + j;
              av[j] = k < e ? (double)lds.va[k] : 0.0;
              xv[j] = k < e ? (double)xf(lds.ci[k]) : 0.0;
            }
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
          epi.row(r0 + tid, t, acc);
        }
      } else {
        const int grp = tid / tpr, lane = tid % tpr, ngrp = kThreads / tpr;
        for (uint32_t rr = grp; rr < nr; rr += ngrp) {
          const uint32_t s = lds.rp[rr], e = lds.rp[rr + 1];
          double t = 0.0;
          for (uint32_t k = s + lane; k < e; k += tpr)
            t = fma((double)lds.va[k], (double)xf(lds.ci[k]), t);
          t = group_sum(t, tpr);
          if (lane == 0) {
            const double v = (MODE == ROW_AX) ? t : (double)bf(r0 + rr) - t;
            epi.row(r0 + rr, v, acc);
          }
        }
      }
    }
  }
  if constexpr (NRED > 0) block_partials<NRED, MAXMASK>(acc, partials);
}

// Functors for the common cases.
template <class VT>
struct XLoad {
  const VT* __restrict__ x;
  __device__ __forceinline__ VT operator()(uint32_t j) const { return x[j]; }
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
    __builtin_nontemporal_store((VT)v, &y[r]);
  }
  __device__ void fin(double (&)[1]) const {}
};

template <class MT, class VT, int MODE, int NRED, unsigned MAXMASK, class XF, class BF, class Epi>
inline int launch_spmv(fx_csr A, hipStream_t st, XF xf, BF bf, Epi epi, double* partials) {
  const unsigned g = spmv_grid(A->nblk, A->ctx->sms);
  hipLaunchKernelGGL((k_spmv<MT, VT, MODE, NRED, MAXMASK, XF, BF, Epi>), dim3(g), dim3(kThreads), 0,
                     st, spmv_in<MT, VT>(A), xf, bf, epi, partials);
  FX_HIP(hipGetLastError());
  if constexpr (NRED > 0) return launch_finish<NRED, MAXMASK>(st, partials, g, epi);
  return 0;
}

}  // namespace fx
