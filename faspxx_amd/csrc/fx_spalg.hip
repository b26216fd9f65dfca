// fx_spalg.hip -- setup-time sparse algebra on CSR matrices in HBM, each
// result a new matrix built on the device entry for entry as the reference
// builds it:
//   * triangular parts, MAT::GetLowerTri / GetUpperTri (src/MAT.cxx:290-345):
//     row i keeps its entries up to and including the diagonal entry
//     (lower) or from it to the row's end (upper), by position;
//   * a A + b B, MAT::Add (src/MAT.cxx:652-719): per row the union of the
//     two patterns with columns ascending; an entry in both is
//     fma(b, B_ij, a * A_ij) (the reference adds b * B_ij to the stored
//     a * A_ij, contracted), an entry in one of them a * A_ij or b * B_ij;
//   * L R, MAT::Mult (src/MAT.cxx:722-800): row i holds the distinct columns
//     of the rows of R that row i of L selects, in order of first
//     appearance (L's entries in order, then R's row in order), not sorted;
//     each value is the fma chain over those products in the same order,
//     starting from 0.
// diagPtr is formed as FormDiagPtr does (first match; 0 for a row without a
// diagonal entry).  Structure-only operands (no values) give a structure-
// only result for the triangular parts and the product; a + b needs values.
//
// These run once per setup, so each kernel gives a row to a thread.  The
// product accumulates each row in an open-addressing table (twice the
// row's candidate count, so every probe sequence ends) kept in HBM scratch.
#include <vector>

#include "fx_internal.hpp"

namespace fx {
namespace {

constexpr uint32_t kEmpty = 0xffffffffu;

inline unsigned row_blocks(uint32_t nrow) { return (nrow + kThreads - 1) / kThreads; }

// ---- triangular parts -------------------------------------------------------
__global__ void k_tri_count(const uint32_t* __restrict__ rp, const uint32_t* __restrict__ ci,
                            const uint32_t* __restrict__ dp, uint32_t nrow, int upper,
                            uint32_t* __restrict__ cnt, uint32_t* __restrict__ bad) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= nrow) return;
  const uint32_t s = rp[i], e = rp[i + 1], d = dp[i];
  if (d < s || d >= e || ci[d] != i) {  // no diagonal entry in row i
    bad[0] = 1u;
    cnt[i] = 0;
    return;
  }
  cnt[i] = upper ? e - d : d - s + 1;
}

__global__ void k_tri_copy(const uint32_t* __restrict__ rp, const uint32_t* __restrict__ ci,
                           const double* __restrict__ va, const uint32_t* __restrict__ dp,
                           uint32_t nrow, int upper, const uint32_t* __restrict__ orp,
                           uint32_t* __restrict__ oci, double* __restrict__ ova) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= nrow) return;
  const uint32_t s = upper ? dp[i] : rp[i], e = upper ? rp[i + 1] : dp[i] + 1;
  const uint32_t o = orp[i];
  for (uint32_t k = s; k < e; ++k) {
    oci[o + (k - s)] = ci[k];
    if (va) ova[o + (k - s)] = va[k];
  }
}

// ---- a A + b B ------------------------------------------------------------
__device__ __forceinline__ uint32_t find_col(const uint32_t* __restrict__ ci, uint32_t s,
                                             uint32_t e, uint32_t col) {
  for (uint32_t k = s; k < e; ++k)
    if (ci[k] == col) return k;
  return kEmpty;
}

__global__ void k_add_count(const uint32_t* __restrict__ arp, const uint32_t* __restrict__ aci,
                            const uint32_t* __restrict__ brp, const uint32_t* __restrict__ bci,
                            uint32_t nrow, uint32_t* __restrict__ cnt) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= nrow) return;
  const uint32_t as = arp[i], ae = arp[i + 1];
  uint32_t c = ae - as;
  for (uint32_t k = brp[i]; k < brp[i + 1]; ++k)
    if (find_col(aci, as, ae, bci[k]) == kEmpty) ++c;
  cnt[i] = c;
}

__global__ void k_add_fill(double a, const uint32_t* __restrict__ arp,
                           const uint32_t* __restrict__ aci, const double* __restrict__ ava,
                           double b, const uint32_t* __restrict__ brp,
                           const uint32_t* __restrict__ bci, const double* __restrict__ bva,
                           uint32_t nrow, const uint32_t* __restrict__ orp,
                           uint32_t* __restrict__ oci, double* __restrict__ ova) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= nrow) return;
  const uint32_t as = arp[i], ae = arp[i + 1], o = orp[i];
  uint32_t end = o;
  for (uint32_t k = as; k < ae; ++k, ++end) {  // A's row scaled, in A's order
    oci[end] = aci[k];
    ova[end] = a * ava[k];
  }
  for (uint32_t k = brp[i]; k < brp[i + 1]; ++k) {  // B's entries merged or appended
    const uint32_t col = bci[k];
    const uint32_t ka = find_col(aci, as, ae, col);
    if (ka != kEmpty) {
      const uint32_t p = o + (ka - as);
      ova[p] = fma(b, bva[k], ova[p]);
    } else {
      oci[end] = col;
      ova[end] = b * bva[k];
      ++end;
    }
  }
  // columns ascending (SortCSRRow; the columns of a row are distinct)
  for (uint32_t p = o + 1; p < end; ++p) {
    const uint32_t c = oci[p];
    const double v = ova[p];
    uint32_t q = p;
    while (q > o && oci[q - 1] > c) {
      oci[q] = oci[q - 1];
      ova[q] = ova[q - 1];
      --q;
    }
    oci[q] = c;
    ova[q] = v;
  }
}

// s X with X's structure (MAT::Scale, src/MAT.cxx:351-356: v *= s)
__global__ void k_scaled_copy(const uint32_t* __restrict__ ci, const double* __restrict__ va,
                              double s, uint32_t nnz, uint32_t* __restrict__ oci,
                              double* __restrict__ ova) {
  const uint32_t stride = gridDim.x * kThreads;
  for (uint32_t k = blockIdx.x * kThreads + threadIdx.x; k < nnz; k += stride) {
    oci[k] = ci[k];
    ova[k] = va[k] * s;
  }
}

// MAT::Shift (src/MAT.cxx:359-364): values[diagPtr[j]] += a for every row
// j, in the reference's loop over diagPtr.  A row without a diagonal entry
// keeps diagPtr 0 (FormDiagPtr), so it adds a to values[0] once more.  Adds
// that land on one entry each add the same a, so the entry goes through the
// same sequence v, v + a, (v + a) + a, ... whichever row adds first: an
// atomic compare-and-swap loop per row gives the reference's bits exactly.
__global__ void k_shift_diag(const uint32_t* __restrict__ dp, double* __restrict__ va,
                             uint32_t nrow, uint32_t nnz, double a) {
  const uint32_t j = blockIdx.x * kThreads + threadIdx.x;
  if (j >= nrow) return;
  const uint32_t d = dp[j];
  if (d >= nnz) return;
  unsigned long long* p = reinterpret_cast<unsigned long long*>(va + d);
  unsigned long long old = *p, seen;
  do {
    seen = old;
    const double nv = __longlong_as_double((long long)seen) + a;
    old = atomicCAS(p, seen, (unsigned long long)__double_as_longlong(nv));
  } while (old != seen);
}

// ---- L R ------------------------------------------------------------------
// Candidate products of row i: the summed lengths of the rows of R that row
// i of L selects.
__global__ void k_mult_count(const uint32_t* __restrict__ lrp, const uint32_t* __restrict__ lci,
                             const uint32_t* __restrict__ rrp, uint32_t nrow,
                             uint32_t* __restrict__ cand) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= nrow) return;
  uint32_t c = 0;
  for (uint32_t k = lrp[i]; k < lrp[i + 1]; ++k) {
    const uint32_t r = lci[k];
    c += rrp[r + 1] - rrp[r];
  }
  cand[i] = c;
}

__device__ __forceinline__ uint32_t col_hash(uint32_t c, uint32_t cap) {
  return (uint32_t)(((uint64_t)(c * 2654435761u) * cap) >> 32);
}

// Row i: its candidates (L's entries in order, then each R row in order)
// through an open-addressing table of 2 * cand_i slots at keys[2 * off_i];
// a column's first candidate appends it to the row's list at tcol/tval +
// off_i, later ones fma into its value.
__global__ void k_mult_rows(const uint32_t* __restrict__ lrp, const uint32_t* __restrict__ lci,
                            const double* __restrict__ lva, const uint32_t* __restrict__ rrp,
                            const uint32_t* __restrict__ rci, const double* __restrict__ rva,
                            uint32_t nrow, const uint32_t* __restrict__ off,
                            uint32_t* __restrict__ keys, uint32_t* __restrict__ slot,
                            uint32_t* __restrict__ tcol, double* __restrict__ tval,
                            uint32_t* __restrict__ cnt) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= nrow) return;
  const uint32_t base = off[i], P = off[i + 1] - base;
  uint32_t c = 0;
  if (P) {
    const size_t hb = 2 * (size_t)base;
    const uint32_t cap = 2 * P;
    for (uint32_t k = lrp[i]; k < lrp[i + 1]; ++k) {
      const uint32_t r = lci[k];
      const double a = lva ? lva[k] : 1.0;
      for (uint32_t l = rrp[r]; l < rrp[r + 1]; ++l) {
        const uint32_t col = rci[l];
        const double bv = rva ? rva[l] : 1.0;
        uint32_t h = col_hash(col, cap);
        for (uint32_t probe = 0; probe < cap; ++probe) {  // at most P keys in 2P slots
          const uint32_t key = keys[hb + h];
          if (key == col) {
            const uint32_t p = base + slot[hb + h];
            if (tval) tval[p] = fma(a, bv, tval[p]);
            break;
          }
          if (key == kEmpty) {
            keys[hb + h] = col;
            slot[hb + h] = c;
            tcol[base + c] = col;
            if (tval) tval[base + c] = fma(a, bv, 0.0);
            ++c;
            break;
          }
          h = (h + 1 == cap) ? 0u : h + 1;
        }
      }
    }
  }
  cnt[i] = c;
}

// Compaction of the product rows into CSR: 32 lanes per row, so that a
// wave's loads and stores are consecutive (one thread per row took 3.1 ms
// for the 51.7M entries of the 128^3 stencil's square).
constexpr uint32_t kCopyLanes = 32;
__global__ void k_mult_copy(const uint32_t* __restrict__ off, const uint32_t* __restrict__ tcol,
                            const double* __restrict__ tval, uint32_t nrow,
                            const uint32_t* __restrict__ orp, uint32_t* __restrict__ oci,
                            double* __restrict__ ova) {
  const uint64_t g = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  const uint64_t i = g / kCopyLanes;
  const uint32_t lane = (uint32_t)(g % kCopyLanes);
  if (i >= nrow) return;
  const uint32_t s = off[i], o = orp[i], len = orp[i + 1] - o;
  for (uint32_t j = lane; j < len; j += kCopyLanes) {
    oci[o + j] = tcol[s + j];
    if (tval) ova[o + j] = tval[s + j];
  }
}

// Device scratch released on every exit path.
struct Scratch {
  hipStream_t st;
  std::vector<void*> bufs;
  explicit Scratch(hipStream_t s) : st(s) {}
  template <class T>
  T* get(size_t count) {
    void* p = nullptr;
    if (hipMalloc(&p, sizeof(T) * (count ? count : 1)) != hipSuccess) return nullptr;
    bufs.push_back(p);
    return (T*)p;
  }
  ~Scratch() {
    (void)hipStreamSynchronize(st);
    for (void* p : bufs) (void)hipFree(p);
  }
};

// A new matrix handle of the given shape whose rp is `rp` (device, nrow + 1
// entries, already scanned): reads nnz = rp[nrow], allocates the rest.
int new_from_rowptr(fx_ctx ctx, uint32_t nrow, uint32_t mcol, bool values, uint32_t* rp,
                    fx_csr* out) {
  hipStream_t st = ctx->stream;
  uint32_t nnz = 0;
  FX_HIP(hipMemcpyAsync(&nnz, rp + nrow, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  FX_HIP(hipStreamSynchronize(st));
  fx_csr C = new fx_csr_s();
  C->ctx = ctx;
  if (nnz == 0) {  // MAT::Empty, src/MAT.cxx:26-29
    *out = C;
    return 0;
  }
  C->nrow = nrow;
  C->mcol = mcol;
  C->nnz = nnz;
  C->has_values = values;
  const int rc = csr_storage(C);
  if (rc) {
    csr_release(C);
    return rc;
  }
  if (hipMemcpyAsync(C->rp, rp, sizeof(uint32_t) * ((size_t)nrow + 1), hipMemcpyDeviceToDevice,
                     st) != hipSuccess) {
    csr_release(C);
    set_error("row pointer copy failed");
    return FX_ERROR_DEVICE;
  }
  *out = C;
  return 0;
}

int finish(fx_csr C, fx_csr* out) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    csr_release(C);
    set_error("sparse algebra kernel: %s", hipGetErrorString(e));
    return FX_ERROR_DEVICE;
  }
  const int rc = csr_finish_device(C);
  if (rc) {
    csr_release(C);
    return rc;
  }
  *out = C;
  return 0;
}

int scaled_copy(fx_ctx ctx, double s, fx_csr X, fx_csr* out) {
  fx_csr C = nullptr;
  FX_TRY(new_from_rowptr(ctx, X->nrow, X->mcol, true, X->rp, &C));
  hipLaunchKernelGGL(k_scaled_copy, dim3(stream_blocks(X->nnz)), dim3(kThreads), 0, ctx->stream,
                     X->ci, X->va, s, X->nnz, C->ci, C->va);
  return finish(C, out);
}

int tri_part(fx_ctx ctx, fx_csr A, int upper, fx_csr* out) {
  clear_error();
  FX_CHECK(ctx && A && out, FX_ERROR_INPUT_PAR, "null argument");
  *out = nullptr;
  if (A->nnz == 0) {
    fx_csr C = new fx_csr_s();
    C->ctx = ctx;
    *out = C;
    return 0;
  }
  hipStream_t st = ctx->stream;
  const uint32_t n = A->nrow;
  Scratch S(st);
  uint32_t* cnt = S.get<uint32_t>(n);
  uint32_t* orp = S.get<uint32_t>((size_t)n + 1);
  uint32_t* bad = S.get<uint32_t>(1);
  FX_CHECK(cnt && orp && bad, FX_ERROR_ALLOC_MEM, "triangular part: allocation failed");
  FX_HIP(hipMemsetAsync(bad, 0, sizeof(uint32_t), st));
  hipLaunchKernelGGL(k_tri_count, dim3(row_blocks(n)), dim3(kThreads), 0, st, A->rp, A->ci, A->dp,
                     n, upper, cnt, bad);
  FX_HIP(hipGetLastError());
  uint32_t hbad = 0;
  FX_HIP(hipMemcpyAsync(&hbad, bad, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  FX_HIP(hipStreamSynchronize(st));
  FX_CHECK(!hbad, FX_ERROR_MAT_DATA, "triangular part needs a diagonal entry in every row");
  FX_TRY(scan_u32(st, cnt, n, orp));
  fx_csr C = nullptr;
  FX_TRY(new_from_rowptr(ctx, n, A->mcol, A->has_values, orp, &C));
  hipLaunchKernelGGL(k_tri_copy, dim3(row_blocks(n)), dim3(kThreads), 0, st, A->rp, A->ci,
                     A->has_values ? A->va : nullptr, A->dp, n, upper, C->rp, C->ci,
                     C->has_values ? C->va : nullptr);
  return finish(C, out);
}

}  // namespace
}  // namespace fx

using namespace fx;

extern "C" {

int fx_csr_lower_tri(fx_ctx ctx, fx_csr A, fx_csr* L) { return tri_part(ctx, A, 0, L); }

int fx_csr_upper_tri(fx_ctx ctx, fx_csr A, fx_csr* U) { return tri_part(ctx, A, 1, U); }

int fx_csr_add(fx_ctx ctx, double a, fx_csr A, double b, fx_csr B, fx_csr* C) {
  clear_error();
  FX_CHECK(ctx && A && B && C, FX_ERROR_INPUT_PAR, "null argument");
  *C = nullptr;
  FX_CHECK(A->nnz == 0 || B->nnz == 0 || (A->nrow == B->nrow && A->mcol == B->mcol),
           FX_ERROR_NONMATCH_SIZE, "a A + b B needs matrices of one shape");
  FX_CHECK((A->nnz == 0 || A->has_values) && (B->nnz == 0 || B->has_values), FX_ERROR_INPUT_PAR,
           "a A + b B needs matrices with values");
  hipStream_t st = ctx->stream;
  if (A->nnz == 0 || B->nnz == 0) {  // one operand empty: the other, scaled
    fx_csr X = A->nnz ? A : B;
    const double s = A->nnz ? a : b;
    if (X->nnz == 0) {
      fx_csr E = new fx_csr_s();
      E->ctx = ctx;
      *C = E;
      return 0;
    }
    return scaled_copy(ctx, s, X, C);
  }
  const uint32_t n = A->nrow;
  Scratch S(st);
  uint32_t* cnt = S.get<uint32_t>(n);
  uint32_t* orp = S.get<uint32_t>((size_t)n + 1);
  FX_CHECK(cnt && orp, FX_ERROR_ALLOC_MEM, "a A + b B: allocation failed");
  hipLaunchKernelGGL(k_add_count, dim3(row_blocks(n)), dim3(kThreads), 0, st, A->rp, A->ci, B->rp,
                     B->ci, n, cnt);
  FX_HIP(hipGetLastError());
  FX_TRY(scan_u32(st, cnt, n, orp));
  fx_csr R = nullptr;
  FX_TRY(new_from_rowptr(ctx, n, A->mcol, true, orp, &R));
  hipLaunchKernelGGL(k_add_fill, dim3(row_blocks(n)), dim3(kThreads), 0, st, a, A->rp, A->ci,
                     A->va, b, B->rp, B->ci, B->va, n, R->rp, R->ci, R->va);
  return finish(R, C);
}

int fx_csr_shift_diag(fx_ctx ctx, fx_csr A, double a) {
  clear_error();
  FX_CHECK(ctx && A, FX_ERROR_INPUT_PAR, "null argument");
  if (A->nnz == 0 || !A->has_values) return 0;  // structure only: src/MAT.cxx:361
  hipLaunchKernelGGL(k_shift_diag, dim3(row_blocks(A->nrow)), dim3(kThreads), 0, ctx->stream,
                     A->dp, A->va, A->nrow, A->nnz, a);
  FX_HIP(hipGetLastError());
  return fx_csr_set_values_dev(A, A->va);
}

int fx_csr_mult(fx_ctx ctx, fx_csr L, fx_csr R, fx_csr* C) {
  clear_error();
  FX_CHECK(ctx && L && R && C, FX_ERROR_INPUT_PAR, "null argument");
  *C = nullptr;
  if (L->nnz == 0 || R->nnz == 0) {
    fx_csr E = new fx_csr_s();
    E->ctx = ctx;
    *C = E;
    return 0;
  }
  FX_CHECK(L->mcol == R->nrow, FX_ERROR_NONMATCH_SIZE, "L R needs L's columns = R's rows");
  hipStream_t st = ctx->stream;
  const uint32_t n = L->nrow;
  const bool values = L->has_values && R->has_values;
  Scratch S(st);
  uint32_t* cand = S.get<uint32_t>(n);
  uint32_t* off = S.get<uint32_t>((size_t)n + 1);
  uint32_t* cnt = S.get<uint32_t>(n);
  uint32_t* orp = S.get<uint32_t>((size_t)n + 1);
  FX_CHECK(cand && off && cnt && orp, FX_ERROR_ALLOC_MEM, "L R: allocation failed");
  hipLaunchKernelGGL(k_mult_count, dim3(row_blocks(n)), dim3(kThreads), 0, st, L->rp, L->ci, R->rp,
                     n, cand);
  FX_HIP(hipGetLastError());
  // the candidate total must fit the u32 offsets
  std::vector<uint32_t> hc(n);
  FX_HIP(hipMemcpyAsync(hc.data(), cand, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, st));
  FX_HIP(hipStreamSynchronize(st));
  uint64_t total = 0;
  for (uint32_t v : hc) total += v;
  FX_CHECK(total < 0x80000000ull, FX_ERROR_ALLOC_MEM,
           "L R: %llu candidate products exceed the 2^31 scratch limit",
           (unsigned long long)total);
  FX_TRY(scan_u32(st, cand, n, off));
  uint32_t* keys = S.get<uint32_t>(2 * total);
  uint32_t* slot = S.get<uint32_t>(2 * total);
  uint32_t* tcol = S.get<uint32_t>(total);
  double* tval = values ? S.get<double>(total) : nullptr;
  FX_CHECK(keys && slot && tcol && (tval || !values), FX_ERROR_ALLOC_MEM,
           "L R: scratch allocation failed");
  FX_HIP(hipMemsetAsync(keys, 0xff, sizeof(uint32_t) * 2 * (total ? total : 1), st));
  hipLaunchKernelGGL(k_mult_rows, dim3(row_blocks(n)), dim3(kThreads), 0, st, L->rp, L->ci,
                     values ? L->va : nullptr, R->rp, R->ci, values ? R->va : nullptr, n, off,
                     keys, slot, tcol, tval, cnt);
  FX_HIP(hipGetLastError());
  FX_TRY(scan_u32(st, cnt, n, orp));
  fx_csr P = nullptr;
  FX_TRY(new_from_rowptr(ctx, n, R->mcol, values, orp, &P));
  const uint64_t copy_threads = (uint64_t)n * kCopyLanes;
  hipLaunchKernelGGL(k_mult_copy, dim3((unsigned)((copy_threads + kThreads - 1) / kThreads)),
                     dim3(kThreads), 0, st, off, tcol, tval, n, P->rp, P->ci,
                     values ? P->va : nullptr);
  return finish(P, C);
}

}  // extern "C"
