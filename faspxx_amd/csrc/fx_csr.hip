// fx_csr.hip -- CSRx matrices in HBM: creation, SpMV (MAT::Apply),
// residual (MAT::Residual), diagonal (MAT::GetDiag) and transpose
// (MAT::TransInPlace), src/MAT.cxx.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "fx_spmv.hpp"

namespace fx {

namespace {

// Row-block table (see fx_csr_s).
std::vector<uint32_t> build_row_blocks(uint32_t nrow, const uint32_t* rp, uint32_t* max_row,
                                       uint32_t rows_cap = kBlockRows,
                                       uint32_t nnz_cap = kBlockNnz) {
  std::vector<uint32_t> blk;
  blk.reserve(nrow / 128 + 2);
  blk.push_back(0);
  uint32_t r = 0, mx = 0;
  while (r < nrow) {
    const uint32_t start = r;
    const uint32_t len = rp[r + 1] - rp[r];
    if (len > mx) mx = len;
    if (len > nnz_cap) {
      ++r;
    } else {
      uint32_t nz = 0;
      while (r < nrow && r - start < rows_cap) {
        const uint32_t l = rp[r + 1] - rp[r];
        if (l > mx) mx = l;
        if (nz + l > nnz_cap) break;
        nz += l;
        ++r;
      }
    }
    blk.push_back(r);
  }
  *max_row = mx;
  return blk;
}

// ---- kernels ---------------------------------------------------------------

// diagPtr as MAT::FormDiagPtr (src/MAT.cxx:917-928): first k with col == row;
// rows without a diagonal keep 0 (the reference's freshly resized vector).
__global__ void k_form_diagptr(const uint32_t* __restrict__ rp, const uint32_t* __restrict__ ci,
                               uint32_t* __restrict__ dp, uint32_t nrow) {
  const uint32_t j = blockIdx.x * kThreads + threadIdx.x;
  if (j >= nrow) return;
  uint32_t d = 0;
  for (uint32_t k = rp[j]; k < rp[j + 1]; ++k) {
    if (ci[k] == j) {
      d = k;
      break;
    }
  }
  dp[j] = d;
}

// v_j = values[diagPtr[j]] (or 1 for a structure-only matrix), src/MAT.cxx:256-267
template <class T>
__global__ void k_get_diag(const double* __restrict__ va, const uint32_t* __restrict__ dp,
                           T* __restrict__ d, uint32_t n) {
  const uint32_t j = blockIdx.x * kThreads + threadIdx.x;
  if (j >= n) return;
  d[j] = va ? (T)va[dp[j]] : (T)1;
}

// cd[k] = ci[k] - r0 for the entries of row block b (one workgroup per
// block); *bad is set when a delta does not fit in 16 bits.
__global__ void k_col_delta(const uint32_t* __restrict__ ci, const uint32_t* __restrict__ blk,
                            uint32_t nblk,
                            int16_t* __restrict__ cd, uint32_t* bad) {
  const uint32_t b = blockIdx.x;
  if (b >= nblk) return;
  const uint32_t r0 = blk[2 * b];
  const uint32_t k0 = blk[2 * b + 1], k1 = blk[2 * b + 3];
  bool fits = true;
  for (uint32_t k = k0 + threadIdx.x; k < k1; k += kThreads) {
    const int64_t d = (int64_t)ci[k] - (int64_t)r0;
    fits = fits && d >= -32768 && d <= 32767;
    cd[k] = (int16_t)d;
  }
  if (!fits) *bad = 1u;
}

__global__ void k_f64_to_f32(const double* __restrict__ s, float* __restrict__ d, uint32_t n) {
  const uint32_t stride = gridDim.x * kThreads;
  for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) d[i] = (float)s[i];
}

// max |x_i| into *out (one reduction)
template <class T>
__global__ __launch_bounds__(kThreads) void k_norminf(const T* __restrict__ x, uint32_t n,
                                                      double* partials) {
  double acc[1] = {0.0};
  const uint32_t stride = gridDim.x * kThreads;
  for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < n; i += stride)
    acc[0] = fmax(acc[0], fabs((double)x[i]));
  block_partials<1, 1u>(acc, partials);
}

// r = b if ||x||_inf < CLOSE_ZERO, else the row chain b - A x
// (MAT::Residual, src/MAT.cxx:459-469).
template <class VT>
struct EpiResidual {
  VT* __restrict__ r;
  const VT* __restrict__ b;
  const double* __restrict__ xinf;
  __device__ bool skip() const { return false; }
  __device__ __forceinline__ void row(uint32_t i, double v, double (&)[1]) const {
    r[i] = (*xinf < kCloseZero) ? b[i] : (VT)v;
  }
  __device__ void fin(double (&)[1]) const {}
};

// ---- transpose (MAT::TransInPlace, src/MAT.cxx:479-538) -------------------
// Counting sort by column.  Placement within a transposed row is first done
// with atomics (any order), then each row's entries are ranked by their
// source position k, which restores the reference's order (source rows
// ascending, ties by position) deterministically.

__global__ void k_col_count(const uint32_t* __restrict__ ci, uint32_t nnz, uint32_t* cnt) {
  const uint32_t stride = gridDim.x * kThreads;
  for (uint32_t k = blockIdx.x * kThreads + threadIdx.x; k < nnz; k += stride)
    atomicAdd(&cnt[ci[k]], 1u);
}

// Exclusive scan of cnt[0..m) into out[0..m], single workgroup, chunked.
__global__ __launch_bounds__(1024) void k_scan(const uint32_t* __restrict__ cnt, uint32_t m,
                                               uint32_t* __restrict__ out) {
  __shared__ uint32_t s[1024];
  __shared__ uint32_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (uint32_t base = 0; base < m; base += 1024) {
    const uint32_t i = base + threadIdx.x;
    const uint32_t v = i < m ? cnt[i] : 0u;
    s[threadIdx.x] = v;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) {
      const uint32_t t = threadIdx.x >= off ? s[threadIdx.x - off] : 0u;
      __syncthreads();
      s[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < m) out[i] = carry + s[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == 1023) carry += s[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[m] = carry;
}

// Multi-workgroup exclusive scan (integers, so any order gives the same
// sums): each workgroup scans a tile of kScanTile counts into out and writes
// the tile's total; k_scan scans the totals; each tile then adds its offset.
constexpr uint32_t kScanPer = 8, kScanTile = 1024 * kScanPer;

__global__ __launch_bounds__(1024) void k_scan_tiles(const uint32_t* __restrict__ cnt, uint32_t m,
                                                     uint32_t* __restrict__ out,
                                                     uint32_t* __restrict__ sums) {
  __shared__ uint32_t ws[16];
  const uint32_t tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const uint32_t base = blockIdx.x * kScanTile + tid * kScanPer;
  uint32_t v[kScanPer], run = 0;
#pragma unroll
  for (uint32_t j = 0; j < kScanPer; ++j) {
    v[j] = base + j < m ? cnt[base + j] : 0u;
    run += v[j];
  }
  uint32_t x = run;  // inclusive scan of the threads' sums within the wave
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(x, off, 64);
    if ((int)lane >= off) x += y;
  }
  if (lane == 63) ws[wid] = x;
  __syncthreads();
  if (wid == 0) {
    uint32_t w = lane < 16 ? ws[lane] : 0u;
#pragma unroll
    for (int off = 1; off < 16; off <<= 1) {
      const uint32_t y = __shfl_up(w, off, 64);
      if ((int)lane >= off) w += y;
    }
    if (lane < 16) ws[lane] = w;
  }
  __syncthreads();
  uint32_t acc = (wid ? ws[wid - 1] : 0u) + x - run;
#pragma unroll
  for (uint32_t j = 0; j < kScanPer; ++j) {
    if (base + j < m) out[base + j] = acc;
    acc += v[j];
  }
  if (tid == 0) sums[blockIdx.x] = ws[15];
}

// out[i] += offs[tile of i] for i < m; out[m] = the total.
__global__ void k_scan_add(uint32_t* __restrict__ out, uint32_t m,
                           const uint32_t* __restrict__ offs, uint32_t ntile) {
  const uint32_t b = blockIdx.x;
  const uint32_t add = offs[b];
  for (uint32_t i = b * kScanTile + threadIdx.x; i < m && i < (b + 1) * kScanTile; i += blockDim.x)
    out[i] += add;
  if (b == 0 && threadIdx.x == 0) out[m] = offs[ntile];
}

__global__ void k_row_ids(const uint32_t* __restrict__ rp, uint32_t nrow, uint32_t* __restrict__ rid) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= nrow) return;
  for (uint32_t k = rp[i]; k < rp[i + 1]; ++k) rid[k] = i;
}

__global__ void k_scatter(const uint32_t* __restrict__ ci, uint32_t nnz, uint32_t* cursor,
                          uint32_t* __restrict__ perm) {
  const uint32_t stride = gridDim.x * kThreads;
  for (uint32_t k = blockIdx.x * kThreads + threadIdx.x; k < nnz; k += stride) {
    const uint32_t pos = atomicAdd(&cursor[ci[k]], 1u);
    perm[pos] = k;
  }
}

// One wave per transposed row: entry e's final slot is the number of entries
// of the row with a smaller source position.
__global__ __launch_bounds__(kThreads) void k_rank_rows(const uint32_t* __restrict__ rpT, uint32_t m,
                                                        const uint32_t* __restrict__ perm,
                                                        uint32_t* __restrict__ sorted) {
  const uint32_t row = blockIdx.x * kWaves + (threadIdx.x >> 6);
  const uint32_t lane = threadIdx.x & 63;
  if (row >= m) return;
  const uint32_t s = rpT[row], e = rpT[row + 1], L = e - s;
  for (uint32_t a = lane; a < L; a += 64) {
    const uint32_t key = perm[s + a];
    uint32_t rank = 0;
    for (uint32_t f = 0; f < L; ++f) rank += perm[s + f] < key ? 1u : 0u;
    sorted[s + rank] = key;
  }
}

__global__ void k_gather_t(const uint32_t* __restrict__ sorted, uint32_t nnz,
                           const uint32_t* __restrict__ rid, const double* __restrict__ va,
                           uint32_t* __restrict__ ciT, double* __restrict__ vaT) {
  const uint32_t stride = gridDim.x * kThreads;
  for (uint32_t p = blockIdx.x * kThreads + threadIdx.x; p < nnz; p += stride) {
    const uint32_t k = sorted[p];
    ciT[p] = rid[k];
    if (va) vaT[p] = va[k];
  }
}

__global__ void k_gather_values(const uint32_t* __restrict__ src, const double* __restrict__ va,
                                double* __restrict__ vaT, uint32_t nnz) {
  const uint32_t stride = gridDim.x * kThreads;
  for (uint32_t p = blockIdx.x * kThreads + threadIdx.x; p < nnz; p += stride) vaT[p] = va[src[p]];
}

__global__ void k_pattern_grad(const uint32_t* __restrict__ rp, const uint32_t* __restrict__ ci,
                               const double* __restrict__ lam, const double* __restrict__ x,
                               double* __restrict__ g, uint32_t nrow) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= nrow) return;
  const double li = lam[i];
  for (uint32_t k = rp[i]; k < rp[i + 1]; ++k) g[k] = -(li * x[ci[k]]);
}

int csr_alloc(fx_csr A) {
  const size_t n1 = (size_t)A->nrow + 1, nz = A->nnz ? A->nnz : 1, nd = A->nrow ? A->nrow : 1;
  FX_HIP(hipMalloc(&A->rp, sizeof(uint32_t) * n1));
  FX_HIP(hipMalloc(&A->ci, sizeof(uint32_t) * nz));
  FX_HIP(hipMalloc(&A->dp, sizeof(uint32_t) * nd));
  FX_HIP(hipMalloc(&A->va, sizeof(double) * nz));
  return 0;
}

// Row patterns of a matrix given on the host: pid[i] and a dictionary of
// plen-strided offset lists (col - row).  False when some row is longer than
// kPatMaxLen, there are more than kPatMax distinct patterns, or the
// dictionary exceeds kPatDict offsets.  Each
// pattern is found by a hash of (length, offsets) and then compared entry by
// entry, so a hash collision cannot merge two patterns.
bool build_patterns(uint32_t nrow, const uint32_t* rp, const uint32_t* ci, uint32_t max_row,
                    std::vector<uint8_t>& pid, std::vector<int32_t>& dict, uint32_t& plen,
                    uint32_t& npat) {
  if (max_row > kPatMaxLen || nrow == 0) return false;
  plen = max_row ? max_row : 1;
  std::vector<std::vector<int32_t>> pats;
  std::unordered_map<uint64_t, std::vector<uint32_t>> byhash;
  pid.assign(nrow, 0);
  std::vector<int32_t> off(kPatMaxLen);
  for (uint32_t i = 0; i < nrow; ++i) {
    const uint32_t len = rp[i + 1] - rp[i];
    uint64_t h = 1469598103934665603ull ^ len;  // FNV-1a over the offsets
    for (uint32_t j = 0; j < len; ++j) {
      off[j] = (int32_t)((int64_t)ci[rp[i] + j] - (int64_t)i);
      h = (h ^ (uint32_t)off[j]) * 1099511628211ull;
    }
    auto& ids = byhash[h];
    int found = -1;
    for (uint32_t id : ids) {
      const auto& q = pats[id];
      if (q.size() == len && std::equal(q.begin(), q.end(), off.begin())) {
        found = (int)id;
        break;
      }
    }
    if (found < 0) {
      if (pats.size() >= kPatMax) return false;
      found = (int)pats.size();
      pats.emplace_back(off.begin(), off.begin() + len);
      ids.push_back((uint32_t)found);
    }
    pid[i] = (uint8_t)found;
  }
  npat = (uint32_t)pats.size();
  if ((size_t)npat * plen > kPatDict) return false;
  dict.assign((size_t)npat * plen, 0);
  for (uint32_t p = 0; p < npat; ++p)
    std::copy(pats[p].begin(), pats[p].end(), dict.begin() + (size_t)p * plen);
  return true;
}

bool patterns_disabled() {
  const char* e = std::getenv("FASPXX_FX_NOPAT");
  return e && e[0] == '1';
}

// Matrices whose columns take neither row patterns nor 16-bit deltas get
// rows sorted by length (default; FASPXX_FX_NOSORT=1: not), else sliced
// rows for the f64 kernels (FASPXX_FX_NOSELL=1: not).  FASPXX_FX_SELL=1
// gives every matrix made from host arrays sliced rows, and no sorting.
// sell_mode: 0 off, 1 where the columns are u32 and rows are not sorted,
// 2 for every matrix.
int sell_mode() {
  const char* e = std::getenv("FASPXX_FX_NOSELL");
  if (e && e[0] == '1') return 0;
  e = std::getenv("FASPXX_FX_SELL");
  return (e && e[0] == '1') ? 2 : 1;
}

bool sort_enabled() {
  const char* e = std::getenv("FASPXX_FX_NOSORT");
  return !(e && e[0] == '1');
}

}  // namespace

// The values of the slices from the CSR values: dst[p] = src[smap[p]]
// (0 in padding slots).
template <class T>
__global__ void k_sell_fill(const uint32_t* __restrict__ smap, const T* __restrict__ src,
                            T* __restrict__ dst, uint32_t slots) {
  const uint32_t stride = gridDim.x * kThreads;
  for (uint32_t p = blockIdx.x * kThreads + threadIdx.x; p < slots; p += stride) {
    const uint32_t m = smap[p];
    dst[p] = m == kSellNoRow ? (T)0 : src[m];
  }
}

// The sorted copy's values: dst[orp[s] + j] = src[rp[omap[s]] + j], one
// thread per sorted row.
template <class T>
__global__ void k_sorted_fill(const uint32_t* __restrict__ rp, const uint32_t* __restrict__ orp,
                              const uint32_t* __restrict__ omap, const T* __restrict__ src,
                              T* __restrict__ dst, uint32_t n) {
  const uint32_t stride = gridDim.x * kThreads;
  for (uint32_t s = blockIdx.x * kThreads + threadIdx.x; s < n; s += stride) {
    const uint32_t k0 = rp[omap[s]], d0 = orp[s], len = orp[s + 1] - d0;
    for (uint32_t j = 0; j < len; ++j) dst[d0 + j] = src[k0 + j];
  }
}

// Refresh the sorted copy's f64 values from va (fx_csr_make_f32 converts
// them for the f32 copy).
int sorted_fill(fx_csr A) {
  if (!A->sorted || !A->has_values || A->nnz == 0) return 0;
  if (!A->oval) FX_HIP(hipMalloc(&A->oval, sizeof(double) * A->nnz));
  hipLaunchKernelGGL(k_sorted_fill<double>, dim3(stream_blocks(A->nrow)), dim3(kThreads), 0,
                     A->ctx->stream, A->rp, A->orp, A->omap, A->va, A->oval, A->nrow);
  FX_HIP(hipGetLastError());
  return 0;
}

// Values of a sorted matrix's copy from values src in the matrix's order
// (same precision), into dst.
int sorted_fill_f32(fx_csr A, const float* src, float* dst) {
  if (!A->sorted || A->nnz == 0) return 0;
  hipLaunchKernelGGL(k_sorted_fill<float>, dim3(stream_blocks(A->nrow)), dim3(kThreads), 0,
                     A->ctx->stream, A->rp, A->orp, A->omap, src, dst, A->nrow);
  FX_HIP(hipGetLastError());
  return 0;
}

int sell_fill64(fx_csr A, const double* va, double* dst) {
  if (!A->sslots) return 0;
  hipLaunchKernelGGL(k_sell_fill<double>, dim3(stream_blocks(A->sslots)), dim3(kThreads), 0,
                     A->ctx->stream, A->smap, va, dst, A->sslots);
  FX_HIP(hipGetLastError());
  return 0;
}

namespace {

// Build the sliced form (for the f64 kernels) from the host row pointers
// and columns.  The values are filled on the device from A->va, which must
// be uploaded.
int build_sell(fx_csr A, const uint32_t* hrp, const uint32_t* hci) {
  const uint32_t n = A->nrow;
  const uint32_t nwin = (n + kSellRows - 1) / kSellRows;
  std::vector<uint32_t> swin(nwin + 1), lwin(nwin + 1), soff(1, 0), slane, lrow, scol, smap;
  std::vector<uint32_t> sh, lo;
  for (uint32_t w = 0; w < nwin; ++w) {
    const uint32_t r0 = w * kSellRows, r1 = std::min(n, r0 + kSellRows);
    swin[w] = (uint32_t)(soff.size() - 1);
    lwin[w] = (uint32_t)lrow.size();
    sh.clear();
    lo.clear();
    for (uint32_t r = r0; r < r1; ++r) (hrp[r + 1] - hrp[r] <= kSellMax ? sh : lo).push_back(r);
    auto longer = [&](uint32_t a, uint32_t b) { return hrp[a + 1] - hrp[a] > hrp[b + 1] - hrp[b]; };
    std::stable_sort(sh.begin(), sh.end(), longer);
    std::stable_sort(lo.begin(), lo.end(), longer);
    lrow.insert(lrow.end(), lo.begin(), lo.end());
    for (size_t i = 0; i < sh.size(); i += 64) {
      const uint32_t width = hrp[sh[i] + 1] - hrp[sh[i]];  // the slice's longest row
      const uint32_t base = soff.back();
      soff.push_back(base + width * 64);
      scol.resize((size_t)base + width * 64, 0u);
      smap.resize((size_t)base + width * 64, kSellNoRow);
      for (uint32_t lane = 0; lane < 64; ++lane) {
        if (i + lane >= sh.size()) {
          slane.push_back(kSellNoRow);
          continue;
        }
        const uint32_t r = sh[i + lane], len = hrp[r + 1] - hrp[r];
        slane.push_back(((r - r0) << 16) | len);
        for (uint32_t j = 0; j < len; ++j) {
          const size_t p = (size_t)base + (size_t)j * 64 + lane;
          scol[p] = hci[hrp[r] + j];
          smap[p] = hrp[r] + j;
        }
      }
    }
  }
  swin[nwin] = (uint32_t)(soff.size() - 1);
  lwin[nwin] = (uint32_t)lrow.size();
  A->nslice = (uint32_t)(soff.size() - 1);
  A->sslots = soff.back();
  hipStream_t st = A->ctx->stream;
  auto up = [&](uint32_t** d, const std::vector<uint32_t>& h) -> int {
    FX_HIP(hipMalloc(d, sizeof(uint32_t) * (h.empty() ? 1 : h.size())));
    if (!h.empty())
      FX_HIP(hipMemcpyAsync(*d, h.data(), sizeof(uint32_t) * h.size(), hipMemcpyHostToDevice, st));
    return 0;
  };
  FX_TRY(up(&A->swin, swin));
  FX_TRY(up(&A->lwin, lwin));
  FX_TRY(up(&A->soff, soff));
  FX_TRY(up(&A->slane, slane));
  FX_TRY(up(&A->lrow, lrow));
  FX_TRY(up(&A->scol, scol));
  FX_TRY(up(&A->smap, smap));
  FX_HIP(hipMalloc(&A->sval, sizeof(double) * (A->sslots ? A->sslots : 1)));
  if (A->has_values) FX_TRY(sell_fill64(A, A->va, A->sval));
  // the windows' block table (first row, first entry), as blk
  std::vector<uint32_t> blk(2 * ((size_t)nwin + 1));
  for (uint32_t w = 0; w <= nwin; ++w) {
    const uint32_t r = std::min(n, w * kSellRows);
    blk[2 * w] = r;
    blk[2 * w + 1] = hrp[r];
  }
  FX_TRY(up(&A->sblk, blk));
  A->nwin = nwin;
  A->sell = true;
  FX_HIP(hipStreamSynchronize(st));  // the host vectors go out of scope
  return 0;
}

// Build the sorted copy from the host row pointers and columns; its values
// are filled on the device from A->va.  The rows are cut into kXcds
// segments of about equal entry counts, and each segment's rows are sorted
// by decreasing length (ties in row order).  Each segment's row blocks are
// padded with empty blocks to the same count, so that xcd_block() runs
// segment s on XCD s.  An XCD then reads and writes the vectors only in its
// own rows' range: a global sort would scatter every XCD's writes over the
// whole result.
int build_sorted(fx_csr A, const uint32_t* hrp, const uint32_t* hci) {
  const uint32_t n = A->nrow;
  std::vector<uint32_t> seg(kXcds + 1, n);
  seg[0] = 0;
  for (uint32_t s = 1, r = 0; s < (uint32_t)kXcds; ++s) {
    const uint64_t target = (uint64_t)A->nnz * s / kXcds;
    while (r < n && hrp[r] < target) ++r;
    seg[s] = r;
  }
  std::vector<uint32_t> omap(n);
  for (uint32_t r = 0; r < n; ++r) omap[r] = r;
  for (int s = 0; s < kXcds; ++s)
    std::stable_sort(omap.begin() + seg[s], omap.begin() + seg[s + 1],
                     [&](uint32_t a, uint32_t b) {
                       return hrp[a + 1] - hrp[a] > hrp[b + 1] - hrp[b];
                     });
  std::vector<uint32_t> orp(n + 1), ocol(A->nnz ? A->nnz : 1);
  orp[0] = 0;
  for (uint32_t s = 0; s < n; ++s) {
    const uint32_t r = omap[s], len = hrp[r + 1] - hrp[r];
    std::copy(hci + hrp[r], hci + hrp[r + 1], ocol.begin() + orp[s]);
    orp[s + 1] = orp[s] + len;
  }
  // per segment: blocks of the rows of more than kSeqRowLen entries, then of
  // the shorter ones (never mixed, so that the short rows are summed one per
  // thread, left to right)
  std::vector<std::vector<uint32_t>> starts(kXcds);
  size_t per = 0;
  for (int s = 0; s < kXcds; ++s) {
    const uint32_t s0 = seg[s], s1 = seg[s + 1];
    uint32_t split = s0;
    while (split < s1 && orp[split + 1] - orp[split] > kSeqRowLen) ++split;
    uint32_t mx = 0;
    for (const auto& range : {std::make_pair(s0, split), std::make_pair(split, s1)}) {
      const std::vector<uint32_t> rb =
          build_row_blocks(range.second - range.first, orp.data() + range.first, &mx);
      for (size_t i = 0; i + 1 < rb.size(); ++i) starts[s].push_back(range.first + rb[i]);
    }
    per = std::max(per, starts[s].size());
  }
  if (per == 0) per = 1;
  std::vector<uint32_t> rows;
  rows.reserve(per * kXcds + 1);
  for (int s = 0; s < kXcds; ++s) {
    rows.insert(rows.end(), starts[s].begin(), starts[s].end());
    rows.insert(rows.end(), per - starts[s].size(), seg[s + 1]);  // empty blocks
  }
  rows.push_back(n);
  std::vector<uint32_t> blk(2 * rows.size());
  for (size_t b = 0; b < rows.size(); ++b) {
    blk[2 * b] = rows[b];
    blk[2 * b + 1] = orp[rows[b]];
  }
  std::vector<uint16_t> rpb(n ? n : 1, 0);
  for (size_t b = 0; b + 1 < rows.size(); ++b)
    for (uint32_t r = rows[b]; r < rows[b + 1]; ++r) rpb[r] = (uint16_t)(orp[r] - orp[rows[b]]);
  hipStream_t st = A->ctx->stream;
  auto up = [&](auto** d, const auto& h) -> int {
    using T = std::remove_reference_t<decltype(h[0])>;
    FX_HIP(hipMalloc(d, sizeof(T) * h.size()));
    FX_HIP(hipMemcpyAsync(*d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, st));
    return 0;
  };
  FX_TRY(up(&A->omap, omap));
  FX_TRY(up(&A->orp, orp));
  FX_TRY(up(&A->ocol, ocol));
  FX_TRY(up(&A->oblk, blk));
  FX_TRY(up(&A->orpb, rpb));
  A->onblk = (uint32_t)rows.size() - 1;
  A->sorted = true;
  FX_TRY(sorted_fill(A));
  FX_HIP(hipStreamSynchronize(st));  // the host vectors go out of scope
  return 0;
}

int csr_set_blocks(fx_csr A, const uint32_t* host_rp, const uint32_t* host_ci) {
  // Row patterns for a matrix given on the host (FASPXX_FX_NOPAT=1: none;
  // FASPXX_FX_COL32=1 keeps u32 columns); they get their own row blocks.
  // Sliced rows where neither patterns nor 16-bit deltas apply, or for
  // every matrix given on the host with FASPXX_FX_SELL=1.
  // FASPXX_FX_COL32=1 keeps the plain u32-column blocks (no slices).
  const int sell = host_ci && A->nnz ? sell_mode() : 0;
  const char* e = std::getenv("FASPXX_FX_COL32");
  const bool col32_env = e && e[0] == '1';
  const bool col32 = col32_env || sell == 2;
  std::vector<uint8_t> pid;
  std::vector<int32_t> dict;
  uint32_t plen = 0, npat = 0;
  bool pat = false;
  if (host_ci && A->nnz && !col32 && !patterns_disabled()) {
    uint32_t longest = 0;
    for (uint32_t r = 0; r < A->nrow; ++r) longest = std::max(longest, host_rp[r + 1] - host_rp[r]);
    pat = build_patterns(A->nrow, host_rp, host_ci, longest, pid, dict, plen, npat);
  }
  uint32_t mx = 0;
  const std::vector<uint32_t> rows =
      pat ? build_row_blocks(A->nrow, host_rp, &mx, kPatRows, kPatNnz)
          : build_row_blocks(A->nrow, host_rp, &mx);
  A->nblk = (uint32_t)rows.size() - 1;
  A->max_row = mx;
  std::vector<uint32_t> blk(2 * rows.size());
  for (size_t b = 0; b < rows.size(); ++b) {
    blk[2 * b] = rows[b];
    blk[2 * b + 1] = host_rp[rows[b]];
  }
  FX_HIP(hipMalloc(&A->blk, sizeof(uint32_t) * blk.size()));
  hipStream_t st = A->ctx->stream;
  FX_HIP(hipMemcpyAsync(A->blk, blk.data(), sizeof(uint32_t) * blk.size(), hipMemcpyHostToDevice,
                        st));
  std::vector<uint16_t> rpb(A->nrow ? A->nrow : 1, 0);
  for (size_t b = 0; b + 1 < rows.size(); ++b)
    for (uint32_t r = rows[b]; r < rows[b + 1]; ++r)
      rpb[r] = (uint16_t)(host_rp[r] - host_rp[rows[b]]);
  FX_HIP(hipMalloc(&A->rpb, sizeof(uint16_t) * rpb.size()));
  FX_HIP(hipMemcpyAsync(A->rpb, rpb.data(), sizeof(uint16_t) * rpb.size(), hipMemcpyHostToDevice,
                        st));
  if (pat) {
    FX_HIP(hipMalloc(&A->pid, pid.size()));
    FX_HIP(hipMalloc(&A->pdict, sizeof(int32_t) * dict.size()));
    FX_HIP(hipMemcpyAsync(A->pid, pid.data(), pid.size(), hipMemcpyHostToDevice, st));
    FX_HIP(hipMemcpyAsync(A->pdict, dict.data(), sizeof(int32_t) * dict.size(),
                          hipMemcpyHostToDevice, st));
    A->plen = plen;
    A->npat = npat;
    FX_HIP(hipStreamSynchronize(st));  // the host vectors go out of scope
    return 0;
  }
  // 2-byte column deltas when every block admits them
  if (A->nnz && A->nblk && !col32) {
    uint32_t* bad = nullptr;
    FX_HIP(hipMalloc(&A->cd, sizeof(int16_t) * A->nnz));
    FX_HIP(hipMalloc(&bad, sizeof(uint32_t)));
    uint32_t hbad = 0;
    hipError_t err = hipMemsetAsync(bad, 0, sizeof(uint32_t), st);
    if (err == hipSuccess) {
      hipLaunchKernelGGL(k_col_delta, dim3(A->nblk), dim3(kThreads), 0, st, A->ci, A->blk,
                         A->nblk, A->cd, bad);
      err = hipGetLastError();
    }
    if (err == hipSuccess) err = hipMemcpyAsync(&hbad, bad, sizeof(uint32_t), hipMemcpyDeviceToHost, st);
    if (err == hipSuccess) err = hipStreamSynchronize(st);
    (void)hipFree(bad);
    if (err != hipSuccess || hbad) {
      (void)hipFree(A->cd);
      A->cd = nullptr;
      FX_HIP(err);
    }
  }
  if (host_ci && !A->cd && sell != 2 && !col32_env && sort_enabled() && A->nrow && A->nnz) {
    FX_TRY(build_sorted(A, host_rp, host_ci));
  } else if (!A->cd && (sell == 2 || (sell == 1 && !col32_env))) {
    FX_TRY(build_sell(A, host_rp, host_ci));
  }
  FX_HIP(hipStreamSynchronize(st));
  return 0;
}

void csr_free(fx_csr A) {
  if (!A) return;
  if (A->ctx && A->ctx->stream) (void)hipStreamSynchronize(A->ctx->stream);
  (void)hipFree(A->rp);
  (void)hipFree(A->ci);
  (void)hipFree(A->dp);
  (void)hipFree(A->va);
  (void)hipFree(A->va32);
  (void)hipFree(A->blk);
  (void)hipFree(A->cd);
  (void)hipFree(A->rpb);
  (void)hipFree(A->pid);
  (void)hipFree(A->pdict);
  (void)hipFree(A->src);
  void* sell_arrays[] = {A->sblk, A->swin, A->soff, A->slane, A->scol,
                         A->smap, A->sval, A->lwin, A->lrow};
  for (void* q : sell_arrays) (void)hipFree(q);
  void* sorted_arrays[] = {A->oblk, A->orpb, A->orp, A->omap, A->ocol, A->oval, A->oval32};
  for (void* q : sorted_arrays) (void)hipFree(q);
  delete A;
}

}  // namespace

int csr_storage(fx_csr A) { return csr_alloc(A); }

int csr_pattern_blocks512(fx_csr A, uint32_t** blk, uint16_t** rpb, uint32_t* nblk) {
  hipStream_t st = A->ctx->stream;
  std::vector<uint32_t> hrp((size_t)A->nrow + 1);
  FX_HIP(hipMemcpyAsync(hrp.data(), A->rp, sizeof(uint32_t) * hrp.size(), hipMemcpyDeviceToHost,
                        st));
  FX_HIP(hipStreamSynchronize(st));
  uint32_t mx = 0;
  const std::vector<uint32_t> rows = build_row_blocks(A->nrow, hrp.data(), &mx, kPat2Rows, kPat2Nnz);
  std::vector<uint32_t> hb(2 * rows.size());
  for (size_t b = 0; b < rows.size(); ++b) {
    hb[2 * b] = rows[b];
    hb[2 * b + 1] = hrp[rows[b]];
  }
  std::vector<uint16_t> hr(A->nrow ? A->nrow : 1, 0);
  for (size_t b = 0; b + 1 < rows.size(); ++b)
    for (uint32_t r = rows[b]; r < rows[b + 1]; ++r) hr[r] = (uint16_t)(hrp[r] - hrp[rows[b]]);
  FX_HIP(hipMalloc(blk, sizeof(uint32_t) * hb.size()));
  FX_HIP(hipMalloc(rpb, sizeof(uint16_t) * hr.size()));
  FX_HIP(hipMemcpyAsync(*blk, hb.data(), sizeof(uint32_t) * hb.size(), hipMemcpyHostToDevice, st));
  FX_HIP(hipMemcpyAsync(*rpb, hr.data(), sizeof(uint16_t) * hr.size(), hipMemcpyHostToDevice, st));
  FX_HIP(hipStreamSynchronize(st));
  *nblk = (uint32_t)rows.size() - 1;
  return 0;
}

void csr_release(fx_csr A) { csr_free(A); }

int scan_u32(hipStream_t st, const uint32_t* cnt, uint32_t m, uint32_t* out) {
  if (m <= kScanTile) {
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, st, cnt, m, out);
    FX_HIP(hipGetLastError());
    return 0;
  }
  // 2M counts: 3.7 ms in one workgroup before
  const uint32_t nt = (m + kScanTile - 1) / kScanTile;
  uint32_t* tmp = nullptr;  // tile totals [0, nt), their scan [nt, 2 nt + 1)
  FX_HIP(hipMallocAsync((void**)&tmp, sizeof(uint32_t) * (2 * (size_t)nt + 1), st));
  hipLaunchKernelGGL(k_scan_tiles, dim3(nt), dim3(1024), 0, st, cnt, m, out, tmp);
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, st, (const uint32_t*)tmp, nt, tmp + nt);
  hipLaunchKernelGGL(k_scan_add, dim3(nt), dim3(1024), 0, st, out, m, (const uint32_t*)(tmp + nt),
                     nt);
  const hipError_t e = hipGetLastError();
  FX_HIP(hipFreeAsync(tmp, st));
  FX_HIP(e);
  return 0;
}

int csr_finish_device(fx_csr A) {
  hipStream_t st = A->ctx->stream;
  if (A->nrow) {
    hipLaunchKernelGGL(k_form_diagptr, dim3((A->nrow + kThreads - 1) / kThreads), dim3(kThreads),
                       0, st, A->rp, A->ci, A->dp, A->nrow);
    FX_HIP(hipGetLastError());
  }
  std::vector<uint32_t> hrp((size_t)A->nrow + 1);
  FX_HIP(hipMemcpyAsync(hrp.data(), A->rp, sizeof(uint32_t) * hrp.size(), hipMemcpyDeviceToHost,
                        st));
  FX_HIP(hipStreamSynchronize(st));
  return csr_set_blocks(A, hrp.data(), nullptr);
}

int csr_spmv_dispatch(fx_ctx ctx, fx_csr A, const void* x, void* y, int dtype, hipEvent_t e0,
                      hipEvent_t e1) {
  if (A->nrow == 0) return 0;
  if (dtype == FX_F64)
    return launch_spmv<double, double, ROW_AX, 0, 0u>(A, ctx->stream, XLoad<double>{(const double*)x},
                                                      BNone{}, EpiStore<double>{(double*)y}, nullptr,
                                                      e0, e1);
  FX_CHECK(A->va32 || !A->has_values, FX_ERROR_INPUT_PAR,
           "f32 SpMV needs the f32 value copy (fx_csr_make_f32)");
  return launch_spmv<float, float, ROW_AX, 0, 0u>(A, ctx->stream, XLoad<float>{(const float*)x},
                                                  BNone{}, EpiStore<float>{(float*)y}, nullptr, e0,
                                                  e1);
}

}  // namespace fx

using namespace fx;

extern "C" {

int fx_csr_create(fx_ctx ctx, uint32_t nrow, uint32_t mcol, uint32_t nnz, const uint32_t* rowPtr,
                  const uint32_t* colInd, const double* values, const uint32_t* diagPtr,
                  fx_csr* out) {
  clear_error();
  FX_CHECK(ctx && out, FX_ERROR_INPUT_PAR, "null argument");
  fx_csr A = new fx_csr_s();
  A->ctx = ctx;
  if (nrow == 0 || mcol == 0 || nnz == 0) {  // MAT::Empty, src/MAT.cxx:26-29
    A->nrow = A->mcol = A->nnz = 0;
    *out = A;
    return 0;
  }
  auto fail = [&](int code) {
    csr_free(A);
    return code;
  };
  if (!rowPtr || !colInd) {
    set_error("null rowPtr / colInd");
    return fail(FX_ERROR_INPUT_PAR);
  }
  // Structure validation (the device kernels index with these arrays).
  if (rowPtr[0] != 0 || rowPtr[nrow] != nnz) {
    set_error("rowPtr[0] = %u, rowPtr[nrow] = %u, nnz = %u", rowPtr[0], rowPtr[nrow], nnz);
    return fail(FX_ERROR_MAT_DATA);
  }
  for (uint32_t j = 0; j < nrow; ++j) {
    if (rowPtr[j] > rowPtr[j + 1]) {
      set_error("rowPtr decreases at row %u", j);
      return fail(FX_ERROR_MAT_DATA);
    }
  }
  for (uint32_t k = 0; k < nnz; ++k) {
    if (colInd[k] >= mcol) {
      set_error("colInd[%u] = %u >= mcol = %u", k, colInd[k], mcol);
      return fail(FX_ERROR_MAT_DATA);
    }
  }
  const uint32_t ndiag = nrow < mcol ? nrow : mcol;
  if (diagPtr) {
    for (uint32_t j = 0; j < ndiag; ++j) {
      if (diagPtr[j] >= nnz) {
        set_error("diagPtr[%u] = %u >= nnz", j, diagPtr[j]);
        return fail(FX_ERROR_MAT_DATA);
      }
    }
  }
  A->nrow = nrow;
  A->mcol = mcol;
  A->nnz = nnz;
  A->has_values = values != nullptr;
  int rc = csr_alloc(A);
  if (rc) return fail(rc);
  hipStream_t st = ctx->stream;
  if (hipMemcpyAsync(A->rp, rowPtr, sizeof(uint32_t) * (nrow + 1), hipMemcpyHostToDevice, st) ||
      hipMemcpyAsync(A->ci, colInd, sizeof(uint32_t) * nnz, hipMemcpyHostToDevice, st) ||
      (values &&
       hipMemcpyAsync(A->va, values, sizeof(double) * nnz, hipMemcpyHostToDevice, st))) {
    set_error("upload failed");
    return fail(FX_ERROR_DEVICE);
  }
  if (diagPtr) {
    if (hipMemsetAsync(A->dp, 0, sizeof(uint32_t) * nrow, st) ||
        hipMemcpyAsync(A->dp, diagPtr, sizeof(uint32_t) * ndiag, hipMemcpyHostToDevice, st)) {
      set_error("upload failed");
      return fail(FX_ERROR_DEVICE);
    }
  } else {
    hipLaunchKernelGGL(k_form_diagptr, dim3((nrow + kThreads - 1) / kThreads), dim3(kThreads), 0,
                       st, A->rp, A->ci, A->dp, nrow);
  }
  rc = csr_set_blocks(A, rowPtr, colInd);
  if (rc) return fail(rc);
  *out = A;
  return 0;
}

int fx_csr_destroy(fx_csr A) {
  csr_free(A);
  return 0;
}

int fx_csr_index_bytes(fx_csr A, int* bytes) {
  clear_error();
  FX_CHECK(A && bytes, FX_ERROR_INPUT_PAR, "null argument");
  *bytes = A->pid ? 0 : A->cd ? 2 : 4;  // sliced rows keep u32 columns
  return 0;
}

int fx_csr_layout(fx_csr A, int* layout) {
  clear_error();
  FX_CHECK(A && layout, FX_ERROR_INPUT_PAR, "null argument");
  *layout = A->pid      ? FX_LAYOUT_PATTERN
            : A->cd     ? FX_LAYOUT_D16
            : A->sorted ? FX_LAYOUT_SORTED
            : A->sell   ? FX_LAYOUT_SLICED
                        : FX_LAYOUT_U32;
  return 0;
}

int fx_csr_info(fx_csr A, uint32_t* nrow, uint32_t* mcol, uint32_t* nnz) {
  clear_error();
  FX_CHECK(A, FX_ERROR_INPUT_PAR, "null matrix");
  if (nrow) *nrow = A->nrow;
  if (mcol) *mcol = A->mcol;
  if (nnz) *nnz = A->nnz;
  return 0;
}

int fx_csr_download(fx_csr A, uint32_t* rowPtr, uint32_t* colInd, double* values,
                    uint32_t* diagPtr) {
  clear_error();
  FX_CHECK(A, FX_ERROR_INPUT_PAR, "null matrix");
  if (A->nrow == 0) return 0;
  hipStream_t st = A->ctx->stream;
  if (rowPtr)
    FX_HIP(hipMemcpyAsync(rowPtr, A->rp, sizeof(uint32_t) * (A->nrow + 1), hipMemcpyDeviceToHost, st));
  if (colInd)
    FX_HIP(hipMemcpyAsync(colInd, A->ci, sizeof(uint32_t) * A->nnz, hipMemcpyDeviceToHost, st));
  if (values && A->has_values)
    FX_HIP(hipMemcpyAsync(values, A->va, sizeof(double) * A->nnz, hipMemcpyDeviceToHost, st));
  if (diagPtr) {
    const uint32_t nd = A->nrow < A->mcol ? A->nrow : A->mcol;
    FX_HIP(hipMemcpyAsync(diagPtr, A->dp, sizeof(uint32_t) * nd, hipMemcpyDeviceToHost, st));
  }
  FX_HIP(hipStreamSynchronize(st));
  return 0;
}

int fx_csr_make_f32(fx_csr A) {
  clear_error();
  FX_CHECK(A, FX_ERROR_INPUT_PAR, "null matrix");
  if (A->nnz == 0 || !A->has_values) return 0;
  if (!A->va32) FX_HIP(hipMalloc(&A->va32, sizeof(float) * A->nnz));
  hipLaunchKernelGGL(k_f64_to_f32, dim3(stream_blocks(A->nnz)), dim3(kThreads), 0, A->ctx->stream,
                     A->va, A->va32, A->nnz);
  FX_HIP(hipGetLastError());
  if (A->sorted) {  // the sorted copy's f32 values, converted from its f64 ones
    if (!A->oval32) FX_HIP(hipMalloc(&A->oval32, sizeof(float) * A->nnz));
    hipLaunchKernelGGL(k_f64_to_f32, dim3(stream_blocks(A->nnz)), dim3(kThreads), 0,
                       A->ctx->stream, A->oval, A->oval32, A->nnz);
    FX_HIP(hipGetLastError());
  }
  return 0;
}

int fx_csr_set_values_dev(fx_csr A, const double* dev_values) {
  clear_error();
  FX_CHECK(A && dev_values, FX_ERROR_INPUT_PAR, "null argument");
  if (A->nnz == 0) return 0;
  if (dev_values != A->va)
    FX_HIP(hipMemcpyAsync(A->va, dev_values, sizeof(double) * A->nnz, hipMemcpyDeviceToDevice,
                          A->ctx->stream));
  // a structure-only matrix that gains values is another operator to a
  // solver set up on it (its replay graphs read no values): a new id
  if (!A->has_values) A->gen = next_handle_gen();
  A->has_values = true;
  if (A->sell) FX_TRY(sell_fill64(A, A->va, A->sval));
  if (A->sorted) FX_TRY(sorted_fill(A));
  if (A->va32) return fx_csr_make_f32(A);
  return 0;
}

double* fx_csr_values_dev(fx_csr A) { return A ? A->va : nullptr; }

int fx_spmv(fx_ctx ctx, fx_csr A, fx_vec v, fx_vec w) {
  clear_error();
  FX_CHECK(ctx && A && v && w, FX_ERROR_INPUT_PAR, "null argument");
  FX_CHECK(v->n == A->mcol && w->n == A->nrow, FX_ERROR_NONMATCH_SIZE,
           "SpMV sizes: A %ux%u, v %u, w %u", A->nrow, A->mcol, v->n, w->n);
  FX_CHECK(v->dtype == w->dtype, FX_ERROR_INPUT_PAR, "SpMV vector dtypes differ");
  FX_CHECK(v->d != w->d || A->nrow == 0, FX_ERROR_INPUT_PAR, "SpMV in place is not supported");
  return csr_spmv_dispatch(ctx, A, v->d, w->d, v->dtype);
}

int fx_spmv_timed(fx_ctx ctx, fx_csr A, fx_vec v, fx_vec w, float* ms) {
  clear_error();
  FX_CHECK(ctx && A && v && w && ms, FX_ERROR_INPUT_PAR, "null argument");
  FX_CHECK(v->n == A->mcol && w->n == A->nrow, FX_ERROR_NONMATCH_SIZE,
           "SpMV sizes: A %ux%u, v %u, w %u", A->nrow, A->mcol, v->n, w->n);
  FX_CHECK(v->dtype == w->dtype, FX_ERROR_INPUT_PAR, "SpMV vector dtypes differ");
  FX_CHECK(v->d != w->d, FX_ERROR_INPUT_PAR, "SpMV in place is not supported");
  FX_CHECK(A->nrow > 0, FX_ERROR_INPUT_PAR, "empty matrix: nothing to time");
  for (auto& e : ctx->tev)
    if (!e) FX_HIP(hipEventCreate(&e));
  FX_TRY(csr_spmv_dispatch(ctx, A, v->d, w->d, v->dtype, ctx->tev[0], ctx->tev[1]));
  FX_HIP(hipEventSynchronize(ctx->tev[1]));
  FX_HIP(hipEventElapsedTime(ms, ctx->tev[0], ctx->tev[1]));
  return 0;
}

int fx_residual(fx_ctx ctx, fx_csr A, fx_vec b, fx_vec x, fx_vec r) {
  clear_error();
  FX_CHECK(ctx && A && b && x && r, FX_ERROR_INPUT_PAR, "null argument");
  FX_CHECK(x->n == A->mcol && b->n == A->nrow && r->n == A->nrow, FX_ERROR_NONMATCH_SIZE,
           "Residual sizes mismatch");
  FX_CHECK(b->dtype == x->dtype && x->dtype == r->dtype, FX_ERROR_INPUT_PAR, "dtypes differ");
  FX_CHECK(r->d != x->d, FX_ERROR_INPUT_PAR, "Residual with r aliasing x is not supported");
  if (A->nrow == 0) return 0;
  hipStream_t st = ctx->stream;
  const unsigned g = stream_blocks(x->n);
  double* xinf = ctx->dscal + 1;
  if (x->dtype == FX_F64) {
    hipLaunchKernelGGL(k_norminf<double>, dim3(g), dim3(kThreads), 0, st, (const double*)x->d, x->n,
                       ctx->ws.partials);
    FX_TRY((launch_finish<1, 1u>(st, ctx->ws.partials, g, FinStore<1>{xinf, 0})));
    return launch_spmv<double, double, ROW_B_MINUS_AX, 0, 0u>(
        A, st, XLoad<double>{(const double*)x->d}, BLoad<double>{(const double*)b->d},
        EpiResidual<double>{(double*)r->d, (const double*)b->d, xinf}, nullptr);
  }
  FX_CHECK(A->va32 || !A->has_values, FX_ERROR_INPUT_PAR, "f32 residual needs fx_csr_make_f32");
  hipLaunchKernelGGL(k_norminf<float>, dim3(g), dim3(kThreads), 0, st, (const float*)x->d, x->n,
                     ctx->ws.partials);
  FX_TRY((launch_finish<1, 1u>(st, ctx->ws.partials, g, FinStore<1>{xinf, 0})));
  return launch_spmv<float, float, ROW_B_MINUS_AX, 0, 0u>(
      A, st, XLoad<float>{(const float*)x->d}, BLoad<float>{(const float*)b->d},
      EpiResidual<float>{(float*)r->d, (const float*)b->d, xinf}, nullptr);
}

int fx_csr_get_diag(fx_ctx ctx, fx_csr A, fx_vec v) {
  clear_error();
  FX_CHECK(ctx && A && v, FX_ERROR_INPUT_PAR, "null argument");
  FX_CHECK(v->n == A->nrow, FX_ERROR_NONMATCH_SIZE, "GetDiag: vector size %u != nrow %u", v->n,
           A->nrow);
  FX_CHECK(A->nrow <= A->mcol, FX_ERROR_MAT_SIZE, "GetDiag needs nrow <= mcol");
  if (A->nrow == 0) return 0;
  const unsigned g = (A->nrow + kThreads - 1) / kThreads;
  const double* va = A->has_values ? A->va : nullptr;
  if (v->dtype == FX_F64)
    hipLaunchKernelGGL(k_get_diag<double>, dim3(g), dim3(kThreads), 0, ctx->stream, va, A->dp,
                       (double*)v->d, A->nrow);
  else
    hipLaunchKernelGGL(k_get_diag<float>, dim3(g), dim3(kThreads), 0, ctx->stream, va, A->dp,
                       (float*)v->d, A->nrow);
  FX_HIP(hipGetLastError());
  return 0;
}

int fx_csr_transpose(fx_ctx ctx, fx_csr A, fx_csr* At) {
  clear_error();
  FX_CHECK(ctx && A && At, FX_ERROR_INPUT_PAR, "null argument");
  fx_csr T = new fx_csr_s();
  T->ctx = ctx;
  if (A->nnz == 0) {
    *At = T;
    return 0;
  }
  T->nrow = A->mcol;
  T->mcol = A->nrow;
  T->nnz = A->nnz;
  T->has_values = A->has_values;
  int rc = csr_alloc(T);
  if (rc) {
    csr_free(T);
    return rc;
  }
  hipStream_t st = ctx->stream;
  const uint32_t m = A->mcol, nnz = A->nnz;
  uint32_t *cnt = nullptr, *perm = nullptr, *sorted = nullptr, *rid = nullptr;
  auto cleanup = [&]() {
    (void)hipStreamSynchronize(st);
    (void)hipFree(cnt);
    (void)hipFree(perm);
    (void)hipFree(sorted);
    (void)hipFree(rid);
  };
  auto keep_sorted = [&]() {  // becomes T->src
    T->src = sorted;
    sorted = nullptr;
  };
  if (hipMalloc(&cnt, sizeof(uint32_t) * (m + 1)) || hipMalloc(&perm, sizeof(uint32_t) * nnz) ||
      hipMalloc(&sorted, sizeof(uint32_t) * nnz) || hipMalloc(&rid, sizeof(uint32_t) * nnz)) {
    cleanup();
    csr_free(T);
    set_error("transpose: allocation failed");
    return FX_ERROR_ALLOC_MEM;
  }
  const unsigned gz = stream_blocks(nnz);
  (void)hipMemsetAsync(cnt, 0, sizeof(uint32_t) * (m + 1), st);
  hipLaunchKernelGGL(k_col_count, dim3(gz), dim3(kThreads), 0, st, A->ci, nnz, cnt);
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, st, cnt, m, T->rp);
  (void)hipMemcpyAsync(cnt, T->rp, sizeof(uint32_t) * m, hipMemcpyDeviceToDevice, st);
  hipLaunchKernelGGL(k_scatter, dim3(gz), dim3(kThreads), 0, st, A->ci, nnz, cnt, perm);
  hipLaunchKernelGGL(k_rank_rows, dim3((m + kWaves - 1) / kWaves), dim3(kThreads), 0, st, T->rp, m,
                     perm, sorted);
  hipLaunchKernelGGL(k_row_ids, dim3((A->nrow + kThreads - 1) / kThreads), dim3(kThreads), 0, st,
                     A->rp, A->nrow, rid);
  hipLaunchKernelGGL(k_gather_t, dim3(gz), dim3(kThreads), 0, st, sorted, nnz, rid,
                     A->has_values ? A->va : nullptr, T->ci, T->has_values ? T->va : nullptr);
  hipLaunchKernelGGL(k_form_diagptr, dim3((m + kThreads - 1) / kThreads), dim3(kThreads), 0, st,
                     T->rp, T->ci, T->dp, m);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    cleanup();
    csr_free(T);
    set_error("transpose kernels: %s", hipGetErrorString(e));
    return FX_ERROR_DEVICE;
  }
  std::vector<uint32_t> hrp(m + 1);
  if (hipMemcpyAsync(hrp.data(), T->rp, sizeof(uint32_t) * (m + 1), hipMemcpyDeviceToHost, st) ||
      hipStreamSynchronize(st)) {
    cleanup();
    csr_free(T);
    set_error("transpose: download failed");
    return FX_ERROR_DEVICE;
  }
  keep_sorted();
  cleanup();
  rc = csr_set_blocks(T, hrp.data(), nullptr);
  if (rc) {
    csr_free(T);
    return rc;
  }
  if (A->va32) {
    rc = fx_csr_make_f32(T);
    if (rc) {
      csr_free(T);
      return rc;
    }
  }
  *At = T;
  return 0;
}

// Refresh a transpose's values from its source: At.va[p] = A.va[src[p]].
int fx_csr_transpose_update(fx_ctx ctx, fx_csr A, fx_csr At) {
  clear_error();
  FX_CHECK(ctx && A && At, FX_ERROR_INPUT_PAR, "null argument");
  FX_CHECK(At->src && At->nnz == A->nnz && At->nrow == A->mcol && At->mcol == A->nrow,
           FX_ERROR_INPUT_PAR, "At is not a transpose of A");
  if (A->nnz == 0) return 0;
  FX_CHECK(A->has_values, FX_ERROR_INPUT_PAR, "A has no values");
  hipLaunchKernelGGL(k_gather_values, dim3(stream_blocks(A->nnz)), dim3(kThreads), 0, ctx->stream,
                     At->src, A->va, At->va, A->nnz);
  FX_HIP(hipGetLastError());
  // the copies of the values (sliced, sorted, f32) follow At->va
  return fx_csr_set_values_dev(At, At->va);
}

// g_k = -lam[row(k)] * x[col(k)] for every stored entry k: the gradient of a
// loss through x = A^{-1} b with respect to A's stored values, given the
// adjoint solution lam of A^T lam = dL/dx.  Device pointers, nnz doubles out.
int fx_csr_pattern_grad(fx_ctx ctx, fx_csr A, fx_vec lam, fx_vec x, double* dev_out) {
  clear_error();
  FX_CHECK(ctx && A && lam && x && dev_out, FX_ERROR_INPUT_PAR, "null argument");
  FX_CHECK(lam->n == A->nrow && x->n == A->mcol, FX_ERROR_NONMATCH_SIZE, "size mismatch");
  FX_CHECK(lam->dtype == FX_F64 && x->dtype == FX_F64, FX_ERROR_INPUT_PAR, "f64 vectors required");
  if (A->nnz == 0) return 0;
  hipLaunchKernelGGL(k_pattern_grad, dim3((A->nrow + kThreads - 1) / kThreads), dim3(kThreads), 0,
                     ctx->stream, A->rp, A->ci, (const double*)lam->d, (const double*)x->d,
                     dev_out, A->nrow);
  FX_HIP(hipGetLastError());
  return 0;
}

}  // extern "C"
