// fx_mcgs.hip -- multicolor Gauss-Seidel (SOR) smoother.  No reference
// counterpart: FASP++ has no Gauss-Seidel, and a sequential one has no
// parallel form, so the device offers this separately validated smoother
// instead of a drop-in.
//
// Coloring: rows in natural order, each taking the smallest color not used by
// an already-colored row adjacent to it in the pattern of A + A^T (5- and
// 7-point stencils come out red-black).  Within one color no row reads
// another, so a color is one parallel pass:
//     r_i = b_i - sum_k a_ik x_k   (row order, the Residual kernel's sum)
//     x_i = fma(r_i, dinv_i, x_i)  dinv_i = weight * (1 / a_ii)
// A sweep runs colors 0..C-1 (then C-1..0 for a symmetric sweep).  The
// results do not depend on the order inside a color, so they equal the
// sequential-by-color definition bit for bit wherever rows are summed by one
// thread (rows of up to 32 entries on average per block).
//
// Layout: each color's rows are stored as their own CSR matrix (rows in
// ascending original order, original column indices), with the original row
// index of each; the pass is the row-blocked SpMV kernel in b - A x mode over
// that matrix, with an epilogue that updates x in place.
#include <algorithm>
#include <vector>

#include "fx_solver.hpp"
#include "fx_spmv.hpp"

struct fx_mcgs_s {
  fx_ctx ctx = nullptr;
  fx_csr A = nullptr;
  fx_pc jac = nullptr;  // its dinv = weight / diag
  uint32_t ncolors = 0;
  std::vector<uint32_t> color;      // host, per row
  std::vector<fx_csr> sub;          // per color
  std::vector<uint32_t*> rows;      // per color, device: original row of each sub-row
};

namespace fx {
namespace {

struct XLoadInPlace {  // x is updated by the same kernel (other rows): no restrict
  const double* x;
  __device__ __forceinline__ double operator()(uint32_t j) const { return x[j]; }
};
struct BRows {
  const double* __restrict__ b;
  const uint32_t* __restrict__ rows;
  __device__ __forceinline__ double operator()(uint32_t r) const { return b[rows[r]]; }
};
struct EpiGS {
  double* x;
  const double* __restrict__ dinv;
  const uint32_t* __restrict__ rows;
  __device__ bool skip() const { return false; }
  __device__ __forceinline__ void row(uint32_t r, double v, double (&)[1]) const {
    const uint32_t i = rows[r];
    x[i] = fma(v, dinv[i], x[i]);
  }
  __device__ void fin(double (&)[1]) const {}
};

// Greedy natural-order coloring of the pattern of A + A^T.
uint32_t greedy_color(uint32_t n, const std::vector<uint32_t>& rp, const std::vector<uint32_t>& ci,
                      std::vector<uint32_t>& color) {
  const uint32_t nnz = rp[n];
  std::vector<uint32_t> tp(n + 1, 0), tr(nnz);
  for (uint32_t k = 0; k < nnz; ++k)
    if (ci[k] < n) ++tp[ci[k] + 1];
  for (uint32_t i = 0; i < n; ++i) tp[i + 1] += tp[i];
  std::vector<uint32_t> cur(tp.begin(), tp.end() - 1);
  for (uint32_t i = 0; i < n; ++i)
    for (uint32_t k = rp[i]; k < rp[i + 1]; ++k)
      if (ci[k] < n) tr[cur[ci[k]]++] = i;
  const uint32_t kNone = 0xffffffffu;
  color.assign(n, kNone);
  std::vector<uint32_t> mark;  // mark[c] == i: color c is taken by a neighbour of i
  uint32_t nc = 0;
  for (uint32_t i = 0; i < n; ++i) {
    auto take = [&](uint32_t j) {
      if (j == i || j >= n || color[j] == kNone) return;
      if (color[j] >= mark.size()) mark.resize(color[j] + 1, kNone);
      mark[color[j]] = i;
    };
    for (uint32_t k = rp[i]; k < rp[i + 1]; ++k) take(ci[k]);
    for (uint32_t k = tp[i]; k < tp[i + 1]; ++k) take(tr[k]);
    uint32_t c = 0;
    while (c < mark.size() && mark[c] == i) ++c;
    color[i] = c;
    nc = std::max(nc, c + 1);
  }
  return nc;
}

void mcgs_free(fx_mcgs s) {
  if (!s) return;
  if (s->ctx) (void)hipStreamSynchronize(s->ctx->stream);
  for (fx_csr m : s->sub) fx_csr_destroy(m);
  for (uint32_t* r : s->rows) (void)hipFree(r);
  fx_pc_destroy(s->jac);
  delete s;
}

}  // namespace
}  // namespace fx

using namespace fx;

extern "C" {

int fx_mcgs_create(fx_ctx ctx, fx_csr A, double weight, fx_mcgs* out) {
  clear_error();
  FX_CHECK(ctx && A && out, FX_ERROR_INPUT_PAR, "null argument");
  FX_CHECK(A->nrow == A->mcol && A->nrow > 0, FX_ERROR_MAT_SIZE, "needs a square, non-empty matrix");
  FX_CHECK(A->has_values, FX_ERROR_INPUT_PAR, "needs a matrix with values");
  const uint32_t n = A->nrow, nnz = A->nnz;
  std::vector<uint32_t> rp(n + 1), ci(nnz);
  std::vector<double> va(nnz);
  FX_TRY(fx_csr_download(A, rp.data(), ci.data(), va.data(), nullptr));
  fx_mcgs s = new fx_mcgs_s();
  s->ctx = ctx;
  s->A = A;
  auto fail = [&](int rc) {
    mcgs_free(s);
    return rc;
  };
  int rc = fx_pc_create_jacobi(ctx, A, weight, 1, &s->jac);
  if (rc) return fail(rc);
  s->ncolors = greedy_color(n, rp, ci, s->color);
  // per color: rows in ascending order, their entries, their original indices
  std::vector<uint32_t> count(s->ncolors, 0);
  for (uint32_t i = 0; i < n; ++i) ++count[s->color[i]];
  for (uint32_t c = 0; c < s->ncolors; ++c) {
    std::vector<uint32_t> rows, srp(1, 0), sci, sdp;
    std::vector<double> sva;
    rows.reserve(count[c]);
    for (uint32_t i = 0; i < n; ++i) {
      if (s->color[i] != c) continue;
      rows.push_back(i);
      for (uint32_t k = rp[i]; k < rp[i + 1]; ++k) {
        sci.push_back(ci[k]);
        sva.push_back(va[k]);
      }
      srp.push_back((uint32_t)sci.size());
    }
    const uint32_t nr = (uint32_t)rows.size(), nz = (uint32_t)sci.size();
    sdp.assign(nr, 0);  // unused by the sweep (the diagonal comes from dinv)
    fx_csr m = nullptr;
    rc = fx_csr_create(ctx, nr, n, nz, srp.data(), sci.data(), sva.data(), nz ? sdp.data() : nullptr,
                       &m);
    if (rc) return fail(rc);
    s->sub.push_back(m);
    uint32_t* d = nullptr;
    if (hipMalloc(&d, sizeof(uint32_t) * (nr ? nr : 1)) != hipSuccess) {
      set_error("hipMalloc failed");
      return fail(FX_ERROR_ALLOC_MEM);
    }
    s->rows.push_back(d);
    if (nr && hipMemcpy(d, rows.data(), sizeof(uint32_t) * nr, hipMemcpyHostToDevice) != hipSuccess) {
      set_error("upload failed");
      return fail(FX_ERROR_DEVICE);
    }
  }
  *out = s;
  return 0;
}

int fx_mcgs_info(fx_mcgs s, uint32_t* ncolors, uint32_t* color) {
  clear_error();
  FX_CHECK(s, FX_ERROR_INPUT_PAR, "null smoother");
  if (ncolors) *ncolors = s->ncolors;
  if (color) std::copy(s->color.begin(), s->color.end(), color);
  return 0;
}

int fx_mcgs_sweep(fx_mcgs s, fx_vec b, fx_vec x, uint32_t sweeps, int symmetric) {
  clear_error();
  FX_CHECK(s && b && x, FX_ERROR_INPUT_PAR, "null argument");
  FX_CHECK(b->n == s->A->nrow && x->n == s->A->nrow, FX_ERROR_NONMATCH_SIZE,
           "smoother size mismatch");
  FX_CHECK(b->dtype == FX_F64 && x->dtype == FX_F64, FX_ERROR_INPUT_PAR, "f64 vectors only");
  FX_CHECK(b->d != x->d, FX_ERROR_INPUT_PAR, "b and x must differ");
  hipStream_t st = s->ctx->stream;
  auto pass = [&](uint32_t c) -> int {
    fx_csr m = s->sub[c];
    if (m->nrow == 0) return 0;
    const uint32_t* rows = s->rows[c];
    return launch_spmv<double, double, ROW_B_MINUS_AX, 0, 0u>(
        m, st, XLoadInPlace{(const double*)x->d}, BRows{(const double*)b->d, rows},
        EpiGS{(double*)x->d, s->jac->dinv, rows}, nullptr);
  };
  for (uint32_t sw = 0; sw < sweeps; ++sw) {
    for (uint32_t c = 0; c < s->ncolors; ++c) FX_TRY(pass(c));
    if (symmetric)
      for (uint32_t c = s->ncolors; c-- > 0;) FX_TRY(pass(c));
  }
  return 0;
}

int fx_mcgs_destroy(fx_mcgs s) {
  mcgs_free(s);
  return 0;
}

}  // extern "C"
