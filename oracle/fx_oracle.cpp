// fx_oracle.cpp -- CPU restatement of FASP++'s iterative-solve path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: it is used by the
// pytest suite, by the smoke check and by the timed CPU-baseline leg of the
// benchmark, and by nothing else.  The product (faspxx_amd/, include/) never
// links, loads or calls it.
//
// Every routine below restates one reference routine in plain C-style C++ and
// keeps the reference's floating-point operation order, so that a build with
// the same flags (-O2, FMA contraction on, as g++ does for C++17) produces the
// same bits.  The reference citation is given on each routine.  Deviations from
// the reference are stated where they occur; each one removes undefined
// behaviour and changes nothing on inputs where the reference is well defined:
//   D1  GMRES/FGMRES back-substitution and solution-update loops use a signed
//       index (reference: `USI k = count-2; k >= 0` never ends,
//       src/GMRES.cxx:214,224,386,396, src/FGMRES.cxx:209,219).
//   D2  CG's false-convergence restart zeroes p at full length (reference
//       calls `pk.SetValues(0.0)`, which resizes p to length 0, src/CG.cxx:213).
//   D3  CG's breakdown branch starts from alpha = 0 if it fires on the first
//       iteration (reference reads an uninitialised alpha, src/CG.cxx:109).
//   D4  An empty CSR row yields 0 in SpMV (reference reads the next row's
//       first entry, src/MAT.cxx:415; CSRx forbids empty rows anyway).
//   D5  MG::SetupLevel with level == numLevelsCoarse is rejected (reference:
//       writes past the end of infoHL, src/MG.cxx:27-31).
//   D6  Row sorting after MAT::Add uses a signed insertion index (reference
//       SortCSRRow: `USI l; l >= begin` wraps when an entry moves to the front
//       of row 0, src/MATUtil.cxx:454); MAT::Add's merge window, which may
//       read one slot past its buffer, gets that slot allocated.
//
// Pinning: see tests/test_oracle_*.py -- reference unit-test golden values,
// the compiled reference (oracle/_ref) on the same inputs, closed forms on
// stencil matrices, and textbook identities (CG A-orthogonality, Arnoldi).

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

// include/Faspxx.hxx:31-45
constexpr double kCloseZero = 1e-20;
constexpr double kSmallDbl = -1e60;  // negative: makes `> kSmallDbl` tests always true
constexpr double kChkRatio = 0.95;
constexpr uint32_t kMaxStag = 20;

// src/RetCode.hxx:24-42
enum {
  RC_SUCCESS = 0,
  RC_DIVIDE_ZERO = -18,
  RC_ALLOC_MEM = -20,
  RC_SOLVER_TYPE = -30,
  RC_SOLVER_STAG = -32,
  RC_SOLVER_SOLSTAG = -33,
  RC_SOLVER_TOLSMALL = -34,
  RC_SOLVER_MAXIT = -39,
  RC_INPUT_PAR = -12,
};

}  // namespace

extern "C" {

// A CSRx matrix as the reference stores it (src/MAT.hxx:43-47).  `va` may be
// null: the matrix is then a pure sparsity structure.
struct orc_csr {
  uint32_t nrow, mcol, nnz;
  const uint32_t* rp;
  const uint32_t* ci;
  const double* va;
  const uint32_t* dp;
};

// SOLParams (src/Param.hxx:57-80) minus the name/verbosity fields.
struct orc_params {
  uint32_t maxIter, minIter, savIter, restart;
  double relTol, absTol;
  // GMRES/FGMRES only (src/GMRES.hxx:61-62 defaults 30 / 10)
  uint32_t maxRestart, minRestart;
  int leftPrecond;  // GMRES only: 0 right (default), 1 left
};

// Preconditioner: 0 = Identity (src/Iter.cxx:16-20),
// 1 = Jacobi used with minIter == maxIter == iters (example/TestPCG.cxx:71-76).
struct orc_pc {
  int kind;            // 0 identity, 1 Jacobi, 2 symmetric multicolor Gauss-Seidel
  uint32_t iters;
  const double* dinv;  // omega / diag, formed as in src/Iter.cxx:45-47
  const uint32_t* color;  // kind 2: coloring (orc_mc_color) and its weight
  uint32_t ncolors;
  double weight;
};

struct orc_result {
  int retcode;
  uint32_t numIter;
  double norm2, normInf;
};

// Per-cycle trace of GMRES / FGMRES (test instrumentation; it changes no
// arithmetic).  For cycle k < cap: len[k] inner iterations, planned[k] the
// restart length the cycle started with (len[k] is shorter when the cycle
// ends on convergence or maxIter), est[k] the Givens residual estimate
// |g[len]| at the cycle's end, rtrue[k] the norm of the true residual
// b - A x after the cycle's x update (for left-preconditioned GMRES, of its
// preconditioned form).  count = cycles run.
struct orc_cycles {
  uint32_t cap, count;
  uint32_t* len;
  uint32_t* planned;
  double* est;
  double* rtrue;
};

// ---------------------------------------------------------------------------
// VEC operations -- src/VEC.cxx.  Elementwise ops keep the reference's
// expression shape; reductions keep its four strided accumulators.
// ---------------------------------------------------------------------------

// VEC::Dot, src/VEC.cxx:468-485
double orc_dot(uint32_t n, const double* a, const double* b) {
  const uint32_t len = n - n % 4;
  double d1 = 0.0, d2 = 0.0, d3 = 0.0, d4 = 0.0;
  for (uint32_t i = 0; i < len; i += 4) {
    d1 += a[i] * b[i];
    d2 += a[i + 1] * b[i + 1];
    d3 += a[i + 2] * b[i + 2];
    d4 += a[i + 3] * b[i + 3];
  }
  for (uint32_t i = len; i < n; ++i) d1 += a[i] * b[i];
  return d1 + d2 + d3 + d4;
}

// VEC::Norm2, src/VEC.cxx:424-438 (std::pow(v,2) is evaluated as v*v)
double orc_norm2(uint32_t n, const double* a) {
  const uint32_t len = n - n % 4;
  double s1 = 0.0, s2 = 0.0, s3 = 0.0, s4 = 0.0;
  for (uint32_t i = 0; i < len; i += 4) {
    s1 += a[i] * a[i];
    s2 += a[i + 1] * a[i + 1];
    s3 += a[i + 2] * a[i + 2];
    s4 += a[i + 3] * a[i + 3];
  }
  for (uint32_t i = len; i < n; ++i) s1 += a[i] * a[i];
  return std::sqrt(s1 + s2 + s3 + s4);
}

// VEC::NormInf, src/VEC.cxx:441-465
double orc_norminf(uint32_t n, const double* a) {
  double m = 0.0;
  for (uint32_t i = 0; i < n; ++i) {
    const double t = std::fabs(a[i]);
    if (t > m) m = t;
  }
  return m;
}

// VEC::AXPY, src/VEC.cxx:245-256:  y += a x
void orc_axpy(uint32_t n, double a, const double* x, double* y) {
  for (uint32_t i = 0; i < n; ++i) y[i] += a * x[i];
}

// VEC::XPAY, src/VEC.cxx:259-271:  y = a y + x
void orc_xpay(uint32_t n, double a, const double* x, double* y) {
  for (uint32_t i = 0; i < n; ++i) y[i] = a * y[i] + x[i];
}

// VEC::AXPBY, src/VEC.cxx:274-326:  y = a y + b x  (with a==1 / b==1 forms)
void orc_axpby(uint32_t n, double a, double b, const double* x, double* y) {
  const int form = (a == 1.0) + 2 * (b == 1.0);
  for (uint32_t i = 0; i < n; ++i) {
    switch (form) {
      case 0: y[i] = a * y[i] + b * x[i]; break;
      case 1: y[i] += b * x[i]; break;
      case 2: y[i] = a * y[i] + x[i]; break;
      default: y[i] += x[i]; break;
    }
  }
}

// VEC::WAXPBY, src/VEC.cxx:329-379:  w = a v1 + b v2
void orc_waxpby(uint32_t n, double a, const double* v1, double b, const double* v2,
                double* w) {
  const int form = (a == 1.0) + 2 * (b == 1.0);
  for (uint32_t i = 0; i < n; ++i) {
    switch (form) {
      case 0: w[i] = a * v1[i] + b * v2[i]; break;
      case 1: w[i] = v1[i] + b * v2[i]; break;
      case 2: w[i] = a * v1[i] + v2[i]; break;
      default: w[i] = v1[i] + v2[i]; break;
    }
  }
}

// VEC::PointwiseMult, src/VEC.cxx:168-179
void orc_pmul(uint32_t n, const double* v, double* y) {
  for (uint32_t i = 0; i < n; ++i) y[i] *= v[i];
}

// VEC::Scale, src/VEC.cxx:154-165
void orc_scale(uint32_t n, double a, double* y) {
  for (uint32_t i = 0; i < n; ++i) y[i] *= a;
}

// VEC::operator+=, src/VEC.cxx:66-78
void orc_add(uint32_t n, const double* v, double* y) {
  for (uint32_t i = 0; i < n; ++i) y[i] += v[i];
}

// ---------------------------------------------------------------------------
// MAT operations -- src/MAT.cxx
// ---------------------------------------------------------------------------

// MAT::Apply, src/MAT.cxx:373-452.  Row sums run left to right starting from
// the first product; the structure-only form sums v[col].  (D4 for empty rows.)
void orc_spmv(const orc_csr* A, const double* v, double* w) {
  for (uint32_t i = 0; i < A->nrow; ++i) {
    const uint32_t b = A->rp[i], e = A->rp[i + 1];
    if (b == e) {
      w[i] = 0.0;
      continue;
    }
    if (A->va) {
      double t = A->va[b] * v[A->ci[b]];
      for (uint32_t k = b + 1; k < e; ++k) t += A->va[k] * v[A->ci[k]];
      w[i] = t;
    } else {
      double t = v[A->ci[b]];
      for (uint32_t k = b + 1; k < e; ++k) t += v[A->ci[k]];
      w[i] = t;
    }
  }
}

// MAT::Residual, src/MAT.cxx:455-476:  r = b - A x, or r = b when
// ||x||_inf < CLOSE_ZERO (the preconditioning shortcut, :459-462).
void orc_residual(const orc_csr* A, const double* b, const double* x, double* r) {
  if (orc_norminf(A->mcol, x) < kCloseZero) {
    std::memcpy(r, b, sizeof(double) * A->nrow);
    return;
  }
  for (uint32_t i = 0; i < A->nrow; ++i) {
    const uint32_t s = A->rp[i], e = A->rp[i + 1];
    double t = b[i] - A->va[s] * x[A->ci[s]];
    for (uint32_t k = s + 1; k < e; ++k) t -= A->va[k] * x[A->ci[k]];
    r[i] = t;
  }
}

// MAT::FormDiagPtr, src/MAT.cxx:917-928: first k in row j with col == j.
// Rows without a diagonal keep the value already in dp (the reference's
// freshly resized vector holds 0 there).
void orc_form_diagptr(uint32_t nrow, const uint32_t* rp, const uint32_t* ci,
                      uint32_t* dp) {
  for (uint32_t j = 0; j < nrow; ++j) {
    for (uint32_t k = rp[j]; k < rp[j + 1]; ++k) {
      if (ci[k] == j) {
        dp[j] = k;
        break;
      }
    }
  }
}

// MAT::GetDiag, src/MAT.cxx:256-267 (ones for a structure-only matrix)
void orc_get_diag(const orc_csr* A, double* d) {
  for (uint32_t j = 0; j < A->nrow; ++j) d[j] = A->va ? A->va[A->dp[j]] : 1.0;
}

// Jacobi::Setup's diagonal, src/Iter.cxx:45-47:
// GetDiag, then Reciprocal (1.0/d, src/VEC.cxx:182-193), then Scale(weight).
void orc_jacobi_dinv(const orc_csr* A, double weight, double* dinv) {
  orc_get_diag(A, dinv);
  for (uint32_t j = 0; j < A->nrow; ++j) dinv[j] = 1.0 / dinv[j];
  for (uint32_t j = 0; j < A->nrow; ++j) dinv[j] *= weight;
}

// MAT::TransInPlace, src/MAT.cxx:479-538: counting sort by column; rows of the
// result list source rows in ascending order.  Outputs: rpT[mcol+1],
// ciT[nnz], vaT[nnz] (if A has values), dpT[mcol] (FormDiagPtr of the result).
void orc_transpose(const orc_csr* A, uint32_t* rpT, uint32_t* ciT, double* vaT,
                   uint32_t* dpT) {
  const uint32_t n = A->nrow, m = A->mcol, nnz = A->nnz;
  // rpT[c + 2] counts column c; prefix sums from index 2 give row starts shifted by one
  std::memset(rpT, 0, sizeof(uint32_t) * (m + 1));
  for (uint32_t k = 0; k < nnz; ++k) {
    const uint32_t c = A->ci[k];
    if (c + 1 < m) ++rpT[c + 2];
  }
  for (uint32_t c = 2; c <= m; ++c) rpT[c] += rpT[c - 1];
  for (uint32_t i = 0; i < n; ++i) {
    for (uint32_t p = A->rp[i]; p < A->rp[i + 1]; ++p) {
      const uint32_t slot = A->ci[p] + 1;
      const uint32_t k = rpT[slot];
      ciT[k] = i;
      if (A->va) vaT[k] = A->va[p];
      rpT[slot] = k + 1;
    }
  }
  std::memset(dpT, 0, sizeof(uint32_t) * m);
  orc_form_diagptr(m, rpT, ciT, dpT);
}

// CheckCSRx, src/MATUtil.cxx:177-249: 0 (SUCCESS) or -12 (ERROR_INPUT_PAR).
int orc_check_csrx(const orc_csr* A) {
  const uint32_t row = A->nrow, col = A->mcol, nnz = A->nnz;
  if (row == 0 || col == 0 || nnz == 0) return RC_SUCCESS;
  if (A->rp[row] != nnz) return RC_INPUT_PAR;
  for (uint32_t j = 0; j < row; ++j)
    if (A->rp[j] >= A->rp[j + 1]) return RC_INPUT_PAR;
  for (uint32_t j = 0; j < row; ++j) {
    const uint32_t b = A->rp[j], e = A->rp[j + 1];
    if (e == b + 1 && A->ci[b] != j) return RC_INPUT_PAR;
    if (e > b + 1) {
      for (uint32_t k = b; k + 1 < e; ++k)
        if (A->ci[k] >= A->ci[k + 1]) return RC_INPUT_PAR;
      if (A->ci[e - 1] >= col) return RC_INPUT_PAR;
    }
  }
  const uint32_t ndiag = row < col ? row : col;
  uint32_t count = 0;
  for (uint32_t j = 0; j < row; ++j) {
    for (uint32_t k = A->rp[j]; k < A->rp[j + 1]; ++k) {
      if (A->ci[k] == j) {
        if (count >= ndiag || A->dp[count] != k) return RC_INPUT_PAR;
        ++count;
      }
    }
  }
  return count == ndiag ? RC_SUCCESS : RC_INPUT_PAR;
}

// ---- Multicolor Gauss-Seidel (no reference counterpart) ---------------------
// The reference has no Gauss-Seidel; the device offers a multicolor one as a
// separately validated smoother.  This is its definition.
//
// Coloring: rows in natural order, each taking the smallest color not used by
// an already-colored row adjacent to it in the pattern of A + A^T (self loops
// ignored).  Returns the number of colors.
uint32_t orc_mc_color(const orc_csr* A, uint32_t* color) {
  const uint32_t n = A->nrow;
  // column lists of A (rows k with a_ki != 0), for the A^T half of the pattern
  std::vector<uint32_t> tp(n + 1, 0), tr(A->nnz);
  for (uint32_t k = 0; k < A->nnz; ++k)
    if (A->ci[k] < n) ++tp[A->ci[k] + 1];
  for (uint32_t i = 0; i < n; ++i) tp[i + 1] += tp[i];
  std::vector<uint32_t> cur(tp.begin(), tp.end() - 1);
  for (uint32_t i = 0; i < n; ++i)
    for (uint32_t k = A->rp[i]; k < A->rp[i + 1]; ++k)
      if (A->ci[k] < n) tr[cur[A->ci[k]]++] = i;
  const uint32_t kNone = 0xffffffffu;
  for (uint32_t i = 0; i < n; ++i) color[i] = kNone;
  std::vector<uint32_t> mark;  // mark[c] == i: color c is taken by a neighbour of i
  uint32_t ncolors = 0;
  for (uint32_t i = 0; i < n; ++i) {
    auto take = [&](uint32_t j) {
      if (j == i || j >= n || color[j] == kNone) return;
      if (color[j] >= mark.size()) mark.resize(color[j] + 1, kNone);
      mark[color[j]] = i;
    };
    for (uint32_t k = A->rp[i]; k < A->rp[i + 1]; ++k) take(A->ci[k]);
    for (uint32_t k = tp[i]; k < tp[i + 1]; ++k) take(tr[k]);
    uint32_t c = 0;
    while (c < mark.size() && mark[c] == i) ++c;
    color[i] = c;
    if (c + 1 > ncolors) ncolors = c + 1;
  }
  return ncolors;
}

// Sweeps of multicolor Gauss-Seidel (SOR with relaxation `weight`): for each
// color in turn (0..C-1; for a symmetric sweep then C-1..0), every row i of
// that color is updated from the current x:
//     r_i = b_i - sum_k a_ik x_k      (row order, t = b - a*x then t -= a*x)
//     x_i = fma(r_i, dinv_i, x_i)     dinv_i = weight * (1 / a_ii)
// Rows of one color do not read each other, so the order inside a color does
// not matter.
void orc_mcgs(const orc_csr* A, const uint32_t* color, uint32_t ncolors, double weight,
              const double* b, double* x, uint32_t sweeps, int symmetric) {
  const uint32_t n = A->nrow;
  std::vector<double> dinv(n);
  orc_jacobi_dinv(A, weight, dinv.data());
  auto pass = [&](uint32_t c) {
    for (uint32_t i = 0; i < n; ++i) {
      if (color[i] != c) continue;
      const uint32_t s = A->rp[i], e = A->rp[i + 1];
      double t = b[i];
      if (e > s) {
        t = b[i] - A->va[s] * x[A->ci[s]];
        for (uint32_t k = s + 1; k < e; ++k) t -= A->va[k] * x[A->ci[k]];
      }
      x[i] = std::fma(t, dinv[i], x[i]);
    }
  };
  for (uint32_t sw = 0; sw < sweeps; ++sw) {
    for (uint32_t c = 0; c < ncolors; ++c) pass(c);
    if (symmetric)
      for (uint32_t c = ncolors; c-- > 0;) pass(c);
  }
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Solvers
// ---------------------------------------------------------------------------
namespace {

using Vec = std::vector<double>;

struct Ctx {
  const orc_csr* A;
  const orc_pc* pc;
  orc_params prm;
  uint32_t n;
  double* hist;  // optional residual history (resAbs per reported iteration)
  uint32_t hist_cap;
  orc_cycles* cyc = nullptr;  // optional GMRES / FGMRES cycle trace
  double* vout = nullptr;     // optional: the basis V[0..maxR] of the last cycle
  // optional GMRES / FGMRES restart length kept in the solver object (the
  // reference's member `restart`): 0 on entry = as Setup leaves it; on exit
  // the length the solve ended with
  uint32_t* restart_io = nullptr;
  void rec(uint32_t it, double r) const {
    if (hist && it < hist_cap) hist[it] = r;
  }
  void cycle(uint32_t len, uint32_t planned, double est, double rtrue) const {
    if (!cyc) return;
    if (cyc->count < cyc->cap) {
      cyc->len[cyc->count] = len;
      cyc->planned[cyc->count] = planned;
      cyc->est[cyc->count] = est;
      cyc->rtrue[cyc->count] = rtrue;
    }
    ++cyc->count;
  }
  void basis(const std::vector<std::vector<double>>& V) const {
    if (!vout) return;
    for (size_t j = 0; j < V.size(); ++j)
      std::memcpy(vout + j * n, V[j].data(), sizeof(double) * n);
  }
};

// Preconditioner application z <- P(r) with z as the initial guess.
// Identity::Solve (src/Iter.cxx:16-20) copies; Jacobi::Solve with
// minIter == maxIter (src/Iter.cxx:56-112) runs `iters` sweeps and computes no
// norms: work = b - A z (shortcut to b when ||z||_inf < 1e-20),
// work *= dinv, z += work.
void apply_pc(const Ctx& c, const double* r, double* z, Vec& work) {
  const uint32_t n = c.n;
  if (c.pc->kind == 0) {
    std::memcpy(z, r, sizeof(double) * n);
    return;
  }
  if (c.pc->kind == 2) {  // no reference counterpart: the device's multicolor GS
    orc_mcgs(c.A, c.pc->color, c.pc->ncolors, c.pc->weight, r, z, c.pc->iters, 1);
    return;
  }
  for (uint32_t it = 0; it < c.pc->iters; ++it) {
    orc_residual(c.A, r, z, work.data());
    orc_pmul(n, c.pc->dinv, work.data());
    orc_add(n, work.data(), z);
  }
}

// b - A x as the solvers form it: Apply then XPAY(-1, b) (e.g. src/CG.cxx:70-71)
void true_residual(const Ctx& c, const double* b, const double* x, double* r) {
  orc_spmv(c.A, x, r);
  orc_xpay(c.n, -1.0, b, r);
}

// CG::Solve, src/CG.cxx:53-249 (with D2, D3)
int cg(const Ctx& c, const double* b, double* x, orc_result* res) {
  const uint32_t n = c.n;
  const orc_params& P = c.prm;
  int err = RC_SUCCESS;
  const double solStagTol = 1e-4 * P.relTol;
  uint32_t stagStep = 0, moreStep = 0, numIter = 0;
  double resAbs = 1.0, resRel = 1.0, denAbs = 1.0, ratio = 0.0, resAbsOld = 1.0;
  double alpha = 0.0, beta, tmpa, tmpb;
  Vec r(n), p(n), z(n), q(n), safe(n, 0.0), work(n);

  true_residual(c, b, x, r.data());
  std::fill(z.begin(), z.end(), 0.0);
  apply_pc(c, r.data(), z.data(), work);
  p = z;
  tmpa = orc_dot(n, z.data(), r.data());

  while (numIter < P.maxIter) {
    if (numIter == P.minIter) {
      resAbs = orc_norm2(n, r.data());
      resAbsOld = resAbs;
      denAbs = (kCloseZero > resAbs) ? kCloseZero : resAbs;
      resRel = resAbs / denAbs;
      if (resRel < P.relTol || resAbs < P.absTol) break;
    }
    if (numIter >= P.minIter) c.rec(numIter, resAbs);

    ++numIter;
    orc_spmv(c.A, p.data(), q.data());
    tmpb = orc_dot(n, q.data(), p.data());
    if (std::fabs(tmpb) > kCloseZero * kCloseZero) {
      alpha = tmpa / tmpb;
    } else {
      orc_axpy(n, alpha, p.data(), x);
      orc_axpy(n, -alpha, q.data(), r.data());
      resAbs = orc_norm2(n, r.data());
      resRel = resAbs / denAbs;
      if (resRel > P.relTol && resAbs > P.absTol) err = RC_DIVIDE_ZERO;
      break;
    }
    orc_axpy(n, alpha, p.data(), x);
    orc_axpy(n, -alpha, q.data(), r.data());

    if (numIter >= P.minIter) {
      resAbs = orc_norm2(n, r.data());
      resRel = resAbs / denAbs;
      ratio = resAbs / resAbsOld;
      if (numIter >= P.savIter && resAbs < resAbsOld) safe.assign(x, x + n);

      if (ratio > kChkRatio) {
        // Check I
        if (orc_norminf(n, x) < kCloseZero) {
          err = RC_SOLVER_SOLSTAG;
          break;
        }
        // Check II
        const double xRelDiff = std::fabs(alpha) * orc_norm2(n, p.data()) / orc_norm2(n, x);
        if (stagStep <= kMaxStag && xRelDiff < solStagTol) {
          true_residual(c, b, x, r.data());
          resAbs = orc_norm2(n, r.data());
          resRel = resAbs / denAbs;
          if (resRel < P.relTol || resAbs < P.absTol) break;
          if (stagStep >= kMaxStag) {
            err = RC_SOLVER_STAG;
            break;
          }
          std::fill(p.begin(), p.end(), 0.0);
          ++stagStep;
        }
      }
      // Check III
      if (resRel < P.relTol) {
        true_residual(c, b, x, r.data());
        resAbs = orc_norm2(n, r.data());
        resRel = resAbs / denAbs;
        if (resRel < P.relTol || resAbs < P.absTol) break;
        if (moreStep >= P.restart) {
          err = RC_SOLVER_TOLSMALL;
          break;
        }
        std::fill(p.begin(), p.end(), 0.0);  // D2
        ++moreStep;
      }
    }

    if (numIter < P.maxIter) {
      resAbsOld = resAbs;
      std::fill(z.begin(), z.end(), 0.0);
      apply_pc(c, r.data(), z.data(), work);
      tmpb = orc_dot(n, z.data(), r.data());
      beta = tmpb / tmpa;
      tmpa = tmpb;
      orc_xpay(n, beta, z.data(), p.data());
    }
  }

  res->numIter = numIter;
  if (!(numIter == P.minIter && numIter == P.maxIter)) {
    res->norm2 = resAbs;
    res->normInf = orc_norminf(n, r.data());
  }
  if (numIter > P.savIter) std::memcpy(x, safe.data(), sizeof(double) * n);
  res->retcode = err;
  return err;
}

// BiCGStab::Solve, src/BiCGStab.cxx:68-270.  Note the preconditioned
// recurrence: the residual is updated with P A p and P A s, the solution with
// P p and P s (:114-146).
int bicgstab(const Ctx& c, const double* b, double* x, orc_result* res) {
  const uint32_t n = c.n;
  const orc_params& P = c.prm;
  int err = RC_SUCCESS;
  const double solStagTol = 1e-4 * P.relTol;
  uint32_t stagStep = 0, moreStep = 0, numIter = 0;
  double resAbs = 1.0, resRel = 1.0, denAbs = 1.0, ratio = 0.0, resAbsOld = 1.0;
  double alpha = 0.0, beta, rr0, rr0old, omega, t12;
  Vec r0s(n), tmp(n), apj(n), asj(n), pj(n), rj(n), sj(n), ptmp(n), stmp(n), mp(n),
      ms(n), safe(n, 0.0), work(n);

  orc_spmv(c.A, x, tmp.data());
  orc_waxpby(n, 1.0, b, -1.0, tmp.data(), rj.data());
  r0s = rj;
  pj = rj;

  while (numIter < P.maxIter) {
    if (numIter == P.minIter) {
      resAbs = orc_norm2(n, rj.data());
      denAbs = (kCloseZero > resAbs) ? kCloseZero : resAbs;
      resRel = resAbs / denAbs;
      if (resRel < P.relTol || resAbs < P.absTol) break;
    }
    if (numIter >= P.minIter) c.rec(numIter, resAbs);

    ++numIter;
    rr0 = orc_dot(n, rj.data(), r0s.data());
    orc_spmv(c.A, pj.data(), apj.data());
    std::fill(ptmp.begin(), ptmp.end(), 0.0);
    apply_pc(c, apj.data(), ptmp.data(), work);
    t12 = orc_dot(n, ptmp.data(), r0s.data());
    if (std::fabs(t12) > 1e-40) {
      alpha = rr0 / t12;
    } else {
      err = RC_DIVIDE_ZERO;
      break;
    }
    orc_waxpby(n, 1.0, rj.data(), -alpha, ptmp.data(), sj.data());
    orc_spmv(c.A, sj.data(), asj.data());
    std::fill(stmp.begin(), stmp.end(), 0.0);
    apply_pc(c, asj.data(), stmp.data(), work);
    omega = orc_dot(n, stmp.data(), sj.data()) / orc_dot(n, stmp.data(), stmp.data());

    std::fill(mp.begin(), mp.end(), 0.0);
    apply_pc(c, pj.data(), mp.data(), work);
    std::fill(ms.begin(), ms.end(), 0.0);
    apply_pc(c, sj.data(), ms.data(), work);
    orc_waxpby(n, alpha, mp.data(), omega, ms.data(), tmp.data());
    orc_xpay(n, 1.0, tmp.data(), x);
    orc_waxpby(n, 1.0, sj.data(), -omega, stmp.data(), rj.data());

    if (numIter >= P.minIter) {
      resAbs = orc_norm2(n, rj.data());
      resRel = resAbs / denAbs;
      ratio = resAbs / resAbsOld;
      if (numIter >= P.savIter && resAbs < resAbsOld) safe.assign(x, x + n);

      if (ratio > kChkRatio && numIter > P.minIter) {
        // Check I uses the stagnation tolerance here (src/BiCGStab.cxx:169)
        if (orc_norminf(n, x) < solStagTol) {
          err = RC_SOLVER_SOLSTAG;
          break;
        }
        const double xRelDiff = std::fabs(alpha) * orc_norm2(n, pj.data()) / orc_norm2(n, x);
        if (stagStep <= kMaxStag && xRelDiff < solStagTol) {
          true_residual(c, b, x, rj.data());
          resAbs = orc_norm2(n, rj.data());
          resRel = resAbs / denAbs;
          if (resRel < P.relTol || resAbs < P.absTol) break;
          if (stagStep >= kMaxStag) {
            err = RC_SOLVER_STAG;
            break;
          }
          std::fill(pj.begin(), pj.end(), 0.0);
          ++stagStep;
        }
      }
      if (resRel < P.relTol) {
        true_residual(c, b, x, rj.data());
        resAbs = orc_norm2(n, rj.data());
        resRel = resAbs / denAbs;
        if (resRel < P.relTol || resAbs < P.absTol) break;
        if (moreStep >= P.restart) {
          err = RC_SOLVER_TOLSMALL;
          break;
        }
        std::fill(pj.begin(), pj.end(), 0.0);
        ++moreStep;
      }
    }

    if (numIter < P.maxIter) {
      resAbsOld = resAbs;
      rr0old = rr0;
      rr0 = orc_dot(n, rj.data(), r0s.data());
      beta = rr0 / rr0old * alpha / omega;
      orc_waxpby(n, 1.0, pj.data(), -omega, ptmp.data(), tmp.data());
      orc_waxpby(n, 1.0, rj.data(), beta, tmp.data(), pj.data());
    }
  }

  res->numIter = numIter;
  if (!(numIter == P.minIter && numIter == P.maxIter)) {
    res->norm2 = resAbs;
    res->normInf = orc_norminf(n, rj.data());
  }
  if (numIter > P.savIter) std::memcpy(x, safe.data(), sizeof(double) * n);
  res->retcode = err;
  return err;
}

// Restart-length clamping of GMRES::Setup / FGMRES::Setup
// (src/GMRES.cxx:43-46,75-79; src/FGMRES.cxx:29-32,67-71).
void clamp_restart(const orc_params& P, uint32_t& maxR, uint32_t& minR, uint32_t& restart) {
  maxR = P.maxRestart > P.maxIter ? P.maxIter : P.maxRestart;
  minR = P.minRestart > maxR ? maxR : P.minRestart;
  restart = P.restart;
  if (restart > maxR)
    restart = maxR;
  else if (restart < minR)
    restart = minR;
}

// One Arnoldi column's Givens update, shared by GMRES and FGMRES
// (src/GMRES.cxx:182-203).  h is (maxR+1) x maxR row-major with row stride ld.
void givens_column(std::vector<double>& h, uint32_t ld, std::vector<double>& hc,
                   std::vector<double>& hs, std::vector<double>& g, uint32_t cnt) {
  const uint32_t c1 = cnt - 1;
  auto H = [&](uint32_t i, uint32_t j) -> double& { return h[(size_t)i * ld + j]; };
  for (uint32_t j = 1; j < cnt; ++j) {
    const double t = H(j - 1, c1);
    H(j - 1, c1) = hs[j - 1] * H(j, c1) + hc[j - 1] * t;
    H(j, c1) = -hs[j - 1] * t + hc[j - 1] * H(j, c1);
  }
  double t = H(cnt, c1) * H(cnt, c1);
  t += H(c1, c1) * H(c1, c1);
  t = std::sqrt(t);
  const double gamma = t > kSmallDbl ? t : kSmallDbl;
  hc[c1] = H(c1, c1) / gamma;
  hs[c1] = H(cnt, c1) / gamma;
  H(c1, c1) = hs[c1] * H(cnt, c1) + hc[c1] * H(c1, c1);
  g[cnt] = -hs[c1] * g[c1];
  g[c1] = hc[c1] * g[c1];
}

// Upper-triangular solve of the least-squares system (D1: signed index),
// src/GMRES.cxx:213-219.
void back_substitute(const std::vector<double>& h, uint32_t ld, std::vector<double>& g,
                     uint32_t cnt) {
  auto H = [&](uint32_t i, uint32_t j) -> double { return h[(size_t)i * ld + j]; };
  const uint32_t c1 = cnt - 1;
  g[c1] = g[c1] / H(c1, c1);
  for (int k = (int)cnt - 2; k >= 0; --k) {
    double t = 0.0;
    for (uint32_t j = (uint32_t)k + 1; j < cnt; ++j) t -= H(k, j) * g[j];
    t += g[k];
    g[k] = t / H(k, k);
  }
}

// Variable restart rule after a cycle.  GMRES uses strict comparisons
// (src/GMRES.cxx:256-265), FGMRES non-strict (src/FGMRES.cxx:247-256).  The
// subtraction is unsigned, as in the reference.
void adapt_restart(double cr, bool inclusive, uint32_t maxR, uint32_t minR,
                   uint32_t& restart) {
  const double max_cr = 0.99026806874157040, min_cr = 0.17364817766693041;
  const uint32_t dec = 3;
  const bool mid = inclusive ? (cr <= max_cr && cr >= min_cr) : (cr < max_cr && cr > min_cr);
  if (cr > max_cr)
    restart = maxR;
  else if (mid) {
    if (restart - dec > minR)
      restart -= dec;
    else
      restart = maxR;
  }
}

// GMRES::RSolve (src/GMRES.cxx:114-280) and GMRES::LSolve (:283-453), with D1.
int gmres(const Ctx& c, const double* b, double* x, orc_result* res) {
  const uint32_t n = c.n;
  const orc_params& P = c.prm;
  const bool left = P.leftPrecond != 0;
  uint32_t maxR, minR, restart;
  clamp_restart(P, maxR, minR, restart);
  if (c.restart_io && *c.restart_io) restart = *c.restart_io < maxR ? *c.restart_io : maxR;
  int err = RC_SUCCESS;
  double t, ri, rj, cr;
  double resAbs = 1.0, resRel = 1.0, denAbs = 1.0, ratio = 0.0, resAbsOld = 1.0;
  uint32_t count = 1, count_1 = 2, numIter = 0;
  const uint32_t ld = maxR;
  std::vector<double> h((size_t)(maxR + 1) * maxR, 0.0), hc(maxR + 1, 0.0), hs(maxR, 0.0),
      g(maxR + 2, 0.0);
  std::vector<Vec> V(maxR + 1, Vec(n, 0.0));
  Vec wk(n), tmp(n), safe(x, x + n), work(n);
  auto H = [&](uint32_t i, uint32_t j) -> double& { return h[(size_t)i * ld + j]; };

  if (!left) {
    true_residual(c, b, x, V[0].data());
    resAbs = ri = orc_norm2(n, V[0].data());
  } else {
    true_residual(c, b, x, tmp.data());
    ri = orc_norm2(n, tmp.data());
    std::fill(V[0].begin(), V[0].end(), 0.0);
    apply_pc(c, tmp.data(), V[0].data(), work);
    resAbs = orc_norm2(n, V[0].data());
  }
  denAbs = (kCloseZero > resAbs) ? kCloseZero : resAbs;

  while (numIter < P.maxIter) {
    if (numIter == P.minIter) {
      resRel = resAbs / denAbs;
      resAbsOld = resAbs;
      if (resRel < P.relTol || resAbs < P.absTol) break;
    }
    if (resAbs < kSmallDbl) break;  // never true: kSmallDbl is negative
    g[0] = resAbs;
    orc_scale(n, 1 / resAbs, V[0].data());

    count = 0;
    while (count < restart && numIter < P.maxIter) {
      if (numIter >= P.minIter) c.rec(numIter, resAbs);
      ++numIter;
      count_1 = count++;
      if (!left) {
        std::fill(tmp.begin(), tmp.end(), 0.0);
        apply_pc(c, V[count_1].data(), tmp.data(), work);
        orc_spmv(c.A, tmp.data(), V[count].data());
      } else {
        orc_spmv(c.A, V[count_1].data(), tmp.data());
        std::fill(V[count].begin(), V[count].end(), 0.0);
        apply_pc(c, tmp.data(), V[count].data(), work);
      }
      for (uint32_t j = 0; j < count; ++j) {
        H(j, count_1) = orc_dot(n, V[count].data(), V[j].data());
        orc_axpby(n, 1.0, -H(j, count_1), V[j].data(), V[count].data());
      }
      t = orc_norm2(n, V[count].data());
      H(count, count_1) = t;
      if (std::fabs(t) > kSmallDbl)
        orc_scale(n, 1.0 / t, V[count].data());
      else
        break;

      givens_column(h, ld, hc, hs, g, count);
      resAbs = std::fabs(g[count]);
      ratio = resAbs / resAbsOld;
      if (!left) {
        resRel = resAbs / denAbs;
        resAbsOld = resAbs;
      } else {
        resAbsOld = resAbs;
        resRel = resAbs / denAbs;
      }
      if ((resAbs < P.absTol || resRel < P.relTol) && numIter > P.minIter) break;
    }

    back_substitute(h, ld, g, count);
    wk = V[count_1];
    orc_scale(n, g[count_1], wk.data());
    for (int j = (int)count - 2; j >= 0; --j) orc_axpby(n, 1.0, g[j], V[j].data(), wk.data());

    if (!left) {
      std::fill(tmp.begin(), tmp.end(), 0.0);
      apply_pc(c, wk.data(), tmp.data(), work);
      orc_axpy(n, 1.0, tmp.data(), x);
    } else {
      orc_axpy(n, 1.0, wk.data(), x);
    }
    c.basis(V);
    const double est = resAbs;
    if (numIter >= P.savIter && resAbs < resAbsOld) safe.assign(x, x + n);
    if (numIter > P.maxIter) {
      err = RC_SOLVER_MAXIT;
      break;
    }

    if (!left) {
      true_residual(c, b, x, V[0].data());
      resAbs = rj = orc_norm2(n, V[0].data());
    } else {
      true_residual(c, b, x, tmp.data());
      rj = orc_norm2(n, tmp.data());
      std::fill(V[0].begin(), V[0].end(), 0.0);
      apply_pc(c, tmp.data(), V[0].data(), work);
      resAbs = orc_norm2(n, V[0].data());
    }
    c.cycle(count, restart, est, resAbs);
    resRel = resAbs / denAbs;
    if ((resRel < P.relTol || resAbs < P.absTol) && numIter >= P.minIter) break;

    cr = rj / ri;
    ri = rj;
    adapt_restart(cr, false, maxR, minR, restart);
  }
  if (c.restart_io) *c.restart_io = restart;

  res->numIter = numIter;
  if (!(numIter == P.minIter && numIter == P.maxIter)) {
    const double* rv = left ? tmp.data() : V[0].data();
    res->norm2 = orc_norm2(n, rv);
    res->normInf = orc_norminf(n, rv);
  }
  if (numIter > P.savIter) std::memcpy(x, safe.data(), sizeof(double) * n);
  res->retcode = err;
  return err;
}

// FGMRES::Solve, src/FGMRES.cxx:108-271, with D1.  Z persists across cycles
// and is the preconditioner's initial guess (:158-159).  Z is the solver's
// workspace: zero at setup, kept across calls on the same workspace.
int fgmres(const Ctx& c, const double* b, double* x, orc_result* res, std::vector<Vec>& Z) {
  const uint32_t n = c.n;
  const orc_params& P = c.prm;
  uint32_t maxR, minR, restart;
  clamp_restart(P, maxR, minR, restart);
  if (c.restart_io && *c.restart_io) restart = *c.restart_io < maxR ? *c.restart_io : maxR;
  int err = RC_SUCCESS;
  double t, ri, rj, cr;
  double resAbs = 1.0, resRel = 1.0, denAbs = 1.0, ratio = 0.0, resAbsOld = 1.0;
  uint32_t count = 1, count_1 = 2, numIter = 0;
  const uint32_t ld = maxR;
  std::vector<double> h((size_t)(maxR + 1) * maxR, 0.0), hc(maxR + 1, 0.0), hs(maxR, 0.0),
      g(maxR + 2, 0.0);
  std::vector<Vec> V(maxR + 1, Vec(n, 0.0));
  if (Z.size() != maxR + 1) Z.assign(maxR + 1, Vec(n, 0.0));
  Vec wk(n), safe(x, x + n), work(n);
  auto H = [&](uint32_t i, uint32_t j) -> double& { return h[(size_t)i * ld + j]; };

  true_residual(c, b, x, V[0].data());
  resAbs = ri = orc_norm2(n, V[0].data());
  denAbs = (resAbs > kCloseZero) ? resAbs : kCloseZero;

  while (numIter < P.maxIter) {
    if (numIter == P.minIter) {
      resRel = resAbs / denAbs;
      resAbsOld = resAbs;
      if (resRel < P.relTol || resAbs < P.absTol) break;
    }
    if (resAbs < kSmallDbl) break;
    g[0] = resAbs;
    orc_scale(n, 1 / resAbs, V[0].data());

    count = 0;
    while (count < restart && numIter < P.maxIter) {
      if (numIter >= P.minIter) c.rec(numIter, resAbs);
      ++numIter;
      count_1 = count++;
      apply_pc(c, V[count_1].data(), Z[count_1].data(), work);
      orc_spmv(c.A, Z[count_1].data(), V[count].data());
      for (uint32_t j = 0; j < count; ++j) {
        H(j, count_1) = orc_dot(n, V[count].data(), V[j].data());
        orc_axpby(n, 1.0, -H(j, count_1), V[j].data(), V[count].data());
      }
      t = orc_norm2(n, V[count].data());
      H(count, count_1) = t;
      if (std::fabs(t) > kSmallDbl)
        orc_scale(n, 1.0 / t, V[count].data());
      else
        break;
      givens_column(h, ld, hc, hs, g, count);
      resAbs = std::fabs(g[count]);
      ratio = resAbs / resAbsOld;
      resAbsOld = resAbs;
      resRel = resAbs / denAbs;
      if ((resAbs < P.absTol || resRel < P.relTol) && numIter > P.minIter) break;
    }

    back_substitute(h, ld, g, count);
    wk = Z[count_1];
    orc_scale(n, g[count_1], wk.data());
    for (int j = (int)count - 2; j >= 0; --j) orc_axpby(n, 1.0, g[j], Z[j].data(), wk.data());
    orc_axpy(n, 1.0, wk.data(), x);
    c.basis(V);
    const double est = resAbs;
    if (numIter >= P.savIter && resAbs < resAbsOld) safe.assign(x, x + n);
    if (numIter > P.maxIter) {
      err = RC_SOLVER_MAXIT;
      break;
    }

    true_residual(c, b, x, V[0].data());
    resAbs = rj = orc_norm2(n, V[0].data());
    c.cycle(count, restart, est, resAbs);
    resRel = resAbs / denAbs;
    if ((resRel < P.relTol || resAbs < P.absTol) && numIter >= P.minIter) break;

    cr = rj / ri;
    ri = rj;
    adapt_restart(cr, true, maxR, minR, restart);
  }
  if (c.restart_io) *c.restart_io = restart;

  res->numIter = numIter;
  if (!(numIter == P.minIter && numIter == P.maxIter)) {
    res->norm2 = orc_norm2(n, V[0].data());
    res->normInf = orc_norminf(n, V[0].data());
  }
  if (numIter > P.savIter) std::memcpy(x, safe.data(), sizeof(double) * n);
  res->retcode = err;
  return err;
}

// Jacobi::Solve as a solver (minIter < maxIter), src/Iter.cxx:56-112.
int jacobi(const Ctx& c, const double* dinv, const double* b, double* x, orc_result* res) {
  const uint32_t n = c.n;
  const orc_params& P = c.prm;
  double resAbs = 1.0, resRel = 1.0, denAbs = 1.0, resAbsOld = 1.0;
  uint32_t numIter = 0;
  Vec work(n);
  while (numIter < P.maxIter) {
    orc_residual(c.A, b, x, work.data());
    if (numIter >= P.minIter) {
      resAbs = orc_norm2(n, work.data());
      if (numIter == P.minIter) denAbs = (kCloseZero > resAbs) ? kCloseZero : resAbs;
      resRel = resAbs / denAbs;
      if (resRel < P.relTol || resAbs < P.absTol) break;
      resAbsOld = resAbs;
      c.rec(numIter, resAbs);
    }
    orc_pmul(n, dinv, work.data());
    orc_add(n, work.data(), x);
    ++numIter;
  }
  (void)resAbsOld;
  res->numIter = numIter;
  if (!(numIter == P.minIter && numIter == P.maxIter)) {
    orc_residual(c.A, b, x, work.data());
    res->norm2 = orc_norm2(n, work.data());
    res->normInf = orc_norminf(n, work.data());
  }
  res->retcode = RC_SUCCESS;
  return RC_SUCCESS;
}

}  // namespace

extern "C" {

// Solver kinds follow SOLType (src/Param.hxx:36-43).
enum { ORC_CG = 1, ORC_BICGSTAB = 2, ORC_GMRES = 4, ORC_FGMRES = 5, ORC_JACOBI = 11 };

// One solve.  x is the initial guess on entry and the result on exit.
// For ORC_JACOBI the preconditioner's dinv is the solver's diagInv.
// hist (optional, hist_cap entries) receives the residual norm reported at
// each iteration (the value SOL::PrintInfo would print).
int orc_solve(int kind, const orc_csr* A, const orc_pc* pc, const orc_params* prm,
              const double* b, double* x, orc_result* res, double* hist, uint32_t hist_cap) {
  Ctx c{A, pc, *prm, A->nrow, hist, hist_cap};
  res->retcode = 0;
  res->numIter = 0;
  res->norm2 = 1e60;   // SOL() initialises norm2 / normInf to LARGE_DBL
  res->normInf = 1e60;
  switch (kind) {
    case ORC_CG: return cg(c, b, x, res);
    case ORC_BICGSTAB: return bicgstab(c, b, x, res);
    case ORC_GMRES: return gmres(c, b, x, res);
    case ORC_FGMRES: {
      std::vector<Vec> Z;
      return fgmres(c, b, x, res, Z);
    }
    case ORC_JACOBI: return jacobi(c, pc->dinv, b, x, res);
    default: res->retcode = RC_SOLVER_TYPE; return RC_SOLVER_TYPE;
  }
}

// GMRES / FGMRES with test instrumentation: the per-cycle trace (cyc, may be
// null), the last cycle's basis V[0..maxR] after its x update (vout,
// (maxR+1) n doubles, may be null), for FGMRES the solver's kept basis Z
// (z_io, (maxR+1) n doubles: the stale Z on entry -- zeros for a freshly set
// up solver -- and the kept Z on exit; null starts from zeros), and the
// restart length the solver object keeps between Solve calls (restart_io:
// 0 on entry = freshly set up; may be null).  With z_io and restart_io a
// sequence of Solve calls on one reference solver object is replayed.
int orc_krylov_traced(int kind, const orc_csr* A, const orc_pc* pc, const orc_params* prm,
                      const double* b, double* x, orc_result* res, double* hist,
                      uint32_t hist_cap, orc_cycles* cyc, double* vout, double* z_io,
                      uint32_t* restart_io) {
  Ctx c{A, pc, *prm, A->nrow, hist, hist_cap};
  c.cyc = cyc;
  c.vout = vout;
  c.restart_io = restart_io;
  if (cyc) cyc->count = 0;
  res->retcode = 0;
  res->numIter = 0;
  res->norm2 = 1e60;
  res->normInf = 1e60;
  if (kind == ORC_GMRES) return gmres(c, b, x, res);
  if (kind != ORC_FGMRES) {
    res->retcode = RC_SOLVER_TYPE;
    return RC_SOLVER_TYPE;
  }
  uint32_t maxR, minR, restart;
  clamp_restart(*prm, maxR, minR, restart);
  const uint32_t n = A->nrow;
  std::vector<Vec> Z(maxR + 1, Vec(n, 0.0));
  if (z_io)
    for (uint32_t j = 0; j <= maxR; ++j) std::memcpy(Z[j].data(), z_io + (size_t)j * n, sizeof(double) * n);
  const int rc = fgmres(c, b, x, res, Z);
  if (z_io)
    for (uint32_t j = 0; j <= maxR; ++j) std::memcpy(z_io + (size_t)j * n, Z[j].data(), sizeof(double) * n);
  return rc;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Multigrid -- MG<MAT>, src/MG.cxx:16-211, as example/TestMG.cxx:81-99 sets it
// up: every level's smoother and coarse solver is a Jacobi solver on the
// fine matrix A (Jacobi::Setup(mat), src/Iter.cxx:29-53), and its transfer
// operator serves as both restriction and prolongation through Apply.  The
// cycle forms every level's residual with the fine A into the one fine-size
// work vector r (src/MG.cxx:135,156), so the transfer operators must be
// square (n x n); an IdentityMatrix transfer (src/MAT.cxx:1088-1095) copies.
// ---------------------------------------------------------------------------
extern "C" {

struct orc_mg_level {
  const orc_csr* transfer;  // null: IdentityMatrix
  orc_params smooth;        // pre- and post-smoother (one Jacobi object)
  orc_params coarse;        // coarse solver, used on the last level
};

}  // extern "C"

namespace {

struct MgRun {
  const orc_csr* A;
  const double* dinv;
  uint32_t nlev;
  const orc_mg_level* lv;
  const uint32_t* ncycles;
  uint32_t n;
  int level = -1;
  Vec r;
  std::vector<Vec> hb, hx;  // infoHL[l].b, infoHL[l].x

  void jacobi_solve(const orc_params& p, const double* b, double* x) const {
    Ctx c{A, nullptr, p, n, nullptr, 0};
    orc_result rr{};
    jacobi(c, dinv, b, x, &rr);
  }
  void apply(const orc_mg_level& L, const double* in, double* out) const {
    if (L.transfer)
      orc_spmv(L.transfer, in, out);
    else
      std::memcpy(out, in, sizeof(double) * n);
  }
  // MG::MGCycle, src/MG.cxx:117-172
  void cycle(const double* b, double* x) {
    if (nlev == 0) {  // "No multilevel info available": x = b
      std::memcpy(x, b, sizeof(double) * n);
      return;
    }
    ++level;
    if (level == (int)nlev) return;
    const orc_mg_level& L = lv[level];
    double* bl = hb[level].data();
    double* xl = hx[level].data();
    jacobi_solve(L.smooth, b, x);
    orc_residual(A, b, x, r.data());
    apply(L, r.data(), bl);
    std::fill(hx[level].begin(), hx[level].end(), 0.0);
    if (level + 1 == (int)nlev)
      jacobi_solve(L.coarse, bl, xl);
    else
      for (uint32_t i = 0; i < ncycles[level]; ++i) cycle(bl, xl);
    apply(L, xl, r.data());
    orc_axpy(n, 1.0, r.data(), x);
    jacobi_solve(L.smooth, b, x);
    --level;
  }
};

}  // namespace

extern "C" {

// MG::Solve, src/MG.cxx:175-229.  dinv: the smoothers' diagInv (weight / diag
// of A).  hist (optional) receives resAbs at each checked iteration.
int orc_mg(const orc_csr* A, const double* dinv, uint32_t nlev, const orc_mg_level* lv,
           const uint32_t* ncycles, const orc_params* prm, const double* b, double* x,
           orc_result* res, double* hist, uint32_t hist_cap) {
  const uint32_t n = A->nrow;
  for (uint32_t l = 0; l < nlev; ++l) {
    const orc_csr* T = lv[l].transfer;
    if (T && (T->nrow != n || T->mcol != n)) return RC_INPUT_PAR;
  }
  MgRun run{A, dinv, nlev, lv, ncycles, n};
  run.r.assign(n, 0.0);
  run.hb.assign(nlev, Vec(n, 0.0));
  run.hx.assign(nlev, Vec(n, 0.0));
  const orc_params& P = *prm;
  double resAbs = 1.0, denAbs = 1.0;
  uint32_t numIter = 0;
  res->retcode = RC_SUCCESS;
  res->norm2 = 1e60;
  res->normInf = 1e60;
  while (numIter < P.maxIter) {
    orc_residual(A, b, x, run.r.data());
    if (numIter >= P.minIter) {
      resAbs = orc_norm2(n, run.r.data());
      if (numIter == P.minIter) denAbs = (kCloseZero > resAbs) ? kCloseZero : resAbs;
      if (resAbs / denAbs < P.relTol || resAbs < P.absTol) break;
      if (hist && numIter < hist_cap) hist[numIter] = resAbs;
    }
    run.cycle(b, x);
    ++numIter;
  }
  res->numIter = numIter;
  if (!(numIter == P.minIter && numIter == P.maxIter)) {
    orc_residual(A, b, x, run.r.data());
    res->norm2 = orc_norm2(n, run.r.data());
    res->normInf = orc_norminf(n, run.r.data());
  }
  return RC_SUCCESS;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Setup-time sparse algebra -- src/MAT.cxx:290-345 (GetLowerTri, GetUpperTri),
// :652-719 (Add), :722-800 (Mult).  Results are returned in malloc'd arrays
// (orc_mat_free releases them).
// ---------------------------------------------------------------------------
extern "C" {

struct orc_mat_out {
  uint32_t nrow, mcol, nnz;
  uint32_t* rp;
  uint32_t* ci;
  double* va;  // null for a structure-only result
  uint32_t* dp;
};

void orc_mat_free(orc_mat_out* m) {
  std::free(m->rp);
  std::free(m->ci);
  std::free(m->va);
  std::free(m->dp);
  m->rp = m->ci = m->dp = nullptr;
  m->va = nullptr;
}

}  // extern "C"

namespace {

void emit(uint32_t nrow, uint32_t mcol, const std::vector<uint32_t>& rp,
          const std::vector<uint32_t>& ci, const std::vector<double>* va, orc_mat_out* out) {
  const uint32_t nnz = rp[nrow];
  out->nrow = nrow;
  out->mcol = mcol;
  out->nnz = nnz;
  out->rp = (uint32_t*)std::malloc(sizeof(uint32_t) * (nrow + 1));
  out->ci = (uint32_t*)std::malloc(sizeof(uint32_t) * (nnz ? nnz : 1));
  out->va = va ? (double*)std::malloc(sizeof(double) * (nnz ? nnz : 1)) : nullptr;
  out->dp = (uint32_t*)std::calloc(nrow ? nrow : 1, sizeof(uint32_t));
  std::memcpy(out->rp, rp.data(), sizeof(uint32_t) * (nrow + 1));
  std::memcpy(out->ci, ci.data(), sizeof(uint32_t) * nnz);
  if (va) std::memcpy(out->va, va->data(), sizeof(double) * nnz);
  orc_form_diagptr(nrow, out->rp, out->ci, out->dp);  // FormDiagPtr, src/MAT.cxx:917-928
}

}  // namespace

extern "C" {

// MAT::GetLowerTri / GetUpperTri: row j keeps [rowPtr[j], diagPtr[j]] or
// [diagPtr[j], rowPtr[j+1]) by position.
void orc_mat_tri(const orc_csr* A, int upper, orc_mat_out* out) {
  const uint32_t n = A->nrow;
  std::vector<uint32_t> rp(n + 1, 0), ci;
  std::vector<double> va;
  for (uint32_t j = 0; j < n; ++j) {
    const uint32_t s = upper ? A->dp[j] : A->rp[j];
    const uint32_t e = upper ? A->rp[j + 1] : A->dp[j] + 1;
    for (uint32_t k = s; k < e; ++k) {
      ci.push_back(A->ci[k]);
      if (A->va) va.push_back(A->va[k]);
    }
    rp[j + 1] = (uint32_t)ci.size();
  }
  emit(n, A->mcol, rp, ci, A->va ? &va : nullptr, out);
}

// MAT::Add: returns 1 (target unchanged) when either operand has no entries,
// as the reference returns before assigning its result (src/MAT.cxx:659-669).
int orc_mat_add(double a, const orc_csr* A, double b, const orc_csr* B, orc_mat_out* out) {
  if (A->nnz == 0 || B->nnz == 0) return 1;
  const uint32_t n = A->nrow, cap = A->nnz + B->nnz;
  std::vector<uint32_t> rp(n + 1, 0), ci(cap + 1, 0xffffffffu);
  std::vector<double> va(cap + 1, 0.0);
  uint32_t count = 0;
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t countrow = 0;
    for (uint32_t j = A->rp[i]; j < A->rp[i + 1]; ++j) {
      va[count] = a * A->va[j];
      ci[count] = A->ci[j];
      ++rp[i + 1];
      ++count;
      ++countrow;
    }
    for (uint32_t k = B->rp[i]; k < B->rp[i + 1]; ++k) {
      bool added = false;
      // the reference's window: this row's entries from A plus one slot
      for (uint32_t l = rp[i]; l < rp[i] + countrow + 1; ++l) {
        if (B->ci[k] == ci[l]) {
          va[l] = std::fma(b, B->va[k], va[l]);
          added = true;
          break;
        }
      }
      if (!added) {
        va[count] = b * B->va[k];
        ci[count] = B->ci[k];
        ++rp[i + 1];
        ++count;
      }
    }
    rp[i + 1] += rp[i];
  }
  // SortCSRRow (src/MATUtil.cxx:439-465): insertion sort of each row (D6)
  for (uint32_t j = 0; j < n; ++j) {
    for (int64_t k = (int64_t)rp[j] + 1; k < (int64_t)rp[j + 1]; ++k) {
      const uint32_t c = ci[k];
      const double v = va[k];
      int64_t l = k - 1;
      for (; l >= (int64_t)rp[j] && c < ci[l]; --l) {
        ci[l + 1] = ci[l];
        va[l + 1] = va[l];
      }
      ci[l + 1] = c;
      va[l + 1] = v;
    }
  }
  emit(n, A->mcol, rp, ci, &va, out);
  return 0;
}

// MAT::Mult: columns in order of first appearance, values summed over the
// (k, l) products in order (structure-only operands: no values).
void orc_mat_mult(const orc_csr* L, const orc_csr* R, orc_mat_out* out) {
  const uint32_t n = L->nrow, m = R->mcol;
  std::vector<uint32_t> rp(n + 1, 0), ci, seen;
  seen.reserve(m);
  for (uint32_t i = 0; i < n; ++i) {
    seen.clear();
    for (uint32_t k = L->rp[i]; k < L->rp[i + 1]; ++k) {
      const uint32_t r = L->ci[k];
      for (uint32_t j = R->rp[r]; j < R->rp[r + 1]; ++j) {
        const uint32_t c = R->ci[j];
        uint32_t l = 0;
        while (l < seen.size() && seen[l] != c) ++l;
        if (l == seen.size()) seen.push_back(c);
      }
    }
    ci.insert(ci.end(), seen.begin(), seen.end());
    rp[i + 1] = (uint32_t)ci.size();
  }
  const bool values = L->va && R->va;
  std::vector<double> va(values ? ci.size() : 0);
  if (values) {
    for (uint32_t i = 0; i < n; ++i) {
      for (uint32_t j = rp[i]; j < rp[i + 1]; ++j) {
        double v = 0.0;
        for (uint32_t k = L->rp[i]; k < L->rp[i + 1]; ++k) {
          const uint32_t r = L->ci[k];
          for (uint32_t l = R->rp[r]; l < R->rp[r + 1]; ++l)
            if (R->ci[l] == ci[j]) v = std::fma(L->va[k], R->va[l], v);
        }
        va[j] = v;
      }
    }
  }
  emit(n, m, rp, ci, values ? &va : nullptr, out);
}

// MAT::Shift, src/MAT.cxx:359-364: for every row j in order,
// values[diagPtr[j]] += a (a structure-only matrix is left alone).  A row
// without a diagonal entry has diagPtr 0 from FormDiagPtr and adds to
// values[0].  va is updated in place; diagPtr entries past nnz are skipped
// (the reference would index out of bounds there).
void orc_mat_shift(const orc_csr* A, double a, double* va) {
  if (!A->va || A->nnz == 0) return;
  for (uint32_t j = 0; j < A->nrow; ++j) {
    const uint32_t d = A->dp[j];
    if (d < A->nnz) va[d] += a;
  }
}

}  // extern "C"
