// ref_shim.cxx -- C entry points over the compiled FASP++ reference.
//
// TEST INFRASTRUCTURE ONLY.  Compiled together with the reference's own
// sources (src/*.cxx, read where they lie) into oracle/_ref/libfaspxx_ref.so
// by oracle/Makefile.  Used by tests to pin the CPU restatement
// (oracle/fx_oracle.cpp) and, optionally, as the timed CPU baseline.
//
// The shim builds the reference's MAT / VEC / solver objects from plain
// arrays and calls their public methods.  GMRES and FGMRES are not exposed:
// their back-substitution loops never terminate (unsigned `k >= 0`,
// src/GMRES.cxx:214, src/FGMRES.cxx:209) and index out of bounds.

#include <chrono>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "BiCGStab.hxx"
#include "CG.hxx"
#include "Iter.hxx"
#include "MAT.hxx"
#include "MATUtil.hxx"
#include "ReadData.hxx"
#include "VEC.hxx"

namespace {

MAT make_mat(uint32_t nrow, uint32_t mcol, uint32_t nnz, const uint32_t* rp,
             const uint32_t* ci, const double* va, const uint32_t* dp) {
  std::vector<USI> r(rp, rp + nrow + 1), c(ci, ci + nnz);
  std::vector<USI> d;
  if (dp) d.assign(dp, dp + (nrow < mcol ? nrow : mcol));
  if (va) {
    std::vector<DBL> v(va, va + nnz);
    if (dp) return MAT(nrow, mcol, nnz, v, c, r, d);
    return MAT(nrow, mcol, nnz, v, c, r);
  }
  if (dp) return MAT(nrow, mcol, nnz, c, r, d);
  return MAT(nrow, mcol, nnz, c, r);
}

void copy_out(const VEC& v, double* out) {
  const double* p = nullptr;
  v.GetArray(&p);
  std::memcpy(out, p, sizeof(double) * v.GetSize());
}

}  // namespace

extern "C" {

// Threads the reference's OpenMP loops use (1 in the default serial build).
int ref_omp_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

struct ref_params {
  uint32_t maxIter, minIter, savIter, restart;
  double relTol, absTol;
};

struct ref_result {
  int retcode;
  uint32_t numIter;
  double norm2, normInf;
  double solve_ms;  // steady_clock around Solve only, as example/TestCG.cxx:86-88 times it
};

void ref_spmv(uint32_t nrow, uint32_t mcol, uint32_t nnz, const uint32_t* rp,
              const uint32_t* ci, const double* va, const uint32_t* dp, const double* v,
              double* w) {
  MAT A = make_mat(nrow, mcol, nnz, rp, ci, va, dp);
  VEC x(mcol, v), y(nrow, 0.0);
  A.Apply(x, y);
  copy_out(y, w);
}

void ref_residual(uint32_t nrow, uint32_t mcol, uint32_t nnz, const uint32_t* rp,
                  const uint32_t* ci, const double* va, const uint32_t* dp, const double* b,
                  const double* x, double* r) {
  MAT A = make_mat(nrow, mcol, nnz, rp, ci, va, dp);
  VEC bb(nrow, b), xx(mcol, x), rr(nrow, 0.0);
  A.Residual(bb, xx, rr);
  copy_out(rr, r);
}

// MAT::TransInPlace, then the transposed arrays.
void ref_transpose(uint32_t nrow, uint32_t mcol, uint32_t nnz, const uint32_t* rp,
                   const uint32_t* ci, const double* va, const uint32_t* dp, uint32_t* rpT,
                   uint32_t* ciT, double* vaT) {
  MAT A = make_mat(nrow, mcol, nnz, rp, ci, va, dp);
  A.TransInPlace();
  // GetRowPtr / GetColInd / GetValues return new[]-allocated copies
  int* r = A.GetRowPtr();
  int* c = A.GetColInd();
  std::memcpy(rpT, r, sizeof(uint32_t) * (mcol + 1));
  std::memcpy(ciT, c, sizeof(uint32_t) * nnz);
  delete[] r;
  delete[] c;
  if (va) {
    double* v = A.GetValues();
    std::memcpy(vaT, v, sizeof(double) * nnz);
    delete[] v;
  }
}

int ref_check_csrx(uint32_t nrow, uint32_t mcol, uint32_t nnz, const uint32_t* rp,
                   const uint32_t* ci, const double* va, const uint32_t* dp) {
  std::vector<USI> r(rp, rp + nrow + 1), c(ci, ci + nnz);
  std::vector<USI> d(dp, dp + (nrow < mcol ? nrow : mcol));
  std::vector<DBL> v(va, va + nnz);
  return CheckCSRx(nrow, mcol, nnz, v, c, r, d);
}

double ref_dot(uint32_t n, const double* a, const double* b) {
  VEC x(n, a), y(n, b);
  return x.Dot(y);
}
double ref_norm2(uint32_t n, const double* a) {
  VEC x(n, a);
  return x.Norm2();
}
double ref_norminf(uint32_t n, const double* a) {
  VEC x(n, a);
  return x.NormInf();
}
void ref_axpy(uint32_t n, double a, const double* x, double* y) {
  VEC xx(n, x), yy(n, y);
  yy.AXPY(a, xx);
  copy_out(yy, y);
}
void ref_xpay(uint32_t n, double a, const double* x, double* y) {
  VEC xx(n, x), yy(n, y);
  yy.XPAY(a, xx);
  copy_out(yy, y);
}
void ref_axpby(uint32_t n, double a, double b, const double* x, double* y) {
  VEC xx(n, x), yy(n, y);
  yy.AXPBY(a, b, xx);
  copy_out(yy, y);
}
void ref_waxpby(uint32_t n, double a, const double* v1, double b, const double* v2, double* w) {
  VEC x1(n, v1), x2(n, v2), ww(n, 0.0);
  ww.WAXPBY(a, x1, b, x2);
  copy_out(ww, w);
}

// kind: 1 CG, 2 BiCGStab, 11 Jacobi (solver).  pc_kind: 0 Identity,
// 1 Jacobi(weight) with minIter = maxIter = pc_iters (example/TestPCG.cxx:71-76).
// For kind 11 the Jacobi solver itself uses `weight`.
int ref_solve(int kind, uint32_t nrow, uint32_t nnz, const uint32_t* rp, const uint32_t* ci,
              const double* va, const uint32_t* dp, int pc_kind, uint32_t pc_iters,
              double weight, const ref_params* prm, const double* b, double* x,
              ref_result* res) {
  MAT A = make_mat(nrow, nrow, nnz, rp, ci, va, dp);
  VEC bb(nrow, b), xx(nrow, x);
  class Jacobi jac;
  Identity ident;
  SOL* pcd = &ident;
  if (pc_kind == 1) {
    jac.SetOutput(PRINT_NONE);
    jac.SetWeight(weight);
    jac.SetMaxIter(pc_iters);
    jac.SetMinIter(pc_iters);
    jac.Setup(A);
    pcd = &jac;
  }
  FaspRetCode rc = FaspRetCode::SUCCESS;
  auto configure = [&](SOL& s) {
    s.SetOutput(PRINT_NONE);
    s.SetMaxIter(prm->maxIter);
    s.SetMinIter(prm->minIter);
    s.SetSavIter(prm->savIter);
    s.SetRestart(prm->restart);
    s.SetRelTol(prm->relTol);
    s.SetAbsTol(prm->absTol);
  };
  if (kind == 1) {
    class CG s;
    configure(s);
    s.SetupPCD(*pcd);
    s.Setup(A);
    const auto t0 = std::chrono::steady_clock::now();
    rc = s.Solve(bb, xx);
    res->solve_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    res->numIter = s.GetIterations();
    res->norm2 = s.GetNorm2();
    res->normInf = s.GetInfNorm();
  } else if (kind == 2) {
    class BiCGStab s;
    configure(s);
    s.SetupPCD(*pcd);
    s.Setup(A);
    const auto t0 = std::chrono::steady_clock::now();
    rc = s.Solve(bb, xx);
    res->solve_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    res->numIter = s.GetIterations();
    res->norm2 = s.GetNorm2();
    res->normInf = s.GetInfNorm();
  } else if (kind == 11) {
    class Jacobi s;
    configure(s);
    s.SetWeight(weight);
    s.Setup(A);
    const auto t0 = std::chrono::steady_clock::now();
    rc = s.Solve(bb, xx);
    res->solve_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    res->numIter = s.GetIterations();
    res->norm2 = s.GetNorm2();
    res->normInf = s.GetInfNorm();
  } else {
    return -30;
  }
  res->retcode = rc;
  copy_out(xx, x);
  return rc;
}

}  // extern "C"

// ---- setup-time sparse algebra: MAT::GetLowerTri / GetUpperTri / Add / Mult.
// The result's arrays are malloc'd into *out (ref_mat_free releases them).
extern "C" {

struct ref_mat_out {
  uint32_t nrow, mcol, nnz;
  uint32_t* rp;
  uint32_t* ci;
  double* va;
};

void ref_mat_free(ref_mat_out* m) {
  std::free(m->rp);
  std::free(m->ci);
  std::free(m->va);
  m->rp = m->ci = nullptr;
  m->va = nullptr;
}

}  // extern "C"

namespace {

void mat_out(const MAT& M, ref_mat_out* out) {
  out->nrow = M.GetRowSize();
  out->mcol = M.GetColSize();
  out->nnz = M.GetNNZ();
  const uint32_t n1 = out->nrow + 1, nz = out->nnz ? out->nnz : 1;
  out->rp = (uint32_t*)std::malloc(sizeof(uint32_t) * n1);
  out->ci = (uint32_t*)std::malloc(sizeof(uint32_t) * nz);
  out->va = (double*)std::malloc(sizeof(double) * nz);
  int* r = M.GetRowPtr();
  int* c = M.GetColInd();
  double* v = M.GetValues();
  std::memcpy(out->rp, r, sizeof(uint32_t) * n1);
  std::memcpy(out->ci, c, sizeof(uint32_t) * out->nnz);
  std::memcpy(out->va, v, sizeof(double) * out->nnz);
  delete[] r;
  delete[] c;
  delete[] v;
}

}  // namespace

extern "C" {

void ref_mat_tri(uint32_t nrow, uint32_t mcol, uint32_t nnz, const uint32_t* rp,
                 const uint32_t* ci, const double* va, const uint32_t* dp, int upper,
                 ref_mat_out* out) {
  MAT A = make_mat(nrow, mcol, nnz, rp, ci, va, dp), T;
  if (upper)
    A.GetUpperTri(T);
  else
    A.GetLowerTri(T);
  mat_out(T, out);
}

// Returns 1 when the reference left the target unchanged (an operand empty).
int ref_mat_add(double a, uint32_t nrow, uint32_t mcol, uint32_t annz, const uint32_t* arp,
                const uint32_t* aci, const double* ava, double b, uint32_t bnnz,
                const uint32_t* brp, const uint32_t* bci, const double* bva, ref_mat_out* out) {
  MAT A = make_mat(nrow, mcol, annz, arp, aci, ava, nullptr);
  MAT B = make_mat(nrow, mcol, bnnz, brp, bci, bva, nullptr);
  MAT C;
  C.Add(a, A, b, B);
  if (C.GetNNZ() == 0) return 1;
  mat_out(C, out);
  return 0;
}

void ref_mat_mult(uint32_t lrow, uint32_t lcol, uint32_t lnnz, const uint32_t* lrp,
                  const uint32_t* lci, const double* lva, uint32_t rcol, uint32_t rnnz,
                  const uint32_t* rrp, const uint32_t* rci, const double* rva,
                  ref_mat_out* out) {
  MAT L = make_mat(lrow, lcol, lnnz, lrp, lci, lva, nullptr);
  MAT R = make_mat(lcol, rcol, rnnz, rrp, rci, rva, nullptr);
  MAT C;
  C.Mult(L, R);
  mat_out(C, out);
}

// ReadMat (src/ReadData.cxx:361-446): a .csr or .mtx file read and
// normalised to CSRx by the reference (SortCSRRow, CSRtoMAT / MTXtoMAT).
// Returns the reference's FaspRetCode; out is filled on success.
int ref_read_mat(const char* file, ref_mat_out* out) {
  MAT A;
  const FaspRetCode rc = ReadMat(file, A);
  if (rc != FaspRetCode::SUCCESS) return (int)rc;
  mat_out(A, out);
  return 0;
}

// ReadCSR (src/ReadData.cxx:221-357): the raw arrays as the file holds them
// (0-based), before any normalisation.
int ref_read_csr(const char* file, ref_mat_out* out) {
  USI row = 0, col = 0, nnz = 0;
  std::vector<USI> rp, ci;
  std::vector<DBL> va;
  const FaspRetCode rc = ReadCSR(file, row, col, nnz, rp, ci, va);
  if (rc != FaspRetCode::SUCCESS) return (int)rc;
  out->nrow = row;
  out->mcol = col;
  out->nnz = nnz;
  out->rp = (uint32_t*)std::malloc(sizeof(uint32_t) * (row + 1));
  out->ci = (uint32_t*)std::malloc(sizeof(uint32_t) * (nnz ? nnz : 1));
  out->va = (double*)std::malloc(sizeof(double) * (nnz ? nnz : 1));
  for (USI i = 0; i <= row; ++i) out->rp[i] = rp[i];
  for (USI k = 0; k < nnz; ++k) {
    out->ci[k] = ci[k];
    out->va[k] = va[k];
  }
  return 0;
}

// MAT::Shift on a matrix whose diagPtr the reference forms itself
// (FormDiagPtr in the constructor); the shifted values go to va_out.
void ref_mat_shift(uint32_t nrow, uint32_t mcol, uint32_t nnz, const uint32_t* rp,
                   const uint32_t* ci, const double* va, double a, double* va_out) {
  MAT A = make_mat(nrow, mcol, nnz, rp, ci, va, nullptr);
  A.Shift(a);
  double* v = A.GetValues();
  std::memcpy(va_out, v, sizeof(double) * nnz);
  delete[] v;
}

}  // extern "C"
