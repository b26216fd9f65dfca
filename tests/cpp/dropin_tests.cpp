// dropin_tests.cpp -- the reference's unit tests, written against the drop-in
// classes (faspxx_amd/host), run on the GPU.  Mirrors
// utest/src/UnitTestsVEC.cxx, utest/src/UnitTestsMAT.cxx and
// utest/src/UnitTestsJacobi.cxx (same inputs, same tolerances), plus the
// solver call sequences of example/TestPCG.cxx and example/TestKrylov.cxx on
// a 2D Poisson system.  Prints one line per check; exit status = failures.
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "CG.hxx"
#include "Iter.hxx"
#include "Krylov.hxx"
#include "MAT.hxx"
#include "ReadData.hxx"
#include "VEC.hxx"

static int g_fail = 0;
#define REQUIRE(cond)                                                   \
  do {                                                                  \
    if (!(cond)) {                                                      \
      std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond);       \
      ++g_fail;                                                         \
    }                                                                   \
  } while (0)

static void test_vec() {
  const double TOL = 1e-10;  // UnitTestsVEC.cxx:20
  const double q[100] = {85, 89, 23, 39, 99, 41, 48, 68, 8,  16, 24, 28, 94, 93, 9,  49, 97,
                         40, 60, 27, 92, 100, 95, 5, 81, 14, 21, 22, 7, 37, 46, 71, 34, 90,
                         33, 54, 29, 19, 31, 64, 74, 12, 58, 59, 83, 88, 15, 52, 10, 18, 66,
                         38, 53, 57, 73, 1,  70, 67, 72, 78, 36, 47, 51, 84, 76, 45, 26, 43,
                         65, 42, 80, 56, 91, 11, 30, 82, 62, 6,  87, 13, 32, 3,  77, 4,  61,
                         35, 50, 69, 20, 98, 86, 96, 75, 25, 79, 17, 2,  55, 44, 63};
  VEC v7(100, q);
  REQUIRE(v7.GetSize() == 100);
  // Dot: exact (UnitTestsVEC.cxx:239-252)
  double dot1 = 0.0;
  for (USI i = 0; i < v7.GetSize(); i++) dot1 += v7[i] * v7[i];
  REQUIRE(dot1 == v7.Dot(v7));
  // Norm2 (:204-216)
  REQUIRE(std::abs(std::sqrt(dot1) - v7.Norm2()) < TOL);
  // AXPBY / WAXPBY (:218-237)
  {
    VEC v = v7;
    VEC v1(v), v2(v), v3(v), v4(v.GetSize());
    const double a = 1.0, b = 3.14159;
    for (USI i = 0; i < v.GetSize(); i++) v[i] = a * v1[i] + b * v[i];
    v2.AXPBY(a, b, v1);
    v4.WAXPBY(a, v1, b, v3);
    for (USI i = 0; i < v.GetSize(); i++) {
      REQUIRE(std::abs(v[i] - v2[i]) < TOL);
      REQUIRE(std::abs(v2[i] - v4[i]) < TOL);
    }
  }
  // PointwiseMult / PointwiseDivide exact (:254-286)
  {
    VEC v1 = v7, v2 = v7, v3(v1);
    for (USI i = 0; i < v1.GetSize(); i++) v3[i] = v1[i] * v2[i];
    v1.PointwiseMult(v2);
    for (USI i = 0; i < v1.GetSize(); i++) REQUIRE(v1[i] == v3[i]);
    VEC w1 = v7, w2 = v7, w3(w1);
    for (USI i = 0; i < w1.GetSize(); i++) w3[i] = w1[i] / w2[i];
    w1.PointwiseDivide(w2);
    for (USI i = 0; i < w1.GetSize(); i++) REQUIRE(w1[i] == w3[i]);
  }
  // Reciprocal, Scale, Shift, Abs, Max, Min, +=, -=
  {
    VEC v = v7, vv = v7;
    for (USI i = 0; i < vv.GetSize(); i++) vv[i] = 1.0 / vv[i];
    v.Reciprocal();
    for (USI i = 0; i < v.GetSize(); i++) REQUIRE(std::abs(v[i] - vv[i]) < TOL);
    VEC s = v7;
    s.Scale(-2.0);
    s.Shift(1.0);
    REQUIRE(s[0] == -169.0);
    s.Abs();
    REQUIRE(s[0] == 169.0);
    REQUIRE(v7.Max() == 100.0 && v7.Min() == 1.0);
    VEC t = v7;
    t += v7;
    REQUIRE(t[3] == 78.0);
    t -= v7;
    REQUIRE(t[3] == 39.0);
  }
  std::printf("ok vec\n");
}

static void test_mat() {
  const double TOL = 1e-14;  // UnitTestsMAT.cxx:21
  const std::vector<DBL> values1 = {0, 1, 2, 3, 4, 5, 2, 1, 2, 1};
  const std::vector<USI> colInd1 = {0, 2, 0, 1, 3, 0, 2, 3, 2, 3};
  const std::vector<USI> rowPtr1 = {0, 2, 5, 8, 10};
  const std::vector<USI> diagPtr1 = {0, 3, 6, 9};
  const MAT mat1(4, 4, 10, values1, colInd1, rowPtr1, diagPtr1);
  const VEC vec2(4, 0.314);
  const double p1[4] = {0.314, 2.826, 2.512, 0.942};
  const VEC vec3(4, p1);
  VEC r(vec2.GetSize());
  mat1.Apply(vec2, r);  // UnitTestsMAT.cxx:152-161
  for (USI i = 0; i < r.GetSize(); i++) REQUIRE(std::abs(r[i] - vec3[i]) < TOL);
  REQUIRE(mat1.GetNNZ() == 10 && mat1.GetRowSize() == 4);
  REQUIRE(mat1.GetValue(1, 3) == 4.0 && mat1.GetValue(0, 1) == 0.0);
  VEC d;
  mat1.GetDiag(d);
  REQUIRE(d.GetSize() == 4 && d[0] == 0.0 && d[1] == 3.0 && d[2] == 2.0 && d[3] == 1.0);
  // copy, scale (UnitTestsMAT.cxx:59-62: mat1 * 2.87)
  MAT mat7 = mat1;
  mat7.Scale(2.87);
  REQUIRE(std::abs(mat7.GetValue(1, 0) - 5.74) < TOL);
  // residual with x = 0 returns b (src/MAT.cxx:459-462)
  VEC zero(4, 0.0), res(4);
  mat1.Residual(vec3, zero, res);
  for (USI i = 0; i < 4; i++) REQUIRE(res[i] == vec3[i]);
  // transpose twice is the identity on the structure and values
  MAT t = mat1;
  t.TransInPlace();
  REQUIRE(t.GetValue(2, 0) == 1.0 && t.GetValue(0, 2) == 5.0);
  t.TransInPlace();
  for (USI i = 0; i < 4; i++)
    for (USI j = 0; j < 4; j++) REQUIRE(t.GetValue(i, j) == mat1.GetValue(i, j));
  std::printf("ok mat\n");
}

// Setup-time algebra on the UnitTestsMAT.cxx:42-46 matrix: Zero (its
// :179-187 section) and the operations it has no section for, against the
// dense 4 x 4 arithmetic (small integers: exact).
static void test_mat_algebra() {
  const std::vector<DBL> values1 = {0, 1, 2, 3, 4, 5, 2, 1, 2, 1};
  const std::vector<USI> colInd1 = {0, 2, 0, 1, 3, 0, 2, 3, 2, 3};
  const std::vector<USI> rowPtr1 = {0, 2, 5, 8, 10};
  const std::vector<USI> diagPtr1 = {0, 3, 6, 9};
  const MAT mat1(4, 4, 10, values1, colInd1, rowPtr1, diagPtr1);
  double A[4][4];
  for (USI i = 0; i < 4; i++)
    for (USI j = 0; j < 4; j++) A[i][j] = mat1.GetValue(i, j);
  {
    MAT m(mat1);
    m.Zero();
    for (USI i = 0; i < 4; i++)
      for (USI j = 0; j < 4; j++) REQUIRE(m.GetValue(i, j) == 0);
  }
  MAT lo, up;
  mat1.GetLowerTri(lo);
  mat1.GetUpperTri(up);
  REQUIRE(lo.GetNNZ() == 7 && up.GetNNZ() == 7);  // (0,0) is stored: both keep it
  for (USI i = 0; i < 4; i++)
    for (USI j = 0; j < 4; j++) {
      REQUIRE(lo.GetValue(i, j) == (j <= i ? A[i][j] : 0.0));
      REQUIRE(up.GetValue(i, j) == (j >= i ? A[i][j] : 0.0));
    }
  {
    MAT m(mat1);
    m.Shift(1.5);
    for (USI i = 0; i < 4; i++)
      for (USI j = 0; j < 4; j++) REQUIRE(m.GetValue(i, j) == A[i][j] + (i == j ? 1.5 : 0.0));
  }
  {
    MAT s;
    s.Add(2.0, up, -1.0, lo);  // union of the patterns: mat1's; the diagonal 2d - d
    REQUIRE(s.GetNNZ() == 10);
    for (USI i = 0; i < 4; i++)
      for (USI j = 0; j < 4; j++)
        REQUIRE(s.GetValue(i, j) == (j > i ? 2 * A[i][j] : j < i ? -A[i][j] : A[i][j]));
    MAT e, keep(mat1);
    keep.Add(1.0, e, 1.0, mat1);  // an empty operand leaves the target unchanged
    REQUIRE(keep.GetNNZ() == 10 && keep.GetValue(1, 3) == 4.0);
  }
  {
    MAT p;
    p.Mult(mat1, mat1);
    for (USI i = 0; i < 4; i++)
      for (USI j = 0; j < 4; j++) {
        double want = 0;
        for (USI k = 0; k < 4; k++) want += A[i][k] * A[k][j];
        REQUIRE(p.GetValue(i, j) == want);
      }
    MAT q(mat1);
    q.MultLeft(mat1);
    for (USI i = 0; i < 4; i++)
      for (USI j = 0; j < 4; j++) REQUIRE(q.GetValue(i, j) == p.GetValue(i, j));
    int* rp = p.GetRowPtr();
    REQUIRE(rp[0] == 0 && (USI)rp[4] == p.GetNNZ());
    delete[] rp;
  }
  {
    MAT dinv;
    mat1.GetDiagInv(dinv);
    REQUIRE(dinv.GetNNZ() == 4 && dinv.GetValue(1, 1) == 1.0 / 3.0 && dinv.GetValue(3, 3) == 1.0);
  }
  {
    MAT m(mat1);
    m.Empty();
    REQUIRE(m.GetRowSize() == 0 && m.GetNNZ() == 0);
  }
  std::printf("ok mat algebra\n");
}

// example/TestMatFree.cxx's pattern: a LOP subclass whose Apply works on
// host copies (GetArray) -- the 5-point Laplacian on an m x m grid -- solved
// by the device CG; against the same solve with the assembled MAT.
class MatFree : public LOP {
 public:
  explicit MatFree(USI m) : LOP(m * m), m(m) {}
  void Apply(const VEC& x, VEC& y) const override {
    const DBL* xv;
    x.GetArray(&xv);
    y.SetValues(x.GetSize(), 0.0);
    DBL* yv;
    y.GetArray(&yv);
    for (USI r = 0; r < m; ++r)
      for (USI c = 0; c < m; ++c) {
        const USI k = r * m + c;
        DBL v = 4.0 * xv[k];
        if (r > 0) v -= xv[k - m];
        if (c > 0) v -= xv[k - 1];
        if (c + 1 < m) v -= xv[k + 1];
        if (r + 1 < m) v -= xv[k + m];
        yv[k] = v;
      }
  }

 private:
  USI m;
};

static void test_matfree() {
  const USI m = 31, n = m * m;
  std::vector<DBL> va;
  std::vector<USI> ci, rp(1, 0);
  for (USI r = 0; r < m; ++r)
    for (USI c = 0; c < m; ++c) {
      const USI k = r * m + c;
      if (r > 0) ci.push_back(k - m), va.push_back(-1.0);
      if (c > 0) ci.push_back(k - 1), va.push_back(-1.0);
      ci.push_back(k), va.push_back(4.0);
      if (c + 1 < m) ci.push_back(k + 1), va.push_back(-1.0);
      if (r + 1 < m) ci.push_back(k + m), va.push_back(-1.0);
      rp.push_back((USI)ci.size());
    }
  const MAT A(n, n, (USI)ci.size(), va, ci, rp);
  std::vector<DBL> bh(n);
  for (USI k = 0; k < n; ++k) bh[k] = std::sin(0.37 * k);
  const VEC b(bh);
  VEC x1(n, 0.25), x2(n, 0.25);  // TestMatFree's initial guess
  Identity pcd;
  MatFree op(m);
  USI it[2];
  for (int pass = 0; pass < 2; ++pass) {
    CG cg;
    cg.SetMaxIter(10000);
    cg.SetRelTol(1e-8);
    cg.SetAbsTol(1e-12);
    cg.SetupPCD(pcd);
    if (pass == 0)
      cg.Setup(op);
    else
      cg.Setup(A);
    REQUIRE(cg.Solve(b, pass == 0 ? x1 : x2) == SUCCESS);
    it[pass] = cg.GetIterations();
  }
  REQUIRE(it[0] > 10 && (it[0] > it[1] ? it[0] - it[1] : it[1] - it[0]) <= 2);
  VEC d(x1);
  d.AXPY(-1.0, x2);
  REQUIRE(d.Norm2() <= 1e-8 * x2.Norm2());
  std::printf("ok matfree (%u / %u iterations)\n", it[0], it[1]);
}

static void test_jacobi() {
  // UnitTestsJacobi.cxx:85-157 (3x3, Octave solution, tol 1e-5)
  std::vector<DBL> value = {4, 2, 1, 3, 8, 2, 2, 1, 9};
  std::vector<USI> colInd = {0, 1, 2, 0, 1, 2, 0, 1, 2};
  std::vector<USI> rowPtr = {0, 3, 6, 9};
  std::vector<USI> diagPtr = {0, 4, 8};
  MAT mat(3, 3, 9, value, colInd, rowPtr, diagPtr);
  VEC f(std::vector<DBL>{1, 1, 1});
  const double xstar[3] = {0.221719, 0.027149, 0.058824};
  for (double w : {1.0, 0.5}) {
    VEC x(3, 0.0);
    class Jacobi J;
    J.SetRelTol(1e-9);
    J.SetMaxIter(100);
    J.SetWeight(w);
    J.Setup(mat);
    J.Solve(f, x);
    for (USI i = 0; i < 3; i++) REQUIRE(std::abs(x[i] - xstar[i]) < 1e-5);
  }
  // 2x2 (UnitTestsJacobi.cxx:23-83)
  MAT m2(2, 2, 4, std::vector<DBL>{4, 2, 1, 7}, std::vector<USI>{0, 1, 0, 1},
         std::vector<USI>{0, 2, 4}, std::vector<USI>{0, 3});
  VEC f2(std::vector<DBL>{1, 1});
  for (double w : {1.0, 0.5}) {
    VEC x(2, 0.0);
    class Jacobi J;
    J.SetRelTol(1e-9);
    J.SetMaxIter(100);
    J.SetWeight(w);
    J.Setup(m2);
    J.Solve(f2, x);
    REQUIRE(std::abs(x[0] - 0.19231) < 1e-5 && std::abs(x[1] - 0.11538) < 1e-5);
  }
  std::printf("ok jacobi\n");
}

// 2D Poisson (5-point, m x m): the TestPCG / TestKrylov call sequences.
static void test_solvers() {
  const USI m = 64, n = m * m;
  std::vector<USI> rp(1, 0), ci, dp;
  std::vector<DBL> va;
  for (USI i = 0; i < n; ++i) {
    const USI ix = i % m, iy = i / m;
    if (iy > 0) { ci.push_back(i - m); va.push_back(-1); }
    if (ix > 0) { ci.push_back(i - 1); va.push_back(-1); }
    dp.push_back((USI)ci.size()); ci.push_back(i); va.push_back(4);
    if (ix < m - 1) { ci.push_back(i + 1); va.push_back(-1); }
    if (iy < m - 1) { ci.push_back(i + m); va.push_back(-1); }
    rp.push_back((USI)ci.size());
  }
  MAT mat(n, n, (USI)ci.size(), va, ci, rp, dp);
  VEC xs(n), b(n);
  for (USI i = 0; i < n; ++i) xs[i] = std::sin(0.01 * i);
  mat.Apply(xs, b);

  // TestPCG: Jacobi preconditioner with minIter = maxIter
  class Jacobi pcd;
  pcd.SetOutput(PRINT_NONE);
  pcd.SetWeight(0.8);
  pcd.SetMaxIter(1);
  pcd.SetMinIter(1);
  pcd.Setup(mat);
  class CG solver;
  solver.SetMaxIter(1000);
  solver.SetRelTol(1e-10);
  solver.SetAbsTol(0.0);
  solver.SetupPCD(pcd);
  solver.Setup(mat);
  VEC x(n, 0.0);
  FaspRetCode rc = solver.Solve(b, x);
  VEC r(n);
  mat.Apply(x, r);
  r.XPAY(-1.0, b);
  REQUIRE(rc == SUCCESS);
  REQUIRE(r.Norm2() / b.Norm2() < 1e-9);
  REQUIRE(solver.GetIterations() > 10 && solver.GetNorm2() < 1e-9 * b.Norm2());
  std::printf("ok pcg %u iterations\n", solver.GetIterations());

  // TestKrylov: every name through Krylov()
  for (const char* name : {"cg", "BiCGStab", "gmres", "fgmres"}) {
    SOLParams prm;
    prm.algName = name;
    prm.maxIter = 2000;
    prm.relTol = 1e-9;
    prm.absTol = 0.0;
    Identity id;
    VEC y(n, 0.0);
    rc = Krylov(mat, b, y, id, prm);
    VEC e = y;
    e.AXPY(-1.0, xs);
    REQUIRE(rc == SUCCESS);
    REQUIRE(e.Norm2() / xs.Norm2() < 1e-6);
    std::printf("krylov %s: rc %d\n", name, (int)rc);
  }
}

// One CG object and one Jacobi preconditioner kept across a rebuilt MAT:
// the matrix is assigned anew (another size, then the same size with other
// values), Jacobi::Setup and CG::Setup run again, and each solve must still
// reach the tolerance on the new system.
static MAT poisson2d(USI m, double scale) {
  std::vector<USI> rp(1, 0), ci;
  std::vector<DBL> va;
  const USI n = m * m;
  for (USI i = 0; i < n; ++i) {
    const USI ix = i % m, iy = i / m;
    if (iy > 0) { ci.push_back(i - m); va.push_back(-scale); }
    if (ix > 0) { ci.push_back(i - 1); va.push_back(-scale); }
    ci.push_back(i); va.push_back(4 * scale + 0.01 * (i % 7));
    if (ix < m - 1) { ci.push_back(i + 1); va.push_back(-scale); }
    if (iy < m - 1) { ci.push_back(i + m); va.push_back(-scale); }
    rp.push_back((USI)ci.size());
  }
  return MAT(n, n, (USI)ci.size(), va, ci, rp);
}

static void test_rebuild() {
  MAT mat = poisson2d(64, 1.0);
  class Jacobi pcd;
  pcd.SetOutput(PRINT_NONE);
  pcd.SetMaxIter(1);
  pcd.SetMinIter(1);
  class CG solver;
  solver.SetMaxIter(2000);
  solver.SetRelTol(1e-10);
  solver.SetAbsTol(0.0);
  solver.SetupPCD(pcd);
  const struct { USI m; double scale; } steps[] = {{64, 1.0}, {48, 1.0}, {48, 2.5}, {64, 1.0}};
  for (const auto& st : steps) {
    mat = poisson2d(st.m, st.scale);
    pcd.Setup(mat);
    solver.Setup(mat);
    const USI n = st.m * st.m;
    VEC xs(n), b(n), x(n, 0.0), r(n);
    for (USI i = 0; i < n; ++i) xs[i] = std::cos(0.02 * i);
    mat.Apply(xs, b);
    REQUIRE(solver.Solve(b, x) == SUCCESS);
    mat.Apply(x, r);
    r.XPAY(-1.0, b);
    REQUIRE(r.Norm2() / b.Norm2() < 1e-9);
    x.AXPY(-1.0, xs);
    REQUIRE(x.Norm2() / xs.Norm2() < 1e-7);
    std::printf("rebuild %ux%u scale %.1f: %u iterations\n", st.m, st.m, st.scale,
                solver.GetIterations());
  }
  std::printf("ok rebuild\n");
}

// ReadMat (src/ReadData.cxx:361-446) into a device MAT, then CG on it:
// nos7 as stored is its lower triangle, so solve with the normalised
// fdm_10X10 system, which is SPD.  The data directory is passed in argv[1].
static void test_readmat(const char* dir) {
  const std::string f = std::string(dir) + "/fdm_10X10.csr";
  MAT A;
  REQUIRE(ReadMat(f.c_str(), A) == SUCCESS);
  REQUIRE(A.GetRowSize() == 100 && A.GetNNZ() == 460);
  VEC ones(100, 1.0), y(100);
  A.Apply(ones, y);
  USI row, col, nnz;
  std::vector<USI> rp, ci;
  std::vector<DBL> va;
  REQUIRE(ReadCSRx(f.c_str(), row, col, nnz, rp, ci, va) == SUCCESS);
  for (USI i = 0; i < row; ++i) {
    double t = va[rp[i]];
    for (USI k = rp[i] + 1; k < rp[i + 1]; ++k) t += va[k];
    REQUIRE(y[i] == t);
  }
  class CG s;
  s.SetMaxIter(500);
  s.SetRelTol(1e-10);
  s.SetAbsTol(0.0);
  Identity id;
  s.SetupPCD(id);
  s.Setup(A);
  VEC x(100, 0.0);
  REQUIRE(s.Solve(y, x) == SUCCESS);
  VEC e = x;
  e.AXPY(-1.0, ones);
  REQUIRE(e.Norm2() < 1e-7);
  MAT B;
  REQUIRE(ReadMat((std::string(dir) + "/nos7-.mtx").c_str(), B) == SUCCESS);
  REQUIRE(B.GetRowSize() == 729 && B.GetNNZ() == 2673);  // the missing (1,1) inserted as 0
  REQUIRE(B.GetValue(0, 0) == 0.0);
  REQUIRE(ReadMat((std::string(dir) + "/missing.csr").c_str(), B) == ERROR_OPEN_FILE);
  std::printf("ok readmat\n");
}

int main(int argc, char** argv) {
  if (argc > 1) test_readmat(argv[1]);
  test_vec();
  test_mat();
  test_mat_algebra();
  test_matfree();
  test_jacobi();
  test_solvers();
  test_rebuild();
  std::printf("%s (%d failures)\n", g_fail ? "FAILED" : "PASSED", g_fail);
  return g_fail;
}
