// TestKrylov.cpp -- the reference's Krylov driver (example/TestKrylov.cxx:21-80)
// built against the drop-in classes, so it runs on the GPU unchanged: read the
// options (command line over parameter file over defaults), read A (.csr or
// .mtx, normalised to CSRx), b (default 0) and x0 (default 1), and solve with
// Krylov() and the identity preconditioner.  Options:
//   -mat <file> -rhs <file> -xin <file> -par <file>
//   -algName cg|bicgstab|gmres|fgmres -maxIter -minIter -savIter -restart
//   -resRel -resAbs -verbose
// It ends with one line "retcode numIter norm2 ms" for scripts.
#include <chrono>
#include <cstdio>
#include <cstring>
#include <iostream>

#include "Iter.hxx"
#include "Krylov.hxx"
#include "LOP.hxx"
#include "Param.hxx"
#include "ReadData.hxx"

int main(int argc, const char* args[]) {
  std::string parFile = "tests/golden/input.param";
  std::string matFile = "tests/golden/fdm_10X10.csr";
  std::string rhsFile, xinFile;

  Parameters params(argc, args);
  params.AddParam("-par", "Solver parameter file", &parFile);
  params.AddParam("-mat", "Coefficient matrix A", &matFile);
  params.AddParam("-rhs", "Right-hand-side b", &rhsFile);
  params.AddParam("-xin", "Initial guess for iteration", &xinFile);
  SOLParams solParam;
  params.SetSOLParams(solParam);
  params.Parse();
  params.Print();

  MAT mat;
  FaspRetCode retCode = ReadMat(matFile.c_str(), mat);
  if (retCode < 0) return retCode;
  const USI nrow = mat.GetRowSize(), mcol = mat.GetColSize();
  std::cout << "nrow: " << nrow << ", mcol: " << mcol << std::endl;

  VEC b;
  b.SetValues(nrow, 0.0);
  if (!rhsFile.empty()) {
    retCode = ReadVEC(rhsFile.c_str(), b);
    if (retCode < 0) return retCode;
  }
  VEC x;
  x.SetValues(mcol, 1.0);
  if (!xinFile.empty()) {
    retCode = ReadVEC(xinFile.c_str(), x);
    if (retCode < 0) return retCode;
  }

  Identity pc;
  const auto t0 = std::chrono::steady_clock::now();
  retCode = Krylov(mat, b, x, pc, solParam);
  const double ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  std::cout << "Solving linear system costs " << ms << "ms" << std::endl;
  // the solver's results, read back through a second solve object would
  // repeat the work: report the true residual of x instead
  VEC r;
  r.SetValues(nrow, 0.0);
  mat.Residual(b, x, r);
  std::printf("result %d %.17g %.3f\n", (int)retCode, r.Norm2(), ms);
  return retCode;
}
