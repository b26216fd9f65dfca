// mg_tests.cpp -- drives the drop-in MG<MAT> the way example/TestMG.cxx:66-99
// does, on the device, and prints what it computed for tests/test_gpu_mg.py:
//   mg_tests <A file> <levels> <ncycles> <smoother: default|one> <maxIter>
//            <minIter> <relTol> <absTol> [<transfer .mtx> [<b file> [<x0 file>]]]
// Every level gets its own Jacobi smoother and Jacobi coarse solver set up on
// A (default SOLParams, or "one": one sweep, tolerances 0) and the same
// transfer operator (IdentityMatrix unless a file is given).  b defaults to
// 0 and x0 to 1, as in TestMG.  Output: "retcode numIter norm2 normInf",
// then x, one value per line.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <vector>

#include "Iter.hxx"
#include "MG.hxx"
#include "ReadData.hxx"

int main(int argc, char** argv) {
  if (argc < 9) return 100;
  MAT mat;
  if (ReadMat(argv[1], mat) < 0) return 101;
  const USI levels = (USI)std::atoi(argv[2]);
  const USI ncycles = (USI)std::atoi(argv[3]);
  const bool one = std::strcmp(argv[4], "one") == 0;
  const USI n = mat.GetRowSize();

  std::unique_ptr<MAT> transfer;
  if (argc > 9 && std::strcmp(argv[9], "-") != 0) {
    transfer.reset(new MAT());
    if (ReadMat(argv[9], *transfer) < 0) return 102;
  } else {
    transfer.reset(new IdentityMatrix(n));
  }
  VEC b, x;
  b.SetValues(n, 0.0);
  x.SetValues(n, 1.0);
  if (argc > 10 && ReadVEC(argv[10], b) < 0) return 103;
  if (argc > 11 && ReadVEC(argv[11], x) < 0) return 104;

  MG<MAT> solver;
  solver.SetMaxIter((USI)std::atoi(argv[5]));
  solver.SetMinIter((USI)std::atoi(argv[6]));
  solver.SetRelTol(std::atof(argv[7]));
  solver.SetAbsTol(std::atof(argv[8]));

  std::vector<Jacobi> smoothers(levels), coarse(levels);
  solver.SetupALL(mat, levels);
  for (USI i = 0; i < levels; ++i) {
    for (Jacobi* j : {&smoothers[i], &coarse[i]}) {
      j->Setup(mat);
      if (one) {
        j->SetMaxIter(1);
        j->SetMinIter(0);
        j->SetRelTol(0.0);
        j->SetAbsTol(0.0);
      }
    }
    solver.SetupLevel(mat, i, transfer.get(), &smoothers[i], &coarse[i]);
  }
  solver.SetNumCycles(ncycles);
  const FaspRetCode rc = solver.Solve(b, x);
  std::printf("%d %u %.17g %.17g\n", (int)rc, solver.GetIterations(), solver.GetNorm2(),
              solver.GetInfNorm());
  for (USI i = 0; i < n; ++i) std::printf("%.17g\n", x[i]);
  return 0;
}
