// readdata_tests.cpp -- prints what the drop-in's file readers produce, for
// tests/test_readdata.py to compare (no device needed):
//   readdata_tests csrx <file>   the normalised CSRx arrays ReadMat uploads
//   readdata_tests raw <file>    ReadCSR's arrays before normalisation
//   readdata_tests vec <file>    ReadVEC is not used (it builds a device VEC)
// Output: "retcode row col nnz", then rowPtr, colInd, values, one per line.
#include <cstdio>
#include <cstring>

#include "ReadData.hxx"

int main(int argc, char** argv) {
  if (argc != 3) return 2;
  USI row = 0, col = 0, nnz = 0;
  std::vector<USI> rp, ci;
  std::vector<DBL> va;
  FaspRetCode rc;
  if (std::strcmp(argv[1], "csrx") == 0)
    rc = ReadCSRx(argv[2], row, col, nnz, rp, ci, va);
  else if (std::strcmp(argv[1], "raw") == 0)
    rc = ReadCSR(argv[2], row, col, nnz, rp, ci, va);
  else
    return 2;
  std::printf("%d %u %u %u\n", (int)rc, row, col, nnz);
  for (USI v : rp) std::printf("%u\n", v);
  for (USI v : ci) std::printf("%u\n", v);
  for (DBL v : va) std::printf("%.17g\n", v);
  return 0;
}
