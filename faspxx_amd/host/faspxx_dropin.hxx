// faspxx_dropin.hxx -- C++ classes with FASP++'s MAT / VEC / LOP / SOL
// interface (names, signatures, argument meaning, return codes), backed by
// the gfx950 C ABI (include/faspxx_hip.h).  A program written against the
// reference's headers compiles against these and runs its vectors, SpMV and
// solvers on the GPU.
//
// Interface sources: include/Faspxx.hxx:23-46 (types, constants),
// src/RetCode.hxx:24-55 (FaspRetCode), src/Param.hxx:27-80 (Output, SOLType,
// SOLParams), src/VEC.hxx:25-146, src/LOP.hxx:22-66, src/MAT.hxx:39-213,
// src/SOL.hxx:24-126, src/Iter.hxx:26-71, src/CG.hxx:24-55,
// src/BiCGStab.hxx:23-68, src/GMRES.hxx:26-93, src/FGMRES.hxx:26-80,
// src/Krylov.hxx:25.
//
// Storage: a VEC's values live in device memory; a host mirror is created on
// demand for element access (operator[], GetArray, GetValue) and written back
// before the next device operation.  All objects use one process-wide device
// context (faspxx_amd::SetDevice selects the GPU before first use).
#pragma once

#include <cstdint>
#include <iostream>
#include <map>
#include <string>
#include <vector>

#include "faspxx_hip.h"

typedef int INT;
typedef unsigned int USI;
typedef double DBL;

const DBL SMALL_TOL = 1e-14;
const DBL LARGE_DBL = 1e+60;
const DBL SMALL_DBL = -1e+60;
const DBL CLOSE_ZERO = 1e-20;
const USI MAX_ITER_NUM = 100000;
const DBL KSM_CHK_RATIO = 0.95;
const USI MAX_STAG_NUM = 20;
const USI PRT_STEP_NUM = 20;

enum FaspRetCode {
  SUCCESS = 0,
  ERROR_OPEN_FILE = -10,
  ERROR_INPUT_FILE = -11,
  ERROR_INPUT_PAR = -12,
  ERROR_VEC_SIZE = -14,
  ERROR_MAT_SIZE = -15,
  ERROR_NONMATCH_SIZE = -16,
  ERROR_MAT_DATA = -17,
  ERROR_DIVIDE_ZERO = -18,
  ERROR_MAT_ZERODIAG = -19,
  ERROR_ALLOC_MEM = -20,
  ERROR_DUMMY_VAR = -23,
  ERROR_SOLVER_TYPE = -30,
  ERROR_SOLVER_PCD_TYPE = -31,
  ERROR_SOLVER_STAG = -32,
  ERROR_SOLVER_SOLSTAG = -33,
  ERROR_SOLVER_TOLSMALL = -34,
  ERROR_SOLVER_MAXIT = -39,
  ERROR_AMG_INTERP_TYPE = -40,
  ERROR_AMG_SMOOTH_TYPE = -41,
  ERROR_AMG_COARSE_TYPE = -42,
  ERROR_AMG_COARSEING = -43,
  ERROR_AMG_SETUP = -49,
  ERROR_ILU_TYPE = -50,
  ERROR_ILU_SETUP = -59,
  ERROR_SWZ_TYPE = -60,
  ERROR_SWZ_SETUP = -69,
  ERROR_DEVICE = -90,  // HIP runtime failure (no reference counterpart)
  ERROR_DSOLVER_SETUP = -91,
  ERROR_DSOLVER_SOLVE = -92,
  PRINT_HELP = 0
};

enum Output { PRINT_NONE = 0, PRINT_MIN = 2, PRINT_SOME = 4, PRINT_MORE = 6, PRINT_MAX = 8 };

enum SOLType {
  SOLVER_CG = 1,
  SOLVER_BICGSTAB = 2,
  SOLVER_MINRES = 3,
  SOLVER_GMRES = 4,
  SOLVER_FGMRES = 5,
  SOLVER_VFGMRES = 6,
  SOLVER_JACOBI = 11,
  SOLVER_GS = 12,
  SOLVER_SGS = 13,
  SOLVER_SOR = 14,
  SOLVER_SSOR = 15,
  SOLVER_MG = 21,
  SOLVER_FMG = 22,
  SOLVER_UMFPACK = 91,
  SOLVER_MUMPS = 92,
  SOLVER_SUPERLU = 93,
  SOLVER_PARDISO = 94
};

struct SOLParams {
  SOLType type = SOLVER_CG;
  std::string algName = "cg";
  USI maxIter = 1;
  USI minIter = 0;
  USI savIter = MAX_ITER_NUM;
  USI restart = 30;
  double relTol = 1e-6;
  double absTol = 1e-8;
  Output verbose = PRINT_NONE;
};

// Parameters: program defaults < parameter file < command line
// (src/Param.hxx:82-226; defined in faspxx_param.cpp).  marker: 0 optional,
// 1 required (only shown in the help), 2 names the parameter file.
class Parameters {
 public:
  Parameters(int argc, const char* argv[]) : argc(argc), argv(argv) {}
  void AddParam(const std::string& name, const std::string& help, bool* ptr, int marker = 0);
  void AddParam(const std::string& name, const std::string& help, USI* ptr, int marker = 0);
  void AddParam(const std::string& name, const std::string& help, INT* ptr, int marker = 0);
  void AddParam(const std::string& name, const std::string& help, double* ptr, int marker = 0);
  void AddParam(const std::string& name, const std::string& help, std::string* ptr,
                int marker = 0);
  void AddParam(const std::string& name, const std::string& help, Output* ptr, int marker = 0);
  void Parse();
  void SetSOLParams(SOLParams& solParam);
  void SetSOLParams(SOLParams& solParam, SOLParams& pcdParam);
  void PrintUserParams(std::ostream& out) const;
  void PrintFileParams(std::ostream& out) const;
  void PrintCommandLineParams(std::ostream& out) const;
  void Print(std::ostream& out = std::cout) const;
  void PrintHelp(std::ostream& out = std::cout) const;

 private:
  enum Kind { KBool = 0, KInt = 1, KDouble = 2, KString = 3, KOutput = 4 };
  struct Entry {
    Kind kind;
    std::string name, help;
    void* ptr;
    int marker;
  };
  int argc;
  const char** argv;
  std::vector<Entry> entries;
  std::map<std::string, std::string> fromCommandLine, fromFile;
  std::string original;
  void Add(Kind kind, const std::string& name, const std::string& help, void* ptr, int marker);
  void ReadCommandLine();
  void ReadFile();
  static std::string ValueText(const Entry& e);
  static void Assign(Entry& e, const std::string& text);
};

namespace faspxx_amd {
// Select the GPU used by every object created afterwards (default 0).
void SetDevice(int device);
// The process-wide context (created on first use).
fx_ctx Context();
}  // namespace faspxx_amd

class MAT;

class VEC {
 public:
  friend class MAT;

  explicit VEC() : size(0) {}
  explicit VEC(const USI& size, const DBL& value = 0.0);
  explicit VEC(const std::vector<DBL>& src);
  explicit VEC(const USI& size, const DBL* src);
  VEC(const VEC& src);
  ~VEC();
  VEC& operator=(const VEC& v);

  DBL& operator[](const USI& position);
  const DBL& operator[](const USI& position) const;
  VEC& operator+=(const VEC& v);
  VEC& operator-=(const VEC& v);

  void Reserve(const USI& size);
  void SetValues(const USI& size, const DBL& value = 0.0);
  void SetValues(const std::vector<DBL>& src);
  void SetValues(const USI& size, const DBL* array);
  DBL GetValue(const USI& position) const;
  void GetValues(const USI& size, const USI* index, DBL* array) const;
  void GetArray(DBL** array);
  void GetArray(const DBL** array) const;
  USI GetSize() const { return size; }

  void Scale(const DBL& a);
  void Reciprocal();
  void PointwiseMult(const VEC& v);
  void PointwiseDivide(const VEC& v);
  void CopyTo(VEC& dst) const;
  void Shift(const DBL& a);
  void Abs();
  void AXPY(const DBL& a, const VEC& x);
  void XPAY(const DBL& a, const VEC& x);
  void AXPBY(const DBL& a, const DBL& b, const VEC& y);
  void WAXPBY(const DBL& a, const VEC& x, const DBL& b, const VEC& y);
  DBL Max() const;
  DBL Min() const;
  DBL Norm2() const;
  DBL NormInf() const;
  DBL Dot(const VEC& v) const;

  // Device access (no reference counterpart): the device copy, current.
  fx_vec Device() const;
  // The device copy was written through Device(): drop the host mirror.
  void DeviceWritten();
  // The device copy, current, about to be overwritten (drops the mirror).
  fx_vec DeviceOut();
  // A view of a device vector owned elsewhere (no reference counterpart):
  // the solvers hand these to a LOP's Apply.  Its size cannot change.
  VEC(fx_vec view, bool borrowed);

 private:
  USI size;
  bool borrowed = false;
  mutable fx_vec dev = nullptr;
  mutable std::vector<DBL> host;
  mutable bool host_valid = false;    // mirror holds the current values
  mutable bool device_stale = false;  // mirror was modified after the last upload
  void Resize(USI n);
  void SyncToHost() const;
  void SyncToDevice() const;
};

class LOP {
 protected:
  USI nrow;
  USI mcol;

 public:
  explicit LOP() : nrow(0), mcol(0) {}
  explicit LOP(const USI& nrow, const USI& mcol) : nrow(nrow), mcol(mcol) {}
  explicit LOP(const USI& nrow) : nrow(nrow), mcol(nrow) {}
  LOP(const LOP& lop) = default;
  LOP& operator=(const LOP& lop) = default;
  virtual ~LOP() = default;
  USI GetRowSize() const { return nrow; }
  USI GetColSize() const { return mcol; }
  virtual void Apply(const VEC& x, VEC& y) const;
  virtual void Residual(const VEC& b, const VEC& x, VEC& r) const;
};

class MAT : public LOP {
 public:
  MAT() = default;
  explicit MAT(const USI& nrow, const USI& mcol, const USI& nnz, const std::vector<DBL>& values,
               const std::vector<USI>& colInd, const std::vector<USI>& rowPtr,
               const std::vector<USI>& diagPtr);
  explicit MAT(const USI& nrow, const USI& mcol, const USI& nnz, const std::vector<DBL>& values,
               const std::vector<USI>& colInd, const std::vector<USI>& rowPtr);
  explicit MAT(const USI& nrow, const USI& mcol, const USI& nnz, const std::vector<USI>& colInd,
               const std::vector<USI>& rowPtr);
  explicit MAT(const USI& nrow, const USI& mcol, const USI& nnz, const std::vector<USI>& colInd,
               const std::vector<USI>& rowPtr, const std::vector<USI>& diagPtr);
  explicit MAT(const VEC& v);
  explicit MAT(const std::vector<DBL>& v);
  MAT(const MAT& mat);
  ~MAT() override;
  MAT& operator=(const MAT& mat);

  void SetValues(const USI& nrow, const USI& mcol, const USI& nnz, const std::vector<DBL>& values,
                 const std::vector<USI>& colInd, const std::vector<USI>& rowPtr,
                 const std::vector<USI>& diagPtr);
  void SetValues(const USI& nrow, const USI& mcol, const USI& nnz, const std::vector<DBL>& values,
                 const std::vector<USI>& colInd, const std::vector<USI>& rowPtr);
  USI GetNNZ() const { return nnz; }
  void GetDiag(VEC& v) const;
  void CopyTo(MAT& mat) const { mat = *this; }
  void Scale(const DBL a);
  void Apply(const VEC& v, VEC& w) const override;
  void Residual(const VEC& b, const VEC& x, VEC& r) const override;
  void TransInPlace();
  DBL GetValue(const USI& row, const USI& col) const;

  // Setup-time algebra (src/MAT.cxx:270-372, 621-800, 917-935), on the device.
  void GetDiagInv(MAT& m) const;
  void GetLowerTri(MAT& lTri) const;
  void GetUpperTri(MAT& uTri) const;
  void Shift(const DBL a);
  void Zero();
  // new[]-allocated host copies, as the reference returns them
  double* GetValues() const;
  int* GetRowPtr() const;
  int* GetColInd() const;
  // *this = a mat1 + b mat2; left unchanged when either has no entries, as
  // the reference leaves it (src/MAT.cxx:659-669)
  void Add(const DBL a, const MAT& mat1, const DBL b, const MAT& mat2);
  void Mult(const MAT& matl, const MAT& matr);
  void MultLeft(const MAT& mat);   // *this = *this mat
  void MultRight(const MAT& mat);  // *this = mat *this
  void FormDiagPtr() {}            // formed on the device whenever the structure changes
  void Empty();

  // Device handle (no reference counterpart).
  fx_csr Device() const { return csr; }

 private:
  USI nnz = 0;
  bool has_values = false;
  fx_csr csr = nullptr;
  void Build(USI nrow, USI mcol, USI nnz, const DBL* values, const USI* colInd,
             const USI* rowPtr, const USI* diagPtr);
  void Release();
  void Adopt(fx_csr h, bool values);  // take over a handle made by the C ABI
};

// IdentityMatrix (src/MAT.hxx:215-232): a MAT of the given size with no
// entries whose Apply copies (src/MAT.cxx:1088-1095).
class IdentityMatrix : public MAT {
 public:
  IdentityMatrix() = default;
  explicit IdentityMatrix(USI size);
  void Apply(const VEC& b, VEC& x) const override { x = b; }
};

class SOL {
 protected:
  const LOP* A;
  SOL* pcd;
  double norm2;
  double normInf;
  USI numIter;
  SOLParams params;

 public:
  SOL() : A(nullptr), pcd(nullptr), norm2(LARGE_DBL), normInf(LARGE_DBL), numIter(0) {}
  virtual ~SOL();
  void SetOutput(Output verbose) { params.verbose = verbose; }
  void SetMaxIter(USI maxIter) { params.maxIter = maxIter; }
  void SetMinIter(USI minIter) { params.minIter = minIter; }
  void SetSavIter(USI savIter) { params.savIter = savIter; }
  void SetRestart(USI restart) { params.restart = restart; }
  void SetRelTol(double relTol) { params.relTol = relTol; }
  void SetAbsTol(double absTol) { params.absTol = absTol; }
  void SetSolType(SOLType type) { params.type = type; }
  static void SetSolTypeFromName(SOLParams& params);
  const char* GetSolType(SOLType type) const;
  double GetNorm2() const { return norm2; }
  double GetInfNorm() const { return normInf; }
  USI GetIterations() const { return numIter; }
  virtual void SetupPCD(SOL& pc) { pcd = &pc; }
  virtual FaspRetCode Setup(const LOP& A);
  virtual FaspRetCode Solve(const VEC& b, VEC& x);
  virtual void Clean() {}

  // Device preconditioner equivalent to this solver used as a preconditioner
  // (no reference counterpart): Identity -> identity, Jacobi -> its sweeps.
  virtual fx_pc DevicePC(const MAT& A);

 protected:
  fx_pc pc_cache = nullptr;
  fx_params DeviceParams() const;
  void Record(const fx_result& r);
};

class Identity : public SOL {
 public:
  FaspRetCode Solve(const VEC& b, VEC& x) override;
  fx_pc DevicePC(const MAT& A) override;
};

class Jacobi : public SOL {
 public:
  Jacobi() : weight(1.0) {}
  Jacobi(DBL weight) : weight(weight) {}
  ~Jacobi() override;
  void SetWeight(const DBL w) { weight = w; }
  FaspRetCode Setup(const MAT& A);
  FaspRetCode Solve(const VEC& b, VEC& x) override;
  void Clean() override {}
  fx_pc DevicePC(const MAT& A) override;

 private:
  double weight;
  const MAT* mat = nullptr;
  fx_pc own = nullptr;
  double own_weight = 0.0;
  uint32_t own_iters = 0;
  fx_solver solver = nullptr;
};

// Device-backed Krylov solver base (CG, BiCGStab, GMRES, FGMRES).
class KrylovSOL : public SOL {
 public:
  explicit KrylovSOL(SOLType t) : kind(t) {}
  ~KrylovSOL() override;
  FaspRetCode Setup(const LOP& A) override;
  FaspRetCode Solve(const VEC& b, VEC& x) override;
  void Clean() override;

 protected:
  SOLType kind;
  bool setup_pending = true;  // Setup ran since the last Solve
  const MAT* mat = nullptr;
  const LOP* lop = nullptr;  // an operator that is not a MAT: its Apply is called back
  fx_op op = nullptr;
  fx_solver solver = nullptr;
  uint32_t maxRestart = 30, minRestart = 10;
  bool rightPrecond = true;
};

class CG : public KrylovSOL {
 public:
  CG() : KrylovSOL(SOLVER_CG) {}
};

class BiCGStab : public KrylovSOL {
 public:
  BiCGStab() : KrylovSOL(SOLVER_BICGSTAB) {}
};

class GMRES : public KrylovSOL {
 public:
  GMRES() : KrylovSOL(SOLVER_GMRES) {}
  void SetRightPrecond() { rightPrecond = true; }
  void SetLeftPrecond() { rightPrecond = false; }
  void SetMaxMinRestart(const USI maxR, const USI minR) {
    maxRestart = maxR;
    minRestart = minR;
  }
};

class FGMRES : public KrylovSOL {
 public:
  FGMRES() : KrylovSOL(SOLVER_FGMRES) {}
  void SetMaxMinRestart(const USI maxR, const USI minR) {
    maxRestart = maxR;
    minRestart = minR;
  }
};

FaspRetCode Krylov(LOP& A, VEC& b, VEC& x, SOL& pcd, SOLParams& params);
