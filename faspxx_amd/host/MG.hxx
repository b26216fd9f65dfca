// MG.hxx -- the reference's abstract multigrid solver MG<TTT> (src/MG.hxx:24-109,
// src/MG.cxx:16-229) over the drop-in classes: every vector operation,
// residual, transfer-operator Apply and smoother Solve of the cycle runs on
// the device through the objects the caller passes in.
//
// The cycle is the reference's:
//   pre-smooth; r = b - A x with the fine operator A and the solver's one
//   work vector r; b_l = T r (the level's transfer operator through Apply);
//   x_l = 0; on the last level the coarse solver, else numCycles[l] cycles
//   one level down; r = T x_l; x += r; post-smooth.
// Restriction and prolongation are the same operator applied the same way,
// and every level's residual uses the fine A, so transfer operators must be
// square (an IdentityMatrix, as example/TestMG.cxx:85-86 sets them up).
// SetupLevel rejects level >= numLevelsCoarse (the reference writes past the
// end of infoHL for level == numLevelsCoarse).
#pragma once

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "faspxx_dropin.hxx"

template <class TTT>
struct HL {
  TTT* restriction = nullptr;
  TTT* prolongation = nullptr;
  SOL* preSolver = nullptr;
  SOL* coarseSolver = nullptr;
  SOL* postSolver = nullptr;
  USI fineSpaceSize = 0;
  USI coarSpaceSize = 0;
  VEC b;
  VEC x;
  VEC r;
};

template <class TTT>
class MG : public SOL {
 private:
  INT level = -1;
  USI numLevelsMax = 20;
  USI numLevelsCoarse = 0;
  bool useSymmOper = true;
  std::vector<USI> numCycles;
  VEC r;

 public:
  std::vector<HL<TTT>> infoHL;

  MG() { SetNumCycles(1); }
  ~MG() override = default;

  // src/MG.cxx:18-23: one count per coarse level in use
  void SetNumCycles(USI ncycle) { numCycles.assign(numLevelsCoarse, ncycle); }

  // src/MG.cxx:87-106 (the reference's Setup only records A)
  FaspRetCode Setup(const TTT& A) {
    SetSolType(SOLVER_MG);
    this->A = &A;
    return SUCCESS;
  }

  // src/MG.cxx:26-49
  FaspRetCode SetupLevel(const TTT& A, const USI lev, TTT* tranOpers, SOL* smoothers,
                         SOL* coarseSolvers) {
    (void)A;
    if (lev >= numLevelsCoarse) {
      std::fprintf(stderr, "### ABORT: Too many levels specified!\n");
      std::abort();
    }
    HL<TTT>& L = infoHL[lev];
    L.fineSpaceSize = tranOpers->GetColSize();
    L.coarSpaceSize = tranOpers->GetRowSize();
    L.b.SetValues(L.coarSpaceSize, 0.0);
    L.x.SetValues(L.coarSpaceSize, 0.0);
    L.r.SetValues(L.coarSpaceSize, 0.0);
    L.prolongation = tranOpers;
    L.restriction = tranOpers;
    L.preSolver = smoothers;
    L.postSolver = smoothers;
    L.coarseSolver = coarseSolvers;
    return SUCCESS;
  }

  // src/MG.cxx:52-84
  FaspRetCode SetupALL(const TTT& A, const USI numLevels) {
    useSymmOper = true;
    numLevelsCoarse = numLevels;
    SetSolType(SOLVER_MG);
    infoHL.resize(numLevelsCoarse);
    r.SetValues(A.GetColSize(), 0.0);
    this->A = &A;
    return SUCCESS;
  }

  // src/MG.cxx:175-219; always returns SUCCESS, as the reference does
  FaspRetCode Solve(const VEC& b, VEC& x) override {
    double resAbs = 1.0, denAbs = 1.0;
    numIter = 0;
    while (numIter < params.maxIter) {
      this->A->Residual(b, x, r);
      if (numIter >= params.minIter) {
        resAbs = r.Norm2();
        if (numIter == params.minIter) denAbs = (CLOSE_ZERO > resAbs) ? CLOSE_ZERO : resAbs;
        if (resAbs / denAbs < params.relTol || resAbs < params.absTol) break;
      }
      Cycle(b, x);
      ++numIter;
    }
    if (!(numIter == params.minIter && numIter == params.maxIter)) {
      this->A->Residual(b, x, r);
      fx_result res{};
      res.numIter = numIter;
      res.norm2 = r.Norm2();
      res.normInf = r.NormInf();
      Record(res);
    }
    return SUCCESS;
  }

  void Clean() override {}

 private:
  // MG::MGCycle, src/MG.cxx:117-172
  void Cycle(const VEC& b, VEC& x) {
    if (numLevelsCoarse == 0) {
      std::printf("### WARNING: No multilevel info available! Do nothing!!!\n");
      x = b;
      return;
    }
    ++level;
    if (level == (INT)numLevelsCoarse) return;
    HL<TTT>& L = infoHL[level];
    L.preSolver->Solve(b, x);
    this->A->Residual(b, x, r);
    L.restriction->Apply(r, L.b);
    L.x.SetValues(L.coarSpaceSize, 0.0);
    if (level + 1 == (INT)numLevelsCoarse) {
      L.coarseSolver->Solve(L.b, L.x);
    } else {
      for (USI i = 0; i < numCycles[level]; ++i) Cycle(L.b, L.x);
    }
    L.prolongation->Apply(L.x, r);
    x.AXPY(1.0, r);
    L.postSolver->Solve(b, x);
    --level;
  }
};
