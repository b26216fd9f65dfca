// fx_solver.hpp -- solver object and the per-operation device backend.
#pragma once

#include <vector>

#include "fx_internal.hpp"
#include "fx_reuse.hpp"

namespace fx {

struct FastWs;  // fused-path workspace (fx_fast.hip)

}  // namespace fx

struct fx_op_s {
  fx_ctx ctx = nullptr;
  uint32_t nrow = 0, mcol = 0;
  fx_op_apply_fn apply = nullptr;
  void* user = nullptr;
  // apply only enqueues device work on the context's stream, the same work
  // on every call: the fused solvers capture it in their replay graphs
  bool device_only = false;
};

struct fx_solver_s {
  fx_ctx ctx = nullptr;
  int type = FX_SOLVER_CG;
  fx_csr A = nullptr;
  fx_op op = nullptr;  // set instead of A by fx_solver_setup_op
  fx_pc pc = nullptr;
  uint64_t A_gen = 0, pc_gen = 0;  // the handles' ids at setup
  fx_params prm{};
  std::vector<double> hist_host;
  uint32_t hist_count = 0;
  int last_path = FX_PATH_OPS;  // the form the last Solve ran
  // GMRES / FGMRES: the restart length the last Solve ended with (0: as set
  // up).  The reference keeps it in the solver: Setup resets it
  // (src/GMRES.cxx:44,75-79), Solve adapts it per cycle (:255-265) and the
  // next Solve starts from there.
  uint32_t restart_cur = 0;
  std::vector<fx_vec> Z;  // FGMRES preconditioned basis, kept across calls
  fx::FastWs* fast = nullptr;
  void record(uint32_t it, double v);
  void release_fast();
};

namespace fx {

struct OpsError {
  int code;
};

// Vector operations of the reference solvers over device vectors of one
// dtype.  Scalars are returned to the host (each reduction synchronises).
struct Ops {
  fx_ctx ctx;
  fx_csr A;
  fx_pc pc;
  int dt;
  uint32_t n;
  fx_op op = nullptr;  // the operator's action, when it is not a matrix
  std::vector<fx_vec> owned;

  Ops(fx_ctx c, fx_csr a, fx_pc p, int dtype);
  Ops(fx_ctx c, fx_op a, fx_pc p, int dtype);
  // the operator is op when it is set, else A
  Ops(fx_ctx c, fx_csr a, fx_op op, fx_pc p, int dtype);
  Ops(const Ops&) = delete;
  Ops& operator=(const Ops&) = delete;
  ~Ops();
  fx_vec vec();
  static void chk(int rc);
  void spmv(fx_vec x, fx_vec y);
  void residual(fx_vec b, fx_vec x, fx_vec r);
  void axpy(double a, fx_vec x, fx_vec y);
  void xpay(double a, fx_vec x, fx_vec y);
  void axpby(double a, double b, fx_vec x, fx_vec y);
  void waxpby(double a, fx_vec x, double b, fx_vec y, fx_vec w);
  void scale(double a, fx_vec y);
  void set(fx_vec y, double v);
  void copy(fx_vec s, fx_vec d);
  double dot(fx_vec x, fx_vec y);
  double norm2(fx_vec x);
  double norminf(fx_vec x);
  void precond(fx_vec r, fx_vec z);
  void true_residual(fx_vec b, fx_vec x, fx_vec r);
};

// Hessenberg matrix, Givens rotations and least-squares right-hand side of
// GMRES / FGMRES, kept in f64 on the host.
struct Givens {
  uint32_t ld = 0;
  std::vector<double> h, hc, hs, g;
  void init(uint32_t maxR);
  double& H(uint32_t i, uint32_t j) { return h[(size_t)i * ld + j]; }
  void column(uint32_t cnt);
  void solve(uint32_t cnt);
};

void clamp_restart(const fx_params& P, uint32_t& maxR, uint32_t& minR, uint32_t& restart);
// The restart bounds and the length a Solve starts with (the one the last
// Solve on S ended with, if any since setup).
void solve_restart(fx_solver S, uint32_t& maxR, uint32_t& minR, uint32_t& restart);
void adapt_restart(double cr, bool inclusive, uint32_t maxR, uint32_t minR, uint32_t& restart);

int ops_cg(fx_solver S, Ops& o, fx_vec b, fx_vec x, fx_result* res);
int ops_bicgstab(fx_solver S, Ops& o, fx_vec b, fx_vec x, fx_result* res);
int ops_gmres(fx_solver S, Ops& o, fx_vec b, fx_vec x, fx_result* res);
int ops_fgmres(fx_solver S, Ops& o, fx_vec b, fx_vec x, fx_result* res);
int ops_jacobi(fx_solver S, Ops& o, fx_vec b, fx_vec x, fx_result* res);

// Fused, graph-replayed solvers (fx_fast.hip).
bool fast_eligible(fx_solver S);
int fast_solve(fx_solver S, fx_vec b, fx_vec x, fx_result* res);
// Force the per-operation path (FASPXX_FX_NO_FAST=1 in the environment).
bool fast_disabled();

}  // namespace fx
