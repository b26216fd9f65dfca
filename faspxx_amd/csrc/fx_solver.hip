// fx_solver.hip -- preconditioners, the solver object, the Krylov() front end
// and the host-driven ("per-operation") solvers.
//
// The per-operation solvers state the reference algorithms step by step
// (src/CG.cxx, src/BiCGStab.cxx, src/GMRES.cxx, src/FGMRES.cxx, src/Iter.cxx)
// over device vectors: every vector operation is a device kernel and every
// scalar the reference branches on is read back to the host.  They serve any
// preconditioner and any parameter set.  The fused, graph-replayed solvers in
// fx_fast.hip take over for the common configurations (fx_solver_solve picks).
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstring>
#include <string>

#include "fx_solver.hpp"

namespace fx {

namespace {

__global__ void k_jacobi_dinv(const double* __restrict__ va, const uint32_t* __restrict__ dp,
                              double w, double* __restrict__ dinv, uint32_t n) {
  const uint32_t j = blockIdx.x * kThreads + threadIdx.x;
  if (j >= n) return;
  // GetDiag, Reciprocal, Scale(weight): src/Iter.cxx:45-47
  const double d = va ? va[dp[j]] : 1.0;
  const double inv = 1.0 / d;
  dinv[j] = inv * w;
}

__global__ void k_d2f(const double* __restrict__ s, float* __restrict__ d, uint32_t n) {
  const uint32_t j = blockIdx.x * kThreads + threadIdx.x;
  if (j < n) d[j] = (float)s[j];
}

}  // namespace

int ensure_dinv32(fx_pc pc) {
  if (pc->kind != 1 || pc->dinv32) return 0;
  const uint32_t n = pc->A->nrow;
  // padded to whole 16-byte packs (the streaming kernels load packs)
  FX_HIP(hipMalloc(&pc->dinv32, sizeof(float) * pack_round(n)));
  if (n)
    hipLaunchKernelGGL(k_d2f, dim3((n + kThreads - 1) / kThreads), dim3(kThreads), 0,
                       pc->ctx->stream, pc->dinv, pc->dinv32, n);
  FX_HIP(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------------
// Per-operation device backend
// ---------------------------------------------------------------------------

Ops::Ops(fx_ctx c, fx_csr a, fx_pc p, int dtype) : ctx(c), A(a), pc(p), dt(dtype), n(a->nrow) {}

Ops::Ops(fx_ctx c, fx_op a, fx_pc p, int dtype)
    : ctx(c), A(nullptr), pc(p), dt(dtype), n(a->nrow), op(a) {}

Ops::Ops(fx_ctx c, fx_csr a, fx_op o, fx_pc p, int dtype)
    : ctx(c), A(o ? nullptr : a), pc(p), dt(dtype), n(o ? o->nrow : a->nrow), op(o) {}

Ops::~Ops() {
  for (fx_vec v : owned) fx_vec_destroy(v);
}

fx_vec Ops::vec() {
  fx_vec v = nullptr;
  chk(fx_vec_create(ctx, n, dt, &v));
  owned.push_back(v);
  return v;
}

void Ops::chk(int rc) {
  if (rc != 0) throw OpsError{rc};
}

void Ops::spmv(fx_vec x, fx_vec y) {
  if (!op) {
    chk(fx_spmv(ctx, A, x, y));
    return;
  }
  const int rc = op->apply(op->user, x, y);
  if (rc) {
    if (!fx_last_error()[0]) set_error("operator apply returned %d", rc);
    throw OpsError{rc < 0 ? rc : FX_ERROR_INPUT_PAR};
  }
}
void Ops::residual(fx_vec b, fx_vec x, fx_vec r) {
  if (op) {  // LOP::Residual aborts unless a subclass overrides it (src/LOP.hxx:62-65)
    set_error("the operator has no Residual");
    throw OpsError{FX_ERROR_INPUT_PAR};
  }
  chk(fx_residual(ctx, A, b, x, r));
}
void Ops::axpy(double a, fx_vec x, fx_vec y) { chk(fx_axpy(ctx, a, x, y)); }
void Ops::xpay(double a, fx_vec x, fx_vec y) { chk(fx_xpay(ctx, a, x, y)); }
void Ops::axpby(double a, double b, fx_vec x, fx_vec y) { chk(fx_axpby(ctx, a, b, x, y)); }
void Ops::waxpby(double a, fx_vec x, double b, fx_vec y, fx_vec w) {
  chk(fx_waxpby(ctx, a, x, b, y, w));
}
void Ops::scale(double a, fx_vec y) { chk(fx_scale(ctx, a, y)); }
void Ops::set(fx_vec y, double v) { chk(fx_vec_set(y, v)); }
void Ops::copy(fx_vec s, fx_vec d) { chk(fx_vec_copy(s, d)); }
double Ops::dot(fx_vec x, fx_vec y) {
  double r = 0;
  chk(fx_dot(ctx, x, y, &r));
  return r;
}
double Ops::norm2(fx_vec x) {
  double r = 0;
  chk(fx_norm2(ctx, x, &r));
  return r;
}
double Ops::norminf(fx_vec x) {
  double r = 0;
  chk(fx_norminf(ctx, x, &r));
  return r;
}
void Ops::precond(fx_vec r, fx_vec z) { chk(fx_pc_apply(ctx, pc, r, z)); }
// b - A x as the solvers form it: Apply, then XPAY(-1, b) (e.g. src/CG.cxx:70-71)
void Ops::true_residual(fx_vec b, fx_vec x, fx_vec r) {
  spmv(x, r);
  xpay(-1.0, b, r);
}

// ---------------------------------------------------------------------------
// CG::Solve, src/CG.cxx:53-249 (false-convergence restart zeroes p at full
// length; the reference's pk.SetValues(0.0) at :213 resizes p to 0)
// ---------------------------------------------------------------------------
int ops_cg(fx_solver S, Ops& o, fx_vec b, fx_vec x, fx_result* res) {
  const fx_params& P = S->prm;
  int err = FX_SUCCESS;
  const double solStagTol = 1e-4 * P.relTol;
  uint32_t stagStep = 0, moreStep = 0, numIter = 0;
  double resAbs = 1.0, resRel = 1.0, denAbs = 1.0, ratio = 0.0, resAbsOld = 1.0;
  double alpha = 0.0, beta, tmpa, tmpb;
  fx_vec r = o.vec(), p = o.vec(), z = o.vec(), q = o.vec(), safe = o.vec();

  o.true_residual(b, x, r);
  o.set(z, 0.0);
  o.precond(r, z);
  o.copy(z, p);
  tmpa = o.dot(z, r);

  while (numIter < P.maxIter) {
    if (numIter == P.minIter) {
      resAbs = o.norm2(r);
      resAbsOld = resAbs;
      denAbs = (kCloseZero > resAbs) ? kCloseZero : resAbs;
      resRel = resAbs / denAbs;
      if (resRel < P.relTol || resAbs < P.absTol) break;
    }
    if (numIter >= P.minIter) S->record(numIter, resAbs);

    ++numIter;
    o.spmv(p, q);
    tmpb = o.dot(q, p);
    if (std::fabs(tmpb) > kCloseZero * kCloseZero) {
      alpha = tmpa / tmpb;
    } else {
      o.axpy(alpha, p, x);
      o.axpy(-alpha, q, r);
      resAbs = o.norm2(r);
      resRel = resAbs / denAbs;
      if (resRel > P.relTol && resAbs > P.absTol) err = FX_ERROR_DIVIDE_ZERO;
      break;
    }
    o.axpy(alpha, p, x);
    o.axpy(-alpha, q, r);

    if (numIter >= P.minIter) {
      resAbs = o.norm2(r);
      resRel = resAbs / denAbs;
      ratio = resAbs / resAbsOld;
      if (numIter >= P.savIter && resAbs < resAbsOld) o.copy(x, safe);
      if (ratio > kChkRatio) {
        if (o.norminf(x) < kCloseZero) {  // Check I
          err = FX_ERROR_SOLVER_SOLSTAG;
          break;
        }
        const double xRelDiff = std::fabs(alpha) * o.norm2(p) / o.norm2(x);  // Check II
        if (stagStep <= kMaxStag && xRelDiff < solStagTol) {
          o.true_residual(b, x, r);
          resAbs = o.norm2(r);
          resRel = resAbs / denAbs;
          if (resRel < P.relTol || resAbs < P.absTol) break;
          if (stagStep >= kMaxStag) {
            err = FX_ERROR_SOLVER_STAG;
            break;
          }
          o.set(p, 0.0);
          ++stagStep;
        }
      }
      if (resRel < P.relTol) {  // Check III
        o.true_residual(b, x, r);
        resAbs = o.norm2(r);
        resRel = resAbs / denAbs;
        if (resRel < P.relTol || resAbs < P.absTol) break;
        if (moreStep >= P.restart) {
          err = FX_ERROR_SOLVER_TOLSMALL;
          break;
        }
        o.set(p, 0.0);
        ++moreStep;
      }
    }
    if (numIter < P.maxIter) {
      resAbsOld = resAbs;
      o.set(z, 0.0);
      o.precond(r, z);
      tmpb = o.dot(z, r);
      beta = tmpb / tmpa;
      tmpa = tmpb;
      o.xpay(beta, z, p);
    }
  }
  res->numIter = numIter;
  if (!(numIter == P.minIter && numIter == P.maxIter)) {
    res->norm2 = resAbs;
    res->normInf = o.norminf(r);
  }
  if (numIter > P.savIter) o.copy(safe, x);
  return err;
}

// ---------------------------------------------------------------------------
// BiCGStab::Solve, src/BiCGStab.cxx:68-270 (residual updated with P A p and
// P A s, solution with P p and P s, :114-146)
// ---------------------------------------------------------------------------
int ops_bicgstab(fx_solver S, Ops& o, fx_vec b, fx_vec x, fx_result* res) {
  const fx_params& P = S->prm;
  int err = FX_SUCCESS;
  const double solStagTol = 1e-4 * P.relTol;
  uint32_t stagStep = 0, moreStep = 0, numIter = 0;
  double resAbs = 1.0, resRel = 1.0, denAbs = 1.0, ratio = 0.0, resAbsOld = 1.0;
  double alpha = 0.0, beta, rr0, rr0old, omega, t12;
  fx_vec r0s = o.vec(), tmp = o.vec(), apj = o.vec(), asj = o.vec(), pj = o.vec(), rj = o.vec(),
         sj = o.vec(), ptmp = o.vec(), stmp = o.vec(), mp = o.vec(), ms = o.vec(),
         safe = o.vec();

  o.spmv(x, tmp);
  o.waxpby(1.0, b, -1.0, tmp, rj);
  o.copy(rj, r0s);
  o.copy(rj, pj);

  while (numIter < P.maxIter) {
    if (numIter == P.minIter) {
      resAbs = o.norm2(rj);
      denAbs = (kCloseZero > resAbs) ? kCloseZero : resAbs;
      resRel = resAbs / denAbs;
      if (resRel < P.relTol || resAbs < P.absTol) break;
    }
    if (numIter >= P.minIter) S->record(numIter, resAbs);

    ++numIter;
    rr0 = o.dot(rj, r0s);
    o.spmv(pj, apj);
    o.set(ptmp, 0.0);
    o.precond(apj, ptmp);
    t12 = o.dot(ptmp, r0s);
    if (std::fabs(t12) > 1e-40) {
      alpha = rr0 / t12;
    } else {
      err = FX_ERROR_DIVIDE_ZERO;
      break;
    }
    o.waxpby(1.0, rj, -alpha, ptmp, sj);
    o.spmv(sj, asj);
    o.set(stmp, 0.0);
    o.precond(asj, stmp);
    omega = o.dot(stmp, sj) / o.dot(stmp, stmp);
    o.set(mp, 0.0);
    o.precond(pj, mp);
    o.set(ms, 0.0);
    o.precond(sj, ms);
    o.waxpby(alpha, mp, omega, ms, tmp);
    o.xpay(1.0, tmp, x);
    o.waxpby(1.0, sj, -omega, stmp, rj);

    if (numIter >= P.minIter) {
      resAbs = o.norm2(rj);
      resRel = resAbs / denAbs;
      ratio = resAbs / resAbsOld;
      if (numIter >= P.savIter && resAbs < resAbsOld) o.copy(x, safe);
      if (ratio > kChkRatio && numIter > P.minIter) {
        if (o.norminf(x) < solStagTol) {  // Check I (src/BiCGStab.cxx:169)
          err = FX_ERROR_SOLVER_SOLSTAG;
          break;
        }
        const double xRelDiff = std::fabs(alpha) * o.norm2(pj) / o.norm2(x);
        if (stagStep <= kMaxStag && xRelDiff < solStagTol) {
          o.true_residual(b, x, rj);
          resAbs = o.norm2(rj);
          resRel = resAbs / denAbs;
          if (resRel < P.relTol || resAbs < P.absTol) break;
          if (stagStep >= kMaxStag) {
            err = FX_ERROR_SOLVER_STAG;
            break;
          }
          o.set(pj, 0.0);
          ++stagStep;
        }
      }
      if (resRel < P.relTol) {
        o.true_residual(b, x, rj);
        resAbs = o.norm2(rj);
        resRel = resAbs / denAbs;
        if (resRel < P.relTol || resAbs < P.absTol) break;
        if (moreStep >= P.restart) {
          err = FX_ERROR_SOLVER_TOLSMALL;
          break;
        }
        o.set(pj, 0.0);
        ++moreStep;
      }
    }
    if (numIter < P.maxIter) {
      resAbsOld = resAbs;
      rr0old = rr0;
      rr0 = o.dot(rj, r0s);
      beta = rr0 / rr0old * alpha / omega;
      o.waxpby(1.0, pj, -omega, ptmp, tmp);
      o.waxpby(1.0, rj, beta, tmp, pj);
    }
  }
  res->numIter = numIter;
  if (!(numIter == P.minIter && numIter == P.maxIter)) {
    res->norm2 = resAbs;
    res->normInf = o.norminf(rj);
  }
  if (numIter > P.savIter) o.copy(safe, x);
  return err;
}

// ---------------------------------------------------------------------------
// GMRES / FGMRES shared pieces: Hessenberg, Givens, back-substitution and the
// variable restart rule, on the host in f64 (src/GMRES.cxx:182-265).
// ---------------------------------------------------------------------------
void Givens::init(uint32_t maxR) {
  ld = maxR;
  h.assign((size_t)(maxR + 1) * maxR, 0.0);
  hc.assign(maxR + 1, 0.0);
  hs.assign(maxR, 0.0);
  g.assign(maxR + 2, 0.0);
}

// Apply the previous rotations to column cnt-1, form the new rotation and
// update the right-hand side g (src/GMRES.cxx:182-199).
void Givens::column(uint32_t cnt) {
  const uint32_t c1 = cnt - 1;
  for (uint32_t j = 1; j < cnt; ++j) {
    const double t = H(j - 1, c1);
    H(j - 1, c1) = fma(hs[j - 1], H(j, c1), hc[j - 1] * t);
    H(j, c1) = fma(-hs[j - 1], t, hc[j - 1] * H(j, c1));
  }
  double t = H(cnt, c1) * H(cnt, c1);
  t = fma(H(c1, c1), H(c1, c1), t);
  t = std::sqrt(t);
  const double gamma = t > kSmallDbl ? t : kSmallDbl;
  hc[c1] = H(c1, c1) / gamma;
  hs[c1] = H(cnt, c1) / gamma;
  H(c1, c1) = fma(hs[c1], H(cnt, c1), hc[c1] * H(c1, c1));
  g[cnt] = -hs[c1] * g[c1];
  g[c1] = hc[c1] * g[c1];
}

// Upper-triangular solve (src/GMRES.cxx:213-219) with a signed index: the
// reference's `USI k = count-2; k >= 0` never terminates.
void Givens::solve(uint32_t cnt) {
  const uint32_t c1 = cnt - 1;
  g[c1] = g[c1] / H(c1, c1);
  for (int k = (int)cnt - 2; k >= 0; --k) {
    double t = 0.0;
    for (uint32_t j = (uint32_t)k + 1; j < cnt; ++j) t = fma(-H(k, j), g[j], t);
    t += g[k];
    g[k] = t / H(k, k);
  }
}

void clamp_restart(const fx_params& P, uint32_t& maxR, uint32_t& minR, uint32_t& restart) {
  // GMRES::Setup, src/GMRES.cxx:43-46,75-79
  maxR = P.maxRestart > P.maxIter ? P.maxIter : P.maxRestart;
  minR = P.minRestart > maxR ? maxR : P.minRestart;
  restart = P.restart;
  if (restart > maxR)
    restart = maxR;
  else if (restart < minR)
    restart = minR;
}

// Variable restart after a cycle: GMRES strict (src/GMRES.cxx:256-265),
// FGMRES inclusive (src/FGMRES.cxx:247-256); unsigned subtraction as in the
// reference.
void solve_restart(fx_solver S, uint32_t& maxR, uint32_t& minR, uint32_t& restart) {
  clamp_restart(S->prm, maxR, minR, restart);
  if (S->restart_cur) restart = S->restart_cur < maxR ? S->restart_cur : maxR;
}

void adapt_restart(double cr, bool inclusive, uint32_t maxR, uint32_t minR, uint32_t& restart) {
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

// GMRES::RSolve (src/GMRES.cxx:114-280) and GMRES::LSolve (:283-453).
int ops_gmres(fx_solver S, Ops& o, fx_vec b, fx_vec x, fx_result* res) {
  const fx_params& P = S->prm;
  const bool left = P.leftPrecond != 0;
  uint32_t maxR, minR, restart;
  solve_restart(S, maxR, minR, restart);
  int err = FX_SUCCESS;
  double t, ri, rj, cr;
  double resAbs = 1.0, resRel = 1.0, denAbs = 1.0, ratio = 0.0, resAbsOld = 1.0;
  uint32_t count = 1, count_1 = 2, numIter = 0;
  Givens G;
  G.init(maxR);
  std::vector<fx_vec> V(maxR + 1);
  for (auto& v : V) v = o.vec();
  fx_vec wk = o.vec(), tmp = o.vec(), safe = o.vec();
  o.copy(x, safe);
  (void)ratio;

  if (!left) {
    o.true_residual(b, x, V[0]);
    resAbs = ri = o.norm2(V[0]);
  } else {
    o.true_residual(b, x, tmp);
    ri = o.norm2(tmp);
    o.set(V[0], 0.0);
    o.precond(tmp, V[0]);
    resAbs = o.norm2(V[0]);
  }
  denAbs = (kCloseZero > resAbs) ? kCloseZero : resAbs;

  while (numIter < P.maxIter) {
    if (numIter == P.minIter) {
      resRel = resAbs / denAbs;
      resAbsOld = resAbs;
      if (resRel < P.relTol || resAbs < P.absTol) break;
    }
    if (resAbs < kSmallDbl) break;  // never: kSmallDbl is negative
    G.g[0] = resAbs;
    o.scale(1 / resAbs, V[0]);

    count = 0;
    while (count < restart && numIter < P.maxIter) {
      if (numIter >= P.minIter) S->record(numIter, resAbs);
      ++numIter;
      count_1 = count++;
      if (!left) {
        o.set(tmp, 0.0);
        o.precond(V[count_1], tmp);
        o.spmv(tmp, V[count]);
      } else {
        o.spmv(V[count_1], tmp);
        o.set(V[count], 0.0);
        o.precond(tmp, V[count]);
      }
      for (uint32_t j = 0; j < count; ++j) {
        G.H(j, count_1) = o.dot(V[count], V[j]);
        o.axpby(1.0, -G.H(j, count_1), V[j], V[count]);
      }
      t = o.norm2(V[count]);
      G.H(count, count_1) = t;
      if (std::fabs(t) > kSmallDbl)
        o.scale(1.0 / t, V[count]);
      else
        break;
      G.column(count);
      resAbs = std::fabs(G.g[count]);
      ratio = resAbs / resAbsOld;
      resRel = resAbs / denAbs;
      resAbsOld = resAbs;
      if ((resAbs < P.absTol || resRel < P.relTol) && numIter > P.minIter) break;
    }

    G.solve(count);
    o.copy(V[count_1], wk);
    o.scale(G.g[count_1], wk);
    for (int j = (int)count - 2; j >= 0; --j) o.axpby(1.0, G.g[j], V[j], wk);
    if (!left) {
      o.set(tmp, 0.0);
      o.precond(wk, tmp);
      o.axpy(1.0, tmp, x);
    } else {
      o.axpy(1.0, wk, x);
    }
    if (numIter >= P.savIter && resAbs < resAbsOld) o.copy(x, safe);
    if (numIter > P.maxIter) {
      err = FX_ERROR_SOLVER_MAXIT;
      break;
    }
    if (!left) {
      o.true_residual(b, x, V[0]);
      resAbs = rj = o.norm2(V[0]);
    } else {
      o.true_residual(b, x, tmp);
      rj = o.norm2(tmp);
      o.set(V[0], 0.0);
      o.precond(tmp, V[0]);
      resAbs = o.norm2(V[0]);
    }
    resRel = resAbs / denAbs;
    if ((resRel < P.relTol || resAbs < P.absTol) && numIter >= P.minIter) break;
    cr = rj / ri;
    ri = rj;
    adapt_restart(cr, false, maxR, minR, restart);
  }
  S->restart_cur = restart;
  res->numIter = numIter;
  if (!(numIter == P.minIter && numIter == P.maxIter)) {
    fx_vec rv = left ? tmp : V[0];
    res->norm2 = o.norm2(rv);
    res->normInf = o.norminf(rv);
  }
  if (numIter > P.savIter) o.copy(safe, x);
  return err;
}

// FGMRES::Solve (src/FGMRES.cxx:108-271).  Z persists in the solver between
// cycles and calls and is the preconditioner's initial guess (:158-159).
int ops_fgmres(fx_solver S, Ops& o, fx_vec b, fx_vec x, fx_result* res) {
  const fx_params& P = S->prm;
  uint32_t maxR, minR, restart;
  solve_restart(S, maxR, minR, restart);
  int err = FX_SUCCESS;
  double t, ri, rj, cr;
  double resAbs = 1.0, resRel = 1.0, denAbs = 1.0, ratio = 0.0, resAbsOld = 1.0;
  uint32_t count = 1, count_1 = 2, numIter = 0;
  Givens G;
  G.init(maxR);
  std::vector<fx_vec> V(maxR + 1);
  for (auto& v : V) v = o.vec();
  if (S->Z.size() != maxR + 1 || S->Z[0]->n != o.n || S->Z[0]->dtype != o.dt) {
    for (fx_vec z : S->Z) fx_vec_destroy(z);
    S->Z.assign(maxR + 1, nullptr);
    for (auto& z : S->Z) Ops::chk(fx_vec_create(S->ctx, o.n, o.dt, &z));
  }
  std::vector<fx_vec>& Z = S->Z;
  fx_vec wk = o.vec(), safe = o.vec();
  o.copy(x, safe);
  (void)ratio;

  o.true_residual(b, x, V[0]);
  resAbs = ri = o.norm2(V[0]);
  denAbs = (resAbs > kCloseZero) ? resAbs : kCloseZero;

  while (numIter < P.maxIter) {
    if (numIter == P.minIter) {
      resRel = resAbs / denAbs;
      resAbsOld = resAbs;
      if (resRel < P.relTol || resAbs < P.absTol) break;
    }
    if (resAbs < kSmallDbl) break;
    G.g[0] = resAbs;
    o.scale(1 / resAbs, V[0]);
    count = 0;
    while (count < restart && numIter < P.maxIter) {
      if (numIter >= P.minIter) S->record(numIter, resAbs);
      ++numIter;
      count_1 = count++;
      o.precond(V[count_1], Z[count_1]);
      o.spmv(Z[count_1], V[count]);
      for (uint32_t j = 0; j < count; ++j) {
        G.H(j, count_1) = o.dot(V[count], V[j]);
        o.axpby(1.0, -G.H(j, count_1), V[j], V[count]);
      }
      t = o.norm2(V[count]);
      G.H(count, count_1) = t;
      if (std::fabs(t) > kSmallDbl)
        o.scale(1.0 / t, V[count]);
      else
        break;
      G.column(count);
      resAbs = std::fabs(G.g[count]);
      ratio = resAbs / resAbsOld;
      resAbsOld = resAbs;
      resRel = resAbs / denAbs;
      if ((resAbs < P.absTol || resRel < P.relTol) && numIter > P.minIter) break;
    }
    G.solve(count);
    o.copy(Z[count_1], wk);
    o.scale(G.g[count_1], wk);
    for (int j = (int)count - 2; j >= 0; --j) o.axpby(1.0, G.g[j], Z[j], wk);
    o.axpy(1.0, wk, x);
    if (numIter >= P.savIter && resAbs < resAbsOld) o.copy(x, safe);
    if (numIter > P.maxIter) {
      err = FX_ERROR_SOLVER_MAXIT;
      break;
    }
    o.true_residual(b, x, V[0]);
    resAbs = rj = o.norm2(V[0]);
    resRel = resAbs / denAbs;
    if ((resRel < P.relTol || resAbs < P.absTol) && numIter >= P.minIter) break;
    cr = rj / ri;
    ri = rj;
    adapt_restart(cr, true, maxR, minR, restart);
  }
  S->restart_cur = restart;
  res->numIter = numIter;
  if (!(numIter == P.minIter && numIter == P.maxIter)) {
    res->norm2 = o.norm2(V[0]);
    res->normInf = o.norminf(V[0]);
  }
  if (numIter > P.savIter) o.copy(safe, x);
  return err;
}

// Jacobi::Solve as a solver, src/Iter.cxx:56-112.
int ops_jacobi(fx_solver S, Ops& o, fx_vec b, fx_vec x, fx_result* res) {
  const fx_params& P = S->prm;
  double resAbs = 1.0, resRel = 1.0, denAbs = 1.0;
  uint32_t numIter = 0;
  fx_vec work = o.vec();
  fx_vec dinv = nullptr;
  Ops::chk(fx_vec_wrap(o.ctx, o.n, FX_F64, S->pc->dinv, &dinv));
  o.owned.push_back(dinv);
  while (numIter < P.maxIter) {
    o.residual(b, x, work);
    if (numIter >= P.minIter) {
      resAbs = o.norm2(work);
      if (numIter == P.minIter) denAbs = (kCloseZero > resAbs) ? kCloseZero : resAbs;
      resRel = resAbs / denAbs;
      if (resRel < P.relTol || resAbs < P.absTol) break;
      S->record(numIter, resAbs);
    }
    Ops::chk(fx_pmul(o.ctx, dinv, work));
    Ops::chk(fx_add(o.ctx, work, x));
    ++numIter;
  }
  res->numIter = numIter;
  if (!(numIter == P.minIter && numIter == P.maxIter)) {
    o.residual(b, x, work);
    res->norm2 = o.norm2(work);
    res->normInf = o.norminf(work);
  }
  return FX_SUCCESS;
}

void fx_solver_record_impl(fx_solver S, uint32_t it, double v) {
  if (it >= S->hist_host.size()) S->hist_host.resize((size_t)it + 1, 0.0);
  S->hist_host[it] = v;
  if (it + 1 > S->hist_count) S->hist_count = it + 1;
}

}  // namespace fx

void fx_solver_s::record(uint32_t it, double v) { fx::fx_solver_record_impl(this, it, v); }

using namespace fx;

extern "C" {

int fx_pc_create_identity(fx_ctx ctx, fx_pc* out) {
  clear_error();
  FX_CHECK(ctx && out, FX_ERROR_INPUT_PAR, "null argument");
  fx_pc p = new fx_pc_s();
  p->ctx = ctx;
  p->kind = 0;
  *out = p;
  return 0;
}

int fx_pc_create_jacobi(fx_ctx ctx, fx_csr A, double weight, uint32_t iters, fx_pc* out) {
  clear_error();
  FX_CHECK(ctx && A && out, FX_ERROR_INPUT_PAR, "null argument");
  FX_CHECK(A->nrow <= A->mcol, FX_ERROR_MAT_SIZE, "Jacobi needs nrow <= mcol");
  fx_pc p = new fx_pc_s();
  p->ctx = ctx;
  p->kind = 1;
  p->A = A;
  p->weight = weight;
  p->iters = iters;
  const uint32_t n = A->nrow;
  if (hipMalloc(&p->dinv, sizeof(double) * pack_round(n)) != hipSuccess) {  // whole 16-byte packs
    delete p;
    set_error("hipMalloc failed");
    return FX_ERROR_ALLOC_MEM;
  }
  if (n)
    hipLaunchKernelGGL(k_jacobi_dinv, dim3((n + kThreads - 1) / kThreads), dim3(kThreads), 0,
                       ctx->stream, A->has_values ? A->va : nullptr, A->dp, weight, p->dinv, n);
  if (hipGetLastError() != hipSuccess) {
    fx_pc_destroy(p);
    set_error("Jacobi setup kernel failed");
    return FX_ERROR_DEVICE;
  }
  *out = p;
  return 0;
}

int fx_pc_create_mcgs(fx_ctx ctx, fx_csr A, double weight, uint32_t sweeps, fx_pc* out) {
  clear_error();
  FX_CHECK(ctx && A && out, FX_ERROR_INPUT_PAR, "null argument");
  fx_pc p = new fx_pc_s();
  p->ctx = ctx;
  p->kind = 2;
  p->A = A;
  p->weight = weight;
  p->iters = sweeps;
  const int rc = fx_mcgs_create(ctx, A, weight, &p->mcgs);
  if (rc) {
    delete p;
    return rc;
  }
  *out = p;
  return 0;
}

int fx_pc_destroy(fx_pc p) {
  if (!p) return 0;
  if (p->ctx) (void)hipStreamSynchronize(p->ctx->stream);
  fx_mcgs_destroy(p->mcgs);
  (void)hipFree(p->dinv);
  (void)hipFree(p->dinv32);
  delete p;
  return 0;
}

// Jacobi::Solve with minIter == maxIter == iters (no norms), or Identity.
int fx_pc_apply(fx_ctx ctx, fx_pc pc, fx_vec b, fx_vec x) {
  clear_error();
  FX_CHECK(ctx && pc && b && x, FX_ERROR_INPUT_PAR, "null argument");
  if (pc->kind == 0) return fx_vec_copy(b, x);
  FX_CHECK(b->n == pc->A->nrow && x->n == pc->A->nrow, FX_ERROR_NONMATCH_SIZE,
           "preconditioner size mismatch");
  if (pc->kind == 2) return fx_mcgs_sweep(pc->mcgs, b, x, pc->iters, 1);
  fx_vec work = nullptr, dinv = nullptr;
  int rc = fx_vec_create(ctx, b->n, b->dtype, &work);
  if (rc) return rc;
  if (b->dtype == FX_F64) {
    rc = fx_vec_wrap(ctx, b->n, FX_F64, pc->dinv, &dinv);
  } else {
    rc = ensure_dinv32(pc);
    if (!rc) rc = fx_vec_wrap(ctx, b->n, FX_F32, pc->dinv32, &dinv);
  }
  for (uint32_t it = 0; !rc && it < pc->iters; ++it) {
    rc = fx_residual(ctx, pc->A, b, x, work);
    if (!rc) rc = fx_pmul(ctx, dinv, work);
    if (!rc) rc = fx_add(ctx, work, x);
  }
  fx_vec_destroy(dinv);
  fx_vec_destroy(work);
  return rc;
}

int fx_solver_create(fx_ctx ctx, int type, fx_solver* out) {
  clear_error();
  FX_CHECK(ctx && out, FX_ERROR_INPUT_PAR, "null argument");
  FX_CHECK(type == FX_SOLVER_CG || type == FX_SOLVER_BICGSTAB || type == FX_SOLVER_GMRES ||
               type == FX_SOLVER_FGMRES || type == FX_SOLVER_JACOBI,
           FX_ERROR_SOLVER_TYPE, "unknown solver type %d", type);
  fx_solver s = new fx_solver_s();
  s->ctx = ctx;
  s->type = type;
  fx_params_default(&s->prm);
  *out = s;
  return 0;
}

}  // extern "C"

namespace {

// Bind a matrix, a preconditioner and parameters to a solver.  setup: the
// reference's Setup, which resets the GMRES / FGMRES restart length
// (src/GMRES.cxx:44,75-79); otherwise the reference's Solve reading its A,
// pcd and params as they are now, which keeps it.
int solver_bind(fx_solver s, fx_csr A, fx_pc pc, const fx_params* prm, bool setup) {
  clear_error();
  FX_CHECK(s && A && pc, FX_ERROR_INPUT_PAR, "null argument");
  FX_CHECK(A->nrow == A->mcol, FX_ERROR_MAT_SIZE, "solvers need a square matrix");
  if (s->type == FX_SOLVER_JACOBI)
    FX_CHECK(pc->kind == 1, FX_ERROR_SOLVER_PCD_TYPE, "Jacobi solver needs a Jacobi pc");
  const fx_params np = prm ? *prm : s->prm;
  // Keep the fused-path workspace (and its replay graphs) only when the
  // same matrix and preconditioner handles are bound again with the same
  // parameters.  Handles are compared by their ids, not their addresses: a
  // matrix freed and made again may reuse the address of the old one.
  const bool same = solver_reuse_ok(s->A, s->A_gen, s->pc, s->pc_gen, A, pc) &&
                    std::memcmp(&np, &s->prm, sizeof(fx_params)) == 0;
  s->A = A;
  s->op = nullptr;
  s->pc = pc;
  s->A_gen = A->gen;
  s->pc_gen = pc->gen;
  s->prm = np;
  if (setup) s->restart_cur = 0;
  if (!same) s->release_fast();
  return 0;
}

}  // namespace

extern "C" {

int fx_solver_setup(fx_solver s, fx_csr A, fx_pc pc, const fx_params* prm) {
  return solver_bind(s, A, pc, prm, true);
}

int fx_solver_update(fx_solver s, fx_csr A, fx_pc pc, const fx_params* prm) {
  return solver_bind(s, A, pc, prm, false);
}

int fx_solver_clean(fx_solver s) {
  clear_error();
  FX_CHECK(s, FX_ERROR_INPUT_PAR, "null solver");
  s->restart_cur = 0;
  for (fx_vec z : s->Z) FX_TRY(fx_vec_set(z, 0.0));
  return 0;
}

int fx_op_create(fx_ctx ctx, uint32_t nrow, uint32_t mcol, fx_op_apply_fn apply, void* user,
                 fx_op* out) {
  clear_error();
  FX_CHECK(ctx && apply && out, FX_ERROR_INPUT_PAR, "null argument");
  fx_op op = new fx_op_s();
  op->ctx = ctx;
  op->nrow = nrow;
  op->mcol = mcol;
  op->apply = apply;
  op->user = user;
  *out = op;
  return 0;
}

int fx_op_set_device_only(fx_op op, int device_only) {
  clear_error();
  FX_CHECK(op, FX_ERROR_INPUT_PAR, "null argument");
  op->device_only = device_only != 0;
  return 0;
}

int fx_op_destroy(fx_op op) {
  delete op;
  return 0;
}

}  // extern "C"

namespace {

int solver_bind_op(fx_solver s, fx_op A, fx_pc pc, const fx_params* prm, bool setup) {
  clear_error();
  FX_CHECK(s && A && pc, FX_ERROR_INPUT_PAR, "null argument");
  FX_CHECK(A->nrow == A->mcol, FX_ERROR_MAT_SIZE, "solvers need a square operator");
  FX_CHECK(s->type != FX_SOLVER_JACOBI, FX_ERROR_SOLVER_TYPE,
           "the Jacobi solver needs a matrix (its diagonal)");
  s->release_fast();
  s->A = nullptr;
  s->op = A;
  s->pc = pc;
  s->A_gen = 0;
  s->pc_gen = pc->gen;
  s->prm = prm ? *prm : s->prm;
  if (setup) s->restart_cur = 0;
  return 0;
}

}  // namespace

extern "C" {

int fx_solver_setup_op(fx_solver s, fx_op A, fx_pc pc, const fx_params* prm) {
  return solver_bind_op(s, A, pc, prm, true);
}

int fx_solver_update_op(fx_solver s, fx_op A, fx_pc pc, const fx_params* prm) {
  return solver_bind_op(s, A, pc, prm, false);
}

int fx_solver_solve(fx_solver s, fx_vec b, fx_vec x, fx_result* res) {
  clear_error();
  FX_CHECK(s && (s->A || s->op) && b && x && res, FX_ERROR_INPUT_PAR,
           "solver not set up / null argument");
  const uint32_t n = s->op ? s->op->nrow : s->A->nrow;
  FX_CHECK(b->n == n && x->n == n, FX_ERROR_NONMATCH_SIZE, "Solve: sizes b %u, x %u, A %u", b->n,
           x->n, n);
  FX_CHECK(b->dtype == FX_F64 && x->dtype == FX_F64, FX_ERROR_INPUT_PAR,
           "solvers take f64 vectors (use fx_solve_ir for f32 inner solves)");
  // a bound matrix that has become another operator since set-up (a
  // structure-only one that gained values) gets new replay graphs
  if (s->A && (s->A->gen != s->A_gen || (s->pc && s->pc->gen != s->pc_gen))) {
    s->release_fast();
    s->A_gen = s->A->gen;
    if (s->pc) s->pc_gen = s->pc->gen;
  }
  res->retcode = 0;
  res->numIter = 0;
  res->norm2 = 1e60;  // SOL() initialises norm2 / normInf to LARGE_DBL
  res->normInf = 1e60;
  res->solve_ms = 0.0;
  s->hist_count = 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  FX_HIP(hipEventCreate(&e0));
  FX_HIP(hipEventCreate(&e1));
  FX_HIP(hipEventRecord(e0, s->ctx->stream));
  int rc = 0;
  try {
    s->last_path = fast_eligible(s) ? FX_PATH_FUSED : FX_PATH_OPS;
    if (s->last_path == FX_PATH_FUSED) {
      rc = fast_solve(s, b, x, res);
    } else {
      Ops o = s->op ? Ops(s->ctx, s->op, s->pc, FX_F64) : Ops(s->ctx, s->A, s->pc, FX_F64);
      switch (s->type) {
        case FX_SOLVER_CG: rc = ops_cg(s, o, b, x, res); break;
        case FX_SOLVER_BICGSTAB: rc = ops_bicgstab(s, o, b, x, res); break;
        case FX_SOLVER_GMRES: rc = ops_gmres(s, o, b, x, res); break;
        case FX_SOLVER_FGMRES: rc = ops_fgmres(s, o, b, x, res); break;
        case FX_SOLVER_JACOBI: rc = ops_jacobi(s, o, b, x, res); break;
        default: rc = FX_ERROR_SOLVER_TYPE;
      }
    }
  } catch (const OpsError& e) {
    rc = e.code;
  }
  (void)hipEventRecord(e1, s->ctx->stream);
  (void)hipEventSynchronize(e1);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  res->solve_ms = ms;
  res->retcode = rc;
  return rc;
}

int fx_solver_history(fx_solver s, double* out, uint32_t cap, uint32_t* count) {
  clear_error();
  FX_CHECK(s, FX_ERROR_INPUT_PAR, "null solver");
  const uint32_t c = s->hist_count;
  if (count) *count = c;
  if (out) {
    const uint32_t m = c < cap ? c : cap;
    for (uint32_t i = 0; i < m; ++i) out[i] = i < s->hist_host.size() ? s->hist_host[i] : 0.0;
  }
  return 0;
}

int fx_solver_last_path(fx_solver s, int* path) {
  clear_error();
  FX_CHECK(s && path, FX_ERROR_INPUT_PAR, "null argument");
  *path = s->last_path;
  return 0;
}

int fx_solver_destroy(fx_solver s) {
  if (!s) return 0;
  s->release_fast();
  for (fx_vec z : s->Z) fx_vec_destroy(z);
  delete s;
  return 0;
}

// Krylov(), src/Krylov.cxx:16-53 with SOL::SetSolTypeFromName,
// src/SOL.cxx:127-161: unknown (or non-Krylov) names fall back to CG.
int fx_krylov(fx_ctx ctx, fx_csr A, fx_vec b, fx_vec x, fx_pc pc, const char* algName,
              const fx_params* prm, fx_result* result) {
  clear_error();
  std::string name = algName ? algName : "";
  for (char& c : name) c = (char)std::tolower((unsigned char)c);
  int type = FX_SOLVER_CG;
  if (name == "bicgstab")
    type = FX_SOLVER_BICGSTAB;
  else if (name == "gmres")
    type = FX_SOLVER_GMRES;
  else if (name == "fgmres")
    type = FX_SOLVER_FGMRES;
  fx_solver s = nullptr;
  FX_TRY(fx_solver_create(ctx, type, &s));
  int rc = fx_solver_setup(s, A, pc, prm);
  if (!rc) rc = fx_solver_solve(s, b, x, result);
  fx_solver_destroy(s);
  return rc;
}

}  // extern "C"
