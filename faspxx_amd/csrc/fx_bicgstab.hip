// fx_bicgstab.hip -- fused, graph-replayed BiCGStab (src/BiCGStab.cxx:68-270)
// with the identity or a one-sweep Jacobi preconditioner on the system
// matrix.
//
// The reference's recurrence is the preconditioned P.A form: the residual is
// updated with P A p and P A s, the solution with P p and P s (:114-146).  One
// iteration is five kernels and two SpMVs, with every reduction consumed by
// the next kernel (each of its workgroups sums the partials in the same
// order, so all of them take the same decisions):
//   B1  pt = P (A p);  partials of (pt, r0*)                 (k_spmv + EpiBiP)
//   B2  alpha = rr0 / (pt, r0*);  s = r - alpha pt              (k_bi_s)
//   B3  t = P (A s);  partials of (t, s), (t, t)              (k_spmv + EpiBiS)
//   B4  omega = (t, s) / (t, t);  x += alpha P p + omega P s;
//       r = s - omega t;  partials of (r, r), (r, r0*), max|x|, (p, p), (x, x)
//                                                             (k_bi_update)
//   B5  the end-of-iteration logic of :148-206 on a private copy of the
//       state; p = r + beta (p - omega pt) unless the loop stops; workgroup 0
//       writes the new state                                  (k_bi_pupdate)
// As for CG, the state is double-buffered by the iteration's parity within a
// replay, the last iteration of a replay copies it to a pinned host slot, and
// the rare branches (breakdown, checks I-III, the safe-guard copy) are run
// on the host with the per-operation backend, after which the replay resumes.
// Elementwise expressions are the VEC forms of src/VEC.cxx (WAXPBY picks its
// form from a == 1 / b == 1, XPAY(1, tmp, x) is fma(1, x, tmp)).
#include <cmath>
#include <utility>

#include "fx_fast.hpp"

namespace fx {

namespace {

struct BiState {
  double rr0, t12, alpha, omega, beta, resAbs, resAbsOld, resRel, denAbs, ratio;
  double rr, rr0n, xinf, xx, pp;
  uint32_t numIter, stagStep, moreStep, histCount;
  int status, reason, err, pad;
};

// Workgroups per CU of the streaming kernels B2, B4 and B5: 1 measured best
// on the 1024^2 upwind system (63.4 us per iteration; 2: 70.4 us, with a
// different count, since the partial sums then round differently).
#ifndef FX_BI_PER_CU
#define FX_BI_PER_CU 1
#endif
constexpr int kBiPerCu = FX_BI_PER_CU;

struct BiSide {  // scalars handed from B2 / B4 to the later kernels of an iteration
  double t12, alpha, omega;
  int stop, pad;
};

// VEC::WAXPBY element (src/VEC.cxx:329-379), form = (a == 1) + 2 (b == 1)
template <class T>
__device__ __forceinline__ T waxpby_e(int form, T a, T v1, T b, T v2) {
  switch (form) {
    case 0: return fma(a, v1, b * v2);
    case 1: return fma(b, v2, v1);
    case 2: return fma(a, v1, v2);
    default: return v1 + v2;
  }
}
__host__ __device__ inline int waxpby_form(double a, double b) {
  return (a == 1.0 ? 1 : 0) + (b == 1.0 ? 2 : 0);
}

// B1 epilogue: pt = P (A p), partials of (pt, r0*)
template <class VT, bool JAC>
struct EpiBiP {
  VT* __restrict__ pt;
  const VT* __restrict__ dinv;
  const VT* __restrict__ r0s;
  const BiState* st;
  __device__ bool skip() const { return st->status != ST_RUN; }
  __device__ __forceinline__ void row(uint32_t i, double v, double (&acc)[1]) const {
    const VT a = (VT)v;
    const VT z = JAC ? a * dinv[i] : a;  // one Jacobi sweep from zero: z = (A p) .* dinv
    pt[i] = z;
    acc[0] = fma((double)z, (double)r0s[i], acc[0]);
  }
  __device__ void fin(double (&)[1]) const {}  // consumed by k_bi_s
};

// B3 epilogue: t = P (A s), partials of (t, s) and (t, t)
template <class VT, bool JAC>
struct EpiBiS {
  VT* __restrict__ t;
  const VT* __restrict__ dinv;
  const VT* __restrict__ s;
  const BiState* st;
  const BiSide* side;
  __device__ bool skip() const { return st->status != ST_RUN || side->stop; }
  __device__ __forceinline__ void row(uint32_t i, double v, double (&acc)[2]) const {
    const VT a = (VT)v;
    const VT z = JAC ? a * dinv[i] : a;
    t[i] = z;
    acc[0] = fma((double)z, (double)s[i], acc[0]);
    acc[1] = fma((double)z, (double)z, acc[1]);
  }
  __device__ void fin(double (&)[2]) const {}  // consumed by k_bi_update
};

// B2: alpha, then s = r - alpha pt
template <class VT>
__global__ __launch_bounds__(kWideThreads) void k_bi_s(VT* __restrict__ s, const VT* __restrict__ r,
                                                       const VT* __restrict__ pt, uint32_t n,
                                                       const BiState* __restrict__ st,
                                                       const double* __restrict__ p1, unsigned g1,
                                                       BiSide* __restrict__ side) {
  if (st->status != ST_RUN) return;
  const double rr0 = st->rr0;
  using PK = Pack16<VT>;
  constexpr int W = PK::W;
  const uint32_t nc = n / W, stride = gridDim.x * kWideThreads;
  const uint32_t cfirst = blockIdx.x * kWideThreads + threadIdx.x;
  // the first pass's packs are loaded before alpha is known (vectors are
  // padded to whole packs, so pack 0 is readable even when nc == 0)
  uint32_t c0 = cfirst < nc ? cfirst : 0u, c1 = cfirst + stride < nc ? cfirst + stride : c0;
  PK ra = ld16(r, c0), pa = ld16(pt, c0), rb = ld16(r, c1), pb = ld16(pt, c1);
  double t[1];
  partials_thread<1, 0u, kWideThreads>(p1, g1, threadIdx.x, t);
  partials_combine<1, 0u, kWideThreads>(t);
  const bool brk = !(fabs(t[0]) > 1e-40);  // src/BiCGStab.cxx:119-124
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    side->t12 = t[0];
    side->alpha = brk ? 0.0 : rr0 / t[0];
    side->stop = brk ? 1 : 0;
  }
  if (brk) return;
  const double alpha = rr0 / t[0];
  const int form = waxpby_form(1.0, -alpha);
  const VT one = (VT)1, b = (VT)(-alpha);
  for (uint32_t c = cfirst;;) {
#pragma unroll
    for (int w = 0; w < W; ++w) {
      ra.v[w] = waxpby_e(form, one, ra.v[w], b, pa.v[w]);
      rb.v[w] = waxpby_e(form, one, rb.v[w], b, pb.v[w]);
    }
    if (c < nc) st16(s, c0, ra);
    if (c + stride < nc) st16(s, c1, rb);
    c += 2 * stride;
    if (c >= nc) break;
    c0 = c;
    c1 = c + stride < nc ? c + stride : c0;
    ra = ld16(r, c0);
    pa = ld16(pt, c0);
    rb = ld16(r, c1);
    pb = ld16(pt, c1);
  }
  if (blockIdx.x == 0 && threadIdx.x < n - nc * W) {
    const uint32_t i = nc * W + threadIdx.x;
    s[i] = waxpby_e(form, one, r[i], b, pt[i]);
  }
}

// B4: omega, x += alpha P p + omega P s, r = s - omega t, five partials
template <class VT, bool JAC>
struct BiUpd {
  VT alpha, omega;
  int fx, fr;  // WAXPBY forms of tmp = alpha mp + omega ms and r = s - omega t
  __device__ __forceinline__ void elem(VT pi, VT si, VT ti, VT di, VT r0i, VT& xi, VT& ri,
                                       double (&acc)[5]) const {
    const VT mp = JAC ? pi * di : pi;  // P p
    const VT ms = JAC ? si * di : si;  // P s
    const VT tmp = waxpby_e(fx, alpha, mp, omega, ms);
    xi = fma((VT)1, xi, tmp);  // x.XPAY(1, tmp)
    ri = waxpby_e(fr, (VT)1, si, -omega, ti);
    acc[0] = fma((double)ri, (double)ri, acc[0]);
    acc[1] = fma((double)ri, (double)r0i, acc[1]);
    acc[2] = fmax(acc[2], fabs((double)xi));
    acc[3] = fma((double)pi, (double)pi, acc[3]);
    acc[4] = fma((double)xi, (double)xi, acc[4]);
  }
};

template <class VT, bool JAC>
struct BiPass {
  Pack16<VT> p0, s0, t0, d0, q0, x0, p1, s1, t1, d1, q1, x1;  // q: r0*
  uint32_t c0, c1;
  bool one, two;
  __device__ __forceinline__ void load(const VT* __restrict__ p, const VT* __restrict__ s,
                                       const VT* __restrict__ t, const VT* __restrict__ dinv,
                                       const VT* __restrict__ r0s, const VT* __restrict__ x,
                                       uint32_t c, uint32_t stride, uint32_t nc) {
    one = c < nc;
    two = c + stride < nc;
    c0 = one ? c : 0u;
    c1 = two ? c + stride : c0;
    p0 = ld16(p, c0);
    s0 = ld16(s, c0);
    t0 = ld16(t, c0);
    q0 = ld16(r0s, c0);
    x0 = ld16(x, c0);
    p1 = ld16(p, c1);
    s1 = ld16(s, c1);
    t1 = ld16(t, c1);
    q1 = ld16(r0s, c1);
    x1 = ld16(x, c1);
    if constexpr (JAC) {
      d0 = ld16(dinv, c0);
      d1 = ld16(dinv, c1);
    } else {
      d0 = Pack16<VT>{};
      d1 = Pack16<VT>{};
    }
  }
  __device__ __forceinline__ void run(const BiUpd<VT, JAC>& u, VT* __restrict__ x,
                                      VT* __restrict__ r, double (&acc)[5]) {
    constexpr int W = Pack16<VT>::W;
    double a0[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, a1[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    Pack16<VT> r0v, r1v;
#pragma unroll
    for (int w = 0; w < W; ++w) {
      u.elem(p0.v[w], s0.v[w], t0.v[w], d0.v[w], q0.v[w], x0.v[w], r0v.v[w], a0);
      u.elem(p1.v[w], s1.v[w], t1.v[w], d1.v[w], q1.v[w], x1.v[w], r1v.v[w], a1);
    }
    if (one) {
      st16(x, c0, x0);
      st16(r, c0, r0v);
    }
    if (two) {
      st16(x, c1, x1);
      st16(r, c1, r1v);
    }
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      const double b0 = one ? a0[k] : 0.0, b1 = two ? a1[k] : 0.0;
      acc[k] = k == 2 ? fmax(acc[k], fmax(b0, b1)) : acc[k] + b0 + b1;
    }
  }
};

template <class VT, bool JAC>
__global__ __launch_bounds__(kWideThreads) void k_bi_update(
    VT* __restrict__ x, VT* __restrict__ r, const VT* __restrict__ p, const VT* __restrict__ s,
    const VT* __restrict__ t, const VT* __restrict__ dinv, const VT* __restrict__ r0s, uint32_t n,
    const BiState* __restrict__ st, const double* __restrict__ p3, unsigned g3,
    BiSide* __restrict__ side, double* __restrict__ p4) {
  if (st->status != ST_RUN || side->stop) return;
  const double alpha = side->alpha;
  constexpr int W = Pack16<VT>::W;
  const uint32_t nc = n / W, stride = gridDim.x * kWideThreads;
  const uint32_t cfirst = blockIdx.x * kWideThreads + threadIdx.x;
  BiPass<VT, JAC> pass;
  pass.load(p, s, t, dinv, r0s, x, cfirst, stride, nc);  // before omega is known
  double ts[2];
  partials_thread<2, 0u, kWideThreads>(p3, g3, threadIdx.x, ts);
  partials_combine<2, 0u, kWideThreads>(ts);
  const double omega = ts[0] / ts[1];  // src/BiCGStab.cxx:134 (no guard there)
  if (blockIdx.x == 0 && threadIdx.x == 0) side->omega = omega;
  const BiUpd<VT, JAC> u{(VT)alpha, (VT)omega, waxpby_form(alpha, omega), waxpby_form(1.0, -omega)};
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  pass.run(u, x, r, acc);  // stores only packs below nc
  for (uint32_t c = cfirst + 2 * stride; c < nc; c += 2 * stride) {
    pass.load(p, s, t, dinv, r0s, x, c, stride, nc);
    pass.run(u, x, r, acc);
  }
  if (blockIdx.x == 0 && threadIdx.x < n - nc * W) {
    const uint32_t i = nc * W + threadIdx.x;
    VT xi = x[i], ri;
    u.elem(p[i], s[i], t[i], JAC ? dinv[i] : (VT)0, r0s[i], xi, ri, acc);
    x[i] = xi;
    r[i] = ri;
  }
  block_partials_n<5, 4u, kWideThreads>(acc, p4);
}

// Loop top (src/BiCGStab.cxx:96-104) for iteration S->numIter.
__device__ __forceinline__ void bi_loop_top(BiState* S, const CgPrm& P, double rr, double* hist) {
  const uint32_t it = S->numIter;
  if (it == P.minIter) {
    const double resAbs = sqrt(rr);
    S->resAbs = resAbs;
    S->denAbs = (kCloseZero > resAbs) ? kCloseZero : resAbs;
    S->resRel = resAbs / S->denAbs;
    if (S->resRel < P.relTol || resAbs < P.absTol) {
      S->status = ST_DONE;
      return;
    }
  }
  if (it >= P.minIter) {
    if (hist) hist[it] = S->resAbs;
    S->histCount = it + 1;
  }
}

// End of iteration, src/BiCGStab.cxx:148-206, for the common case; a rare
// branch hands the step to the host, which recomputes it from the state.
__device__ __forceinline__ void bi_after_update(BiState* S, const CgPrm& P, double* hist) {
  const uint32_t it = S->numIter;
  if (it >= P.minIter) {
    const double resAbs = sqrt(S->rr);
    const double resRel = resAbs / S->denAbs;
    const double ratio = resAbs / S->resAbsOld;
    bool host = false;
    if (it >= P.savIter && resAbs < S->resAbsOld) host = true;
    if (ratio > kChkRatio && it > P.minIter) {
      if (S->xinf < P.solStagTol) host = true;
      const double xRelDiff = fabs(S->alpha) * sqrt(S->pp) / sqrt(S->xx);
      if (S->stagStep <= kMaxStag && xRelDiff < P.solStagTol) host = true;
    }
    if (resRel < P.relTol) host = true;
    if (host) {
      S->status = ST_HOST;
      S->reason = R_CHECKS;
      return;
    }
    S->resAbs = resAbs;
    S->resRel = resRel;
    S->ratio = ratio;
  }
  if (it < P.maxIter) {
    S->resAbsOld = S->resAbs;
    const double rr0old = S->rr0;
    S->rr0 = S->rr0n;
    S->beta = S->rr0 / rr0old * S->alpha / S->omega;
  } else {
    S->status = ST_DONE;
    return;
  }
  bi_loop_top(S, P, S->rr, hist);
}

template <class VT>
struct PBiPass {
  Pack16<VT> p0, r0, t0, p1, r1, t1;  // t: pt
  uint32_t c0, c1;
  bool one, two;
  __device__ __forceinline__ void load(const VT* __restrict__ p, const VT* __restrict__ r,
                                       const VT* __restrict__ pt, uint32_t c, uint32_t stride,
                                       uint32_t nc) {
    one = c < nc;
    two = c + stride < nc;
    c0 = one ? c : 0u;
    c1 = two ? c + stride : c0;
    p0 = ld16(p, c0);
    r0 = ld16(r, c0);
    t0 = ld16(pt, c0);
    p1 = ld16(p, c1);
    r1 = ld16(r, c1);
    t1 = ld16(pt, c1);
  }
  __device__ __forceinline__ void run(int f1, int f2, VT mo, VT beta, VT* __restrict__ p) {
    constexpr int W = Pack16<VT>::W;
    const VT one1 = (VT)1;
#pragma unroll
    for (int w = 0; w < W; ++w) {
      // tmp = p - omega pt; p = r + beta tmp   (src/BiCGStab.cxx:214-215)
      p0.v[w] = waxpby_e(f2, one1, r0.v[w], beta, waxpby_e(f1, one1, p0.v[w], mo, t0.v[w]));
      p1.v[w] = waxpby_e(f2, one1, r1.v[w], beta, waxpby_e(f1, one1, p1.v[w], mo, t1.v[w]));
    }
    if (one) st16(p, c0, p0);
    if (two) st16(p, c1, p1);
  }
};

template <class VT>
__global__ __launch_bounds__(kWideThreads) void k_bi_pupdate(
    VT* __restrict__ p, const VT* __restrict__ r, const VT* __restrict__ pt, uint32_t n,
    BiState* st_cur, BiState* st_next, BiState* hout, const double* __restrict__ p4, unsigned g4,
    const BiSide* __restrict__ side, CgPrm P, double* hist) {
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  if (st_cur->status != ST_RUN) {  // stopped: hand the state on unchanged
    if (lead) {
      copy_state(st_next, st_cur);
      if (hout) copy_state(hout, st_cur);
    }
    return;
  }
  constexpr int W = Pack16<VT>::W;
  const uint32_t nc = n / W, stride = gridDim.x * kWideThreads;
  const uint32_t cfirst = blockIdx.x * kWideThreads + threadIdx.x;
  PBiPass<VT> pass;
  pass.load(p, r, pt, cfirst, stride, nc);  // before beta is known
  double a[5];
  partials_thread<5, 4u, kWideThreads>(p4, g4, threadIdx.x, a);
  partials_combine<5, 4u, kWideThreads>(a);
  __shared__ double sh_beta, sh_omega;
  __shared__ int sh_run;
  if (threadIdx.x == 0) {
    BiState S = *st_cur;
    const BiSide sd = *side;
    S.numIter += 1;  // src/BiCGStab.cxx:106
    S.t12 = sd.t12;
    if (sd.stop) {  // (pt, r0*) vanished: src/BiCGStab.cxx:119-124 ends the solve
      S.err = FX_ERROR_DIVIDE_ZERO;
      S.status = ST_DONE;
    } else {
      S.alpha = sd.alpha;
      S.omega = sd.omega;
      S.rr = a[0];
      S.rr0n = a[1];
      S.xinf = a[2];
      S.pp = a[3];
      S.xx = a[4];
      bi_after_update(&S, P, blockIdx.x == 0 ? hist : nullptr);
    }
    sh_beta = S.beta;
    sh_omega = S.omega;
    sh_run = S.status == ST_RUN;
    if (blockIdx.x == 0) {
      *st_next = S;
      if (S.status != ST_RUN) *st_cur = S;  // every later iteration of the replay stops
      if (hout) *hout = S;
    }
  }
  __syncthreads();
  if (!sh_run) return;
  const double beta = sh_beta, omega = sh_omega;
  const int f1 = waxpby_form(1.0, -omega), f2 = waxpby_form(1.0, beta);
  const VT mo = (VT)(-omega), bt = (VT)beta;
  pass.run(f1, f2, mo, bt, p);  // stores only packs below nc
  for (uint32_t c = cfirst + 2 * stride; c < nc; c += 2 * stride) {
    pass.load(p, r, pt, c, stride, nc);
    pass.run(f1, f2, mo, bt, p);
  }
  if (blockIdx.x == 0 && threadIdx.x < n - nc * W) {
    const uint32_t i = nc * W + threadIdx.x;
    p[i] = waxpby_e(f2, (VT)1, r[i], bt, waxpby_e(f1, (VT)1, p[i], mo, pt[i]));
  }
}

// r0* = r, p = r, partials of (r, r0*) and (r, r)
template <class VT>
__global__ __launch_bounds__(kThreads) void k_bi_init(VT* __restrict__ r0s, VT* __restrict__ p,
                                                      const VT* __restrict__ r, uint32_t n,
                                                      double* partials) {
  double acc[2] = {0.0, 0.0};
  const uint32_t stride = gridDim.x * kThreads;
  for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) {
    const VT ri = r[i];
    r0s[i] = ri;
    p[i] = ri;
    acc[0] = fma((double)ri, (double)ri, acc[0]);
    acc[1] = fma((double)ri, (double)ri, acc[1]);
  }
  block_partials<2>(acc, partials);
}

struct FinBiInit {
  BiState* st;
  CgPrm P;
  double* hist;
  __device__ bool skip() const { return false; }
  __device__ void fin(double (&s)[2]) const {
    if (threadIdx.x != 0) return;
    BiState* S = st;
    S->rr0 = s[0];
    S->rr = s[1];
    if (P.maxIter == 0) {
      S->status = ST_DONE;
      return;
    }
    bi_loop_top(S, P, s[1], hist);
  }
};

template <class VT>
struct EpiBMinusAx {  // r = b - (A x): Apply then WAXPBY(1, b, -1, Ax), src/BiCGStab.cxx:82-83
  VT* __restrict__ r;
  const VT* __restrict__ b;
  __device__ bool skip() const { return false; }
  __device__ __forceinline__ void row(uint32_t i, double v, double (&)[1]) const {
    r[i] = fma((VT)(-1), (VT)v, b[i]);
  }
  __device__ void fin(double (&)[1]) const {}
};

// ---------------------------------------------------------------------------
template <class VT>
struct BiRun {
  fx_solver S;
  FastWs& W;
  fx_csr A;
  const VT* dinv;  // null: identity preconditioner
  CgPrm P;
  hipStream_t st;
  fx_op op = nullptr;  // a callback operator in place of A

  BiState* dstate(int c = 0) const { return (BiState*)W.state + (c & 1); }
  BiState* hslot(int slot) const { return (BiState*)W.hstate + slot; }
  double* part13() const { return W.partials; }
  double* part4() const { return W.partials + W.part2_off; }
  VT* vx() const { return (VT*)W.x; }
  VT* vr() const { return (VT*)W.r; }
  VT* vp() const { return (VT*)W.p; }
  VT* vr0s() const { return (VT*)W.x2; }
  VT* vpt() const { return (VT*)W.q; }
  VT* vs() const { return (VT*)W.s; }
  VT* vt() const { return (VT*)W.t; }

  template <bool JAC>
  int launch_iteration_t(int c, BiState* hout) {
    const uint32_t n = W.n;
    const int cus = W.ctx->sms;
    const unsigned gs = fused_grid<VT>(A, op, n, cus);
    const unsigned g = wide_blocks(n, Pack16<VT>::W, cus, 2, kBiPerCu);
    BiSide* side = (BiSide*)W.side;
    FX_TRY((fused_spmv<VT, ROW_AX, 1, 0u>(A, op, W, st, XLoad<VT>{vp()}, BNone{},
                                          EpiBiP<VT, JAC>{vpt(), dinv, vr0s(), dstate(c)},
                                          part13())));
    hipLaunchKernelGGL((k_bi_s<VT>), dim3(g), dim3(kWideThreads), 0, st, vs(), (const VT*)vr(),
                       (const VT*)vpt(), n, (const BiState*)dstate(c), (const double*)part13(), gs,
                       side);
    FX_HIP(hipGetLastError());
    FX_TRY((fused_spmv<VT, ROW_AX, 2, 0u>(
        A, op, W, st, XLoad<VT>{vs()}, BNone{},
        EpiBiS<VT, JAC>{vt(), dinv, vs(), dstate(c), side}, part13())));
    hipLaunchKernelGGL((k_bi_update<VT, JAC>), dim3(g), dim3(kWideThreads), 0, st, vx(), vr(),
                       (const VT*)vp(), (const VT*)vs(), (const VT*)vt(), dinv,
                       (const VT*)vr0s(), n, (const BiState*)dstate(c), (const double*)part13(),
                       gs, side, part4());
    FX_HIP(hipGetLastError());
    hipLaunchKernelGGL((k_bi_pupdate<VT>), dim3(g), dim3(kWideThreads), 0, st, vp(),
                       (const VT*)vr(), (const VT*)vpt(), n, dstate(c), dstate(c + 1), hout,
                       (const double*)part4(), g, (const BiSide*)side, P, W.hist);
    FX_HIP(hipGetLastError());
    return 0;
  }
  int launch_iteration(int c, BiState* hout) {
    return dinv ? launch_iteration_t<true>(c, hout) : launch_iteration_t<false>(c, hout);
  }

  int read_state(BiState& s) {
    FX_HIP(hipStreamSynchronize(st));
    FX_HIP(hipMemcpy(&s, dstate(0), sizeof(BiState), hipMemcpyDeviceToHost));
    return 0;
  }
  int write_state(const BiState& s) {
    FX_HIP(hipMemcpyAsync(dstate(0), &s, sizeof(BiState), hipMemcpyHostToDevice, st));
    FX_HIP(hipStreamSynchronize(st));
    return 0;
  }
  int write_hist(uint32_t it, double v, BiState& s) {
    if (it >= W.hist_cap) return 0;
    FX_HIP(hipMemcpy(W.hist + it, &v, sizeof(double), hipMemcpyHostToDevice));
    s.histCount = it + 1;
    return 0;
  }

  // The rare branches, as src/BiCGStab.cxx:148-216 states them, from the
  // state the device left after updating x and r.  Returns 1 if the solve
  // ends, 0 to resume the device loop.
  int host_step(BiState& s, fx_vec vb, fx_vec vx_, fx_vec vr_, fx_vec vp_, fx_vec vpt_,
                fx_vec vr0s_, fx_vec vtmp, fx_vec vsafe, Ops& o, const fx_params& PP) {
    const double solStagTol = P.solStagTol;
    const uint32_t it = s.numIter;
    double resAbs = s.resAbs, resRel = s.resRel;
    if (it >= PP.minIter) {
      resAbs = std::sqrt(s.rr);
      resRel = resAbs / s.denAbs;
      const double ratio = resAbs / s.resAbsOld;
      s.ratio = ratio;
      if (it >= PP.savIter && resAbs < s.resAbsOld) o.copy(vx_, vsafe);
      if (ratio > kChkRatio && it > PP.minIter) {
        if (s.xinf < solStagTol) {  // Check I (src/BiCGStab.cxx:169)
          s.err = FX_ERROR_SOLVER_SOLSTAG;
          s.resAbs = resAbs;
          s.resRel = resRel;
          return 1;
        }
        const double xRelDiff = std::fabs(s.alpha) * std::sqrt(s.pp) / std::sqrt(s.xx);
        if (s.stagStep <= kMaxStag && xRelDiff < solStagTol) {
          o.true_residual(vb, vx_, vr_);
          resAbs = o.norm2(vr_);
          resRel = resAbs / s.denAbs;
          if (resRel < PP.relTol || resAbs < PP.absTol) {
            s.resAbs = resAbs;
            s.resRel = resRel;
            return 1;
          }
          if (s.stagStep >= kMaxStag) {
            s.err = FX_ERROR_SOLVER_STAG;
            s.resAbs = resAbs;
            s.resRel = resRel;
            return 1;
          }
          o.set(vp_, 0.0);
          ++s.stagStep;
        }
      }
      if (resRel < PP.relTol) {
        o.true_residual(vb, vx_, vr_);
        resAbs = o.norm2(vr_);
        resRel = resAbs / s.denAbs;
        if (resRel < PP.relTol || resAbs < PP.absTol) {
          s.resAbs = resAbs;
          s.resRel = resRel;
          return 1;
        }
        if (s.moreStep >= PP.restart) {
          s.err = FX_ERROR_SOLVER_TOLSMALL;
          s.resAbs = resAbs;
          s.resRel = resRel;
          return 1;
        }
        o.set(vp_, 0.0);
        ++s.moreStep;
      }
      s.resAbs = resAbs;
      s.resRel = resRel;
    }
    if (it < PP.maxIter) {
      s.resAbsOld = s.resAbs;
      const double rr0old = s.rr0;
      s.rr0 = o.dot(vr_, vr0s_);
      s.beta = s.rr0 / rr0old * s.alpha / s.omega;
      o.waxpby(1.0, vp_, -s.omega, vpt_, vtmp);
      o.waxpby(1.0, vr_, s.beta, vtmp, vp_);
    } else {
      return 1;
    }
    if (it == PP.minIter) {  // loop top
      const double ra = o.norm2(vr_);
      s.resAbs = ra;
      s.denAbs = (kCloseZero > ra) ? kCloseZero : ra;
      s.resRel = ra / s.denAbs;
      if (s.resRel < PP.relTol || ra < PP.absTol) return 1;
    }
    if (it >= PP.minIter) FX_TRY(write_hist(it, s.resAbs, s));
    s.status = ST_RUN;
    s.reason = 0;
    return 0;
  }

  int run(const VT* b, VT* x_user, fx_result* res, const fx_params& PP, int pc_dtype) {
    const uint32_t n = W.n;
    FX_HIP(hipMemcpyAsync(vx(), x_user, sizeof(VT) * n, hipMemcpyDeviceToDevice, st));
    BiState s0{};
    s0.resAbs = 1.0;
    s0.resRel = 1.0;
    s0.denAbs = 1.0;
    s0.resAbsOld = 1.0;
    s0.status = ST_RUN;
    FX_TRY(write_state(s0));
    // r = b - A x;  r0* = r;  p = r;  rr0 = (r, r0*);  loop top
    FX_TRY((fused_spmv<VT, ROW_AX, 0, 0u>(A, op, W, st, XLoad<VT>{(const VT*)vx()}, BNone{},
                                          EpiBMinusAx<VT>{vr(), b}, nullptr)));
    const unsigned g = stream_blocks(n);
    hipLaunchKernelGGL((k_bi_init<VT>), dim3(g), dim3(kThreads), 0, st, vr0s(), vp(),
                       (const VT*)vr(), n, W.partials);
    FX_HIP(hipGetLastError());
    FX_TRY((launch_finish<2, 0u>(st, W.partials, g, FinBiInit{dstate(0), P, W.hist})));

    Ops o(W.ctx, A, op, S->pc, pc_dtype);
    fx_vec vb_ = nullptr, vx_ = nullptr, vr_ = nullptr, vp_ = nullptr, vpt_ = nullptr,
           vr0s_ = nullptr, vtmp = nullptr, vsafe = nullptr;
    const int dt = pc_dtype;
    const std::pair<fx_vec*, void*> views[] = {
        {&vb_, (void*)b}, {&vx_, W.x},   {&vr_, W.r},  {&vp_, W.p},
        {&vpt_, W.q},     {&vr0s_, W.x2}, {&vtmp, W.t}, {&vsafe, W.safe}};
    for (const auto& v : views) {
      Ops::chk(fx_vec_wrap(W.ctx, n, dt, v.second, v.first));
      o.owned.push_back(*v.first);
    }

    BiState s{};
    // as for CG: a stretch of replays stops within maxIter / kChunk of them,
    // and there are at most maxIter host steps
    const uint64_t max_chunks = (uint64_t)PP.maxIter / kChunk + 4;
    uint64_t host_steps = 0;
    for (;;) {
      ReplayPace pace;
      pace.hint = host_steps == 0 ? W.last_iters : 0;
      // the iteration that ends the loop is numbered numIter - 1: a
      // prediction of hint iterations is hint launches
      const BiState* hs = nullptr;
      FX_TRY(replay_loop<BiState>(
          W, st, pace, pace.hint, max_chunks + 2,
          [&](int c, int slot) { return launch_iteration(c, slot >= 0 ? hslot(slot) : nullptr); },
          [&](const BiState& h) {
            const double target = std::max(PP.relTol * h.denAbs, PP.absTol);
            return pace.more(h.resAbs, target, h.numIter, PP.minIter == 0 ? h.denAbs : 0.0);
          },
          &hs, "BiCGStab"));
      FX_CHECK(++host_steps <= (uint64_t)PP.maxIter + 2, FX_ERROR_DEVICE,
               "BiCGStab device loop did not stop");
      FX_TRY(read_state(s));
      if (s.status == ST_DONE) break;
      if (host_step(s, vb_, vx_, vr_, vp_, vpt_, vr0s_, vtmp, vsafe, o, PP)) break;
      FX_TRY(write_state(s));
    }
    res->numIter = s.numIter;
    W.last_iters = host_steps == 1 ? s.numIter : 0;  // a prediction only for plain runs
    if (!(s.numIter == PP.minIter && s.numIter == PP.maxIter)) {
      res->norm2 = s.resAbs;
      res->normInf = o.norminf(vr_);
    }
    if (s.numIter > PP.savIter) o.copy(vsafe, vx_);
    FX_HIP(hipMemcpyAsync(x_user, vx(), sizeof(VT) * n, hipMemcpyDeviceToDevice, st));
    const uint32_t hc = s.histCount < W.hist_cap ? s.histCount : W.hist_cap;
    S->hist_host.assign(hc, 0.0);
    if (hc) FX_HIP(hipMemcpy(S->hist_host.data(), W.hist, sizeof(double) * hc, hipMemcpyDeviceToHost));
    S->hist_count = hc;
    FX_HIP(hipStreamSynchronize(st));
    return s.err;
  }
};

}  // namespace

size_t bicgstab_state_bytes() { return sizeof(BiState); }

// BiCGStab on the fused path (f64 vectors); W is allocated by the caller.
int bicgstab_fast(fx_solver S, FastWs& W, const double* b, double* x, fx_result* res) {
  const size_t vb = pack_round(W.n) * sizeof(double);
  if (!W.s) {
    FX_HIP(hipMalloc(&W.s, vb));
    FX_HIP(hipMalloc(&W.t, vb));
  }
  BiRun<double> run{S, W, S->A, S->pc->kind == 1 ? S->pc->dinv : nullptr, make_prm(S->prm),
                    S->ctx->stream};
  run.op = S->op;
  return run.run(b, x, res, S->prm, FX_F64);
}

}  // namespace fx
