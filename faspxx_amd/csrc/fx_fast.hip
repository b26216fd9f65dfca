// fx_fast.hip -- fused, graph-replayed solvers.
//
// Jacobi-preconditioned (one sweep, any weight) or unpreconditioned CG,
// the weighted-Jacobi solver, and the mixed-precision refinement built on
// the f32 CG.  All scalars stay in a device state block; the reference's
// per-iteration decisions (src/CG.cxx:82-236, src/Iter.cxx:69-99) run on the
// device, in the kernel that consumes each reduction (CG) or in the
// reduction's finisher (Jacobi).  The rare branches (CG's breakdown
// guard, stagnation checks I-II, the false-convergence check III, the
// safe-guard copy) stop the device loop and are run on the host with the
// per-operation backend, after which the device loop resumes.
//
// The refinement's f32 inner solve is CG on the diagonally scaled matrix in
// a single-reduction form: 2 kernels and 11 vector passes per iteration
// (see "the scaled inner solver of the refinement" below).
//
// CG iteration = 3 kernels (see "the three-kernel iteration" below), replayed
// as a hipGraph of kChunk iterations; the last iteration of a replay copies
// the state to a pinned host slot.  Per iteration HBM traffic: the SpMV's
// matrix stream + 13 vector passes (K1: p read, q written; K2: x p r q D
// read, x r written; K3: r D p read, p written).  Forming p inside K1
// instead (saving K3) was measured slower: it triples the gathered loads of
// the latency-bound SpMV.
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "fx_fast.hpp"

namespace fx {

bool fast_disabled() {
  const char* e = std::getenv("FASPXX_FX_NO_FAST");
  return e && e[0] == '1';
}

namespace {

struct CgState {
  double tmpa, pq, alpha, beta, resAbs, resAbsOld, resRel, denAbs, ratio;
  double rr, rz, xinf, xx, pp;
  uint32_t numIter, stagStep, moreStep, histCount;
  int status, reason, err, pad;
};

// ---- CG kernels -----------------------------------------------------------

template <class VT>
struct EpiBMinusW {  // r = b - (A x), the Apply + XPAY(-1, b) of src/CG.cxx:70-71
  VT* __restrict__ r;
  const VT* __restrict__ b;
  __device__ bool skip() const { return false; }
  __device__ __forceinline__ void row(uint32_t i, double v, double (&)[1]) const {
    const VT w = (VT)v;
    r[i] = b[i] - w;
  }
  __device__ void fin(double (&)[1]) const {}
};

template <class VT>
struct EpiCgQ {  // q = A p and per-workgroup partials of (q, p)   (src/CG.cxx:99-113)
  VT* __restrict__ q;
  const VT* __restrict__ p;
  const CgState* st;
  __device__ bool skip() const { return st->status != ST_RUN; }
  __device__ __forceinline__ void row(uint32_t r, double v, double (&acc)[1]) const {
    const VT qv = (VT)v;
    q[r] = qv;
    acc[0] = fma((double)qv, (double)p[r], acc[0]);
  }
  __device__ void fin(double (&)[1]) const {}  // the sum is consumed by k_cg_update_w
};

// Loop-top bookkeeping (src/CG.cxx:84-93) for iteration `it`.
__device__ __forceinline__ void cg_loop_top(CgState* S, const CgPrm& P, double rr, double* hist) {
  const uint32_t it = S->numIter;
  if (it == P.minIter) {
    const double resAbs = sqrt(rr);
    S->resAbs = resAbs;
    S->resAbsOld = resAbs;
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

// End-of-iteration logic of src/CG.cxx:130-234 for the common case; any rare
// branch hands the whole step to the host (which recomputes it from S).
__device__ __forceinline__ void cg_after_update(CgState* S, const CgPrm& P, const double (&a)[5], double* hist) {
  S->rr = a[0];
  S->rz = a[1];
  S->xinf = a[2];
  S->xx = a[3];
  S->pp = a[4];
  const uint32_t it = S->numIter;
  if (it >= P.minIter) {
    const double resAbs = sqrt(a[0]);
    const double resRel = resAbs / S->denAbs;
    const double ratio = resAbs / S->resAbsOld;
    bool host = false;
    if (it >= P.savIter && resAbs < S->resAbsOld) host = true;
    if (ratio > kChkRatio) {
      if (a[2] < kCloseZero) host = true;
      const double xRelDiff = fabs(S->alpha) * sqrt(a[4]) / sqrt(a[3]);
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
    S->beta = a[1] / S->tmpa;
    S->tmpa = a[1];
  } else {
    S->status = ST_DONE;
    return;
  }
  cg_loop_top(S, P, a[0], hist);
}

// The streaming kernels below read and write 16-byte packs, two packs per
// thread per pass (all loads of a pass issued together); the n % W tail
// elements are done by block 0.  Per element the arithmetic is the
// reference's (src/CG.cxx:121-126, :223-227).
template <class VT, bool JAC>
struct CgUpd {
  VT a;
  // N = 6 (a multi-sweep Jacobi preconditioner): slot 5 is max |z| of the
  // first sweep, for the next sweep's MAT::Residual shortcut
  template <int N>
  __device__ __forceinline__ void elem(VT pi, VT qi, VT& xi, VT& ri, VT di,
                                       double (&acc)[N]) const {
    xi = fma(a, pi, xi);   // x.AXPY(alpha, p)
    ri = fma(-a, qi, ri);  // r.AXPY(-alpha, q)
    const VT zi = JAC ? ri * di : ri;  // Jacobi sweep from z = 0: z = r .* dinv
    acc[0] = fma((double)ri, (double)ri, acc[0]);
    acc[1] = fma((double)zi, (double)ri, acc[1]);
    acc[2] = fmax(acc[2], fabs((double)xi));
    acc[3] = fma((double)xi, (double)xi, acc[3]);
    acc[4] = fma((double)pi, (double)pi, acc[4]);
    if constexpr (N > 5) acc[5] = fmax(acc[5], fabs((double)zi));
  }
};

// ---- the three-kernel iteration ---------------------------------------------
// K1  q = A p, per-workgroup partials of (q, p)               (k_spmv + EpiCgQ)
// K2  alpha = tmpa / (q, p) from K1's partials, summed by every workgroup in
//     the same order; x += alpha p, r -= alpha q; partials of
//     (r,r) (z,r) max|x| (x,x) (p,p)                           (k_cg_update_w)
// K3  the same sums of K2's partials in every workgroup; the end-of-iteration
//     logic of src/CG.cxx:130-234 on a private copy of the state (identical in
//     every workgroup); p = beta p + z unless the loop stops; workgroup 0
//     writes the new state                                     (k_cg_pupdate_w)
// No finisher kernels: the reductions are consumed where they are needed.
// The state is double-buffered by the iteration's parity within a replay
// (iteration c reads buffer c & 1 and writes buffer (c + 1) & 1), so that no
// workgroup of an iteration can see a state written during it; a stopping
// iteration writes both buffers.
struct CgSide {  // written by K2's workgroup 0 for K3
  double pq, alpha;
  int breakdown, pad;
};

// Workgroups per CU of the CG's K2 and K3: 1 measured best for the FP64 CG
// on the 128^3 stencil (19.7-19.9 ms; 2: 20.9-21.2 ms; with 2 packs per
// pass, 20.3 ms and at 4 per CU 23.5 ms): each workgroup re-sums all of the
// SpMV's 8192 partials.
#ifndef FX_CG_PER_CU
#define FX_CG_PER_CU 1
#endif
constexpr int kCgPerCu = FX_CG_PER_CU;
// Workgroups per CU of K3 alone (it reads K2's partials, whatever its grid):
// 1 measured best (FP64 CG on the 128^3 stencil 19.6-19.9 ms; 2: 20.4 ms;
// 4: 22.4 ms).
#ifndef FX_K3_PER_CU
#define FX_K3_PER_CU 1
#endif
constexpr int kK3PerCu = FX_K3_PER_CU;

// One pass of K2 over packs c + i * stride, i < kPassPacks (all loaded
// before use; packs at or past nc are clamped to pack 0, which the padded
// vectors hold, and their results dropped).
template <class VT, bool JAC>
struct UpdPass {
  Pack16<VT> p[kPassPacks], q[kPassPacks], x[kPassPacks], r[kPassPacks], d[kPassPacks];
  uint32_t c[kPassPacks];
  bool ok[kPassPacks];
  __device__ __forceinline__ void load(const VT* __restrict__ pv, const VT* __restrict__ qv,
                                       const VT* __restrict__ xv, const VT* __restrict__ rv,
                                       const VT* __restrict__ dinv, uint32_t c0, uint32_t stride,
                                       uint32_t nc) {
#pragma unroll
    for (int i = 0; i < kPassPacks; ++i) {
      const uint32_t ci = c0 + i * stride;
      ok[i] = ci < nc;
      c[i] = ok[i] ? ci : 0u;
      p[i] = ld16(pv, c[i]);
      q[i] = ld16(qv, c[i]);
      x[i] = ld16(xv, c[i]);
      r[i] = ld16(rv, c[i]);
      if constexpr (JAC) d[i] = ld16(dinv, c[i]);
      else d[i] = Pack16<VT>{};
    }
  }
  template <int N>
  __device__ __forceinline__ void run(const CgUpd<VT, JAC>& u, VT* __restrict__ xv,
                                      VT* __restrict__ rv, double (&acc)[N]) {
    constexpr int W = Pack16<VT>::W;
#pragma unroll
    for (int i = 0; i < kPassPacks; ++i) {
      double a[N];
#pragma unroll
      for (int k = 0; k < N; ++k) a[k] = 0.0;
#pragma unroll
      for (int w = 0; w < W; ++w) u.elem(p[i].v[w], q[i].v[w], x[i].v[w], r[i].v[w], d[i].v[w], a);
      if (ok[i]) {
        st16(xv, c[i], x[i]);
        st16(rv, c[i], r[i]);
      }
#pragma unroll
      for (int k = 0; k < N; ++k) {
        const double b = ok[i] ? a[k] : 0.0;
        acc[k] = (k == 2 || k == 5) ? fmax(acc[k], b) : acc[k] + b;
      }
    }
  }
};

template <class VT, bool JAC, bool SW = false>
__global__ __launch_bounds__(kWideThreads) void k_cg_update_w(
    VT* __restrict__ x, VT* __restrict__ r, const VT* __restrict__ p, const VT* __restrict__ q,
    const VT* __restrict__ dinv, uint32_t n, const CgState* __restrict__ st,
    const double* __restrict__ p1, unsigned g1, CgSide* __restrict__ side,
    double* __restrict__ p2) {
  if (st->status != ST_RUN) return;
  const double tmpa = st->tmpa;
  constexpr int W = Pack16<VT>::W;
  const uint32_t nc = n / W, stride = gridDim.x * kWideThreads;
  const uint32_t cfirst = blockIdx.x * kWideThreads + threadIdx.x;
  UpdPass<VT, JAC> pass;  // the first pass is loaded before alpha is known (with nc == 0 the
  pass.load(p, q, x, r, dinv, cfirst, stride, nc);  // clamped loads read pack 0 of the padded vectors)
  double pq[1];
  partials_thread<1, 0u, kWideThreads>(p1, g1, threadIdx.x, pq);
  partials_combine<1, 0u, kWideThreads>(pq);
  const bool brk = !(fabs(pq[0]) > kCloseZero * kCloseZero);  // src/CG.cxx:114-119
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    side->pq = pq[0];
    side->alpha = brk ? 0.0 : tmpa / pq[0];
    side->breakdown = brk ? 1 : 0;
  }
  if (brk) return;  // the same decision in every workgroup
  const CgUpd<VT, JAC> u{(VT)(tmpa / pq[0])};
  constexpr int NA = SW ? 6 : 5;
  double acc[NA];
#pragma unroll
  for (int k = 0; k < NA; ++k) acc[k] = 0.0;
  pass.run(u, x, r, acc);  // stores only packs below nc
  for (uint32_t c = cfirst + kPassPacks * stride; c < nc; c += kPassPacks * stride) {
    pass.load(p, q, x, r, dinv, c, stride, nc);
    pass.run(u, x, r, acc);
  }
  if (blockIdx.x == 0 && threadIdx.x < n - nc * W) {
    const uint32_t i = nc * W + threadIdx.x;
    VT xi = x[i], ri = r[i];
    u.elem(p[i], q[i], xi, ri, JAC ? dinv[i] : (VT)0, acc);
    x[i] = xi;
    r[i] = ri;
  }
  block_partials_n<NA, SW ? (4u | 32u) : 4u, kWideThreads>(acc, p2);
}

// Packs per thread per pass of K3: 2 measured 20.1 against 20.3 ms for the
// FP64 CG on the 128^3 stencil (8: 20.06 ms, but 256 VGPRs and spills).
#ifndef FX_PPACKS
#define FX_PPACKS 2
#endif
constexpr int kPPacks = FX_PPACKS;
// SW: z is a vector (the last of several Jacobi sweeps), loaded in place of
// r, and no dinv.
template <class VT, bool JAC, bool SW = false>
struct PPass {
  Pack16<VT> p[kPPacks], r[kPPacks], d[kPPacks];
  uint32_t c[kPPacks];
  bool ok[kPPacks];
  __device__ __forceinline__ void load(const VT* __restrict__ pv, const VT* __restrict__ rv,
                                       const VT* __restrict__ dinv, uint32_t c0, uint32_t stride,
                                       uint32_t nc) {
#pragma unroll
    for (int i = 0; i < kPPacks; ++i) {
      const uint32_t ci = c0 + i * stride;
      ok[i] = ci < nc;
      c[i] = ok[i] ? ci : 0u;
      p[i] = ld16(pv, c[i]);
      r[i] = ld16(rv, c[i]);
      if constexpr (JAC && !SW) d[i] = ld16(dinv, c[i]);
      else d[i] = Pack16<VT>{};
    }
  }
  __device__ __forceinline__ void run(VT beta, VT* __restrict__ pv) {
    constexpr int W = Pack16<VT>::W;
#pragma unroll
    for (int i = 0; i < kPPacks; ++i) {
#pragma unroll
      for (int w = 0; w < W; ++w)
        p[i].v[w] = fma(beta, p[i].v[w],
                        (JAC && !SW) ? r[i].v[w] * d[i].v[w] : r[i].v[w]);  // p.XPAY(beta, z)
      if (ok[i]) st16(pv, c[i], p[i]);
    }
  }
};

// SW (several Jacobi sweeps): r is the last sweep's z, and (z, r) comes
// from that sweep's partials pz (gz of them) instead of K2's slot 1.
template <class VT, bool JAC, bool SW = false>
__global__ __launch_bounds__(kWideThreads) void k_cg_pupdate_w(
    VT* __restrict__ p, const VT* __restrict__ r, const VT* __restrict__ dinv, uint32_t n,
    CgState* st_cur, CgState* st_next, CgState* hout, const double* __restrict__ p2,
    unsigned g2, const CgSide* __restrict__ side, CgPrm P, double* hist,
    const double* __restrict__ pz = nullptr, unsigned gz = 0) {
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
  PPass<VT, JAC, SW> pass;  // loaded before beta is known (nc == 0: pack 0 of the padded vectors)
  pass.load(p, r, dinv, cfirst, stride, nc);
  double a[5];
  partials_thread<5, 4u, kWideThreads>(p2, g2, threadIdx.x, a);
  if constexpr (SW) {
    double zr[1];
    partials_thread<1, 0u, kWideThreads>(pz, gz, threadIdx.x, zr);
    a[1] = zr[0];
  }
  partials_combine<5, 4u, kWideThreads>(a);
  __shared__ double sh_beta;
  __shared__ int sh_run;
  if (threadIdx.x == 0) {
    // src/CG.cxx:99-119 (the (q, p) test), then :121-234, on a private copy
    CgState S = *st_cur;
    const CgSide sd = *side;
    S.numIter += 1;
    S.pq = sd.pq;
    if (sd.breakdown) {
      S.status = ST_HOST;
      S.reason = R_BREAKDOWN;
    } else {
      S.alpha = sd.alpha;
      cg_after_update(&S, P, a, blockIdx.x == 0 ? hist : nullptr);
    }
    sh_beta = S.beta;
    sh_run = S.status == ST_RUN;
    if (blockIdx.x == 0) {
      *st_next = S;
      if (S.status != ST_RUN) *st_cur = S;  // every later iteration of the replay stops
      if (hout) *hout = S;
    }
  }
  __syncthreads();
  if (!sh_run) return;
  const VT beta = (VT)sh_beta;
  pass.run(beta, p);  // stores only packs below nc
  for (uint32_t c = cfirst + kPPacks * stride; c < nc; c += kPPacks * stride) {
    pass.load(p, r, dinv, c, stride, nc);
    pass.run(beta, p);
  }
  if (blockIdx.x == 0 && threadIdx.x < n - nc * W) {
    const uint32_t i = nc * W + threadIdx.x;
    p[i] = fma(beta, p[i], (JAC && !SW) ? r[i] * dinv[i] : r[i]);
  }
}

// ---- a Jacobi preconditioner of several sweeps inside the CG iteration ------
// Jacobi::Solve with minIter == maxIter == k from z = 0 (src/Iter.cxx:56-112,
// the preconditioner of example/TestPCG.cxx:71-76): K2 forms the first sweep
// z1 = r .* dinv inline (with max |z1|); each further sweep is an SpMV in
// Residual form, work = r - A z (or r when ||z||_inf < 1e-20,
// src/MAT.cxx:459-462), z += work .* dinv, preceded by a one-workgroup
// max-norm finisher; the last sweep also gives the partials of (z, r).
template <class VT>
struct XJacZ {  // z1 = r .* dinv, formed where the first extra sweep gathers it
  const VT* __restrict__ r;
  const VT* __restrict__ dinv;
  __device__ __forceinline__ VT operator()(uint32_t j) const { return r[j] * dinv[j]; }
};

template <class VT, bool LAST>
struct EpiCgSweep {
  VT* __restrict__ zn;        // the new z
  const VT* __restrict__ zc;  // the current z (null: z1 = r .* dinv)
  const VT* __restrict__ r;
  const VT* __restrict__ dinv;
  const double* __restrict__ zinf;  // ||z||_inf of the current z
  const CgState* st;
  __device__ bool skip() const { return st->status != ST_RUN; }
  __device__ __forceinline__ void row(uint32_t i, double v, double (&acc)[1]) const {
    const VT w = (*zinf < kCloseZero) ? r[i] : (VT)v;  // MAT::Residual
    const VT zcur = zc ? zc[i] : r[i] * dinv[i];
    const VT z = zcur + w * dinv[i];  // work.PointwiseMult(diagInv); z += work
    zn[i] = z;
    if constexpr (LAST)
      acc[0] = fma((double)z, (double)r[i], acc[0]);
    else
      acc[0] = fmax(acc[0], fabs((double)z));
  }
  __device__ void fin(double (&)[1]) const {}
};

// p = z, and the partials of (r, r), (z, r) for the loop start (z given).
template <class VT>
__global__ __launch_bounds__(kThreads) void k_cg_init_z(VT* __restrict__ p,
                                                        const VT* __restrict__ r,
                                                        const VT* __restrict__ z, uint32_t n,
                                                        double* partials) {
  double acc[2] = {0.0, 0.0};
  const uint32_t stride = gridDim.x * kThreads;
  for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) {
    const VT ri = r[i], zi = z[i];
    p[i] = zi;
    acc[0] = fma((double)ri, (double)ri, acc[0]);
    acc[1] = fma((double)zi, (double)ri, acc[1]);
  }
  block_partials<2>(acc, partials);
}

// z = P(r), p = z, tmpa = (z, r), and the first loop-top check.
template <class VT, bool JAC>
__global__ __launch_bounds__(kThreads) void k_cg_init(VT* __restrict__ p, const VT* __restrict__ r,
                                                      const VT* __restrict__ dinv, uint32_t n,
                                                      double* partials) {
  double acc[2] = {0.0, 0.0};
  const uint32_t stride = gridDim.x * kThreads;
  for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) {
    const VT ri = r[i];
    const VT zi = JAC ? ri * dinv[i] : ri;
    p[i] = zi;
    acc[0] = fma((double)ri, (double)ri, acc[0]);
    acc[1] = fma((double)zi, (double)ri, acc[1]);
  }
  block_partials<2>(acc, partials);
}

struct FinCgInit {
  CgState* st;
  CgPrm P;
  double* hist;
  __device__ bool skip() const { return false; }
  __device__ void fin(double (&s)[2]) const {
    if (threadIdx.x != 0) return;
    CgState* S = st;
    S->tmpa = s[1];
    S->rr = s[0];
    if (P.maxIter == 0) {
      S->status = ST_DONE;
      return;
    }
    cg_loop_top(S, P, s[0], hist);
  }
};

// ---- Jacobi solver kernels (src/Iter.cxx:69-99) ---------------------------

// The Jacobi solver's sweep with its stopping decision one sweep late, so
// that no finisher kernel sits between sweeps.  Sweep m reads x_m (buffer
// m % 3) and the flag "||x_m||_inf >= 1e-20" (flag m % 3: MAT::Residual's
// shortcut, src/MAT.cxx:455-476), writes x_{m+1} (buffer (m+1) % 3), raises
// flag (m+1) % 3 from every workgroup holding an entry of x_{m+1} at or
// above 1e-20 (a flag written only with 1 needs no reduction), and writes
// its partial sums of ||b - A x_m||^2.  Its workgroup 0 meanwhile sums sweep
// m-1's partials and takes that sweep's decision (src/Iter.cxx:69-94) from
// state buffer m % 2 into (m+1) % 2 (both when it stops), and clears flag
// (m+2) % 3 for sweep m+1.  With three x buffers x_{m-1} is still intact
// when sweep m-1 turns out to have converged; sweep m's own x_{m+1} is then
// dropped, and later sweeps skip.
struct JacLagState {
  double resAbs, resRel, denAbs, pad;
  uint32_t numIter, histCount, answer, pending;  // answer: x buffer 0..2
  int status, pad2;
};

struct EpiJacobiLag {
  double* __restrict__ xn;
  const double* __restrict__ x;
  const double* __restrict__ b;
  const double* __restrict__ dinv;
  const int* flag_in;
  int* flag_out;
  int* flag_clear;
  JacLagState* s_in;
  JacLagState* s_out;
  const double* prev;  // sweep m-1's partials (gprev of them)
  unsigned gprev;
  CgPrm P;
  uint32_t buf_prev, buf_cur;  // the x buffers of x_{m-1} and x_m
  double* hist;
  int use_v;  // ||x_m||_inf >= 1e-20: work = b - A x, else work = b
  // Stopped; or, outside workgroup 0, the sweep that decides sweep
  // maxIter - 1: that decision ends the solve whatever it finds, so the
  // sweep's own x_{m+1} and partial sums are never read and only workgroup
  // 0 (the decision) runs.  Workgroup 0 may be writing s_in meanwhile; every
  // mix of the old and new fields it can leave reads as skip here.
  __device__ bool skip() const {
    if (s_in->status != ST_RUN) return true;
    return blockIdx.x != 0 && s_in->pending && s_in->numIter + 1 >= P.maxIter;
  }
  __device__ void init() { use_v = *flag_in; }
  __device__ void block_start() {
    if (blockIdx.x != 0) return;
    if (!s_in->pending) {  // the first sweep: nothing to decide yet
      if (threadIdx.x == 0) {
        JacLagState S = *s_in;
        S.pending = 1;
        *s_out = S;
        *flag_clear = 0;
      }
      return;
    }
    double acc[1];
    partials_thread<1, 0u, kThreads>(prev, gprev, threadIdx.x, acc);
    partials_combine<1, 0u, kThreads>(acc);
    if (threadIdx.x != 0) return;
    JacLagState S = *s_in;
    const uint32_t it = S.numIter;  // the sweep being decided (m - 1)
    *flag_clear = 0;
    if (it >= P.minIter) {
      const double resAbs = sqrt(acc[0]);
      if (it == P.minIter) S.denAbs = (kCloseZero > resAbs) ? kCloseZero : resAbs;
      S.resAbs = resAbs;
      S.resRel = resAbs / S.denAbs;
      if (S.resRel < P.relTol || resAbs < P.absTol) {
        S.status = ST_DONE;
        S.answer = buf_prev;  // x was not updated in sweep m-1
        *s_out = S;
        *s_in = S;
        return;
      }
      hist[it] = resAbs;
      S.histCount = it + 1;
    }
    S.numIter = it + 1;
    S.answer = buf_cur;
    if (S.numIter >= P.maxIter) {
      S.status = ST_DONE;
      *s_in = S;
    }
    *s_out = S;
  }
  __device__ __forceinline__ void row(uint32_t i, double v, double (&acc)[2]) const {
    const double w = use_v ? v : b[i];
    acc[0] = fma(w, w, acc[0]);
    const double wd = w * dinv[i];  // work.PointwiseMult(diagInv)
    const double xv = x[i] + wd;    // x += work
    xn[i] = xv;
    acc[1] = fmax(acc[1], fabs(xv));
  }
  __device__ void block_end(double (&acc)[2]) const {
    const int any = __syncthreads_or(acc[1] >= kCloseZero);
    if (threadIdx.x == 0 && any) *flag_out = 1;
  }
  __device__ void fin(double (&)[2]) const {}
};

// The state after a replay into a pinned host slot (no copy enqueued
// between replays).
__global__ void k_jac_report(const JacLagState* __restrict__ s, JacLagState* hout) {
  if (threadIdx.x == 0) *hout = *s;
}

__global__ void k_jac_flags(int* flags, const double* xinf) {
  flags[0] = (*xinf < kCloseZero) ? 0 : 1;
  flags[1] = 0;
  flags[2] = 0;
}

constexpr int kJacPeriod = 6;  // the sweeps' buffer rotation: 3 x buffers and 2 states

// Final residual norms (src/Iter.cxx:102-109): ||b - A x||_2 and _inf.
struct EpiResNorms {
  const double* __restrict__ b;
  const double* xinf;
  double* out;  // [norm2, normInf]
  __device__ bool skip() const { return false; }
  __device__ __forceinline__ void row(uint32_t i, double v, double (&acc)[2]) const {
    const double w = (*xinf < kCloseZero) ? b[i] : v;
    acc[0] = fma(w, w, acc[0]);
    acc[1] = fmax(acc[1], fabs(w));
  }
  __device__ void fin(double (&s)[2]) const {
    if (threadIdx.x == 0) {
      out[0] = sqrt(s[0]);
      out[1] = s[1];
    }
  }
};

template <class T>
__global__ __launch_bounds__(kThreads) void k_norminf_part(const T* __restrict__ x, uint32_t n,
                                                           double* partials) {
  double acc[1] = {0.0};
  const uint32_t stride = gridDim.x * kThreads;
  for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < n; i += stride)
    acc[0] = fmax(acc[0], fabs((double)x[i]));
  block_partials<1, 1u>(acc, partials);
}

// ||x||_inf into *out: partials kernel + finisher
int norminf_to(hipStream_t st, const double* x, uint32_t n, double* out, double* partials) {
  const unsigned g = stream_blocks(n);
  hipLaunchKernelGGL(k_norminf_part<double>, dim3(g), dim3(kThreads), 0, st, x, n, partials);
  FX_HIP(hipGetLastError());
  return launch_finish<1, 1u>(st, partials, g, FinStore<1>{out, 0});
}

// ---- mixed precision helpers -----------------------------------------------

// r = b - A x in f64 (Apply + XPAY order), stored as f32; ||r||^2 -> out.
struct EpiResidF32 {
  float* __restrict__ r32;
  const double* __restrict__ b;
  double* out;
  __device__ bool skip() const { return false; }
  __device__ __forceinline__ void row(uint32_t i, double v, double (&acc)[1]) const {
    const double w = b[i] - v;
    r32[i] = (float)w;
    acc[0] = fma(w, w, acc[0]);
  }
  __device__ void fin(double (&s)[1]) const {
    if (threadIdx.x == 0) *out = sqrt(s[0]);
  }
};

__global__ void k_add_f32_to_f64(double* __restrict__ x, const float* __restrict__ d, uint32_t n) {
  const uint32_t stride = gridDim.x * kThreads;
  for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < n; i += stride)
    x[i] = x[i] + (double)d[i];
}

// ---- the scaled inner solver of the refinement -----------------------------
// The f32 inner solve runs CG on the symmetrically scaled system
//   Ahat = S A S,  S = diag(1 / sqrt(a_ii)),  Ahat yhat = S r,  d = S yhat,
// which is Jacobi-preconditioned CG in exact arithmetic, with no
// preconditioner vector left in the loop.  It uses the single-reduction
// form of Chronopoulos and Gear (J. Comput. Appl. Math. 25, 1989): with
// w = Ahat r, gamma = (r, r) and delta = (w, r) from one reduction phase,
//   beta = gamma / gamma_old,  alpha = gamma / (delta - beta gamma / alpha_old),
//   p = r + beta p,  s = w + beta s  (s = Ahat p),  y += alpha p,  r -= alpha s,
// so an iteration is two kernels: K1 = SpMV w = Ahat r with partials of
// (w, r); K2 = the scalars, the four updates and partials of the new (r, r).
// Per iteration HBM traffic: the matrix + 11 f32 vector passes (K1: r read,
// w written; K2: r p w s y read, r p s y written), against 13 passes and a
// third kernel for the f32 Jacobi-PCG.
struct IrState {
  double gamma0;      // (r0, r0) of the scaled residual
  double thr;         // stop when gamma < thr: max(relTol^2 gamma0, absTol^2)
  double gamma_prev;  // gamma, alpha and beta of the last update
  double alpha_prev;
  double beta_prev;
  double gamma, delta;  // the last iteration's sums
  uint32_t numIter, maxIter;
  int status, reason;
};
// IR_OUTER: the outer loop met its tolerance at the residual before this
// inner solve, which therefore does not run
enum { IR_CONVERGED = 0, IR_MAXIT = 1, IR_BREAKDOWN = 2, IR_OUTER = 3 };

// True while the inner solve whose state is cs has not run to its end: it
// is running, or it was stopped before it began (IR_OUTER).
__device__ __forceinline__ bool ir_inner_open(const IrState* cs) {
  return cs->status == ST_RUN || cs->reason == IR_OUTER;
}

// The outer loop's record after each residual, in pinned host memory:
// ||b - A x||, the reference value of the relative test, and whether the
// loop stops there.
struct IrOuterRec {
  double resAbs, denAbs;
  int done, pad;
};

// s_i = 1 / sqrt(a_ii); *bad = 1 where row i has no diagonal entry or
// a_ii is not positive and finite.
__global__ void k_ir_scale_diag(const uint32_t* __restrict__ rp, const uint32_t* __restrict__ ci,
                                const uint32_t* __restrict__ dp, const double* __restrict__ va,
                                double* __restrict__ sc, int* __restrict__ bad, uint32_t n) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const uint32_t k = dp[i];
  const bool has = k >= rp[i] && k < rp[i + 1] && ci[k] == i;
  const double d = has ? va[k] : 0.0;
  const bool ok = d > 0.0 && d < INFINITY;
  sc[i] = ok ? 1.0 / sqrt(d) : 0.0;
  if (!ok) *bad = 1;
}

// ahat_k = (float)(a_k s_i s_j) for the entries of row i.
__global__ void k_ir_scale_vals(const uint32_t* __restrict__ rp, const uint32_t* __restrict__ ci,
                                const double* __restrict__ va, const double* __restrict__ sc,
                                float* __restrict__ vs, uint32_t n) {
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const double si = sc[i];
  for (uint32_t k = rp[i]; k < rp[i + 1]; ++k) vs[k] = (float)(va[k] * si * sc[ci[k]]);
}

// Outer residual for the scaled inner solve: t = b - A x in f64 (Apply +
// XPAY order), rhat = (float)(s_i t_i); partials of (t, t) and (rhat, rhat).
// cs: when set, the step runs only if that inner solve has run and stopped
// (it was queued before the host knew).
struct EpiResidScaled {
  float* __restrict__ r32;
  const double* __restrict__ b;
  const double* __restrict__ sc;
  const IrState* cs;
  __device__ bool skip() const { return cs && ir_inner_open(cs); }
  __device__ __forceinline__ void row(uint32_t i, double v, double (&acc)[2]) const {
    const double t = b[i] - v;
    const float rh = (float)(sc[i] * t);
    r32[i] = rh;
    acc[0] = fma(t, t, acc[0]);
    acc[1] = fma((double)rh, (double)rh, acc[1]);
  }
  __device__ void fin(double (&)[2]) const {}
};

// EpiResidScaled for a matrix stored with its rows sorted by length: b and
// S come in that stored order (bo[pos] = b[row]), so that they are read in
// whole lines; the same values as EpiResidScaled's.
struct EpiResidScaledSorted {
  float* __restrict__ r32;
  const double* __restrict__ bo;
  const double* __restrict__ sco;
  const IrState* cs;
  __device__ bool skip() const { return cs && ir_inner_open(cs); }
  __device__ __forceinline__ void row_at(uint32_t pos, uint32_t i, double v,
                                         double (&acc)[2]) const {
    const double t = bo[pos] - v;
    const float rh = (float)(sco[pos] * t);
    r32[i] = rh;
    acc[0] = fma(t, t, acc[0]);
    acc[1] = fma((double)rh, (double)rh, acc[1]);
  }
  __device__ __forceinline__ void row(uint32_t, double, double (&)[2]) const {}
  __device__ void fin(double (&)[2]) const {}
};

// dst[pos] = src[omap[pos]]: a vector in a sorted matrix's stored row order.
__global__ void k_gather_rows(const uint32_t* __restrict__ omap, const double* __restrict__ src,
                              double* __restrict__ dst, uint32_t n) {
  const uint32_t stride = gridDim.x * kThreads;
  for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) dst[i] = src[omap[i]];
}

// The outer loop's record (||t||, its test against the outer tolerance; the
// first residual of a solve sets the reference value), and the inner
// solve's starting state: running, or stopped (IR_OUTER) when the outer
// loop is done, so that an inner solve queued before the host has looked at
// the record does not run.
struct FinIrInit {
  IrOuterRec* rec;  // pinned
  double* den;      // the outer loop's reference value, on the device
  IrState* st;
  double relTol, absTol;    // inner
  double orelTol, oabsTol;  // outer
  uint32_t maxIter;
  int first;
  const IrState* cs;  // see EpiResidScaled
  __device__ bool skip() const { return cs && ir_inner_open(cs); }
  __device__ void fin(double (&s)[2]) const {
    if (threadIdx.x != 0) return;
    const double resAbs = sqrt(s[0]);
    if (first) *den = (kCloseZero > resAbs) ? kCloseZero : resAbs;
    const double denAbs = *den;
    const bool done = resAbs / denAbs < orelTol || resAbs < oabsTol;
    IrState S{};
    S.gamma0 = s[1];
    S.thr = fmax(relTol * relTol * s[1], absTol * absTol);
    S.maxIter = maxIter;
    S.status = done ? ST_DONE : ST_RUN;
    S.reason = done ? IR_OUTER : IR_CONVERGED;
    *st = S;
    IrOuterRec R{};
    R.resAbs = resAbs;
    R.denAbs = denAbs;
    R.done = done ? 1 : 0;
    *rec = R;
  }
};

// K1: w = Ahat r and per-workgroup partials of (w, r).
struct EpiIrW {
  float* __restrict__ w;
  const float* __restrict__ r;
  const IrState* st;
  __device__ bool skip() const { return st->status != ST_RUN; }
  __device__ __forceinline__ void row(uint32_t i, double v, double (&acc)[1]) const {
    const float wv = (float)v;
    w[i] = wv;
    acc[0] = fma((double)wv, (double)r[i], acc[0]);
  }
  __device__ void fin(double (&)[1]) const {}
};

// Packs per thread per pass of the refinement's K2, and its workgroups per
// CU: 2 and 2 (512 workgroups of one pass each on the 128^3 stencil)
// measured 9.90-9.94 ms for the refinement against 10.20-10.26 ms for 4 and
// 1 (256 workgroups); 2 passes of 2 packs, or 1 pack, were slower.
#ifndef FX_IR_PACKS
#define FX_IR_PACKS 2
#endif
#ifndef FX_IR_PER_CU
#define FX_IR_PER_CU 2
#endif
constexpr int kIrPacks = FX_IR_PACKS;
constexpr int kIrPerCu = FX_IR_PER_CU;

// One pass of K2 over packs c + i * stride, i < kIrPacks (all loaded
// before use; packs at or past nc are clamped to pack 0 and dropped).
// odd: an odd iteration, which also reads p, y and the previous
// iteration's r (ro); see ir_elem.
struct IrPass {
  Pack16<float> r[kIrPacks], ro[kIrPacks], w[kIrPacks], p[kIrPacks], s[kIrPacks], y[kIrPacks];
  uint32_t c[kIrPacks];
  bool ok[kIrPacks];
  __device__ __forceinline__ void load(const float* __restrict__ rv, const float* __restrict__ rov,
                                       const float* __restrict__ wv, const float* __restrict__ pv,
                                       const float* __restrict__ sv, const float* __restrict__ yv,
                                       uint32_t c0, uint32_t stride, uint32_t nc, bool odd) {
#pragma unroll
    for (int i = 0; i < kIrPacks; ++i) {
      const uint32_t ci = c0 + i * stride;
      ok[i] = ci < nc;
      c[i] = ok[i] ? ci : 0u;
      r[i] = ld16(rv, c[i]);
      w[i] = ld16(wv, c[i]);
      s[i] = ld16(sv, c[i]);
      if (odd) {
        ro[i] = ld16(rov, c[i]);
        p[i] = ld16(pv, c[i]);
        y[i] = ld16(yv, c[i]);
      }
    }
  }
};

// The updates of one element in iteration it: s and r in every iteration
// (first, it == 0: s = w).  p serves only y += alpha p, so both are formed
// two iterations at a time: an odd iteration (odd) forms
//   p_{it-1} = fma(bp, p_{it-2}, r_{it-1})     (r_0 when it == 1: pair0)
//   p_it     = fma(b, p_{it-1}, r_it)
//   y        = fma(a, p_it, fma(ap, p_{it-1}, y))   (ap * p_0 when it == 1)
// from the previous iteration's r (ro, kept in the other r buffer), alpha
// (ap) and beta (bp): the same roundings, in the same order, as forming p
// and y in each iteration.  Even iterations read and write neither.  After
// an odd number of iterations the last one's p and alpha p are still to be
// formed (k_ir_add_scaled does it).
__device__ __forceinline__ void ir_elem(bool first, bool odd, bool pair0, float a, float b,
                                        float ap, float bp, float& ri, float roi, float wi,
                                        float& pi, float& si, float& yi, double& acc) {
  if (odd) {
    const float pm = pair0 ? roi : fma(bp, pi, roi);
    pi = fma(b, pm, ri);
    yi = fma(a, pi, pair0 ? ap * pm : fma(ap, pm, yi));
  }
  si = first ? wi : fma(b, si, wi);
  ri = fma(-a, si, ri);
  acc = fma((double)ri, (double)ri, acc);
}

// r alternates between two buffers: iteration it reads r_it from ra (even
// it) or rb (odd it) and writes r_{it+1} to the other; an odd iteration
// also reads r_{it-1} from ra.
__global__ __launch_bounds__(kWideThreads) void k_ir_update(
    float* __restrict__ ra, float* __restrict__ rb, const float* __restrict__ w,
    float* __restrict__ p, float* __restrict__ s, float* __restrict__ y, uint32_t n,
    IrState* st_cur, IrState* st_next,
    IrState* hout, const double* __restrict__ p1, unsigned g1, const double* __restrict__ pg_in,
    double* __restrict__ pg_out) {
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  if (st_cur->status != ST_RUN) {  // stopped: hand the state on unchanged
    if (lead) {
      copy_state(st_next, st_cur);
      if (hout) copy_state(hout, st_cur);
    }
    return;
  }
  constexpr int W = Pack16<float>::W;
  const uint32_t nc = n / W, stride = gridDim.x * kWideThreads;
  const uint32_t cfirst = blockIdx.x * kWideThreads + threadIdx.x;
  const uint32_t it = st_cur->numIter;
  const bool odd = (it & 1u) != 0;  // this iteration forms p and y (for it - 1 and it)
  const float* r = odd ? rb : ra;   // r_it
  float* rn = odd ? ra : rb;        // r_{it+1}; over r_{it-1} (ra) in an odd iteration
  IrPass pass;  // the first pass is loaded before the scalars are known
  pass.load(r, ra, w, p, s, y, cfirst, stride, nc, odd);
  double red[2];
  {
    double d1[1], d2[1] = {0.0};
    partials_thread<1, 0u, kWideThreads>(p1, g1, threadIdx.x, d1);
    if (it > 0) partials_thread<1, 0u, kWideThreads>(pg_in, gridDim.x, threadIdx.x, d2);
    red[0] = d1[0];
    red[1] = d2[0];
  }
  partials_combine<2, 0u, kWideThreads>(red);
  const double delta = red[0];
  const double gamma = it > 0 ? red[1] : st_cur->gamma0;
  // the same decisions in every workgroup (the same sums of the same partials)
  int stop = -1;
  double alpha = 0.0, beta = 0.0;
  if (gamma < st_cur->thr || gamma == 0.0) {
    stop = IR_CONVERGED;
  } else if (it >= st_cur->maxIter) {
    stop = IR_MAXIT;
  } else {
    double den = delta;
    if (it > 0) {
      beta = gamma / st_cur->gamma_prev;
      den = delta - beta * gamma / st_cur->alpha_prev;
    }
    alpha = gamma / den;
    if (!(den > 0.0) || !(fabs(alpha) < INFINITY) || !(fabs(beta) < INFINITY)) stop = IR_BREAKDOWN;
  }
  if (stop >= 0) {
    if (lead) {
      copy_state(st_next, st_cur);
      st_next->status = ST_DONE;
      st_next->reason = stop;
      st_next->gamma = gamma;
      st_next->delta = delta;
      st_cur->reason = stop;
      st_cur->status = ST_DONE;  // every later iteration of the replay stops
      if (hout) copy_state(hout, st_next);
    }
    return;
  }
  const bool first = it == 0, pair0 = it == 1;
  const float a = (float)alpha, b = (float)beta;
  // iteration it - 1's alpha and beta (used when odd)
  const float ap = (float)st_cur->alpha_prev, bp = (float)st_cur->beta_prev;
  double acc = 0.0;
  for (uint32_t c = cfirst;;) {
#pragma unroll
    for (int i = 0; i < kIrPacks; ++i) {
      double ai = 0.0;
#pragma unroll
      for (int k = 0; k < W; ++k)
        ir_elem(first, odd, pair0, a, b, ap, bp, pass.r[i].v[k], pass.ro[i].v[k],
                pass.w[i].v[k], pass.p[i].v[k], pass.s[i].v[k], pass.y[i].v[k], ai);
      if (pass.ok[i]) {
        st16(rn, pass.c[i], pass.r[i]);
        st16(s, pass.c[i], pass.s[i]);
        if (odd) {
          st16(p, pass.c[i], pass.p[i]);
          st16(y, pass.c[i], pass.y[i]);
        }
        acc += ai;
      }
    }
    c += kIrPacks * stride;
    if (c >= nc) break;
    pass.load(r, ra, w, p, s, y, c, stride, nc, odd);
  }
  if (blockIdx.x == 0 && threadIdx.x < n - nc * W) {
    const uint32_t i = nc * W + threadIdx.x;
    float ri = r[i], si = s[i];
    float roi = odd ? ra[i] : 0.0f, pi = odd ? p[i] : 0.0f, yi = odd ? y[i] : 0.0f;
    ir_elem(first, odd, pair0, a, b, ap, bp, ri, roi, w[i], pi, si, yi, acc);
    rn[i] = ri;
    s[i] = si;
    if (odd) {
      p[i] = pi;
      y[i] = yi;
    }
  }
  double v[1] = {acc};
  block_partials_n<1, 0u, kWideThreads>(v, pg_out);
  if (lead) {
    copy_state(st_next, st_cur);
    st_next->gamma_prev = gamma;
    st_next->alpha_prev = alpha;
    st_next->beta_prev = beta;
    st_next->gamma = gamma;
    st_next->delta = delta;
    st_next->numIter = it + 1;
    if (hout) copy_state(hout, st_next);
  }
}

// x += s_i d_i: the scaled correction back in the unscaled unknowns.
// x += S y; with cs, only after an inner solve that ran at least one
// iteration to its end (y holds that solve's increment).  fin: the inner
// solve's final state.  After an odd number m of iterations (k_ir_update
// forms p and y two iterations at a time) the last iteration's p and
// alpha p are still to be formed, as ir_elem would: p_{m-1} =
// fma(beta, p_{m-2}, r_{m-1}) with r_{m-1} in ra (r_0 when m == 1), then
// y = fma(alpha, p_{m-1}, y) (alpha p_0 when m == 1).
__global__ void k_ir_add_scaled(double* __restrict__ x, const double* __restrict__ sc,
                                const float* __restrict__ y, const float* __restrict__ p,
                                const float* __restrict__ ra, uint32_t n, const IrState* cs,
                                const IrState* fin) {
  if (cs && (ir_inner_open(cs) || cs->numIter == 0)) return;
  const uint32_t m = fin->numIter;
  const bool pend = (m & 1u) != 0;
  const float ap = (float)fin->alpha_prev, bp = (float)fin->beta_prev;
  const uint32_t stride = gridDim.x * kThreads;
  for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) {
    float yi = m == 1 ? 0.0f : y[i];
    if (pend) {
      const float pm = m == 1 ? ra[i] : fma(bp, p[i], ra[i]);
      yi = m == 1 ? ap * pm : fma(ap, pm, yi);
    }
    x[i] = fma(sc[i], (double)yi, x[i]);
  }
}

}  // namespace

CgPrm make_prm(const fx_params& p) {
  CgPrm c{};
  c.maxIter = p.maxIter;
  c.minIter = p.minIter;
  c.savIter = p.savIter;
  c.relTol = p.relTol;
  c.absTol = p.absTol;
  c.solStagTol = 1e-4 * p.relTol;
  return c;
}

int ws_alloc(FastWs& W, fx_ctx ctx, int dt, uint32_t n, uint32_t hist_cap, size_t state_bytes,
             uint32_t spmv_blocks) {
  W.ctx = ctx;
  W.dt = dt;
  W.n = n;
  const size_t vb = pack_round(n) * (dt == FX_F64 ? 8 : 4);  // whole 16-byte packs
  FX_HIP(hipMalloc(&W.x, vb));
  FX_HIP(hipMalloc(&W.x2, vb));
  FX_HIP(hipMalloc(&W.r, vb));
  FX_HIP(hipMalloc(&W.p, vb));
  FX_HIP(hipMalloc(&W.q, vb));
  FX_HIP(hipMalloc(&W.safe, vb));
  FX_HIP(hipMemsetAsync(W.safe, 0, vb, ctx->stream));
  const size_t np = 8 * (size_t)((spmv_blocks > (uint32_t)kMaxRedBlocks) ? spmv_blocks : kMaxRedBlocks);
  W.part2_off = np;
  FX_HIP(hipMalloc(&W.partials, sizeof(double) * (np + 8 * (size_t)kMaxRedBlocks)));
  FX_HIP(hipMalloc(&W.state, 2 * state_bytes));  // two parity buffers
  FX_HIP(hipMemsetAsync(W.state, 0, 2 * state_bytes, ctx->stream));
  FX_HIP(hipMalloc(&W.side, 64));
  FX_HIP(hipHostMalloc(&W.hstate, 2 * state_bytes, hipHostMallocDefault));
  W.hist_cap = hist_cap + 2;
  FX_HIP(hipMalloc(&W.hist, sizeof(double) * W.hist_cap));
  FX_HIP(hipMalloc(&W.dscal, sizeof(double) * 8));
  FX_HIP(hipEventCreateWithFlags(&W.ev[0], hipEventDisableTiming));
  FX_HIP(hipEventCreateWithFlags(&W.ev[1], hipEventDisableTiming));
  return 0;
}

namespace {

// ---------------------------------------------------------------------------
// CG
// ---------------------------------------------------------------------------
template <class VT>
struct CgRun {
  fx_solver S;
  FastWs& W;
  fx_csr A;
  const VT* dinv;  // null: identity preconditioner
  CgPrm P;
  hipStream_t st;
  uint32_t sweeps = 1;  // Jacobi preconditioner sweeps (z buffers W.s, W.t when > 1)
  fx_op op = nullptr;   // a callback operator in place of A

  // state buffer of parity c (iteration c of a replay reads c & 1)
  CgState* dstate(int c = 0) const { return (CgState*)W.state + (c & 1); }
  double* part1() const { return W.partials; }
  double* part2() const { return W.partials + W.part2_off; }

  template <bool JAC>
  int launch_iteration_t(int c, CgState* hout) {
    VT* p = (VT*)W.p;
    VT* q = (VT*)W.q;
    VT* x = (VT*)W.x;
    VT* r = (VT*)W.r;
    const uint32_t n = W.n;
    const int cus = W.ctx->sms;
    FX_TRY((fused_spmv<VT, ROW_AX, 1, 0u>(A, op, W, st, XLoad<VT>{p}, BNone{},
                                          EpiCgQ<VT>{q, p, dstate(c)}, part1())));
    const unsigned g1 = fused_grid<VT>(A, op, n, cus);
    const unsigned g = wide_blocks(n, Pack16<VT>::W, cus, kPassPacks, kCgPerCu);
    hipLaunchKernelGGL((k_cg_update_w<VT, JAC>), dim3(g), dim3(kWideThreads), 0, st, x, r, p, q,
                       dinv, n, (const CgState*)dstate(c), (const double*)part1(), g1,
                       (CgSide*)W.side, part2());
    FX_HIP(hipGetLastError());
    const unsigned g3 = wide_blocks(n, Pack16<VT>::W, cus, kPPacks, kK3PerCu);
    hipLaunchKernelGGL((k_cg_pupdate_w<VT, JAC>), dim3(g3), dim3(kWideThreads), 0, st, p, r, dinv,
                       n, dstate(c), dstate(c + 1), hout, (const double*)part2(), g,
                       (const CgSide*)W.side, P, W.hist);
    FX_HIP(hipGetLastError());
    return 0;
  }

  // The iteration with a Jacobi preconditioner of `sweeps` > 1 sweeps: K1,
  // K2 (with max |z1|), then per further sweep a max-norm finisher and the
  // sweep's SpMV (z in W.s / W.t alternately; partials in part1, which K2
  // has consumed), then K3 with the last z and its (z, r).
  int launch_iteration_sweeps(int c, CgState* hout) {
    VT* p = (VT*)W.p;
    VT* q = (VT*)W.q;
    VT* x = (VT*)W.x;
    VT* r = (VT*)W.r;
    const uint32_t n = W.n;
    const int cus = W.ctx->sms;
    FX_TRY((fused_spmv<VT, ROW_AX, 1, 0u>(A, op, W, st, XLoad<VT>{p}, BNone{},
                                          EpiCgQ<VT>{q, p, dstate(c)}, part1())));
    const unsigned g1 = fused_grid<VT>(A, op, n, cus);
    const unsigned g = wide_blocks(n, Pack16<VT>::W, cus, kPassPacks, kCgPerCu);
    hipLaunchKernelGGL((k_cg_update_w<VT, true, true>), dim3(g), dim3(kWideThreads), 0, st, x, r,
                       p, q, dinv, n, (const CgState*)dstate(c), (const double*)part1(), g1,
                       (CgSide*)W.side, part2());
    FX_HIP(hipGetLastError());
    VT* zb[2] = {(VT*)W.s, (VT*)W.t};
    const VT* zc = nullptr;  // z1 = r .* dinv, not stored
    double* zinf = W.dscal;
    for (uint32_t j = 1; j < sweeps; ++j) {
      const bool last = j + 1 == sweeps;
      // ||z_j||_inf: K2's slot 5 for z1, else the previous sweep's partials
      if (j == 1)
        FX_TRY((launch_finish<1, 1u>(st, part2() + 5 * (size_t)g, g, FinStore<1>{zinf, 0})));
      else
        FX_TRY((launch_finish<1, 1u>(st, part1(), g1, FinStore<1>{zinf, 0})));
      VT* zn = zb[(j - 1) & 1];
      const CgState* cs = dstate(c);
      if (j == 1 && last)
        FX_TRY((launch_spmv<VT, VT, ROW_B_MINUS_AX, 1, 0u>(
            A, st, XJacZ<VT>{r, dinv}, BLoad<VT>{r}, EpiCgSweep<VT, true>{zn, zc, r, dinv, zinf, cs},
            part1(), nullptr, nullptr, false)));
      else if (j == 1)
        FX_TRY((launch_spmv<VT, VT, ROW_B_MINUS_AX, 1, 1u>(
            A, st, XJacZ<VT>{r, dinv}, BLoad<VT>{r},
            EpiCgSweep<VT, false>{zn, zc, r, dinv, zinf, cs}, part1(), nullptr, nullptr, false)));
      else if (last)
        FX_TRY((launch_spmv<VT, VT, ROW_B_MINUS_AX, 1, 0u>(
            A, st, XLoad<VT>{zc}, BLoad<VT>{r}, EpiCgSweep<VT, true>{zn, zc, r, dinv, zinf, cs},
            part1(), nullptr, nullptr, false)));
      else
        FX_TRY((launch_spmv<VT, VT, ROW_B_MINUS_AX, 1, 1u>(
            A, st, XLoad<VT>{zc}, BLoad<VT>{r}, EpiCgSweep<VT, false>{zn, zc, r, dinv, zinf, cs},
            part1(), nullptr, nullptr, false)));
      zc = zn;
    }
    const unsigned g3 = wide_blocks(n, Pack16<VT>::W, cus, kPPacks, kK3PerCu);
    hipLaunchKernelGGL((k_cg_pupdate_w<VT, true, true>), dim3(g3), dim3(kWideThreads), 0, st, p,
                       zc, dinv, n, dstate(c), dstate(c + 1), hout, (const double*)part2(), g,
                       (const CgSide*)W.side, P, W.hist, (const double*)part1(), g1);
    FX_HIP(hipGetLastError());
    return 0;
  }

  int launch_iteration(int c, CgState* hout) {
    if (dinv && sweeps > 1) return launch_iteration_sweeps(c, hout);
    return dinv ? launch_iteration_t<true>(c, hout) : launch_iteration_t<false>(c, hout);
  }
  CgState* hslot(int slot) const { return (CgState*)W.hstate + slot; }

  int read_state(CgState& s) {
    FX_HIP(hipStreamSynchronize(st));
    FX_HIP(hipMemcpy(&s, W.state, sizeof(CgState), hipMemcpyDeviceToHost));
    return 0;
  }
  int write_state(const CgState& s) {
    FX_HIP(hipMemcpyAsync(W.state, &s, sizeof(CgState), hipMemcpyHostToDevice, st));
    FX_HIP(hipStreamSynchronize(st));
    return 0;
  }
  int write_hist(uint32_t it, double v, CgState& s) {
    if (it >= W.hist_cap) return 0;
    FX_HIP(hipMemcpy(W.hist + it, &v, sizeof(double), hipMemcpyHostToDevice));
    s.histCount = it + 1;
    return 0;
  }

  // The rare branches, stated as in src/CG.cxx:104-234, from the state the
  // device left.  Returns 1 if the solve ends, 0 to resume the device loop.
  int host_step(CgState& s, fx_vec vb, fx_vec vx, fx_vec vr, fx_vec vp, fx_vec vq, fx_vec vsafe,
                Ops& o, const fx_params& PP) {
    const double solStagTol = P.solStagTol;
    if (s.reason == R_BREAKDOWN) {
      o.axpy(s.alpha, vp, vx);
      o.axpy(-s.alpha, vq, vr);
      s.resAbs = o.norm2(vr);
      s.resRel = s.resAbs / s.denAbs;
      if (s.resRel > PP.relTol && s.resAbs > PP.absTol) s.err = FX_ERROR_DIVIDE_ZERO;
      return 1;
    }
    const uint32_t it = s.numIter;
    double resAbs = s.resAbs, resRel = s.resRel;
    if (it >= PP.minIter) {
      resAbs = std::sqrt(s.rr);
      resRel = resAbs / s.denAbs;
      const double ratio = resAbs / s.resAbsOld;
      s.ratio = ratio;
      if (it >= PP.savIter && resAbs < s.resAbsOld) o.copy(vx, vsafe);
      if (ratio > kChkRatio) {
        if (s.xinf < kCloseZero) {
          s.err = FX_ERROR_SOLVER_SOLSTAG;
          s.resAbs = resAbs;
          s.resRel = resRel;
          return 1;
        }
        const double xRelDiff = std::fabs(s.alpha) * std::sqrt(s.pp) / std::sqrt(s.xx);
        if (s.stagStep <= kMaxStag && xRelDiff < solStagTol) {
          o.true_residual(vb, vx, vr);
          resAbs = o.norm2(vr);
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
          o.set(vp, 0.0);
          ++s.stagStep;
        }
      }
      if (resRel < PP.relTol) {
        o.true_residual(vb, vx, vr);
        resAbs = o.norm2(vr);
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
        o.set(vp, 0.0);
        ++s.moreStep;
      }
      s.resAbs = resAbs;
      s.resRel = resRel;
    }
    if (it < PP.maxIter) {
      s.resAbsOld = s.resAbs;
      // z = P(r) into q (q is free until the next SpMV), tmpb = (z, r)
      o.set(vq, 0.0);
      o.precond(vr, vq);
      const double tmpb = o.dot(vq, vr);
      s.beta = tmpb / s.tmpa;
      s.tmpa = tmpb;
      o.xpay(s.beta, vq, vp);
    } else {
      return 1;
    }
    // loop top
    if (it == PP.minIter) {
      const double ra = o.norm2(vr);
      s.resAbs = ra;
      s.resAbsOld = ra;
      s.denAbs = (kCloseZero > ra) ? kCloseZero : ra;
      s.resRel = ra / s.denAbs;
      if (s.resRel < PP.relTol || ra < PP.absTol) return 1;
    }
    if (it >= PP.minIter) FX_TRY(write_hist(it, s.resAbs, s));
    s.status = ST_RUN;
    s.reason = 0;
    return 0;
  }

  // x_user (in/out), b_user: device pointers of dtype VT.
  int run(const VT* b, VT* x_user, fx_result* res, const fx_params& PP, int pc_dtype) {
    const uint32_t n = W.n;
    FX_HIP(hipMemcpyAsync(W.x, x_user, sizeof(VT) * n, hipMemcpyDeviceToDevice, st));
    CgState s0{};
    s0.resAbs = 1.0;
    s0.resRel = 1.0;
    s0.denAbs = 1.0;
    s0.ratio = 0.0;
    s0.resAbsOld = 1.0;
    s0.alpha = 0.0;
    s0.status = ST_RUN;
    FX_TRY(write_state(s0));
    // r = b - A x;  z = P r;  p = z;  tmpa = (z, r);  loop top
    FX_TRY((fused_spmv<VT, ROW_AX, 0, 0u>(A, op, W, st, XLoad<VT>{(const VT*)W.x}, BNone{},
                                          EpiBMinusW<VT>{(VT*)W.r, b}, nullptr)));
    const unsigned g = stream_blocks(n);
    if (dinv && sweeps > 1) {
      // z = P r from z = 0 with every sweep (src/CG.cxx:74-75), into W.s
      fx_vec vr0 = nullptr, vz0 = nullptr;
      int rc = fx_vec_wrap(W.ctx, n, pc_dtype, W.r, &vr0);
      if (!rc) rc = fx_vec_wrap(W.ctx, n, pc_dtype, W.s, &vz0);
      if (!rc) rc = fx_vec_set(vz0, 0.0);
      if (!rc) rc = fx_pc_apply(W.ctx, S->pc, vr0, vz0);
      fx_vec_destroy(vr0);
      fx_vec_destroy(vz0);
      FX_TRY(rc);
      hipLaunchKernelGGL((k_cg_init_z<VT>), dim3(g), dim3(kThreads), 0, st, (VT*)W.p,
                         (const VT*)W.r, (const VT*)W.s, n, W.partials);
    } else if (dinv) {
      hipLaunchKernelGGL((k_cg_init<VT, true>), dim3(g), dim3(kThreads), 0, st, (VT*)W.p,
                         (const VT*)W.r, dinv, n, W.partials);
    } else {
      hipLaunchKernelGGL((k_cg_init<VT, false>), dim3(g), dim3(kThreads), 0, st, (VT*)W.p,
                         (const VT*)W.r, dinv, n, W.partials);
    }
    FX_HIP(hipGetLastError());
    FX_TRY((launch_finish<2, 0u>(st, W.partials, g, FinCgInit{dstate(), P, W.hist})));

    // Device-vector views for the host branches
    Ops o(W.ctx, A, op, S->pc, pc_dtype);
    fx_vec vb = nullptr, vx = nullptr, vr = nullptr, vp = nullptr, vq = nullptr, vs = nullptr;
    const int dt = pc_dtype;
    Ops::chk(fx_vec_wrap(W.ctx, n, dt, (void*)b, &vb));
    o.owned.push_back(vb);
    Ops::chk(fx_vec_wrap(W.ctx, n, dt, W.x, &vx));
    o.owned.push_back(vx);
    Ops::chk(fx_vec_wrap(W.ctx, n, dt, W.r, &vr));
    o.owned.push_back(vr);
    Ops::chk(fx_vec_wrap(W.ctx, n, dt, W.p, &vp));
    o.owned.push_back(vp);

    Ops::chk(fx_vec_wrap(W.ctx, n, dt, W.q, &vq));
    o.owned.push_back(vq);
    Ops::chk(fx_vec_wrap(W.ctx, n, dt, W.safe, &vs));
    o.owned.push_back(vs);

    CgState s{};
    // A replay that finds the loop running advances it by kChunk iterations,
    // so a stretch of replays between host steps stops within maxIter /
    // kChunk replays; every host step is preceded by at least one iteration,
    // so there are at most maxIter of them.
    const uint64_t max_chunks = (uint64_t)PP.maxIter / kChunk + 4;
    uint64_t host_steps = 0;
    for (;;) {
      ReplayPace pace;
      pace.hint = host_steps == 0 ? W.last_iters : 0;
      // the iteration that hands the loop's end to the host is numbered
      // numIter - 1: a prediction of hint iterations is hint launches
      const CgState* hs = nullptr;
      FX_TRY(replay_loop<CgState>(
          W, st, pace, pace.hint, max_chunks + 2,
          [&](int c, int slot) { return launch_iteration(c, slot >= 0 ? hslot(slot) : nullptr); },
          [&](const CgState& h) {
            const double target = std::max(PP.relTol * h.denAbs, PP.absTol);
            return pace.more(h.resAbs, target, h.numIter, PP.minIter == 0 ? h.denAbs : 0.0);
          },
          &hs, "CG"));
      FX_CHECK(++host_steps <= (uint64_t)PP.maxIter + 2, FX_ERROR_DEVICE,
               "CG device loop did not stop");
      FX_TRY(read_state(s));
      if (s.status == ST_DONE) break;
      // ST_HOST
      const int fin = host_step(s, vb, vx, vr, vp, vq, vs, o, PP);
      if (fin) break;
      FX_TRY(write_state(s));
    }
    res->numIter = s.numIter;
    W.last_iters = host_steps == 1 ? s.numIter : 0;  // a prediction only for plain runs
    if (!(s.numIter == PP.minIter && s.numIter == PP.maxIter)) {
      res->norm2 = s.resAbs;
      res->normInf = o.norminf(vr);
    }
    if (s.numIter > PP.savIter) o.copy(vs, vx);
    FX_HIP(hipMemcpyAsync(x_user, W.x, sizeof(VT) * n, hipMemcpyDeviceToDevice, st));
    // history
    const uint32_t hc = s.histCount < W.hist_cap ? s.histCount : W.hist_cap;
    S->hist_host.assign(hc, 0.0);
    if (hc) FX_HIP(hipMemcpy(S->hist_host.data(), W.hist, sizeof(double) * hc, hipMemcpyDeviceToHost));
    S->hist_count = hc;
    FX_HIP(hipStreamSynchronize(st));
    return s.err;
  }
};

// ---------------------------------------------------------------------------
// Jacobi solver
// ---------------------------------------------------------------------------
int jacobi_run(fx_solver S, FastWs& W, const double* b_user, double* x_user, fx_result* res) {
  const fx_params& PP = S->prm;
  const CgPrm P = make_prm(PP);
  hipStream_t st = W.ctx->stream;
  fx_csr A = S->A;
  const uint32_t n = W.n;
  JacLagState* ds = (JacLagState*)W.state;  // two parity buffers
  double* xbuf[3] = {(double*)W.x, (double*)W.x2, (double*)W.r};
  int* flags = (int*)W.side;
  const unsigned G = spmv_grid_of<double>(A, W.ctx->sms);
  double* part[2] = {W.partials, W.partials + 2 * (size_t)G};
  FX_HIP(hipMemcpyAsync(xbuf[0], x_user, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
  // the replay graphs read b from the workspace: a later solve may pass
  // another vector
  const double* b = (const double*)W.p;
  FX_HIP(hipMemcpyAsync(W.p, b_user, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
  JacLagState s0{};
  s0.resAbs = 1.0;
  s0.resRel = 1.0;
  s0.denAbs = 1.0;
  s0.status = P.maxIter == 0 ? ST_DONE : ST_RUN;
  FX_HIP(hipMemcpyAsync(ds, &s0, sizeof(JacLagState), hipMemcpyHostToDevice, st));
  FX_TRY(norminf_to(st, xbuf[0], n, W.dscal, W.partials));
  hipLaunchKernelGGL(k_jac_flags, dim3(1), dim3(1), 0, st, flags, (const double*)W.dscal);
  FX_HIP(hipGetLastError());
  const double* dinv = S->pc->dinv;
  // sweep c: its buffers rotate with c % 3 (x) and c % 2 (state), so every
  // replay holds a multiple of 6 sweeps and starts on x buffer 0 and state
  // buffer 0; the last sweep of a replay is followed by the copy of the
  // state to the replay's host slot
  auto sweep = [&](int c, int slot) -> int {
    const EpiJacobiLag e{xbuf[(c + 1) % 3], xbuf[c % 3], b, dinv, flags + c % 3,
                         flags + (c + 1) % 3, flags + (c + 2) % 3, ds + c % 2, ds + (c + 1) % 2,
                         part[(c + 1) % 2], G, P, (uint32_t)((c + 2) % 3), (uint32_t)(c % 3),
                         W.hist, 0};
    FX_TRY((launch_spmv<double, double, ROW_B_MINUS_AX, 2, 2u>(
        A, st, XLoad<double>{xbuf[c % 3]}, BLoad<double>{b}, e, part[c % 2], nullptr, nullptr,
        false)));
    if (slot >= 0) {
      hipLaunchKernelGGL(k_jac_report, dim3(1), dim3(64), 0, st, (const JacLagState*)ds,
                         (JacLagState*)W.hstate + slot);
      FX_HIP(hipGetLastError());
    }
    return 0;
  };
  // Sweep m decides sweep m - 1, so the loop launches maxIter + 1 sweeps at
  // most, and a run that met the tolerance after numIter sweeps launched
  // numIter + 2: the last run's count predicts this one.
  const uint64_t max_launch = (uint64_t)PP.maxIter + 1;
  const uint64_t need =
      W.last_iters ? std::min<uint64_t>((uint64_t)W.last_iters + 2, max_launch) : max_launch;
  ReplayPace pace;  // two replays in flight while the loop may run on
  const JacLagState* hs = nullptr;
  FX_TRY(replay_loop<JacLagState>(
      W, st, pace, need, max_launch / kJacPeriod + 4, sweep,
      [](const JacLagState&) { return true; }, &hs, "Jacobi", kJacPeriod, max_launch));
  // the state the loop stopped on, from the pinned slot of the replay that
  // reported it: the stopping sweep writes it to both state buffers, and
  // the sweeps after it skip, so it is final
  const JacLagState s = *hs;
  FX_CHECK(s.answer < 3, FX_ERROR_DEVICE, "Jacobi answer buffer %u", s.answer);
  double* ans = xbuf[s.answer];
  res->numIter = s.numIter;
  W.last_iters = s.numIter;
  // the final residual, the answer and the history, queued behind the loop
  // with one wait for all of them
  const bool norms = !(s.numIter == PP.minIter && s.numIter == PP.maxIter);
  if (norms) {
    FX_TRY(norminf_to(st, ans, n, W.dscal, W.partials));
    FX_TRY((launch_spmv<double, double, ROW_B_MINUS_AX, 2, 2u>(
        A, st, XLoad<double>{ans}, BLoad<double>{b}, EpiResNorms{b, W.dscal, W.dscal + 1},
        W.partials)));
  }
  FX_HIP(hipMemcpyAsync(x_user, ans, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
  double h[3] = {0.0, 0.0, 0.0};
  if (norms) FX_HIP(hipMemcpyAsync(h, W.dscal, sizeof(double) * 3, hipMemcpyDeviceToHost, st));
  const uint32_t hc = s.histCount < W.hist_cap ? s.histCount : W.hist_cap;
  S->hist_host.assign(hc, 0.0);
  if (hc)
    FX_HIP(hipMemcpyAsync(S->hist_host.data(), W.hist, sizeof(double) * hc,
                          hipMemcpyDeviceToHost, st));
  S->hist_count = hc;
  FX_HIP(hipStreamSynchronize(st));
  if (norms) {
    res->norm2 = h[1];
    res->normInf = h[2];
  }
  return FX_SUCCESS;
}

// The refinement's scaled inner solve (k_ir_update), replayed in graphs of
// kChunk iterations on an f32 workspace: r = W.r (even iterations' r) and
// W.s (odd iterations'), w = W.q, p = W.p, s = W.x2, y = W.x.  It starts
// from the state FinIrInit wrote to state buffer 0 and r from
// EpiResidScaled (in W.r).
struct IrInner {
  FastWs& W;
  fx_csr As;  // A's structure with the scaled f32 values
  hipStream_t st;

  IrState* dstate(int c) const { return (IrState*)W.state + (c & 1); }
  IrState* hslot(int slot) const { return (IrState*)W.hstate + slot; }
  // partials of (r, r): written by iteration c into buffer (c + 1) & 1
  double* pg(int c) const { return W.partials + W.part2_off + (size_t)(c & 1) * kMaxRedBlocks; }

  int launch_iteration(int c, IrState* hout) {
    float* ra = (float*)W.r;
    float* rb = (float*)W.s;
    float* r = (c & 1) ? rb : ra;  // replays start at even iterations
    float* w = (float*)W.q;
    FX_TRY((launch_spmv<float, float, ROW_AX, 1, 0u>(As, st, XLoad<float>{r}, BNone{},
                                                     EpiIrW{w, r, dstate(c)}, W.partials,
                                                     nullptr, nullptr, false)));
    const unsigned g1 = spmv_grid_of<float>(As, W.ctx->sms);
    const unsigned g = wide_blocks(W.n, Pack16<float>::W, W.ctx->sms, kIrPacks, kIrPerCu);
    hipLaunchKernelGGL(k_ir_update, dim3(g), dim3(kWideThreads), 0, st, ra, rb, (const float*)w,
                       (float*)W.p, (float*)W.x2, (float*)W.x, W.n, dstate(c), dstate(c + 1),
                       hout, (const double*)W.partials, g1, (const double*)pg(c), pg(c + 1));
    FX_HIP(hipGetLastError());
    return 0;
  }

  // Replays until the device loop stops; out = the final state.  A replay
  // that finds the loop running advances it, so it stops within maxIter /
  // kChunk + 1 replays after the predicted ones.
  // hint: the iterations this inner solve is predicted to take (0: none).
  // Iteration hint launches the step that finds the loop stopped, so the
  // loop is predicted to launch hint + 1; the replay predicted to end it is
  // cut to that count (rounded up to even).
  // behind / behind_ran: see replay_loop.
  int run(uint32_t maxIter, IrState& out, uint32_t hint,
          const std::function<int()>& behind = nullptr, bool* behind_ran = nullptr) {
    ReplayPace pace;
    pace.hint = hint;
    const uint64_t need = hint ? (uint64_t)hint + 1 : 0;
    const IrState* hs = nullptr;
    FX_TRY(replay_loop<IrState>(
        W, st, pace, need, (uint64_t)maxIter / kChunk + 4 + need / 2,
        [&](int c, int slot) { return launch_iteration(c, slot >= 0 ? hslot(slot) : nullptr); },
        [&](const IrState& h) { return pace.more(h.gamma, h.thr, h.numIter, h.gamma0); },  // squared
        &hs, "IR inner", 2, UINT64_MAX, behind, behind_ran));
    out = *hs;
    return 0;
  }
};

}  // namespace

bool fast_eligible(fx_solver S) {
  if (fast_disabled()) return false;
  if (S->op) {
    // a callback operator that only enqueues device work, with the identity
    // or a one-sweep Jacobi preconditioner (its diagonal from its own matrix)
    if (!S->op->device_only || S->op->nrow == 0 || S->op->nrow != S->op->mcol) return false;
    if (S->type == FX_SOLVER_JACOBI) return false;
    return S->pc->kind == 0 ||
           (S->pc->kind == 1 && S->pc->iters == 1 && S->pc->A->nrow == S->op->nrow);
  }
  if (S->A->nrow == 0) return false;
  const bool pc_ok =
      S->pc->kind == 0 || (S->pc->kind == 1 && S->pc->iters == 1 && S->pc->A == S->A);
  // CG also takes a Jacobi preconditioner of several sweeps
  const bool pc_cg = pc_ok || (S->pc->kind == 1 && S->pc->iters > 1 && S->pc->A == S->A);
  if (S->type == FX_SOLVER_CG) return pc_cg;
  if (S->type == FX_SOLVER_BICGSTAB) return pc_ok;
  if (S->type == FX_SOLVER_GMRES || S->type == FX_SOLVER_FGMRES) return pc_ok;
  if (S->type == FX_SOLVER_JACOBI) return S->pc->A == S->A;
  return false;
}

int fast_solve(fx_solver S, fx_vec b, fx_vec x, fx_result* res) {
  // a workspace is sized for one matrix: never run it on another size
  const uint32_t n = S->op ? S->op->nrow : S->A->nrow;
  if (S->fast && S->fast->n != n) S->release_fast();
  if (!S->fast) {
    S->fast = new FastWs();
    const size_t sb = S->type == FX_SOLVER_JACOBI     ? sizeof(JacLagState)
                      : S->type == FX_SOLVER_BICGSTAB ? bicgstab_state_bytes()
                      : (S->type == FX_SOLVER_GMRES || S->type == FX_SOLVER_FGMRES)
                          ? gmres_state_bytes()
                                                      : sizeof(CgState);
    int rc = ws_alloc(*S->fast, S->ctx, FX_F64, n, S->prm.maxIter, sb,
                      S->op ? op_grid(n) : spmv_grid_of<double>(S->A, S->ctx->sms));
    if (!rc && S->op) {  // the operator's work vectors, before any graph is captured
      const size_t vb = pack_round(n) * sizeof(double);
      if (hipMalloc(&S->fast->opy, vb) != hipSuccess || hipMalloc(&S->fast->opx, vb) != hipSuccess) {
        set_error("hipMalloc of %zu bytes failed", vb);
        rc = FX_ERROR_ALLOC_MEM;
      }
    }
    if (rc) {
      delete S->fast;
      S->fast = nullptr;
      return rc;
    }
  }
  FastWs& W = *S->fast;
  if (S->type == FX_SOLVER_JACOBI) return jacobi_run(S, W, (const double*)b->d, (double*)x->d, res);
  if (S->type == FX_SOLVER_BICGSTAB)
    return bicgstab_fast(S, W, (const double*)b->d, (double*)x->d, res);
  if (S->type == FX_SOLVER_GMRES || S->type == FX_SOLVER_FGMRES)
    return gmres_fast(S, W, (const double*)b->d, (double*)x->d, res);
  CgRun<double> run{S, W, S->A, S->pc->kind == 1 ? S->pc->dinv : nullptr, make_prm(S->prm),
                    S->ctx->stream};
  run.op = S->op;
  if (S->pc->kind == 1 && S->pc->iters > 1) {
    run.sweeps = S->pc->iters;
    const size_t vb = pack_round(W.n) * sizeof(double);
    if (!W.s) FX_HIP(hipMalloc(&W.s, vb));
    if (!W.t) FX_HIP(hipMalloc(&W.t, vb));
  }
  return run.run((const double*)b->d, (double*)x->d, res, S->prm, FX_F64);
}

}  // namespace fx

void fx_solver_s::release_fast() {
  delete fast;
  fast = nullptr;
}

using namespace fx;

// ---------------------------------------------------------------------------
// Mixed-precision iterative refinement
// ---------------------------------------------------------------------------
struct fx_ir_s {
  fx_ctx ctx = nullptr;
  int inner_type = FX_SOLVER_CG;
  fx_csr A = nullptr;
  fx_pc pc = nullptr;
  fx_solver inner_s = nullptr;
  fx_params outer{}, inner{};
  float* r32 = nullptr;
  float* d32 = nullptr;
  double* dsc = nullptr;
  double* part = nullptr;
  FastWs* W = nullptr;
  // scaled inner solve (CG inner, positive diagonal): s_i = 1 / sqrt(a_ii),
  // the f32 values of S A S, and a matrix handle that shares A's structure
  bool scaled = false;
  double* sc = nullptr;
  float* vs32 = nullptr;
  float* vs32o = nullptr;  // vs32 in a sorted matrix's row order
  std::vector<uint32_t> inner_counts;  // each outer step's inner iterations, last solve
  uint32_t outer_last = 0;             // outer steps of the last solve (0: none)
  IrOuterRec* hrec = nullptr;          // pinned
  double* sco = nullptr;  // sorted matrices: S and b in the stored row order
  double* bo = nullptr;
  int* bad = nullptr;
  uint32_t* blk512 = nullptr;  // row-pattern matrices: As's own 512-row blocks
  uint16_t* rpb512 = nullptr;
  fx_csr_s As{};
  ~fx_ir_s() {
    if (ctx) (void)hipStreamSynchronize(ctx->stream);
    delete W;
    fx_solver_destroy(inner_s);
    fx_pc_destroy(pc);
    (void)hipFree(r32);
    (void)hipFree(d32);
    (void)hipFree(dsc);
    if (hrec) (void)hipHostFree(hrec);
    (void)hipFree(part);
    (void)hipFree(sc);
    (void)hipFree(sco);
    (void)hipFree(bo);
    (void)hipFree(vs32);
    (void)hipFree(vs32o);
    (void)hipFree(bad);
    (void)hipFree(blk512);
    (void)hipFree(rpb512);
  }
};

namespace {

bool ir_refcg_forced() {
  const char* e = std::getenv("FASPXX_FX_IR_REFCG");
  return e && e[0] == '1';
}

// Form S and S A S for the scaled inner solve.  Returns 0 with ir->scaled
// set when every diagonal entry is present, positive and finite, 0 with
// ir->scaled clear otherwise, or an error code.
int ir_scale(fx_ir ir) {
  fx_csr A = ir->A;
  const uint32_t n = A->nrow;
  hipStream_t st = ir->ctx->stream;
  FX_HIP(hipMalloc(&ir->sc, sizeof(double) * pack_round(n)));
  FX_HIP(hipMalloc(&ir->vs32, sizeof(float) * (A->nnz ? A->nnz : 1)));
  FX_HIP(hipMalloc(&ir->bad, sizeof(int)));
  FX_HIP(hipMemsetAsync(ir->bad, 0, sizeof(int), st));
  const unsigned g = (n + kThreads - 1) / kThreads;
  hipLaunchKernelGGL(k_ir_scale_diag, dim3(g), dim3(kThreads), 0, st, A->rp, A->ci, A->dp, A->va,
                     ir->sc, ir->bad, n);
  FX_HIP(hipGetLastError());
  int bad = 0;
  FX_HIP(hipMemcpyAsync(&bad, ir->bad, sizeof(int), hipMemcpyDeviceToHost, st));
  FX_HIP(hipStreamSynchronize(st));
  if (bad) {
    (void)hipFree(ir->sc);
    (void)hipFree(ir->vs32);
    ir->sc = nullptr;
    ir->vs32 = nullptr;
    return 0;
  }
  hipLaunchKernelGGL(k_ir_scale_vals, dim3(g), dim3(kThreads), 0, st, A->rp, A->ci, A->va,
                     ir->sc, ir->vs32, n);
  FX_HIP(hipGetLastError());
  ir->As = *A;  // the same device arrays; only va32 differs, and nothing frees through As
  ir->As.va32 = ir->vs32;
  if (A->sorted) {  // the scaled values in the sorted copy's order too
    FX_HIP(hipMalloc(&ir->vs32o, sizeof(float) * (A->nnz ? A->nnz : 1)));
    FX_TRY(sorted_fill_f32(A, ir->vs32, ir->vs32o));
    ir->As.oval32 = ir->vs32o;
    // S, and room for b, in the stored row order for the f64 residual
    FX_HIP(hipMalloc(&ir->sco, sizeof(double) * pack_round(n)));
    FX_HIP(hipMalloc(&ir->bo, sizeof(double) * pack_round(n)));
    hipLaunchKernelGGL(k_gather_rows, dim3(stream_blocks(n)), dim3(kThreads), 0, st, A->omap,
                       (const double*)ir->sc, ir->sco, n);
    FX_HIP(hipGetLastError());
  }
  if (A->pid) {
    // the f32 inner SpMV of a row-pattern matrix runs on 512-row blocks:
    // IR 10.6 against 11.1 ms on the 128^3 stencil (half the partial sums
    // for k_ir_update; the f64 kernels measured slower with them)
    uint32_t nb = 0;
    FX_TRY(csr_pattern_blocks512(A, &ir->blk512, &ir->rpb512, &nb));
    ir->As.blk = ir->blk512;
    ir->As.rpb = ir->rpb512;
    ir->As.nblk = nb;
    ir->As.pat_rows = kPat2Rows;
  }
  ir->scaled = true;
  return 0;
}

}  // namespace

extern "C" {

int fx_ir_create(fx_ctx ctx, int inner_type, fx_csr A, double weight, const fx_params* outer,
                 const fx_params* inner, fx_ir* out) {
  clear_error();
  FX_CHECK(ctx && A && outer && inner && out, FX_ERROR_INPUT_PAR, "null argument");
  FX_CHECK(A->nrow == A->mcol, FX_ERROR_MAT_SIZE, "IR needs a square matrix");
  FX_CHECK(inner_type == FX_SOLVER_CG || inner_type == FX_SOLVER_GMRES, FX_ERROR_SOLVER_TYPE,
           "IR inner solver must be CG or GMRES");
  fx_ir ir = new fx_ir_s();
  ir->ctx = ctx;
  ir->inner_type = inner_type;
  ir->A = A;
  ir->outer = *outer;
  ir->inner = *inner;
  const uint32_t n = A->nrow;
  auto fail = [&](int rc) {
    delete ir;
    return rc;
  };
  int rc = 0;
  if (inner_type == FX_SOLVER_CG && A->has_values && n > 0 && !ir_refcg_forced()) {
    try {
      rc = ir_scale(ir);
    } catch (const OpsError& e) {
      rc = e.code;
    }
    if (rc) return fail(rc);
  }
  if (!ir->scaled) {
    // f32 copy of A for the Jacobi-preconditioned inner solver
    if (!A->va32 && A->has_values) rc = fx_csr_make_f32(A);
    if (!rc) rc = fx_pc_create_jacobi(ctx, A, weight, 1, &ir->pc);
    if (!rc) rc = ensure_dinv32(ir->pc);
    if (!rc) rc = fx_solver_create(ctx, inner_type, &ir->inner_s);
    if (!rc) rc = fx_solver_setup(ir->inner_s, A, ir->pc, inner);
    if (rc) return fail(rc);
  }
  const size_t nb = n ? n : 1;
  const size_t np = 8 * ((size_t)std::max(spmv_grid_of<double>(A), spmv_grid_of<float>(A)) +
                         kMaxRedBlocks);
  if (hipMalloc(&ir->r32, sizeof(float) * nb) || hipMalloc(&ir->d32, sizeof(float) * nb) ||
      hipMalloc(&ir->dsc, sizeof(double) * 4) ||
      hipHostMalloc(&ir->hrec, sizeof(IrOuterRec), hipHostMallocDefault) ||
      hipMalloc(&ir->part, sizeof(double) * np)) {
    set_error("IR workspace allocation failed");
    return fail(FX_ERROR_ALLOC_MEM);
  }
  if (inner_type == FX_SOLVER_CG) {
    ir->W = new FastWs();
    const size_t sb = sizeof(CgState) > sizeof(IrState) ? sizeof(CgState) : sizeof(IrState);
    rc = ws_alloc(*ir->W, ctx, FX_F32, n, inner->maxIter, sb, spmv_grid_of<float>(A, ctx->sms));
    if (rc) return fail(rc);
    // the scaled inner solve's second r buffer (odd iterations' r)
    if (ir->scaled && hipMalloc(&ir->W->s, sizeof(float) * pack_round(n))) {
      set_error("IR workspace allocation failed");
      return fail(FX_ERROR_ALLOC_MEM);
    }
  }
  *out = ir;
  return 0;
}

int fx_ir_solve(fx_ir ir, fx_vec b, fx_vec x, fx_result* result, uint32_t* inner_iters) {
  clear_error();
  FX_CHECK(ir && b && x && result, FX_ERROR_INPUT_PAR, "null argument");
  fx_csr A = ir->A;
  const uint32_t n = A->nrow;
  FX_CHECK(b->n == n && x->n == n, FX_ERROR_NONMATCH_SIZE, "IR size mismatch");
  FX_CHECK(b->dtype == FX_F64 && x->dtype == FX_F64, FX_ERROR_INPUT_PAR, "IR takes f64 b and x");
  fx_ctx ctx = ir->ctx;
  hipStream_t st = ctx->stream;
  hipEvent_t e0, e1;
  FX_HIP(hipEventCreate(&e0));
  FX_HIP(hipEventCreate(&e1));
  FX_HIP(hipEventRecord(e0, st));
  uint32_t outer_it = 0, inner_total = 0;
  double resAbs = 0.0, denAbs = 1.0;
  int rc = 0, err = 0;
  const fx_params& outer = ir->outer;
  // Scaled: x += S y after an inner solve, then r = b - A x (f64 values,
  // Apply + XPAY order), W.r = (float)(S r), ||r||, the outer decision and
  // the next inner solve's starting state.  cs: see EpiResidScaled.
  auto outer_step = [&](bool add, const IrState* cs) -> int {
    if (add) {
      hipLaunchKernelGGL(k_ir_add_scaled, dim3(stream_blocks(n)), dim3(kThreads), 0, st,
                         (double*)x->d, (const double*)ir->sc, (const float*)ir->W->x,
                         (const float*)ir->W->p, (const float*)ir->W->r, n, cs,
                         (const IrState*)ir->W->state);
      FX_HIP(hipGetLastError());
    }
    if (ir->bo)
      FX_TRY((launch_spmv<double, double, ROW_AX, 2, 0u>(
          A, st, XLoad<double>{(const double*)x->d}, BNone{},
          EpiResidScaledSorted{(float*)ir->W->r, ir->bo, ir->sco, cs}, ir->part, nullptr,
          nullptr, false)));
    else
      FX_TRY((launch_spmv<double, double, ROW_AX, 2, 0u>(
          A, st, XLoad<double>{(const double*)x->d}, BNone{},
          EpiResidScaled{(float*)ir->W->r, (const double*)b->d, ir->sc, cs}, ir->part, nullptr,
          nullptr, false)));
    return launch_finish<2, 0u>(
        st, ir->part, spmv_grid_of<double>(A, ctx->sms),
        FinIrInit{ir->hrec, ir->dsc + 1, (IrState*)ir->W->state, ir->inner.relTol,
                  ir->inner.absTol, outer.relTol, outer.absTol, ir->inner.maxIter,
                  outer_it == 0 ? 1 : 0, cs});
  };
  bool step_queued = false;  // the next residual is already queued
  if (ir->scaled && ir->bo) {  // b in the sorted matrix's stored row order
    hipLaunchKernelGGL(k_gather_rows, dim3(stream_blocks(n)), dim3(kThreads), 0, st, A->omap,
                       (const double*)b->d, ir->bo, n);
    if (hipGetLastError() != hipSuccess) rc = FX_ERROR_DEVICE;
  }
  try {
    while (!rc) {
      if (ir->scaled) {
        if (!step_queued) rc = outer_step(false, nullptr);
        step_queued = false;
        if (rc) break;
        // When the last solve went on past this residual, the outer loop is
        // predicted to go on: the inner solve is queued at once, without
        // waiting for the residual, and FinIrInit stops it on the device if
        // the outer loop is in fact done.
        const bool ahead = outer_it < ir->outer_last && outer_it < outer.maxIter;
        if (!ahead) {
          if (hipStreamSynchronize(st)) {
            rc = FX_ERROR_DEVICE;
            break;
          }
          resAbs = ir->hrec->resAbs;
          denAbs = ir->hrec->denAbs;
          if (ir->hrec->done) break;
          if (outer_it >= outer.maxIter) {
            err = FX_ERROR_SOLVER_MAXIT;
            break;
          }
        }
      } else {
        rc = launch_spmv<double, double, ROW_AX, 1, 0u>(
            A, st, XLoad<double>{(const double*)x->d}, BNone{},
            EpiResidF32{ir->r32, (const double*)b->d, ir->dsc}, ir->part);
        if (rc) break;
        double h = 0;
        if (hipMemcpyAsync(&h, ir->dsc, sizeof(double), hipMemcpyDeviceToHost, st) ||
            hipStreamSynchronize(st)) {
          rc = FX_ERROR_DEVICE;
          break;
        }
        resAbs = h;
        if (outer_it == 0) denAbs = (kCloseZero > resAbs) ? kCloseZero : resAbs;
        if (resAbs / denAbs < outer.relTol || resAbs < outer.absTol) break;
        if (outer_it >= outer.maxIter) {
          err = FX_ERROR_SOLVER_MAXIT;
          break;
        }
      }
      ++outer_it;
      if (ir->scaled) {
        // S A S y = S r in f32 from y = 0; x += S y
        IrInner in{*ir->W, &ir->As, st};
        IrState fs{};
        // the prediction: this outer step's inner count in the last solve
        const size_t step = outer_it - 1;
        const uint32_t hint = step < ir->inner_counts.size() ? ir->inner_counts[step] : 0u;
        // the next outer step is queued right behind the inner solve's
        // predicted last replay, to run only if the inner solve has ended
        // by then
        bool behind_ran = false;
        rc = in.run(ir->inner.maxIter, fs, hint,
                    [&]() { return outer_step(true, (const IrState*)ir->W->state); },
                    &behind_ran);
        if (rc) break;
        if (fs.reason == IR_OUTER) {  // queued ahead, but the outer loop was done
          --outer_it;
          break;
        }
        if (ir->inner_counts.size() <= step) ir->inner_counts.resize(step + 1, 0u);
        ir->inner_counts[step] = fs.reason == IR_CONVERGED ? fs.numIter : 0;
        inner_total += fs.numIter;
        if (behind_ran) {
          step_queued = true;
        } else {
          rc = outer_step(fs.numIter > 0, nullptr);
          if (rc) break;
          step_queued = true;
        }
        continue;
      }
      // inner solve A d = r in f32 from d = 0
      if (hipMemsetAsync(ir->d32, 0, sizeof(float) * n, st)) {
        rc = FX_ERROR_DEVICE;
        break;
      }
      fx_result res_in{};
      if (ir->inner_type == FX_SOLVER_CG) {
        CgRun<float> run{ir->inner_s, *ir->W, A, ir->pc->dinv32, make_prm(ir->inner), st};
        const int irc = run.run(ir->r32, ir->d32, &res_in, ir->inner, FX_F32);
        if (irc == FX_ERROR_DEVICE || irc == FX_ERROR_ALLOC_MEM || irc == FX_ERROR_INPUT_PAR) {
          rc = irc;
          break;
        }
      } else {
        Ops o(ctx, A, ir->pc, FX_F32);
        fx_vec vb = nullptr, vx = nullptr;
        Ops::chk(fx_vec_wrap(ctx, n, FX_F32, ir->r32, &vb));
        o.owned.push_back(vb);
        Ops::chk(fx_vec_wrap(ctx, n, FX_F32, ir->d32, &vx));
        o.owned.push_back(vx);
        ir->inner_s->restart_cur = 0;  // every inner solve starts as set up
        (void)ops_gmres(ir->inner_s, o, vb, vx, &res_in);
      }
      inner_total += res_in.numIter;
      hipLaunchKernelGGL(k_add_f32_to_f64, dim3(stream_blocks(n)), dim3(kThreads), 0, st,
                         (double*)x->d, ir->d32, n);
      if (hipGetLastError() != hipSuccess) {
        rc = FX_ERROR_DEVICE;
        break;
      }
    }
  } catch (const OpsError& e) {
    rc = e.code;
  }
  (void)hipEventRecord(e1, st);
  (void)hipEventSynchronize(e1);
  if (ir->scaled) {
    // the last residual's record (the host did not wait for every one)
    if (ir->hrec) resAbs = ir->hrec->resAbs;
    ir->outer_last = (rc || err) ? 0u : outer_it;
  }
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  result->numIter = outer_it;
  result->norm2 = resAbs;
  result->normInf = 0.0;
  result->solve_ms = ms;
  result->retcode = rc ? rc : err;
  if (inner_iters) *inner_iters = inner_total;
  return result->retcode;
}

int fx_ir_destroy(fx_ir ir) {
  delete ir;
  return 0;
}

int fx_solve_ir(fx_ctx ctx, int inner_type, fx_csr A, double weight, const fx_params* outer,
                const fx_params* inner, fx_vec b, fx_vec x, fx_result* result,
                uint32_t* inner_iters) {
  fx_ir ir = nullptr;
  FX_TRY(fx_ir_create(ctx, inner_type, A, weight, outer, inner, &ir));
  const int rc = fx_ir_solve(ir, b, x, result, inner_iters);
  fx_ir_destroy(ir);
  return rc;
}

}  // extern "C"
