// fx_gmres.hip -- fused, graph-replayed restarted GMRES (GMRES::RSolve,
// src/GMRES.cxx:114-280, and GMRES::LSolve, :283-453) with the identity or a
// one-sweep Jacobi preconditioner on the system matrix.
//
// One restart cycle of length m is captured once as a hipGraph and replayed:
//   start     V0 *= 1 / resAbs, g[0] = resAbs                    (k_gm_start)
//   for k = 0 .. m-1 (kernels skip once the cycle has ended):
//     S_k     w = V[k+1] = A P V[k] (right) or P A V[k] (left);
//             partials of (w, V[0])                     (k_spmv + EpiGmW)
//     M_k,j   j = 0..k: h = (w, V[j]) from the partials (every workgroup
//             sums them in the same order), H(j,k) = h, w = w - h V[j];
//             partials of (w, V[j+1]), or of (w, w) after the last
//                                                                (k_gm_mgs)
//     G_k     t = ||w||, w *= 1/t, H(k+1,k) = t; the Givens update and the
//             cycle's stopping test on the device (one thread); at the end of
//             the cycle the triangular solve                     (k_gm_givens)
//   X         x += P (sum_j y_j V[j]) (right) or x += sum_j y_j V[j] (left)
//                                                                (k_gm_update)
//   R         V0 = b - A x (right) / tmp = b - A x, V0 = P tmp (left), with
//             their norms; the state goes to a pinned host slot   (k_spmv + k_finish)
// The host reads the state once per cycle and applies the variable-restart
// rule of :256-265 to choose the next cycle's graph (one per length, made on
// first use).  FGMRES (FGMRES::Solve, src/FGMRES.cxx:108-271) runs the same
// cycle with the preconditioned vectors kept: Z[k] = P(V[k]) from the stale
// Z[k] (one Jacobi sweep, z + D^-1 (v - A z), with the reference's shortcut
// v when ||z||_inf < 1e-20; identity: Z[k] = V[k]), w = A Z[k], and
// x += sum_j y_j Z[j]; ||Z[k]||_inf rides along as a second reduction slot of
// the SpMV that writes Z[k].  Modified Gram-Schmidt keeps its sequential order: one kernel
// per step, as the method needs one reduction per step.  The Hessenberg
// matrix, the rotations and the right-hand side g stay in f64 on the device;
// the arithmetic is the reference's (src/GMRES.cxx:182-219).
#include <cmath>
#include <map>
#include <vector>

#include "fx_fast.hpp"

namespace fx {

namespace {

enum { GM_RUN = 0, GM_CYCLE_END = 1 };

// Grid of the cycle's streaming kernels (MGS steps, Givens, updates): packs
// per thread, workgroups per CU.  GMRES(30) on the 1024^2 upwind system:
// 2 and 2 (512 workgroups) 109.8 us per iteration; 1 and 2, or 2 and 4, the
// same; 1 and 4 (1024 workgroups, each re-summing the partials) 135 us.
#ifndef FX_GM_PER
#define FX_GM_PER 2
#endif
#ifndef FX_GM_PER_CU
#define FX_GM_PER_CU 2
#endif
constexpr int kGmPer = FX_GM_PER;
constexpr int kGmPerCu = FX_GM_PER_CU;

struct GmState {
  double resAbs, rj, denAbs, pad0;
  uint32_t numIter, count, status, histCount;
  // the cycle's last residual estimate and the reference's resAbsOld after
  // it (src/GMRES.cxx:198-201), for the safe-guard test after the x update
  double est, est_old;
};

// Device-resident least-squares data of one cycle: H is (maxR+1) x maxR,
// row-major with row length ld = maxR.
struct GmLsq {
  double* H;
  double* hc;
  double* hs;
  double* g;
  uint32_t ld;
};

struct GmPrm {
  uint32_t maxIter, minIter, m, pad;
  double relTol, absTol;
};

template <class VT>
struct XScaled {  // the SpMV's input P v for right preconditioning: v .* dinv
  const VT* __restrict__ v;
  const VT* __restrict__ dinv;
  __device__ __forceinline__ VT operator()(uint32_t j) const { return v[j] * dinv[j]; }
};

// S_k epilogue: w = (A P v) or P (A v); partials of (w, V0)
template <class VT, bool LEFTJAC>
struct EpiGmW {
  VT* __restrict__ w;
  const VT* __restrict__ v0;
  const VT* __restrict__ dinv;
  const GmState* st;
  __device__ bool skip() const { return st->status != GM_RUN; }
  __device__ __forceinline__ void row(uint32_t i, double v, double (&acc)[1]) const {
    const VT a = (VT)v;
    const VT z = LEFTJAC ? a * dinv[i] : a;
    w[i] = z;
    acc[0] = fma((double)z, (double)v0[i], acc[0]);
  }
  __device__ void fin(double (&)[1]) const {}
};

template <class VT>
__global__ __launch_bounds__(kWideThreads) void k_gm_start(VT* __restrict__ v0, uint32_t n,
                                                           GmState* st, GmLsq L, GmPrm P,
                                                           double* hist) {
  const GmState S = *st;
  const double a = 1.0 / S.resAbs;  // VEC::Scale(1 / resAbs), src/GMRES.cxx:148
  const uint32_t stride = gridDim.x * kWideThreads;
  for (uint32_t i = blockIdx.x * kWideThreads + threadIdx.x; i < n; i += stride)
    v0[i] = v0[i] * (VT)a;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    L.g[0] = S.resAbs;
    GmState N = S;
    N.count = 0;
    N.status = GM_RUN;
    if (S.numIter >= P.maxIter) N.status = GM_CYCLE_END;
    if (N.status == GM_RUN && S.numIter >= P.minIter && hist) {  // :153
      hist[S.numIter] = S.resAbs;
      N.histCount = S.numIter + 1;
    }
    *st = N;
  }
}

// One MGS step: h = (w, V[j]) from the partials, w -= h V[j], partials of
// (w, next) where next = V[j+1], or w itself after the last step.
template <class VT>
__global__ __launch_bounds__(kWideThreads) void k_gm_mgs(
    VT* __restrict__ w, const VT* __restrict__ vj, const VT* __restrict__ next, uint32_t n,
    const GmState* st, const double* __restrict__ pin, unsigned gin, double* __restrict__ pout,
    GmLsq L, uint32_t j, uint32_t k, double* zinf_out) {
  if (st->status != GM_RUN) return;
  using PK = Pack16<VT>;
  constexpr int W = PK::W;
  const uint32_t nc = n / W, stride = gridDim.x * kWideThreads;
  const uint32_t cfirst = blockIdx.x * kWideThreads + threadIdx.x;
  const bool self = next == nullptr;
  // first pass's packs before h is known (vectors are padded to whole packs)
  uint32_t c0 = cfirst < nc ? cfirst : 0u, c1 = cfirst + stride < nc ? cfirst + stride : c0;
  PK wa = ld16(w, c0), va = ld16(vj, c0), wb = ld16(w, c1), vb = ld16(vj, c1);
  PK na = self ? PK{} : ld16(next, c0), nb = self ? PK{} : ld16(next, c1);
  double h[1];
  if (zinf_out) {  // FGMRES, j = 0: the partials also hold max |Z[k]| (slot 1)
    double hz[2];
    partials_thread<2, 2u, kWideThreads>(pin, gin, threadIdx.x, hz);
    partials_combine<2, 2u, kWideThreads>(hz);
    h[0] = hz[0];
    if (blockIdx.x == 0 && threadIdx.x == 0) *zinf_out = hz[1];
  } else {
    partials_thread<1, 0u, kWideThreads>(pin, gin, threadIdx.x, h);
    partials_combine<1, 0u, kWideThreads>(h);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) L.H[(size_t)j * L.ld + k] = h[0];
  // w = AXPBY(1, -h, V[j], w)   (src/GMRES.cxx:170-171; VEC::AXPBY forms)
  const VT b = (VT)(-h[0]);
  const bool bone = -h[0] == 1.0;
  double acc[1] = {0.0};
  for (uint32_t c = cfirst;;) {
    double a0 = 0.0, a1 = 0.0;
#pragma unroll
    for (int e = 0; e < W; ++e) {
      wa.v[e] = bone ? wa.v[e] + va.v[e] : fma(b, va.v[e], wa.v[e]);
      wb.v[e] = bone ? wb.v[e] + vb.v[e] : fma(b, vb.v[e], wb.v[e]);
      const VT xa = self ? wa.v[e] : na.v[e], xb = self ? wb.v[e] : nb.v[e];
      a0 = fma((double)wa.v[e], (double)xa, a0);
      a1 = fma((double)wb.v[e], (double)xb, a1);
    }
    const bool one = c < nc, two = c + stride < nc;
    if (one) st16(w, c0, wa);
    if (two) st16(w, c1, wb);
    acc[0] += (one ? a0 : 0.0) + (two ? a1 : 0.0);
    c += 2 * stride;
    if (c >= nc) break;
    c0 = c;
    c1 = c + stride < nc ? c + stride : c0;
    wa = ld16(w, c0);
    va = ld16(vj, c0);
    wb = ld16(w, c1);
    vb = ld16(vj, c1);
    if (!self) {
      na = ld16(next, c0);
      nb = ld16(next, c1);
    }
  }
  if (blockIdx.x == 0 && threadIdx.x < n - nc * W) {
    const uint32_t i = nc * W + threadIdx.x;
    const VT wi = bone ? w[i] + vj[i] : fma(b, vj[i], w[i]);
    w[i] = wi;
    acc[0] = fma((double)wi, (double)(self ? wi : next[i]), acc[0]);
  }
  block_partials_n<1, 0u, kWideThreads>(acc, pout);
}

// Givens update of column cnt - 1 (src/GMRES.cxx:182-203)
__device__ __forceinline__ void givens_column(const GmLsq& L, uint32_t cnt) {
  const uint32_t c1 = cnt - 1;
  auto H = [&](uint32_t i, uint32_t jj) -> double& { return L.H[(size_t)i * L.ld + jj]; };
  for (uint32_t jj = 1; jj < cnt; ++jj) {
    const double t = H(jj - 1, c1);
    H(jj - 1, c1) = fma(L.hs[jj - 1], H(jj, c1), L.hc[jj - 1] * t);
    H(jj, c1) = fma(-L.hs[jj - 1], t, L.hc[jj - 1] * H(jj, c1));
  }
  double t = H(cnt, c1) * H(cnt, c1);
  t = fma(H(c1, c1), H(c1, c1), t);
  t = sqrt(t);
  const double gamma = t > kSmallDbl ? t : kSmallDbl;
  L.hc[c1] = H(c1, c1) / gamma;
  L.hs[c1] = H(cnt, c1) / gamma;
  H(c1, c1) = fma(L.hs[c1], H(cnt, c1), L.hc[c1] * H(c1, c1));
  L.g[cnt] = -L.hs[c1] * L.g[c1];
  L.g[c1] = L.hc[c1] * L.g[c1];
}

// Upper-triangular solve in place in g (src/GMRES.cxx:213-219, signed index)
__device__ __forceinline__ void triangular_solve(const GmLsq& L, uint32_t cnt) {
  auto H = [&](uint32_t i, uint32_t jj) -> double { return L.H[(size_t)i * L.ld + jj]; };
  const uint32_t c1 = cnt - 1;
  L.g[c1] = L.g[c1] / H(c1, c1);
  for (int kk = (int)cnt - 2; kk >= 0; --kk) {
    double t = 0.0;
    for (uint32_t jj = (uint32_t)kk + 1; jj < cnt; ++jj) t = fma(-H(kk, jj), L.g[jj], t);
    t += L.g[kk];
    L.g[kk] = t / H(kk, kk);
  }
}

// The same Givens update and triangular solve on LDS copies (one thread;
// every dependent step then costs LDS latency instead of a global-memory
// round trip): columns of up to kGivLds rows, triangles up to kTriLds.
constexpr uint32_t kGivLds = 512;
constexpr uint32_t kTriLds = 64;

__device__ __forceinline__ void givens_column_lds(double* __restrict__ col,
                                                  double* __restrict__ hc,
                                                  double* __restrict__ hs, double& gc1,
                                                  double& gcnt, uint32_t cnt) {
  const uint32_t c1 = cnt - 1;
  for (uint32_t jj = 1; jj < cnt; ++jj) {
    const double t = col[jj - 1];
    col[jj - 1] = fma(hs[jj - 1], col[jj], hc[jj - 1] * t);
    col[jj] = fma(-hs[jj - 1], t, hc[jj - 1] * col[jj]);
  }
  double t = col[cnt] * col[cnt];
  t = fma(col[c1], col[c1], t);
  t = sqrt(t);
  const double gamma = t > kSmallDbl ? t : kSmallDbl;
  hc[c1] = col[c1] / gamma;
  hs[c1] = col[cnt] / gamma;
  col[c1] = fma(hs[c1], col[cnt], hc[c1] * col[c1]);
  gcnt = -hs[c1] * gc1;
  gc1 = hc[c1] * gc1;
}

// G_k: t = ||w||, w *= 1/t, H(k+1,k) = t, Givens, stopping test (workgroup
// 0; it stages column k, the rotations and g[k] in LDS while every
// workgroup scales w), and at the end of the cycle the triangular solve.
template <class VT>
__global__ __launch_bounds__(kWideThreads) void k_gm_givens(
    VT* __restrict__ w, uint32_t n, GmState* st, const double* __restrict__ pin, unsigned gin,
    GmLsq L, GmPrm P, uint32_t k, double* hist) {
  if (st->status != GM_RUN) return;
  double t2[1];
  partials_thread<1, 0u, kWideThreads>(pin, gin, threadIdx.x, t2);
  partials_combine<1, 0u, kWideThreads>(t2);
  const double t = sqrt(t2[0]);  // VEC::Norm2
  const double a = 1.0 / t;      // src/GMRES.cxx:177-178 (t > SMALL_DBL always holds)
  const uint32_t cnt = k + 1;
  const bool staged = cnt < kGivLds;
  __shared__ double col[kGivLds + 1], shc[kGivLds], shs[kGivLds], sg[2];
  __shared__ double tri[kTriLds * kTriLds], tg[kTriLds];
  __shared__ int s_end;
  if (blockIdx.x == 0 && staged) {
    for (uint32_t i = threadIdx.x; i <= cnt; i += kWideThreads)
      col[i] = i == cnt ? t : L.H[(size_t)i * L.ld + k];
    for (uint32_t i = threadIdx.x; i < k; i += kWideThreads) {
      shc[i] = L.hc[i];
      shs[i] = L.hs[i];
    }
    if (threadIdx.x == 0) sg[0] = L.g[k];
  }
  const uint32_t stride = gridDim.x * kWideThreads;
  for (uint32_t i = blockIdx.x * kWideThreads + threadIdx.x; i < n; i += stride)
    w[i] = w[i] * (VT)a;
  if (blockIdx.x != 0) return;
  __syncthreads();
  if (threadIdx.x == 0) {
    GmState S = *st;
    S.numIter += 1;
    S.count = cnt;
    double gcnt;
    if (staged) {
      double gk = sg[0];
      givens_column_lds(col, shc, shs, gk, gcnt, cnt);
      L.hc[k] = shc[k];
      L.hs[k] = shs[k];
      L.g[k] = gk;
      L.g[cnt] = gcnt;
      sg[0] = gk;
    } else {
      L.H[(size_t)cnt * L.ld + k] = t;
      givens_column(L, cnt);
      gcnt = L.g[cnt];
    }
    S.resAbs = fabs(gcnt);
    S.est = S.resAbs;
    S.est_old = S.resAbs;  // resAbsOld = resAbs after the ratio, :198-201
    const double resRel = S.resAbs / S.denAbs;
    const bool conv = (S.resAbs < P.absTol || resRel < P.relTol) && S.numIter > P.minIter;
    const bool end = conv || cnt >= P.m || S.numIter >= P.maxIter;
    if (end) {
      S.status = GM_CYCLE_END;
    } else if (S.numIter >= P.minIter && hist) {  // the next iteration's record, :153
      hist[S.numIter] = S.resAbs;
      S.histCount = S.numIter + 1;
    }
    s_end = end;
    *st = S;
  }
  __syncthreads();
  if (staged)  // the rotated column back to H
    for (uint32_t i = threadIdx.x; i <= cnt; i += kWideThreads) L.H[(size_t)i * L.ld + k] = col[i];
  if (!s_end) return;
  if (staged && cnt <= kTriLds) {  // the cycle's y = H^-1 g on LDS copies
    for (uint32_t e = threadIdx.x; e < cnt * cnt; e += kWideThreads) {
      const uint32_t i = e / cnt, j = e % cnt;
      if (j >= i) tri[i * cnt + j] = j == k ? col[i] : L.H[(size_t)i * L.ld + j];
    }
    for (uint32_t i = threadIdx.x; i < cnt; i += kWideThreads) tg[i] = i == k ? sg[0] : L.g[i];
    __syncthreads();
    if (threadIdx.x == 0) {  // src/GMRES.cxx:213-219, as triangular_solve
      const uint32_t c1 = cnt - 1;
      tg[c1] = tg[c1] / tri[c1 * cnt + c1];
      for (int kk = (int)cnt - 2; kk >= 0; --kk) {
        double u = 0.0;
        for (uint32_t jj = (uint32_t)kk + 1; jj < cnt; ++jj) u = fma(-tri[kk * cnt + jj], tg[jj], u);
        u += tg[kk];
        tg[kk] = u / tri[kk * cnt + kk];
      }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < cnt; i += kWideThreads) L.g[i] = tg[i];
  } else {
    __syncthreads();  // the column written back above
    if (threadIdx.x == 0) triangular_solve(L, cnt);
  }
}

// X: wk = sum_j y_j V[j] in the reference's order, then x += P wk (right) or
// x += wk (left).
template <class VT, bool RIGHTJAC>
__global__ __launch_bounds__(kWideThreads) void k_gm_update(VT* __restrict__ x,
                                                            const VT* const* __restrict__ V,
                                                            const VT* __restrict__ dinv,
                                                            uint32_t n, const GmState* st,
                                                            GmLsq L) {
  const uint32_t cnt = st->count;
  if (cnt == 0) return;
  const uint32_t stride = gridDim.x * kWideThreads;
  for (uint32_t i = blockIdx.x * kWideThreads + threadIdx.x; i < n; i += stride) {
    VT wk = V[cnt - 1][i] * (VT)L.g[cnt - 1];  // copy, then Scale(g[count-1])
    for (int j = (int)cnt - 2; j >= 0; --j) {  // AXPBY(1, g[j], V[j], wk)
      const double gj = L.g[j];
      wk = gj == 1.0 ? wk + V[j][i] : fma((VT)gj, V[j][i], wk);
    }
    const VT d = RIGHTJAC ? wk * dinv[i] : wk;  // P wk from zero
    x[i] = fma((VT)1, d, x[i]);                 // x.AXPY(1, .)
  }
}

// R: V0 = b - A x (right), or tmp = b - A x and V0 = P tmp (left); partials
// of ||tmp||^2 and ||V0||^2.
template <class VT, bool LEFTJAC, bool LEFT>
struct EpiGmRes {
  VT* __restrict__ v0;
  VT* __restrict__ tmp;
  const VT* __restrict__ b;
  const VT* __restrict__ dinv;
  __device__ bool skip() const { return false; }
  __device__ __forceinline__ void row(uint32_t i, double v, double (&acc)[2]) const {
    const VT r = fma((VT)(-1), (VT)v, b[i]);  // Apply, XPAY(-1, b)
    acc[0] = fma((double)r, (double)r, acc[0]);
    if (LEFT) {
      tmp[i] = r;
      const VT z = LEFTJAC ? r * dinv[i] : r;
      v0[i] = z;
      acc[1] = fma((double)z, (double)z, acc[1]);
    } else {
      v0[i] = r;
      acc[1] = acc[0];
    }
  }
  __device__ void fin(double (&)[2]) const {}
};

struct FinGmRes {  // rj = ||b - A x||, resAbs = ||V0||, state -> host slot
  GmState* st;
  GmState* hout;
  __device__ bool skip() const { return false; }
  __device__ void fin(double (&s)[2]) const {
    if (threadIdx.x != 0) return;
    GmState S = *st;
    S.rj = sqrt(s[0]);
    S.resAbs = sqrt(s[1]);
    *st = S;
    if (hout) *hout = S;
  }
};

// ---- FGMRES ----------------------------------------------------------------
// P_k (Jacobi): tmp = z + (work .* dinv), work = v - A z, or v when the stale
// z has ||z||_inf < 1e-20 (Jacobi::Solve as the preconditioner,
// src/Iter.cxx:69-99; MAT::Residual's shortcut, src/MAT.cxx:459-462)
template <class VT>
struct EpiFgmPc {
  VT* __restrict__ tmp;
  const VT* __restrict__ z;
  const VT* __restrict__ v;
  const VT* __restrict__ dinv;
  const double* zinf;
  const GmState* st;
  __device__ bool skip() const { return st->status != GM_RUN; }
  __device__ __forceinline__ void row(uint32_t i, double r, double (&)[1]) const {
    const VT work = (*zinf < kCloseZero) ? v[i] : (VT)r;
    const VT wd = work * dinv[i];  // work.PointwiseMult(diagInv)
    tmp[i] = z[i] + wd;            // x += work
  }
  __device__ void fin(double (&)[1]) const {}
};

// S_k: w = A z with z = tmp (Jacobi) or V[k] (identity); Z[k] = z; partials
// of (w, V0) and max |z|
template <class VT>
struct EpiFgmW {
  VT* __restrict__ w;
  VT* __restrict__ zk;
  const VT* __restrict__ z;
  const VT* __restrict__ v0;
  const GmState* st;
  __device__ bool skip() const { return st->status != GM_RUN; }
  __device__ __forceinline__ void row(uint32_t i, double v, double (&acc)[2]) const {
    const VT a = (VT)v;
    w[i] = a;
    const VT zi = z[i];
    zk[i] = zi;
    acc[0] = fma((double)a, (double)v0[i], acc[0]);
    acc[1] = fmax(acc[1], fabs((double)zi));
  }
  __device__ void fin(double (&)[2]) const {}
};

// X: wk = sum_j y_j Z[j] (src/FGMRES.cxx:216-221), x += wk
template <class VT>
__global__ __launch_bounds__(kWideThreads) void k_fgm_update(VT* __restrict__ x,
                                                             const VT* const* __restrict__ Z,
                                                             uint32_t n, const GmState* st,
                                                             GmLsq L) {
  const uint32_t cnt = st->count;
  if (cnt == 0) return;
  const uint32_t stride = gridDim.x * kWideThreads;
  for (uint32_t i = blockIdx.x * kWideThreads + threadIdx.x; i < n; i += stride) {
    VT wk = Z[cnt - 1][i] * (VT)L.g[cnt - 1];
    for (int j = (int)cnt - 2; j >= 0; --j) {
      const double gj = L.g[j];
      wk = gj == 1.0 ? wk + Z[j][i] : fma((VT)gj, Z[j][i], wk);
    }
    x[i] = fma((VT)1, wk, x[i]);
  }
}

// ||Z[j]||_inf of every stored Z[j] (blockIdx.y = j): per-workgroup maxima,
// then one workgroup per j.
template <class VT>
__global__ __launch_bounds__(kThreads) void k_zinf_part(const VT* const* __restrict__ Z,
                                                        uint32_t n, double* __restrict__ part) {
  const VT* z = Z[blockIdx.y];
  double m = 0.0;
  const uint32_t stride = gridDim.x * kThreads;
  for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < n; i += stride)
    m = fmax(m, fabs((double)z[i]));
  double a[1] = {m};
  block_reduce<1, 1u>(a);
  if (threadIdx.x == 0) part[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = a[0];
}

__global__ __launch_bounds__(kThreads) void k_zinf_fin(const double* __restrict__ part,
                                                       uint32_t g, double* __restrict__ zinf) {
  double m = 0.0;
  for (uint32_t b = threadIdx.x; b < g; b += kThreads)
    m = fmax(m, part[(size_t)blockIdx.x * g + b]);
  double a[1] = {m};
  block_reduce<1, 1u>(a);
  if (threadIdx.x == 0) zinf[blockIdx.x] = a[0];
}

// ---------------------------------------------------------------------------
struct GmWs {
  uint32_t maxR = 0;
  std::vector<void*> V;      // maxR + 1 device vectors
  void** Vdev = nullptr;     // device copy of the pointer table
  void** Zdev = nullptr;     // FGMRES: the solver's Z vectors' pointer table
  double* zinf = nullptr;    // FGMRES: ||Z[j]||_inf, maxR + 1 entries
  double* zpart = nullptr;   // FGMRES: (maxR + 1) x kMaxRedBlocks partial maxima
  void* tmp = nullptr;
  double* lsq = nullptr;     // H, hc, hs, g
  std::map<uint32_t, hipGraphExec_t> cycles;
  ~GmWs() {
    for (auto& c : cycles) (void)hipGraphExecDestroy(c.second);
    for (void* v : V) (void)hipFree(v);
    (void)hipFree(Vdev);
    (void)hipFree(Zdev);
    (void)hipFree(zinf);
    (void)hipFree(zpart);
    (void)hipFree(tmp);
    (void)hipFree(lsq);
  }
};

template <class VT>
struct GmRun {
  fx_solver S;
  FastWs& W;
  GmWs& G;
  fx_csr A;
  const VT* dinv;  // null: identity
  bool left;
  hipStream_t st;
  bool flex = false;  // FGMRES
  fx_op op = nullptr;  // a callback operator in place of A

  GmState* dstate() const { return (GmState*)W.state; }
  VT* zvec(uint32_t j) const { return (VT*)fx_vec_data(S->Z[j]); }
  GmState* hslot() const { return (GmState*)W.hstate; }
  GmLsq lsq() const {
    const uint32_t m = G.maxR;
    GmLsq L;
    L.ld = m;
    L.H = G.lsq;
    L.hc = L.H + (size_t)(m + 1) * m;
    L.hs = L.hc + (m + 1);
    L.g = L.hs + m;
    return L;
  }
  VT* vec(uint32_t j) const { return (VT*)G.V[j]; }
  GmPrm prm(uint32_t m) const {
    GmPrm p{};
    p.maxIter = S->prm.maxIter;
    p.minIter = S->prm.minIter;
    p.m = m;
    p.relTol = S->prm.relTol;
    p.absTol = S->prm.absTol;
    return p;
  }

  // R: residual, its norms, state to the host slot
  int launch_residual(const VT* b, const VT* x, GmState* hout) {
    const int cus = W.ctx->sms;
    const unsigned gs = fused_grid<VT>(A, op, W.n, cus);
    VT* v0 = vec(0);
    VT* tmp = (VT*)G.tmp;
    if (!left) {
      FX_TRY((fused_spmv<VT, ROW_AX, 2, 0u>(A, op, W, st, XLoad<VT>{x}, BNone{},
                                            EpiGmRes<VT, false, false>{v0, tmp, b, dinv}, W.partials)));
    } else if (dinv) {
      FX_TRY((fused_spmv<VT, ROW_AX, 2, 0u>(A, op, W, st, XLoad<VT>{x}, BNone{},
                                            EpiGmRes<VT, true, true>{v0, tmp, b, dinv}, W.partials)));
    } else {
      FX_TRY((fused_spmv<VT, ROW_AX, 2, 0u>(A, op, W, st, XLoad<VT>{x}, BNone{},
                                            EpiGmRes<VT, false, true>{v0, tmp, b, dinv}, W.partials)));
    }
    return launch_finish<2, 0u>(st, W.partials, gs, FinGmRes{dstate(), hout});
  }

  int launch_cycle(uint32_t m, const VT* b, VT* x) {
    const uint32_t n = W.n;
    const int cus = W.ctx->sms;
    const unsigned gs = fused_grid<VT>(A, op, W.n, cus);
    // two workgroups per CU: the MGS steps (which re-sum only the previous
    // step's few hundred partials) keep twice the loads in flight, GMRES(30)
    // on the 1024^2 upwind system 238 against 250 ms
    const unsigned g = wide_blocks(n, Pack16<VT>::W, cus, kGmPer, kGmPerCu);
    const GmLsq L = lsq();
    const GmPrm P = prm(m);
    double* pa = W.partials;
    double* pb = W.partials + W.part2_off;
    hipLaunchKernelGGL((k_gm_start<VT>), dim3(g), dim3(kWideThreads), 0, st, vec(0), n, dstate(),
                       L, P, W.hist);
    FX_HIP(hipGetLastError());
    for (uint32_t k = 0; k < m; ++k) {
      VT* w = vec(k + 1);
      const GmState* cst = dstate();
      if (flex) {
        const VT* z = vec(k);  // identity: Z[k] = V[k]
        if (dinv) {
          // the preconditioner's sweep on its own matrix
          FX_TRY((launch_spmv<VT, VT, ROW_B_MINUS_AX, 0, 0u>(
              S->pc->A, st, XLoad<VT>{zvec(k)}, BLoad<VT>{vec(k)},
              EpiFgmPc<VT>{(VT*)G.tmp, zvec(k), vec(k), dinv, G.zinf + k, cst}, nullptr, nullptr,
              nullptr, false)));
          z = (const VT*)G.tmp;
        }
        FX_TRY((fused_spmv<VT, ROW_AX, 2, 2u>(A, op, W, st, XLoad<VT>{z}, BNone{},
                                              EpiFgmW<VT>{w, zvec(k), z, vec(0), cst}, pa)));
      } else if (!left && dinv) {
        FX_TRY((fused_spmv<VT, ROW_AX, 1, 0u>(A, op, W, st, XScaled<VT>{vec(k), dinv}, BNone{},
                                              EpiGmW<VT, false>{w, vec(0), dinv, cst}, pa)));
      } else if (left && dinv) {
        FX_TRY((fused_spmv<VT, ROW_AX, 1, 0u>(A, op, W, st, XLoad<VT>{vec(k)}, BNone{},
                                              EpiGmW<VT, true>{w, vec(0), dinv, cst}, pa)));
      } else {
        FX_TRY((fused_spmv<VT, ROW_AX, 1, 0u>(A, op, W, st, XLoad<VT>{vec(k)}, BNone{},
                                              EpiGmW<VT, false>{w, vec(0), dinv, cst}, pa)));
      }
      const double* pin = pa;
      unsigned gin = gs;
      for (uint32_t j = 0; j <= k; ++j) {
        double* pout = (pin == pa) ? pb : pa;
        const VT* next = (j < k) ? vec(j + 1) : nullptr;
        const bool zslot = flex && j == 0;  // S_k's partials carry max |Z[k]|
        hipLaunchKernelGGL((k_gm_mgs<VT>), dim3(g), dim3(kWideThreads), 0, st, w,
                           (const VT*)vec(j), next, n, cst, pin, gin, pout, L, j, k,
                           zslot ? G.zinf + k : nullptr);
        FX_HIP(hipGetLastError());
        pin = pout;
        gin = g;
      }
      hipLaunchKernelGGL((k_gm_givens<VT>), dim3(g), dim3(kWideThreads), 0, st, w, n, dstate(),
                         pin, gin, L, P, k, W.hist);
      FX_HIP(hipGetLastError());
    }
    if (flex)
      hipLaunchKernelGGL((k_fgm_update<VT>), dim3(g), dim3(kWideThreads), 0, st, x,
                         (const VT* const*)G.Zdev, n, (const GmState*)dstate(), L);
    else if (!left && dinv)
      hipLaunchKernelGGL((k_gm_update<VT, true>), dim3(g), dim3(kWideThreads), 0, st, x,
                         (const VT* const*)G.Vdev, dinv, n, (const GmState*)dstate(), L);
    else
      hipLaunchKernelGGL((k_gm_update<VT, false>), dim3(g), dim3(kWideThreads), 0, st, x,
                         (const VT* const*)G.Vdev, dinv, n, (const GmState*)dstate(), L);
    FX_HIP(hipGetLastError());
    return launch_residual(b, x, hslot());
  }

  int cycle(uint32_t m, const VT* b, VT* x) {
    if (!st) return launch_cycle(m, b, x);  // the null stream cannot be captured
    auto it = G.cycles.find(m);
    if (it == G.cycles.end()) {
      hipGraph_t gr = nullptr;
      FX_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      const int rc = launch_cycle(m, b, x);
      hipError_t e = hipStreamEndCapture(st, &gr);
      if (rc) return rc;
      FX_HIP(e);
      hipGraphExec_t ex = nullptr;
      e = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
      (void)hipGraphDestroy(gr);
      FX_HIP(e);
      it = G.cycles.emplace(m, ex).first;
    }
    FX_HIP(hipGraphLaunch(it->second, st));
    return 0;
  }

  int run(const VT* b, VT* x_user, fx_result* res) {
    const fx_params& PP = S->prm;
    const uint32_t n = W.n;
    uint32_t maxR, minR, restart;
    solve_restart(S, maxR, minR, restart);
    VT* x = (VT*)W.x;
    FX_HIP(hipMemcpyAsync(x, x_user, sizeof(VT) * n, hipMemcpyDeviceToDevice, st));
    // safe-guard (savIter <= maxIter): safe = x0, then x after a cycle whose
    // estimate improved, and x = safe at the end past savIter
    // (src/GMRES.cxx:115-122,234,277; src/FGMRES.cxx:123,224,268)
    const bool guard = PP.savIter <= PP.maxIter;
    VT* safe = (VT*)W.safe;
    if (guard) FX_HIP(hipMemcpyAsync(safe, x, sizeof(VT) * n, hipMemcpyDeviceToDevice, st));
    if (flex) {  // the stale Z of earlier solves: their max-norms
      const unsigned gz = stream_blocks(n);
      hipLaunchKernelGGL((k_zinf_part<VT>), dim3(gz, maxR + 1), dim3(kThreads), 0, st,
                         (const VT* const*)G.Zdev, n, G.zpart);
      hipLaunchKernelGGL(k_zinf_fin, dim3(maxR + 1), dim3(kThreads), 0, st,
                         (const double*)G.zpart, gz, G.zinf);
      FX_HIP(hipGetLastError());
    }
    // graphs capture b and x: the solve works on the workspace copies
    VT* bw = (VT*)W.r;
    FX_HIP(hipMemcpyAsync(bw, b, sizeof(VT) * n, hipMemcpyDeviceToDevice, st));
    GmState s0{};
    s0.resAbs = 1.0;
    s0.denAbs = 1.0;
    s0.status = GM_CYCLE_END;
    FX_HIP(hipMemcpyAsync(dstate(), &s0, sizeof(GmState), hipMemcpyHostToDevice, st));
    FX_TRY(launch_residual(bw, x, hslot()));
    FX_HIP(hipStreamSynchronize(st));
    GmState s = *hslot();
    double ri = s.rj;  // left: ||b - A x||; right: the same as resAbs
    double resAbs = s.resAbs;
    const double denAbs = (kCloseZero > resAbs) ? kCloseZero : resAbs;
    s.denAbs = denAbs;
    FX_HIP(hipMemcpyAsync(dstate(), &s, sizeof(GmState), hipMemcpyHostToDevice, st));
    uint64_t cycles = 0;
    while (s.numIter < PP.maxIter) {
      if (s.numIter == PP.minIter) {  // :136-141
        const double resRel = resAbs / denAbs;
        if (resRel < PP.relTol || resAbs < PP.absTol) break;
      }
      FX_CHECK(++cycles <= (uint64_t)PP.maxIter + 2, FX_ERROR_DEVICE, "GMRES cycles did not end");
      FX_TRY(cycle(restart, bw, x));
      FX_HIP(hipStreamSynchronize(st));
      s = *hslot();
      if (guard && s.numIter >= PP.savIter && s.est < s.est_old)  // :234
        FX_HIP(hipMemcpyAsync(safe, x, sizeof(VT) * n, hipMemcpyDeviceToDevice, st));
      resAbs = s.resAbs;
      const double rj = s.rj;
      const double resRel = resAbs / denAbs;
      if ((resRel < PP.relTol || resAbs < PP.absTol) && s.numIter >= PP.minIter) break;  // :255
      const double cr = rj / ri;
      ri = rj;
      adapt_restart(cr, flex, maxR, minR, restart);  // FGMRES: <= / >= (src/FGMRES.cxx:251)
    }
    S->restart_cur = restart;
    res->numIter = s.numIter;
    if (!(s.numIter == PP.minIter && s.numIter == PP.maxIter)) {
      res->norm2 = left ? s.rj : resAbs;  // the norm of tmp (left) or V[0] (right), :269-272
      Ops o(W.ctx, A, op, S->pc, FX_F64);
      fx_vec rv = nullptr;
      Ops::chk(fx_vec_wrap(W.ctx, n, FX_F64, left ? G.tmp : G.V[0], &rv));
      o.owned.push_back(rv);
      res->normInf = o.norminf(rv);
    }
    const VT* xout = (guard && s.numIter > PP.savIter) ? safe : x;  // :277
    FX_HIP(hipMemcpyAsync(x_user, xout, sizeof(VT) * n, hipMemcpyDeviceToDevice, st));
    const uint32_t hc = s.histCount < W.hist_cap ? s.histCount : W.hist_cap;
    S->hist_host.assign(hc, 0.0);
    if (hc) FX_HIP(hipMemcpy(S->hist_host.data(), W.hist, sizeof(double) * hc, hipMemcpyDeviceToHost));
    S->hist_count = hc;
    FX_HIP(hipStreamSynchronize(st));
    return FX_SUCCESS;
  }
};

}  // namespace

size_t gmres_state_bytes() { return sizeof(GmState); }

// GMRES on the fused path (f64 vectors); W is allocated by the caller, the
// Krylov basis is kept in W between solves.
int gmres_fast(fx_solver S, FastWs& W, const double* b, double* x, fx_result* res) {
  uint32_t maxR, minR, restart;
  clamp_restart(S->prm, maxR, minR, restart);
  if (!W.gm) W.gm = new GmWs();
  GmWs& G = *(GmWs*)W.gm;
  if (G.maxR != maxR) {
    delete (GmWs*)W.gm;
    W.gm = new GmWs();
    GmWs& N = *(GmWs*)W.gm;
    N.maxR = maxR;
    const size_t vb = pack_round(W.n) * sizeof(double);
    N.V.assign(maxR + 1, nullptr);
    for (auto& v : N.V) FX_HIP(hipMalloc(&v, vb));
    FX_HIP(hipMalloc(&N.Vdev, sizeof(void*) * (maxR + 1)));
    FX_HIP(hipMemcpy(N.Vdev, N.V.data(), sizeof(void*) * (maxR + 1), hipMemcpyHostToDevice));
    FX_HIP(hipMalloc(&N.tmp, vb));
    const size_t nl = (size_t)(maxR + 1) * maxR + (maxR + 1) + maxR + (maxR + 2);
    FX_HIP(hipMalloc(&N.lsq, sizeof(double) * nl));
    FX_HIP(hipMemset(N.lsq, 0, sizeof(double) * nl));
  }
  GmWs& GG = *(GmWs*)W.gm;
  const bool flex = S->type == FX_SOLVER_FGMRES;
  if (flex) {
    // FGMRES keeps Z in the solver (shared with the per-operation path, so
    // the stale Z carries over between solves as in the reference)
    if (S->Z.size() != maxR + 1 || S->Z[0]->n != W.n) {
      for (fx_vec z : S->Z) fx_vec_destroy(z);
      S->Z.assign(maxR + 1, nullptr);
      for (auto& z : S->Z) FX_TRY(fx_vec_create(S->ctx, W.n, FX_F64, &z));
    }
    if (!GG.Zdev) FX_HIP(hipMalloc(&GG.Zdev, sizeof(void*) * (maxR + 1)));
    if (!GG.zinf) FX_HIP(hipMalloc(&GG.zinf, sizeof(double) * (maxR + 1)));
    if (!GG.zpart) FX_HIP(hipMalloc(&GG.zpart, sizeof(double) * (maxR + 1) * kMaxRedBlocks));
    std::vector<void*> zp(maxR + 1);
    for (uint32_t j = 0; j <= maxR; ++j) zp[j] = fx_vec_data(S->Z[j]);
    FX_HIP(hipMemcpy(GG.Zdev, zp.data(), sizeof(void*) * (maxR + 1), hipMemcpyHostToDevice));
  }
  GmRun<double> run{S, W, GG, S->A, S->pc->kind == 1 ? S->pc->dinv : nullptr,
                    !flex && S->prm.leftPrecond != 0, S->ctx->stream, flex};
  run.op = S->op;
  return run.run(b, x, res);
}

void gmres_ws_free(void* p) { delete (GmWs*)p; }

}  // namespace fx
