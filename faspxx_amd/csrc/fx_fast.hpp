// fx_fast.hpp -- pieces shared by the fused, graph-replayed solvers
// (fx_fast.hip: CG, Jacobi, mixed precision; fx_bicgstab.hip: BiCGStab).
#pragma once

#include "fx_solver.hpp"
#include "fx_spmv.hpp"

namespace fx {

constexpr int kChunk = 8;  // iterations per graph replay

enum { ST_RUN = 0, ST_DONE = 1, ST_HOST = 2 };  // device loop status
enum { R_BREAKDOWN = 1, R_CHECKS = 2 };          // why the device loop stopped

// SOLParams as the device loops read them (solStagTol = 1e-4 relTol, as the
// reference's solvers set it, e.g. src/CG.cxx:58).
struct CgPrm {
  uint32_t maxIter, minIter, savIter, pad;
  double relTol, absTol, solStagTol;
};
CgPrm make_prm(const fx_params& p);

// Workgroup size of the streaming kernels that consume reductions, and how
// many of their workgroups a CU may hold (each later workgroup re-reads every
// workgroup's partial sums, so their number stays small).
#ifndef FX_WIDE_THREADS
#define FX_WIDE_THREADS 512
#endif
#ifndef FX_WIDE_PER_CU
#define FX_WIDE_PER_CU 1
#endif
constexpr int kWideThreads = FX_WIDE_THREADS;
constexpr int kWidePerCu = FX_WIDE_PER_CU;

// Copy a solver state struct (a multiple of 8 bytes) word by word, so that
// no private copy of it is made in scratch memory.
template <class T>
__device__ __forceinline__ void copy_state(T* __restrict__ dst, const T* __restrict__ src) {
  static_assert(sizeof(T) % 8 == 0, "state structs are whole 8-byte words");
  const uint64_t* s = reinterpret_cast<const uint64_t*>(src);
  uint64_t* d = reinterpret_cast<uint64_t*>(dst);
#pragma unroll
  for (int i = 0; i < (int)(sizeof(T) / 8); ++i) d[i] = s[i];
}

// Packs per thread per pass in the CG update kernels (all loaded before the
// reduction that the pass waits for).
#ifndef FX_PASS_PACKS
#define FX_PASS_PACKS 4
#endif
constexpr int kPassPacks = FX_PASS_PACKS;

// Workgroups of those kernels over n elements: kWidePerCu per CU at most,
// `per` packs per thread per pass.
inline unsigned wide_blocks(uint32_t n, int W, int cus, int per = 2) {
  const uint64_t packs = ((uint64_t)n + W - 1) / W;
  uint64_t g = (packs + (uint64_t)per * kWideThreads - 1) / ((uint64_t)per * kWideThreads);
  if (g < 1) g = 1;
  if (g > (uint64_t)cus * kWidePerCu) g = (uint64_t)cus * kWidePerCu;
  return (unsigned)g;
}

void gmres_ws_free(void* p);

// ---------------------------------------------------------------------------
// Workspace
// ---------------------------------------------------------------------------
struct FastWs {
  fx_ctx ctx = nullptr;
  int dt = FX_F64;
  uint32_t n = 0;
  void* x = nullptr;  // internal solution / Jacobi ping buffer
  void* x2 = nullptr; // Jacobi pong buffer
  void* r = nullptr;
  void* p = nullptr;
  void* q = nullptr;
  void* safe = nullptr;
  void* s = nullptr;   // BiCGStab: s
  void* t = nullptr;   // BiCGStab: P A s
  double* partials = nullptr;
  void* state = nullptr;  // CgState / JacState
  void* hstate = nullptr; // pinned, 2 slots
  double* hist = nullptr;
  uint32_t hist_cap = 0;
  double* dscal = nullptr;
  void* side = nullptr;        // scalars handed between the kernels of an iteration
  void* gm = nullptr;          // GMRES basis and cycle graphs (fx_gmres.hip)
  size_t part2_off = 0;        // second partials region (doubles)
  hipGraphExec_t chunk[2] = {nullptr, nullptr};  // replay graphs, one per host state slot
  hipEvent_t ev[2] = {nullptr, nullptr};

  ~FastWs();
};

inline FastWs::~FastWs() {
  {
    if (ctx) (void)hipStreamSynchronize(ctx->stream);
    gmres_ws_free(gm);
    void* bufs[] = {x, x2, r, p, q, safe, s, t, partials, state, hist, dscal, side};
    for (void* b : bufs) (void)hipFree(b);
    if (hstate) (void)hipHostFree(hstate);
    for (auto& c : chunk)
      if (c) (void)hipGraphExecDestroy(c);
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
  }
}

// Allocate a workspace: vectors padded to whole 16-byte packs, two partials
// regions, two parity buffers of state_bytes, a pinned two-slot host mirror.
int ws_alloc(FastWs& W, fx_ctx ctx, int dt, uint32_t n, uint32_t hist_cap, size_t state_bytes,
             uint32_t spmv_blocks);

// Fused BiCGStab (fx_bicgstab.hip): its state size, and the solve on a
// workspace allocated with it.
size_t bicgstab_state_bytes();
int bicgstab_fast(fx_solver S, FastWs& W, const double* b, double* x, fx_result* res);

// Fused GMRES (fx_gmres.hip), right or left preconditioned.
size_t gmres_state_bytes();
int gmres_fast(fx_solver S, FastWs& W, const double* b, double* x, fx_result* res);

}  // namespace fx
