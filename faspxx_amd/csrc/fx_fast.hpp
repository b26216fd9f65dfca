// fx_fast.hpp -- pieces shared by the fused, graph-replayed solvers
// (fx_fast.hip: CG, Jacobi, mixed precision; fx_bicgstab.hip: BiCGStab).
#pragma once

#include <cmath>
#include <functional>

#include "fx_solver.hpp"
#include "fx_spmv.hpp"

namespace fx {

// Iterations per graph replay (even: the state's parity buffers restart at
// 0 with every replay).
#ifndef FX_CHUNK
#define FX_CHUNK 32
#endif
constexpr int kChunk = FX_CHUNK;
static_assert(kChunk % 2 == 0, "replays hold an even number of iterations");

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
inline unsigned wide_blocks(uint32_t n, int W, int cus, int per = 2, int per_cu = kWidePerCu) {
  const uint64_t packs = ((uint64_t)n + W - 1) / W;
  uint64_t g = (packs + (uint64_t)per * kWideThreads - 1) / ((uint64_t)per * kWideThreads);
  if (g < 1) g = 1;
  if (g > (uint64_t)cus * per_cu) g = (uint64_t)cus * per_cu;
  return (unsigned)g;
}

void gmres_ws_free(void* p);

// Replay loops keep two replays in flight, so that the GPU never waits for
// the host between them; but the one behind the replay that ends the loop
// runs only skipped kernels (a replay's worth, about 3.5 us per iteration
// for the FP64 CG on the 128^3 stencil).  After each finished replay this
// predicts, from how fast a
// measure that the loop drives below a target (a residual norm) fell over
// the last replay, whether the replay still in flight ends the loop, in
// which case no other is queued yet.
// When the same loop ran before (a solver's earlier solve, the refinement's
// earlier inner solve), its iteration count serves as the prediction
// instead, also for the first replays.  A wrong prediction costs one wait
// for the host between replays.
struct ReplayPace {
  double prev = -1.0;
  uint32_t hint = 0;  // iterations the loop took last time; 0: unknown
  // replays to queue at the start
  int initial() const { return (hint && hint <= (uint32_t)kChunk) ? 1 : 2; }
  // m: the measure after the replay that just finished, target: the value
  // that stops the loop, done: iterations run so far, init: the measure
  // before the loop's first replay (0: unknown).  True when another replay
  // should be queued now.
  bool more(double m, double target, uint32_t done, double init) {
    const double p = prev > 0.0 ? prev : init;
    prev = m;
    if (hint) return (int64_t)hint - (int64_t)done > (int64_t)kChunk;
    if (!(p > 0.0) || !(m > 0.0) || !(target > 0.0) || !(target < m)) return true;
    const double f = m / p;  // per replay
    if (!(f < 1.0)) return true;
    // replays still needed after this one; the one in flight is one of them
    return std::log(target / m) / std::log(f) > 0.9;
  }
};

// The host slots of the replays in flight, oldest first (at most two).  A
// replay is queued on the slot that no replay in flight uses, and the host
// waits for the oldest one, so that it never reads a slot that no replay of
// the current loop has written.
struct SlotQueue {
  int q[2] = {0, 0};
  int n = 0;
  int free_slot() const { return n == 0 ? 0 : q[n - 1] ^ 1; }
  void push(int s) { q[n++] = s; }
  int front() const { return q[0]; }
  void pop() {
    q[0] = q[1];
    --n;
  }
};

// Iterations of the next replay: the most that fit kChunk, or for the
// replay that is predicted to end the loop (need: the iterations the loop
// is predicted to launch, 0: no prediction; queued: those queued so far) the
// count that covers the rest.  Both are multiples of `unit`, the period of
// the loop's buffer rotation (2 for the state's parity buffers), so that
// every replay starts on the same buffers.
inline int replay_len(uint64_t need, uint64_t queued, int unit = 2) {
  const int full = kChunk / unit * unit;
  if (need == 0 || queued >= need || need - queued >= (uint64_t)full) return full;
  return (int)((need - queued + unit - 1) / unit * unit);
}

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
  void* state = nullptr;  // solver state, two parity buffers
  void* hstate = nullptr; // pinned, 2 slots
  double* hist = nullptr;
  uint32_t hist_cap = 0;
  double* dscal = nullptr;
  void* side = nullptr;        // scalars handed between the kernels of an iteration
  void* gm = nullptr;          // GMRES basis and cycle graphs (fx_gmres.hip)
  void* opy = nullptr;         // callback operator: apply's output
  void* opx = nullptr;         // callback operator: an operand formed before apply
  size_t part2_off = 0;        // second partials region (doubles)
  uint32_t last_iters = 0;     // iterations of the last solve on this workspace
  hipGraphExec_t chunk[2] = {nullptr, nullptr};  // replay graphs, one per host state slot
  // shorter replays (2, 4, .. kChunk - 2 iterations) for the replay that is
  // predicted to end a solve, per host state slot; made on first use
  hipGraphExec_t tail[2][kChunk / 2] = {};
  hipEvent_t ev[2] = {nullptr, nullptr};

  ~FastWs();
};

inline FastWs::~FastWs() {
  {
    if (ctx) (void)hipStreamSynchronize(ctx->stream);
    gmres_ws_free(gm);
    void* bufs[] = {x, x2, r, p, q, safe, s, t, partials, state, hist, dscal, side, opy, opx};
    for (void* b : bufs) (void)hipFree(b);
    if (hstate) (void)hipHostFree(hstate);
    for (auto& c : chunk)
      if (c) (void)hipGraphExecDestroy(c);
    for (auto& slot : tail)
      for (auto& c : slot)
        if (c) (void)hipGraphExecDestroy(c);
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
  }
}

// ---------------------------------------------------------------------------
// Replay loops
// ---------------------------------------------------------------------------
// A fused solver's device loop is replayed from graphs of kChunk iterations
// (or fewer, see replay_len), one graph per host slot: iteration c is
// enqueued by iter(c, slot), where slot >= 0 for the replay's last iteration
// (it copies the state to that host slot) and -1 otherwise.

// The replay graph of L iterations for host slot `slot`, made on first use
// and kept in W.
template <class Iter>
int replay_graph(FastWs& W, hipStream_t st, int slot, int L, Iter& iter, hipGraphExec_t* out) {
  hipGraphExec_t& ge = L == kChunk ? W.chunk[slot] : W.tail[slot][L / 2];
  if (!ge) {
    hipGraph_t g = nullptr;
    FX_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    int rc = 0;
    for (int c = 0; c < L && !rc; ++c) rc = iter(c, c == L - 1 ? slot : -1);
    hipError_t e = hipStreamEndCapture(st, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    FX_HIP(e);
    e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    FX_HIP(e);
  }
  *out = ge;
  return 0;
}

// One replay of L iterations on host slot `slot`, then the slot's event.
// The legacy null stream cannot be captured: there the iterations are
// launched directly.
template <class Iter>
int replay_launch(FastWs& W, hipStream_t st, int slot, int L, Iter& iter) {
  if (st) {
    hipGraphExec_t ge = nullptr;
    FX_TRY(replay_graph(W, st, slot, L, iter, &ge));
    FX_HIP(hipGraphLaunch(ge, st));
  } else {
    for (int c = 0; c < L; ++c) FX_TRY(iter(c, c == L - 1 ? slot : -1));
  }
  FX_HIP(hipEventRecord(W.ev[slot], st));
  return 0;
}

// Replays the device loop until the state a replay leaves in its host slot
// is no longer ST_RUN; *out is then that host copy.  pace.initial() replays
// go out at once; after each finished replay another goes out when more(s)
// asks for it (s: the state it left) or none is in flight.  need: the
// iterations the loop is predicted to launch before it stops (0: no
// prediction); the replay predicted to end the loop is cut to that count.
// unit: see replay_len.  max_launch: iterations after which the loop has
// certainly stopped; no replay is queued beyond them while one is in
// flight.  At most max_chunks replays may find the loop running.
// behind: work enqueued right behind the replay predicted to end the loop
// (none when null); *behind_ran is then whether that replay or an earlier
// one ended it, i.e. whether the work found the loop stopped.
template <class State, class Iter, class More>
int replay_loop(FastWs& W, hipStream_t st, const ReplayPace& pace, uint64_t need,
                uint64_t max_chunks, Iter&& iter, More&& more, const State** out,
                const char* what, int unit = 2, uint64_t max_launch = UINT64_MAX,
                const std::function<int()>& behind = nullptr, bool* behind_ran = nullptr) {
  const int full = replay_len(0, 0, unit);
  if (st) {  // the graphs this loop may replay, made before any replay is queued
    const int tail = replay_len(need, need / full * full, unit);
    for (int slot = 0; slot < 2; ++slot) {
      hipGraphExec_t ge = nullptr;
      FX_TRY(replay_graph(W, st, slot, full, iter, &ge));
      FX_TRY(replay_graph(W, st, slot, tail, iter, &ge));
    }
  }
  SlotQueue sq;
  uint64_t queued = 0, chunks = 0, synced = 0, behind_at = 0;
  if (behind_ran) *behind_ran = false;
  auto queue = [&]() -> int {
    const int slot = sq.free_slot(), L = replay_len(need, queued, unit);
    FX_TRY(replay_launch(W, st, slot, L, iter));
    sq.push(slot);
    queued += (uint64_t)L;
    ++chunks;
    if (behind && !behind_at && need && queued >= need) {
      FX_TRY(behind());
      behind_at = chunks;
    }
    return 0;
  };
  FX_TRY(queue());
  if (pace.initial() > 1 && queued < max_launch) FX_TRY(queue());
  for (;;) {
    const int cur = sq.front();
    FX_HIP(hipEventSynchronize(W.ev[cur]));
    sq.pop();
    ++synced;
    const State* hs = (const State*)W.hstate + cur;
    if (hs->status != ST_RUN) {
      *out = hs;
      if (behind_ran) *behind_ran = behind_at && synced <= behind_at;
      return 0;
    }
    FX_CHECK(chunks <= max_chunks, FX_ERROR_DEVICE, "%s device loop did not stop", what);
    if ((more(*hs) && queued < max_launch) || sq.n == 0) FX_TRY(queue());
  }
}

// ---------------------------------------------------------------------------
// The operator of a fused solve
// ---------------------------------------------------------------------------
// A matrix: its SpMV with the solver's epilogue in the same kernel.  A
// callback operator that only enqueues device work (fx_op_s::device_only):
// apply(x) into W.opy (x formed into W.opx first when the SpMV's operand
// functor is not a plain load), then the same epilogue over W.opy in
// k_op_epi, whose partial sums come from op_grid(n) workgroups.  apply is
// called while a replay graph is captured, so its work is replayed with it.
inline unsigned op_grid(uint32_t n) { return stream_blocks(n); }

template <class VT>
inline unsigned fused_grid(fx_csr A, fx_op op, uint32_t n, int cus) {
  return op ? op_grid(n) : spmv_grid_of<VT>(A, cus);
}

template <class XF>
__global__ __launch_bounds__(kThreads) void k_op_operand(double* __restrict__ xs, XF xf0,
                                                         uint32_t n) {
  XF xf = xf0;
  maybe_init(xf);
  const uint32_t stride = gridDim.x * kThreads;
  for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) xs[i] = xf(i);
}

// The epilogue of an SpMV row by row over y = A x: v = y_i (ROW_AX) or
// b_i - y_i (ROW_B_MINUS_AX: the operator's residual, r = b - A x, as the
// per-operation solvers form it with apply and XPAY).
template <int MODE, int NRED, unsigned MAXMASK, class BF, class Epi>
__global__ __launch_bounds__(kThreads) void k_op_epi(const double* __restrict__ y, uint32_t n,
                                                     BF bf, Epi epi0, double* partials) {
  if (epi0.skip()) return;
  Epi epi = epi0;
  maybe_init(epi);
  double acc[NRED > 0 ? NRED : 1];
#pragma unroll
  for (int k = 0; k < (NRED > 0 ? NRED : 1); ++k) acc[k] = 0.0;
  const uint32_t stride = gridDim.x * kThreads;
  for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) {
    const double yi = y[i];
    epi.row(i, MODE == ROW_AX ? yi : bf(i) - yi, acc);
  }
  if constexpr (NRED > 0) block_partials<NRED, MAXMASK>(acc, partials);
}

// W.opy = A x through the callback (x: the operand functor's values).
template <class XF>
int op_apply(fx_op op, FastWs& W, hipStream_t st, XF xf) {
  const uint32_t n = W.n;
  FX_CHECK(W.opy && W.opx, FX_ERROR_DEVICE, "operator work vectors missing");
  const double* x = nullptr;
  if constexpr (std::is_same_v<XF, XLoad<double>>) {
    x = xf.x;
  } else {
    hipLaunchKernelGGL((k_op_operand<XF>), dim3(stream_blocks(n)), dim3(kThreads), 0, st,
                       (double*)W.opx, xf, n);
    FX_HIP(hipGetLastError());
    x = (const double*)W.opx;
  }
  fx_vec vx = nullptr, vy = nullptr;
  int rc = fx_vec_wrap(W.ctx, n, FX_F64, (void*)x, &vx);
  if (!rc) rc = fx_vec_wrap(W.ctx, n, FX_F64, W.opy, &vy);
  if (!rc) {
    rc = op->apply(op->user, vx, vy);
    if (rc) {
      if (!fx_last_error()[0]) set_error("operator apply returned %d", rc);
      rc = rc < 0 ? rc : FX_ERROR_INPUT_PAR;
    }
  }
  fx_vec_destroy(vx);
  fx_vec_destroy(vy);
  return rc;
}

// The fused solvers' SpMV with an epilogue, on a matrix or an operator.
template <class VT, int MODE, int NRED, unsigned MAXMASK, class XF, class BF, class Epi>
int fused_spmv(fx_csr A, fx_op op, FastWs& W, hipStream_t st, XF xf, BF bf, Epi epi,
               double* partials, bool finisher = false) {
  if (!op)
    return launch_spmv<VT, VT, MODE, NRED, MAXMASK>(A, st, xf, bf, epi, partials, nullptr,
                                                    nullptr, finisher);
  if constexpr (std::is_same_v<VT, double>) {
    FX_TRY(op_apply(op, W, st, xf));
    const unsigned g = op_grid(W.n);
    hipLaunchKernelGGL((k_op_epi<MODE, NRED, MAXMASK, BF, Epi>), dim3(g), dim3(kThreads), 0, st,
                       (const double*)W.opy, W.n, bf, epi, partials);
    FX_HIP(hipGetLastError());
    if constexpr (NRED > 0) {
      if (finisher) return launch_finish<NRED, MAXMASK>(st, partials, g, epi);
    }
    return 0;
  } else {
    FX_CHECK(false, FX_ERROR_INPUT_PAR, "callback operators act on f64 vectors");
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
