// fx_noalloc.hip -- fx_pc_apply without device allocation, for callers that
// cannot allocate (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc).
//
// fx_pc_apply runs `iters` Jacobi sweeps z += D^-1 (r - A z) through a work
// vector it allocates on every call (and the max-norm reduction of MAT::
// Residual's shortcut through the context's partials).  This variant keeps
// all of that in one file-scope device array, `scratch`: the max-norm
// partials, the max-norm slot, and the work vector.  Each GPU holds its own
// copy of the array for the life of the process; its address on a device is
// taken once (hipGetSymbolAddress, with that device current) when the first
// context on the device is made, and kernels only ever get that address.
//
// Every call on a GPU shares that GPU's array, and a call leaves it holding
// its own values: calls that overlap on one GPU (two streams, two threads)
// race.  Nothing in it needs to start at zero: each call writes every word
// it reads before reading it.
#include <mutex>

#include "fx_internal.hpp"
#include "fx_spmv.hpp"

namespace fx {

namespace {

// Layout (doubles): partials of the max-norm reduction (one per workgroup of
// its grid, at most kMaxRedBlocks), a slot block, then the work vector of at
// most kNoAllocMaxRows entries.  kNoAllocMaxRows * 8 bytes = 32 MiB.
constexpr uint32_t kNoAllocMaxRows = 1u << 22;
constexpr size_t kScrPartials = 0;
constexpr size_t kScrSlots = kMaxRedBlocks;          // 2048
constexpr size_t kScrWork = kScrSlots + 32;          // 2080: 16640 bytes, 256-byte aligned
constexpr size_t kScrDoubles = kScrWork + kNoAllocMaxRows;
static_assert((kScrWork * sizeof(double)) % 256 == 0, "work vector 256-byte aligned");

static __device__ __align__(256) double scratch[kScrDoubles];

constexpr int kMaxDevices = 64;
double* g_scratch[kMaxDevices] = {};
std::once_flag g_once[kMaxDevices];
int g_rc[kMaxDevices] = {};

// max |x_i| (one partial per workgroup)
__global__ __launch_bounds__(kThreads) void k_xinf(const double* __restrict__ x, uint32_t n,
                                                   double* partials) {
  double acc[1] = {0.0};
  const uint32_t stride = gridDim.x * kThreads;
  for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < n; i += stride)
    acc[0] = fmax(acc[0], fabs(x[i]));
  block_partials<1, 1u>(acc, partials);
}

// work = (b - A z) .* dinv, or b .* dinv when ||z||_inf < CLOSE_ZERO
// (MAT::Residual with its shortcut, src/MAT.cxx:455-476, then
// VEC::PointwiseMult, src/Iter.cxx:72-74)
struct EpiJacobiWork {
  double* __restrict__ work;
  const double* __restrict__ b;
  const double* __restrict__ dinv;
  const double* __restrict__ zinf;
  __device__ bool skip() const { return false; }
  __device__ __forceinline__ void row(uint32_t i, double v, double (&)[1]) const {
    const double r = (*zinf < kCloseZero) ? b[i] : v;
    work[i] = r * dinv[i];
  }
  __device__ void fin(double (&)[1]) const {}
};

// z += work (VEC::operator+=, src/Iter.cxx:75)
__global__ __launch_bounds__(kThreads) void k_add_work(double* __restrict__ z,
                                                       const double* __restrict__ work,
                                                       uint32_t n) {
  const uint32_t stride = gridDim.x * kThreads;
  for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) z[i] = z[i] + work[i];
}

}  // namespace

// The device address of `scratch` on `device` (the current device), taken
// once per device; called when a context is made on it.
int noalloc_init(int device, double** out) {
  FX_CHECK(device >= 0 && device < kMaxDevices, FX_ERROR_INPUT_PAR, "device id %d", device);
  std::call_once(g_once[device], [device]() {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) {
      g_rc[device] = FX_ERROR_DEVICE;
      return;
    }
    g_scratch[device] = static_cast<double*>(p);
  });
  FX_CHECK(g_rc[device] == 0 && g_scratch[device], FX_ERROR_DEVICE,
           "hipGetSymbolAddress(scratch) failed on device %d", device);
  *out = g_scratch[device];
  return 0;
}

}  // namespace fx

using namespace fx;

extern "C" {

int fx_pc_apply_noalloc(fx_ctx ctx, fx_pc pc, fx_vec b, fx_vec x) {
  clear_error();
  FX_CHECK(ctx && pc && b && x, FX_ERROR_INPUT_PAR, "null argument");
  FX_CHECK(ctx->scratch, FX_ERROR_DEVICE, "context has no scratch address");
  FX_CHECK(b->dtype == FX_F64 && x->dtype == FX_F64, FX_ERROR_INPUT_PAR,
           "fx_pc_apply_noalloc takes f64 vectors");
  FX_CHECK(b->n == x->n, FX_ERROR_NONMATCH_SIZE, "sizes b %u, x %u", b->n, x->n);
  hipStream_t st = ctx->stream;
  if (pc->kind == 0) {  // Identity::Solve: x = b
    if (b->n && b->d != x->d)
      FX_HIP(hipMemcpyAsync(x->d, b->d, sizeof(double) * b->n, hipMemcpyDeviceToDevice, st));
    return 0;
  }
  FX_CHECK(pc->kind == 1, FX_ERROR_SOLVER_PCD_TYPE,
           "fx_pc_apply_noalloc: only the identity and Jacobi preconditioners");
  fx_csr A = pc->A;
  const uint32_t n = A->nrow;
  FX_CHECK(b->n == n, FX_ERROR_NONMATCH_SIZE, "preconditioner size mismatch");
  FX_CHECK(n <= kNoAllocMaxRows, FX_ERROR_INPUT_PAR,
           "fx_pc_apply_noalloc: %u rows exceed its scratch of %u rows", n, kNoAllocMaxRows);
  FX_CHECK(b->d != x->d, FX_ERROR_INPUT_PAR, "b and x must differ");
  if (n == 0) return 0;
  double* partials = ctx->scratch + kScrPartials;
  double* zinf = ctx->scratch + kScrSlots;
  double* work = ctx->scratch + kScrWork;
  const double* bd = (const double*)b->d;
  double* xd = (double*)x->d;
  const unsigned g = stream_blocks(n);
  for (uint32_t it = 0; it < pc->iters; ++it) {
    hipLaunchKernelGGL(k_xinf, dim3(g), dim3(kThreads), 0, st, (const double*)xd, n, partials);
    FX_HIP(hipGetLastError());
    FX_TRY((launch_finish<1, 1u>(st, partials, g, FinStore<1>{zinf, 0})));
    FX_TRY((launch_spmv<double, double, ROW_B_MINUS_AX, 0, 0u>(
        A, st, XLoad<double>{xd}, BLoad<double>{bd}, EpiJacobiWork{work, bd, pc->dinv, zinf},
        nullptr)));
    hipLaunchKernelGGL(k_add_work, dim3(g), dim3(kThreads), 0, st, xd, (const double*)work, n);
    FX_HIP(hipGetLastError());
  }
  return 0;
}

}  // extern "C"
