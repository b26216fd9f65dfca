// fx_core.hip -- context, device vectors and the VEC operations
// (src/VEC.cxx) of the gfx950 iterative-solve library.
#include <atomic>
#include <cstdarg>
#include <cstring>

#include "fx_internal.hpp"

namespace fx {

namespace {
thread_local char g_err[1024] = {0};
}

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}
void clear_error() { g_err[0] = 0; }

uint64_t next_handle_gen() {
  static std::atomic<uint64_t> gen{0};
  return gen.fetch_add(1, std::memory_order_relaxed) + 1;
}

// ---------------------------------------------------------------------------
// Elementwise kernels.  Each thread handles elements i = tid, tid + stride, ...
// Expressions follow the reference's contracted forms (src/VEC.cxx).
// ---------------------------------------------------------------------------
enum ElemOp {
  OP_AXPY, OP_XPAY, OP_AXPBY, OP_WAXPBY, OP_SCALE, OP_PMUL, OP_RECIP, OP_ADD, OP_SET,
  OP_SHIFT, OP_ABS, OP_PDIV, OP_SUB
};

template <class T>
__device__ __forceinline__ T elem(int op, int form, double a, double b, T x, T y) {
  // returns the new value of the output element; x is the first input, y the
  // second input / old output as documented per op
  switch (op) {
    case OP_AXPY: return fma((T)a, x, y);   // y += a x
    case OP_XPAY: return fma((T)a, y, x);   // y = a y + x
    case OP_AXPBY:                          // y = a y + b x
      switch (form) {
        case 0: return fma((T)a, y, (T)b * x);
        case 1: return fma((T)b, x, y);
        case 2: return fma((T)a, y, x);
        default: return y + x;
      }
    case OP_SCALE: return y * (T)a;
    case OP_PMUL: return y * x;
    case OP_RECIP: return (T)1 / y;
    case OP_ADD: return y + x;
    case OP_SET: return (T)a;
    case OP_SHIFT: return y + (T)a;  // VEC::Shift, src/VEC.cxx:217-228
    case OP_ABS: return fabs(y);     // VEC::Abs, src/VEC.cxx:231-242
    case OP_PDIV: return y / x;      // VEC::PointwiseDivide, src/VEC.cxx:196-207
    case OP_SUB: return y - x;       // VEC::operator-=, src/VEC.cxx:81-93
    default: return y;
  }
}

template <class T>
__global__ __launch_bounds__(kThreads) void k_elem(int op, int form, double a, double b,
                                                   const T* __restrict__ x, T* __restrict__ y,
                                                   uint32_t n) {
  const uint32_t stride = gridDim.x * kThreads;
  for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) {
    const T xv = x ? x[i] : (T)0;
    y[i] = elem<T>(op, form, a, b, xv, y[i]);
  }
}

// w = a v1 + b v2  (VEC::WAXPBY forms, src/VEC.cxx:329-379)
template <class T>
__global__ __launch_bounds__(kThreads) void k_waxpby(int form, double a, const T* __restrict__ v1,
                                                     double b, const T* __restrict__ v2,
                                                     T* __restrict__ w, uint32_t n) {
  const uint32_t stride = gridDim.x * kThreads;
  for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) {
    const T p = v1[i], q = v2[i];
    T r;
    switch (form) {
      case 0: r = fma((T)a, p, (T)b * q); break;
      case 1: r = fma((T)b, q, p); break;
      case 2: r = fma((T)a, p, q); break;
      default: r = p + q; break;
    }
    w[i] = r;
  }
}

template <class S, class D>
__global__ __launch_bounds__(kThreads) void k_convert(const S* __restrict__ s, D* __restrict__ d,
                                                      uint32_t n) {
  const uint32_t stride = gridDim.x * kThreads;
  for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) d[i] = (D)s[i];
}

// Reductions: kind 0 dot(x,y), 1 sum x^2 (then sqrt), 2 max |x|, 3 max x
// (from SMALL_DBL = -1e60 as VEC::Max does, src/VEC.cxx:382-400), 4 min x
// (from LARGE_DBL = 1e60, VEC::Min, :403-421; computed as -max(-x)).  Optional
// fused update before the reduction (kinds 0/1): upd 1 y += a x (AXPY),
// upd 2 y = a y + x (XPAY).
// Products are formed in double; sums accumulate in double with fma.  Each
// workgroup writes one partial; k_finish combines them.
template <class T>
__global__ __launch_bounds__(kThreads) void k_reduce(int kind, int upd, double a,
                                                     const T* __restrict__ x, T* __restrict__ y,
                                                     const T* __restrict__ z, uint32_t n,
                                                     double* partials) {
  double acc[1] = {kind >= 3 ? -1e60 : 0.0};
  const uint32_t stride = gridDim.x * kThreads;
  for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) {
    if (kind == 2) {
      acc[0] = fmax(acc[0], fabs((double)x[i]));
    } else if (kind == 3) {
      acc[0] = fmax(acc[0], (double)x[i]);
    } else if (kind == 4) {
      acc[0] = fmax(acc[0], -(double)x[i]);
    } else if (upd) {
      const T yn = upd == 1 ? fma((T)a, x[i], y[i]) : fma((T)a, y[i], x[i]);
      y[i] = yn;
      const double u = (double)yn;
      acc[0] = fma(u, kind == 0 ? (double)z[i] : u, acc[0]);
    } else {
      const double u = (double)x[i];
      acc[0] = fma(u, kind == 0 ? (double)y[i] : u, acc[0]);
    }
  }
  if (kind >= 2)
    block_partials<1, 1u>(acc, partials);
  else
    block_partials<1, 0u>(acc, partials);
}

template <class T>
int launch_elem(fx_ctx ctx, int op, int form, double a, double b, const void* x, void* y,
                uint32_t n) {
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_elem<T>, dim3(stream_blocks(n)), dim3(kThreads), 0, ctx->stream, op, form,
                     a, b, (const T*)x, (T*)y, n);
  FX_HIP(hipGetLastError());
  return 0;
}

int elem_dispatch(fx_ctx ctx, int op, int form, double a, double b, fx_vec x, fx_vec y) {
  FX_CHECK(y, FX_ERROR_INPUT_PAR, "null vector");
  if (x) {
    FX_CHECK(x->n == y->n, FX_ERROR_NONMATCH_SIZE, "vector sizes differ (%u vs %u)", x->n, y->n);
    FX_CHECK(x->dtype == y->dtype, FX_ERROR_INPUT_PAR, "vector dtypes differ");
  }
  const void* xp = x ? x->d : nullptr;
  if (y->dtype == FX_F64) return launch_elem<double>(ctx, op, form, a, b, xp, y->d, y->n);
  return launch_elem<float>(ctx, op, form, a, b, xp, y->d, y->n);
}

int reduce_dispatch(fx_ctx ctx, int kind, int upd, double a, fx_vec x, fx_vec y, fx_vec z,
                    double* dev_out) {
  FX_CHECK(x, FX_ERROR_INPUT_PAR, "null vector");
  if (y) FX_CHECK(y->n == x->n && y->dtype == x->dtype, FX_ERROR_NONMATCH_SIZE, "vector mismatch");
  if (z) FX_CHECK(z->n == x->n && z->dtype == x->dtype, FX_ERROR_NONMATCH_SIZE, "vector mismatch");
  const unsigned g = stream_blocks(x->n);
  if (x->dtype == FX_F64)
    hipLaunchKernelGGL(k_reduce<double>, dim3(g), dim3(kThreads), 0, ctx->stream, kind, upd, a,
                       (const double*)x->d, (double*)(y ? y->d : nullptr),
                       (const double*)(z ? z->d : nullptr), x->n, ctx->ws.partials);
  else
    hipLaunchKernelGGL(k_reduce<float>, dim3(g), dim3(kThreads), 0, ctx->stream, kind, upd, a,
                       (const float*)x->d, (float*)(y ? y->d : nullptr),
                       (const float*)(z ? z->d : nullptr), x->n, ctx->ws.partials);
  FX_HIP(hipGetLastError());
  if (kind >= 2)
    return launch_finish<1, 1u>(ctx->stream, ctx->ws.partials, g,
                                FinStore<1>{dev_out, kind == 4 ? 2 : 0});
  return launch_finish<1, 0u>(ctx->stream, ctx->ws.partials, g, FinStore<1>{dev_out, kind == 1});
}

int fetch_scalar(fx_ctx ctx, int slot, double* out) {
  FX_HIP(hipMemcpyAsync(ctx->hpin + slot, ctx->dscal + slot, sizeof(double),
                        hipMemcpyDeviceToHost, ctx->stream));
  FX_HIP(hipStreamSynchronize(ctx->stream));
  *out = ctx->hpin[slot];
  return 0;
}

int convert_copy(fx_ctx ctx, const void* s, int sdt, void* d, int ddt, uint32_t n) {
  if (n == 0) return 0;
  const unsigned g = stream_blocks(n);
  if (sdt == ddt) {
    FX_HIP(hipMemcpyAsync(d, s, (size_t)n * (sdt == FX_F64 ? 8 : 4), hipMemcpyDeviceToDevice,
                          ctx->stream));
  } else if (sdt == FX_F64) {
    hipLaunchKernelGGL((k_convert<double, float>), dim3(g), dim3(kThreads), 0, ctx->stream,
                       (const double*)s, (float*)d, n);
  } else {
    hipLaunchKernelGGL((k_convert<float, double>), dim3(g), dim3(kThreads), 0, ctx->stream,
                       (const float*)s, (double*)d, n);
  }
  FX_HIP(hipGetLastError());
  return 0;
}

}  // namespace fx

using namespace fx;

extern "C" {

const char* fx_last_error(void) { return g_err; }

void fx_params_default(fx_params* p) {
  // SOLParams(), src/Param.hxx:68-79; GMRES() restart bounds src/GMRES.hxx:61-62
  p->maxIter = 1;
  p->minIter = 0;
  p->savIter = 100000;
  p->restart = 30;
  p->relTol = 1e-6;
  p->absTol = 1e-8;
  p->verbose = 0;
  p->maxRestart = 30;
  p->minRestart = 10;
  p->leftPrecond = 0;
}

static int ctx_init(fx_ctx c) {
  FX_HIP(hipMalloc(&c->ws.partials, sizeof(double) * 16 * 65536));
  FX_HIP(hipMalloc(&c->dscal, sizeof(double) * 64));
  FX_HIP(hipHostMalloc(&c->hpin, sizeof(double) * 64, hipHostMallocDefault));
  hipDeviceProp_t prop;
  FX_HIP(hipGetDeviceProperties(&prop, c->device));
  c->sms = prop.multiProcessorCount;
  FX_TRY(noalloc_init(c->device, &c->scratch));  // the device is current here
  return 0;
}

int fx_ctx_create(int device, fx_ctx* out) {
  clear_error();
  FX_CHECK(out, FX_ERROR_INPUT_PAR, "null output");
  FX_HIP(hipSetDevice(device));
  fx_ctx c = new fx_ctx_s();
  c->device = device;
  c->own_stream = true;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    set_error("hipStreamCreate failed");
    return FX_ERROR_DEVICE;
  }
  int rc = ctx_init(c);
  if (rc) {
    fx_ctx_destroy(c);
    return rc;
  }
  *out = c;
  return 0;
}

int fx_ctx_create_on_stream(int device, void* stream, fx_ctx* out) {
  clear_error();
  FX_CHECK(out, FX_ERROR_INPUT_PAR, "null output");
  FX_HIP(hipSetDevice(device));
  fx_ctx c = new fx_ctx_s();
  c->device = device;
  c->stream = (hipStream_t)stream;
  c->own_stream = false;
  int rc = ctx_init(c);
  if (rc) {
    fx_ctx_destroy(c);
    return rc;
  }
  *out = c;
  return 0;
}

int fx_ctx_destroy(fx_ctx c) {
  if (!c) return 0;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  (void)hipFree(c->ws.partials);
  (void)hipFree(c->dscal);
  if (c->hpin) (void)hipHostFree(c->hpin);
  for (auto& e : c->tev)
    if (e) (void)hipEventDestroy(e);
  if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return 0;
}

int fx_ctx_sync(fx_ctx c) {
  FX_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

void* fx_ctx_stream(fx_ctx c) { return c ? (void*)c->stream : nullptr; }

// ---- vectors -------------------------------------------------------------

int fx_vec_create(fx_ctx ctx, uint32_t n, int dtype, fx_vec* out) {
  clear_error();
  FX_CHECK(ctx && out, FX_ERROR_INPUT_PAR, "null argument");
  FX_CHECK(dtype == FX_F64 || dtype == FX_F32, FX_ERROR_INPUT_PAR, "bad dtype %d", dtype);
  fx_vec v = new fx_vec_s();
  v->ctx = ctx;
  v->n = n;
  v->dtype = dtype;
  v->owned = true;
  const size_t bytes = (size_t)(n ? n : 1) * (dtype == FX_F64 ? 8 : 4);
  if (hipMalloc(&v->d, bytes) != hipSuccess) {
    delete v;
    set_error("hipMalloc of %zu bytes failed", bytes);
    return FX_ERROR_ALLOC_MEM;
  }
  if (hipMemsetAsync(v->d, 0, bytes, ctx->stream) != hipSuccess) {
    (void)hipFree(v->d);
    delete v;
    set_error("hipMemsetAsync failed");
    return FX_ERROR_DEVICE;
  }
  *out = v;
  return 0;
}

int fx_vec_wrap(fx_ctx ctx, uint32_t n, int dtype, void* dev_ptr, fx_vec* out) {
  clear_error();
  FX_CHECK(ctx && out && (dev_ptr || n == 0), FX_ERROR_INPUT_PAR, "null argument");
  FX_CHECK(dtype == FX_F64 || dtype == FX_F32, FX_ERROR_INPUT_PAR, "bad dtype %d", dtype);
  fx_vec v = new fx_vec_s();
  v->ctx = ctx;
  v->n = n;
  v->dtype = dtype;
  v->d = dev_ptr;
  v->owned = false;
  *out = v;
  return 0;
}

int fx_vec_destroy(fx_vec v) {
  if (!v) return 0;
  if (v->owned && v->d) {
    (void)hipStreamSynchronize(v->ctx->stream);
    (void)hipFree(v->d);
  }
  delete v;
  return 0;
}

uint32_t fx_vec_size(fx_vec v) { return v ? v->n : 0; }
int fx_vec_dtype(fx_vec v) { return v ? v->dtype : -1; }
void* fx_vec_data(fx_vec v) { return v ? v->d : nullptr; }

int fx_vec_upload(fx_vec v, const void* host) {
  clear_error();
  FX_CHECK(v && (host || v->n == 0), FX_ERROR_INPUT_PAR, "null argument");
  if (v->n == 0) return 0;
  FX_HIP(hipMemcpyAsync(v->d, host, (size_t)v->n * (v->dtype == FX_F64 ? 8 : 4),
                        hipMemcpyHostToDevice, v->ctx->stream));
  FX_HIP(hipStreamSynchronize(v->ctx->stream));
  return 0;
}

int fx_vec_download(fx_vec v, void* host) {
  clear_error();
  FX_CHECK(v && (host || v->n == 0), FX_ERROR_INPUT_PAR, "null argument");
  if (v->n == 0) return 0;
  FX_HIP(hipMemcpyAsync(host, v->d, (size_t)v->n * (v->dtype == FX_F64 ? 8 : 4),
                        hipMemcpyDeviceToHost, v->ctx->stream));
  FX_HIP(hipStreamSynchronize(v->ctx->stream));
  return 0;
}

int fx_vec_set(fx_vec v, double value) {
  clear_error();
  FX_CHECK(v, FX_ERROR_INPUT_PAR, "null vector");
  return elem_dispatch(v->ctx, OP_SET, 0, value, 0.0, nullptr, v);
}

int fx_vec_copy(fx_vec s, fx_vec d) {
  clear_error();
  FX_CHECK(s && d, FX_ERROR_INPUT_PAR, "null vector");
  FX_CHECK(s->n == d->n, FX_ERROR_NONMATCH_SIZE, "vector sizes differ (%u vs %u)", s->n, d->n);
  return convert_copy(d->ctx, s->d, s->dtype, d->d, d->dtype, s->n);
}

int fx_axpy(fx_ctx ctx, double a, fx_vec x, fx_vec y) {
  clear_error();
  return elem_dispatch(ctx, OP_AXPY, 0, a, 0.0, x, y);
}
int fx_xpay(fx_ctx ctx, double a, fx_vec x, fx_vec y) {
  clear_error();
  return elem_dispatch(ctx, OP_XPAY, 0, a, 0.0, x, y);
}
int fx_axpby(fx_ctx ctx, double a, double b, fx_vec x, fx_vec y) {
  clear_error();
  const int form = (a == 1.0) + 2 * (b == 1.0);
  return elem_dispatch(ctx, OP_AXPBY, form, a, b, x, y);
}
int fx_waxpby(fx_ctx ctx, double a, fx_vec x, double b, fx_vec y, fx_vec w) {
  clear_error();
  FX_CHECK(x && y && w, FX_ERROR_INPUT_PAR, "null vector");
  FX_CHECK(x->n == y->n && x->n == w->n, FX_ERROR_NONMATCH_SIZE, "vector sizes differ");
  FX_CHECK(x->dtype == y->dtype && x->dtype == w->dtype, FX_ERROR_INPUT_PAR, "dtypes differ");
  if (x->n == 0) return 0;
  const int form = (a == 1.0) + 2 * (b == 1.0);
  const unsigned g = stream_blocks(x->n);
  if (x->dtype == FX_F64)
    hipLaunchKernelGGL(k_waxpby<double>, dim3(g), dim3(kThreads), 0, ctx->stream, form, a,
                       (const double*)x->d, b, (const double*)y->d, (double*)w->d, x->n);
  else
    hipLaunchKernelGGL(k_waxpby<float>, dim3(g), dim3(kThreads), 0, ctx->stream, form, a,
                       (const float*)x->d, b, (const float*)y->d, (float*)w->d, x->n);
  FX_HIP(hipGetLastError());
  return 0;
}
int fx_scale(fx_ctx ctx, double a, fx_vec y) {
  clear_error();
  return elem_dispatch(ctx, OP_SCALE, 0, a, 0.0, nullptr, y);
}
int fx_pmul(fx_ctx ctx, fx_vec v, fx_vec y) {
  clear_error();
  return elem_dispatch(ctx, OP_PMUL, 0, 0.0, 0.0, v, y);
}
int fx_recip(fx_ctx ctx, fx_vec y) {
  clear_error();
  return elem_dispatch(ctx, OP_RECIP, 0, 0.0, 0.0, nullptr, y);
}
int fx_add(fx_ctx ctx, fx_vec v, fx_vec y) {
  clear_error();
  return elem_dispatch(ctx, OP_ADD, 0, 0.0, 0.0, v, y);
}

int fx_dot_dev(fx_ctx ctx, fx_vec x, fx_vec y, double* dev_out) {
  clear_error();
  FX_CHECK(y, FX_ERROR_INPUT_PAR, "null vector");
  return reduce_dispatch(ctx, 0, 0, 0.0, x, y, nullptr, dev_out);
}
int fx_norm2_dev(fx_ctx ctx, fx_vec x, double* dev_out) {
  clear_error();
  return reduce_dispatch(ctx, 1, 0, 0.0, x, nullptr, nullptr, dev_out);
}
int fx_dot(fx_ctx ctx, fx_vec x, fx_vec y, double* out) {
  FX_TRY(fx_dot_dev(ctx, x, y, ctx->dscal));
  return fetch_scalar(ctx, 0, out);
}
int fx_norm2(fx_ctx ctx, fx_vec x, double* out) {
  FX_TRY(fx_norm2_dev(ctx, x, ctx->dscal));
  return fetch_scalar(ctx, 0, out);
}
int fx_norminf(fx_ctx ctx, fx_vec x, double* out) {
  clear_error();
  FX_TRY(reduce_dispatch(ctx, 2, 0, 0.0, x, nullptr, nullptr, ctx->dscal));
  return fetch_scalar(ctx, 0, out);
}
int fx_max(fx_ctx ctx, fx_vec x, double* out) {
  clear_error();
  FX_TRY(reduce_dispatch(ctx, 3, 0, 0.0, x, nullptr, nullptr, ctx->dscal));
  return fetch_scalar(ctx, 0, out);
}
int fx_min(fx_ctx ctx, fx_vec x, double* out) {
  clear_error();
  FX_TRY(reduce_dispatch(ctx, 4, 0, 0.0, x, nullptr, nullptr, ctx->dscal));
  return fetch_scalar(ctx, 0, out);
}
int fx_shift(fx_ctx ctx, double a, fx_vec y) {
  clear_error();
  return elem_dispatch(ctx, OP_SHIFT, 0, a, 0.0, nullptr, y);
}
int fx_abs(fx_ctx ctx, fx_vec y) {
  clear_error();
  return elem_dispatch(ctx, OP_ABS, 0, 0.0, 0.0, nullptr, y);
}
int fx_pdiv(fx_ctx ctx, fx_vec v, fx_vec y) {
  clear_error();
  return elem_dispatch(ctx, OP_PDIV, 0, 0.0, 0.0, v, y);
}
int fx_sub(fx_ctx ctx, fx_vec v, fx_vec y) {
  clear_error();
  return elem_dispatch(ctx, OP_SUB, 0, 0.0, 0.0, v, y);
}

int fx_axpy_norm2(fx_ctx ctx, double a, fx_vec x, fx_vec y, double* out) {
  clear_error();
  FX_CHECK(y, FX_ERROR_INPUT_PAR, "null vector");
  FX_TRY(reduce_dispatch(ctx, 1, 1, a, x, y, nullptr, ctx->dscal));
  return fetch_scalar(ctx, 0, out);
}
int fx_xpay_norm2(fx_ctx ctx, double a, fx_vec x, fx_vec y, double* out) {
  clear_error();
  FX_CHECK(y, FX_ERROR_INPUT_PAR, "null vector");
  FX_TRY(reduce_dispatch(ctx, 1, 2, a, x, y, nullptr, ctx->dscal));
  return fetch_scalar(ctx, 0, out);
}
int fx_axpy_dot(fx_ctx ctx, double a, fx_vec x, fx_vec y, fx_vec z, double* out) {
  clear_error();
  FX_CHECK(y && z, FX_ERROR_INPUT_PAR, "null vector");
  FX_TRY(reduce_dispatch(ctx, 0, 1, a, x, y, z, ctx->dscal));
  return fetch_scalar(ctx, 0, out);
}

}  // extern "C"
