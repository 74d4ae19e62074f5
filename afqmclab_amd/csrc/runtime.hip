// Context management, error reporting and elementwise helper kernels.
#include <cstdarg>
#include <cstdio>

#include "runtime.hpp"

static thread_local char g_err[1024] = "";

void afqmc_set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

namespace {

template <typename A, typename B>
__global__ void convert_kernel(const cplx<A>* in, cplx<B>* out, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    out[i] = cmk<B>((B)in[i].re, (B)in[i].im);
}

__global__ void logw_shift_kernel(int nw, const zd* logw, zd* out, zd c) {
  int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w < nw) out[w] = cadd(logw[w], c);
}
__global__ void logw_const_add_kernel(int nw, zd* logw, const int* alive, zd c) {
  int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w < nw && alive[w]) logw[w] = cadd(logw[w], c);
}
__global__ void logw_vec_add_kernel(int nw, zd* logw, const int* alive, const zd* v) {
  int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w < nw && alive[w]) logw[w] = cadd(logw[w], v[w]);
}
__global__ void select_alive_kernel(int nw, long long n16, const int* alive, const uint4* src, uint4* dst) {
  const int w = blockIdx.y;
  if (!alive[w]) return;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (long long)gridDim.x * blockDim.x)
    dst[(long long)w * n16 + i] = src[(long long)w * n16 + i];
}
__global__ void select_alive_kernel8(int nw, long long n8, const int* alive, const uint2* src, uint2* dst) {
  const int w = blockIdx.y;
  if (!alive[w]) return;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n8; i += (long long)gridDim.x * blockDim.x)
    dst[(long long)w * n8 + i] = src[(long long)w * n8 + i];
}
__global__ void reset_measure_kernel(DevScalars* sc, zd* site_acc, int L) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    zd z = cmk<double>(0, 0);
    sc->den = z; sc->HNum = z; sc->KNum = z; sc->VNum = z; sc->RNum = z;
    sc->NupTot = z; sc->NdnTot = z; sc->SplusTot = z; sc->SminusTot = z;
  }
  if (i < 4 * L) site_acc[i] = cmk<double>(0, 0);
}

}  // namespace

int launch_convert_f32_f64(const void* in, void* out, long long n, hipStream_t s) {
  unsigned g = (unsigned)std::min<long long>((n + 255) / 256, 65536);
  if (n <= 0) return 0;
  hipLaunchKernelGGL((convert_kernel<float, double>), dim3(g), dim3(256), 0, s, (const zf*)in, (zd*)out, n);
  HIP_LAUNCH_CHECK();
  return 0;
}
int launch_convert_f64_f32(const void* in, void* out, long long n, hipStream_t s) {
  unsigned g = (unsigned)std::min<long long>((n + 255) / 256, 65536);
  if (n <= 0) return 0;
  hipLaunchKernelGGL((convert_kernel<double, float>), dim3(g), dim3(256), 0, s, (const zd*)in, (zf*)out, n);
  HIP_LAUNCH_CHECK();
  return 0;
}
int launch_logw_shift(int nw, const zd* logw, zd* out, zd c, hipStream_t s) {
  hipLaunchKernelGGL(logw_shift_kernel, dim3((nw + 255) / 256), dim3(256), 0, s, nw, logw, out, c);
  HIP_LAUNCH_CHECK();
  return 0;
}
int launch_logw_const_add(int nw, zd* logw, const int* alive, zd c, hipStream_t s) {
  hipLaunchKernelGGL(logw_const_add_kernel, dim3((nw + 255) / 256), dim3(256), 0, s, nw, logw, alive, c);
  HIP_LAUNCH_CHECK();
  return 0;
}
int launch_logw_vec_add(int nw, zd* logw, const int* alive, const zd* v, hipStream_t s) {
  hipLaunchKernelGGL(logw_vec_add_kernel, dim3((nw + 255) / 256), dim3(256), 0, s, nw, logw, alive, v);
  HIP_LAUNCH_CHECK();
  return 0;
}
int launch_select_alive(int nw, long long n, size_t esz, const int* alive, const void* src, void* dst, hipStream_t s) {
  if (esz == 16) {
    long long n16 = n;
    dim3 grid((unsigned)std::min<long long>((n16 + 255) / 256, 1024), nw);
    hipLaunchKernelGGL(select_alive_kernel, grid, dim3(256), 0, s, nw, n16, alive, (const uint4*)src, (uint4*)dst);
  } else {
    long long n8 = n;
    dim3 grid((unsigned)std::min<long long>((n8 + 255) / 256, 1024), nw);
    hipLaunchKernelGGL(select_alive_kernel8, grid, dim3(256), 0, s, nw, n8, alive, (const uint2*)src, (uint2*)dst);
  }
  HIP_LAUNCH_CHECK();
  return 0;
}
int launch_reset_measure(DevScalars* sc, zd* site_acc, int L, hipStream_t s) {
  hipLaunchKernelGGL(reset_measure_kernel, dim3((4 * L + 255) / 256 + (L == 0)), dim3(256), 0, s, sc, site_acc, L);
  HIP_LAUNCH_CHECK();
  return 0;
}

extern "C" {

const char* afqmc_last_error(void) { return g_err; }

int afqmc_version(void) { return 1; }

int afqmc_ctx_create(int device, afqmc_ctx* out) {
  REQUIRE(out, "null argument");
  int n = 0;
  HIP_TRY(hipGetDeviceCount(&n));
  REQUIRE(device >= 0 && device < n, "device %d not available (%d devices)", device, n);
  afqmc_ctx c = new afqmc_ctx_s();
  c->device = device;
  int prev = 0;
  hipGetDevice(&prev);
  hipSetDevice(device);
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  hipSetDevice(prev);
  if (e != hipSuccess) {
    delete c;
    afqmc_set_error("hipStreamCreate failed: %s", hipGetErrorString(e));
    return -2;
  }
  *out = c;
  return 0;
}

int afqmc_ctx_destroy(afqmc_ctx ctx) {
  if (!ctx) return 0;
  ScopedDevice sd(ctx);
  hipStreamSynchronize(ctx->stream);
  hipStreamDestroy(ctx->stream);
  delete ctx;
  return 0;
}

int afqmc_ctx_synchronize(afqmc_ctx ctx) {
  REQUIRE(ctx, "null argument");
  ScopedDevice sd(ctx);
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return 0;
}

static int resolve_events(afqmc_ctx ctx) {
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  for (auto& e : ctx->events) {
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, e.second.first, e.second.second));
    ctx->total_ms[e.first] += ms;
    ctx->count[e.first] += 1;
    hipEventDestroy(e.second.first);
    hipEventDestroy(e.second.second);
  }
  ctx->events.clear();
  return 0;
}

int afqmc_ctx_set_timing(afqmc_ctx ctx, int enable) {
  REQUIRE(ctx, "null argument");
  ScopedDevice sd(ctx);
  int rc = resolve_events(ctx);
  if (rc) return rc;
  ctx->timing = enable ? 1 : 0;
  for (int r = 0; r < REG_COUNT; ++r) { ctx->total_ms[r] = 0.0; ctx->count[r] = 0; }
  return 0;
}

int afqmc_ctx_get_timing(afqmc_ctx ctx, double* total_ms, long long* count, int n) {
  REQUIRE(ctx && total_ms && count, "null argument");
  ScopedDevice sd(ctx);
  int rc = resolve_events(ctx);
  if (rc) return rc;
  for (int r = 0; r < n && r < REG_COUNT; ++r) { total_ms[r] = ctx->total_ms[r]; count[r] = ctx->count[r]; }
  return 0;
}

}  // extern "C"
