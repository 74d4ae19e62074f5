// The comb (popConfiguration, size 1) on device buffers without any allocation: the working
// memory of the comb kernels (the run scalars, the probabilities and the slot counts) lives in a
// file-scope device array instead of hipMalloc'd buffers, for callers that cannot allocate
// (code under HIP graph capture, tight serving loops).  Each device holds its own copy of the
// array; its address is looked up once per device when a context is created on it.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/afqmc_mi355x.h"
#include "runtime.hpp"

namespace {

constexpr size_t kScalarsBytes = 256;
constexpr size_t kScratchBytes =
    kScalarsBytes + (size_t)AFQMC_COMB_NOALLOC_MAX * sizeof(double) + (size_t)AFQMC_COMB_NOALLOC_MAX * sizeof(int);
static_assert(sizeof(DevScalars) <= kScalarsBytes, "DevScalars must fit the scratch header");

static __device__ __align__(256) char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
char* g_ptr[kMaxDevices] = {nullptr};

// run scalars of one comb: stream position 0, uniforms from the caller's buffer of nw values
__global__ void noalloc_init_kernel(DevScalars* sc, int nw) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  DevScalars z{};
  z.ulen = (unsigned long long)nw;
  *sc = z;
}

__global__ void noalloc_stats_kernel(const DevScalars* sc, double* stats) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  stats[0] = sc->wmin;
  stats[1] = sc->wmax;
  stats[2] = sc->wavg;
}

}  // namespace

// Looks up this device's copy of the scratch array (the device must be current).  Called by
// afqmc_ctx_create; later calls for the same device return at once.
int afqmc_noalloc_scratch_init(int device) {
  REQUIRE(device >= 0 && device < kMaxDevices, "device %d beyond the %d devices the scratch table holds", device,
          kMaxDevices);
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_ptr[device]) return 0;
  void* p = nullptr;
  HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  REQUIRE(p != nullptr, "hipGetSymbolAddress returned no address for the comb scratch array");
  g_ptr[device] = static_cast<char*>(p);
  return 0;
}

extern "C" int afqmc_comb_noalloc(afqmc_ctx ctx, int nw, const double* weight, const double* uniforms, int* table,
                                  double* stats) {
  REQUIRE(ctx && weight && uniforms && table, "afqmc_comb_noalloc: null argument");
  REQUIRE(nw > 0 && nw <= AFQMC_COMB_NOALLOC_MAX, "afqmc_comb_noalloc: %d walkers outside 1..%d", nw,
          AFQMC_COMB_NOALLOC_MAX);
  REQUIRE(ctx->device >= 0 && ctx->device < kMaxDevices, "afqmc_comb_noalloc: device %d out of range", ctx->device);
  char* base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    base = g_ptr[ctx->device];
  }
  REQUIRE(base, "afqmc_comb_noalloc: no scratch address for device %d (context not created by afqmc_ctx_create)",
          ctx->device);
  ScopedDevice sd(ctx);
  hipStream_t s = ctx->stream;
  DevScalars* sc = reinterpret_cast<DevScalars*>(base);
  double* prob = reinterpret_cast<double*>(base + kScalarsBytes);
  int* number = reinterpret_cast<int*>(base + kScalarsBytes + (size_t)AFQMC_COMB_NOALLOC_MAX * sizeof(double));
  hipLaunchKernelGGL(noalloc_init_kernel, dim3(1), dim3(64), 0, s, sc, nw);
  HIP_LAUNCH_CHECK();
  int rc = launch_comb(nw, weight, uniforms, sc, table, prob, number, s);
  if (rc) return rc;
  if (stats) {
    hipLaunchKernelGGL(noalloc_stats_kernel, dim3(1), dim3(64), 0, s, sc, stats);
    HIP_LAUNCH_CHECK();
  }
  return 0;
}
