// Test hooks: run one kernel family on host data (used by the GPU unit tests
// that check MFMA operand layouts, the Gauss-Jordan sweep and CholeskyQR).
#include <hip/hip_runtime.h>

#include <vector>

#include "gemm.hpp"
#include "runtime.hpp"
#include "small_dense.hpp"

namespace {

// plain streaming copy: the reference point for HBM traffic counters
__global__ __launch_bounds__(256) void copy_kernel(const float4* __restrict__ src, float4* __restrict__ dst, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

template <typename T>
std::vector<cplx<T>> narrow(const double* p, size_t n) {
  std::vector<cplx<T>> v(n);
  for (size_t i = 0; i < n; ++i) v[i] = cmk<T>((T)p[2 * i], (T)p[2 * i + 1]);
  return v;
}

template <typename T>
int zgemm_t(afqmc_ctx ctx, int opa, int M, int N, int K, int batch, const double* A, long long strideA, const double* B,
            const double* scale, int scale_group, int flags, int tile, double* C) {
  hipStream_t s = ctx->stream;
  const long long lda = (opa == OP_N || opa == OP_NR) ? M : K;
  const size_t nA = (size_t)(strideA ? strideA * batch : (long long)M * K);
  const size_t nB = (size_t)K * N * batch, nC = (size_t)M * N * batch;
  auto hB = narrow<T>(B, nB);
  DevBuf<cplx<T>> dA, dB, dC, dS;
  int rc = dA.alloc(nA);
  if (!rc) rc = dB.alloc(nB);
  if (!rc) rc = dC.alloc(nC);
  if (rc) return rc;
  if (opa == OP_NR) {  // A holds nA real values
    std::vector<T> hr(nA);
    for (size_t i = 0; i < nA; ++i) hr[i] = (T)A[i];
    HIP_TRY(hipMemcpy(dA.p, hr.data(), sizeof(T) * nA, hipMemcpyHostToDevice));
  } else {
    auto hA = narrow<T>(A, nA);
    HIP_TRY(hipMemcpy(dA.p, hA.data(), sizeof(cplx<T>) * nA, hipMemcpyHostToDevice));
  }
  HIP_TRY(hipMemcpy(dB.p, hB.data(), sizeof(cplx<T>) * nB, hipMemcpyHostToDevice));
  GemmArgs a;
  memset(&a, 0, sizeof(a));
  a.M = M; a.N = N; a.K = K; a.batch = batch;
  a.A = dA.p; a.lda = lda; a.strideA = strideA;
  a.B = dB.p; a.ldb = K; a.strideB = (long long)K * N;
  a.C = dC.p; a.ldc = M; a.strideC = (long long)M * N;
  a.alpha_re = 1.0;
  a.scale_group = 1;
  a.herm = flags & 1;
  a.b_upper = (flags >> 1) & 1;
  a.split_bf16 = (flags >> 2) & 1;
  a.tile = tile;
  if (scale) {
    // one K-vector per column group of scale_group columns (per batch: N / scale_group groups)
    const int groups = (N + scale_group - 1) / scale_group;
    const size_t nS = (size_t)K * groups * batch;
    auto hS = narrow<T>(scale, nS);
    if ((rc = dS.alloc(nS))) return rc;
    HIP_TRY(hipMemcpy(dS.p, hS.data(), sizeof(cplx<T>) * nS, hipMemcpyHostToDevice));
    a.scale = dS.p; a.scale_group = scale_group; a.scale_strideG = K; a.scale_strideB = (long long)K * groups;
  }
  if ((rc = launch_cgemm<T>(opa, EPI_STORE, a, s))) return rc;
  std::vector<cplx<T>> hC(nC);
  HIP_TRY(hipMemcpyAsync(hC.data(), dC.p, sizeof(cplx<T>) * nC, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  for (size_t i = 0; i < nC; ++i) { C[2 * i] = hC[i].re; C[2 * i + 1] = hC[i].im; }
  return 0;
}

template <typename T>
int gj_t(afqmc_ctx ctx, int kind, int N, int batch, const double* A, double* Ainv, double* logdet, int* info,
         int layout = 0, const int* alive = nullptr) {
  hipStream_t s = ctx->stream;
  const size_t n = (size_t)N * N * batch;
  auto h = narrow<T>(A, n);
  DevBuf<cplx<T>> d;
  DevBuf<zd> ld;
  DevBuf<int> inf, al;
  DevBuf<char> wk;
  int rc = d.alloc(n);
  if (!rc && kind == 0 && small_dense_gj_work_bytes(N, batch) > 0) rc = wk.alloc_bytes(small_dense_gj_work_bytes(N, batch));
  if (!rc) rc = ld.alloc(batch);
  if (!rc) rc = inf.alloc(batch);
  if (!rc && alive) rc = al.alloc(batch);
  if (rc) return rc;
  HIP_TRY(hipMemcpy(d.p, h.data(), sizeof(cplx<T>) * n, hipMemcpyHostToDevice));
  if (alive) HIP_TRY(hipMemcpy(al.p, alive, sizeof(int) * batch, hipMemcpyHostToDevice));
  // the outputs of skipped matrices keep these values
  HIP_TRY(hipMemcpy(ld.p, logdet, sizeof(zd) * batch, hipMemcpyHostToDevice));
  SmallDenseArgs a;
  memset(&a, 0, sizeof(a));
  a.N = N; a.batch = batch; a.in = d.p; a.ld_in = N; a.stride_in = (long long)N * N;
  a.out = d.p; a.ld_out = N; a.stride_out = a.stride_in; a.logdet = ld.p; a.info = inf.p;
  a.write_inverse = 1; a.nsub = N; a.alive_div = 1; a.layout = layout; a.alive = alive ? al.p : nullptr;
  a.work = wk.p;
  rc = kind == 0 ? launch_gj_inverse<T>(a, s) : launch_chol_inverse<T>(a, s);
  if (rc) return rc;
  std::vector<cplx<T>> o(n);
  HIP_TRY(hipMemcpyAsync(o.data(), d.p, sizeof(cplx<T>) * n, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(logdet, ld.p, sizeof(zd) * batch, hipMemcpyDeviceToHost, s));
  if (info) HIP_TRY(hipMemcpyAsync(info, inf.p, sizeof(int) * batch, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  for (size_t i = 0; i < n; ++i) { Ainv[2 * i] = o[i].re; Ainv[2 * i + 1] = o[i].im; }
  return 0;
}

}  // namespace

extern "C" {

int afqmc_test_zgemm(afqmc_ctx ctx, int precision, int opa, int M, int N, int K, int batch, const double* A,
                     long long strideA, const double* B, const double* scale, int scale_group, int flags, int tile,
                     double* C) {
  REQUIRE(ctx && A && B && C, "null argument");
  REQUIRE(flags >= 0 && flags <= 7, "flags: bit 0 Hermitian result, bit 1 upper-triangular B, bit 2 split-BF16 products");
  REQUIRE(M > 0 && N > 0 && K > 0 && batch > 0, "sizes must be positive");
  REQUIRE(opa >= 0 && opa <= 3, "opa must be 0 (N), 1 (T), 2 (C) or 3 (N with real A)");
  REQUIRE(!scale || opa == OP_N, "row scaling is supported with op(A) = N");
  REQUIRE(opa != OP_NR || !strideA || strideA > 0, "invalid strideA");
  ScopedDevice sd(ctx);
  return precision ? zgemm_t<float>(ctx, opa, M, N, K, batch, A, strideA, B, scale, scale_group, flags, tile, C)
                   : zgemm_t<double>(ctx, opa, M, N, K, batch, A, strideA, B, scale, scale_group, flags, tile, C);
}

int afqmc_test_gj_inverse(afqmc_ctx ctx, int precision, int N, int batch, const double* A, double* Ainv,
                          double* logdet, int* info) {
  REQUIRE(ctx && A && Ainv && logdet, "null argument");
  ScopedDevice sd(ctx);
  return precision ? gj_t<float>(ctx, 0, N, batch, A, Ainv, logdet, info)
                   : gj_t<double>(ctx, 0, N, batch, A, Ainv, logdet, info);
}

int afqmc_test_gj_inverse_layout(afqmc_ctx ctx, int precision, int N, int batch, int layout, const double* A,
                                 double* Ainv, double* logdet, int* info) {
  REQUIRE(ctx && A && Ainv && logdet, "null argument");
  ScopedDevice sd(ctx);
  return precision ? gj_t<float>(ctx, 0, N, batch, A, Ainv, logdet, info, layout)
                   : gj_t<double>(ctx, 0, N, batch, A, Ainv, logdet, info, layout);
}

int afqmc_test_gj_inverse_masked(afqmc_ctx ctx, int precision, int N, int batch, int layout, const int* alive,
                                 const double* A, double* Ainv, double* logdet, int* info) {
  REQUIRE(ctx && alive && A && Ainv && logdet, "null argument");
  ScopedDevice sd(ctx);
  return precision ? gj_t<float>(ctx, 0, N, batch, A, Ainv, logdet, info, layout, alive)
                   : gj_t<double>(ctx, 0, N, batch, A, Ainv, logdet, info, layout, alive);
}

int afqmc_test_copy(afqmc_ctx ctx, long long bytes, int reps, double* ms) {
  REQUIRE(ctx && bytes > 0 && bytes % 16 == 0 && reps > 0, "afqmc_test_copy: bytes must be a positive multiple of 16");
  ScopedDevice sd(ctx);
  hipStream_t s = ctx->stream;
  DevBuf<float4> a, b;
  const long long n = bytes / 16;
  int rc = a.alloc((size_t)n);
  if (!rc) rc = b.alloc((size_t)n);
  if (rc) return rc;
  HIP_TRY(hipMemsetAsync(a.p, 0, (size_t)bytes, s));
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  HIP_TRY(hipEventRecord(e0, s));
  for (int r = 0; r < reps; ++r) {
    hipLaunchKernelGGL(copy_kernel, dim3(4096), dim3(256), 0, s, a.p, b.p, n);
    HIP_LAUNCH_CHECK();
  }
  HIP_TRY(hipEventRecord(e1, s));
  HIP_TRY(hipEventSynchronize(e1));
  float t = 0.f;
  HIP_TRY(hipEventElapsedTime(&t, e0, e1));
  if (ms) *ms = t / reps;
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  return 0;
}

int afqmc_test_chol_inverse(afqmc_ctx ctx, int N, int batch, const double* S, double* Rinv, double* logsum,
                            int* info) {
  REQUIRE(ctx && S && Rinv && logsum, "null argument");
  ScopedDevice sd(ctx);
  return gj_t<double>(ctx, 1, N, batch, S, Rinv, logsum, info);
}

}  // extern "C"
