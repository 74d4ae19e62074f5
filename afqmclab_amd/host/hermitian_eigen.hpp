// Hermitian eigendecomposition for the driver's one-time set-up (the reference calls LAPACK's
// zheev through BL_NAME(eigen), tensor_2d_bl_cpu.cpp): hipSOLVER's zheevd on the GPU.  Eigenvalues
// ascending, eigenvectors as the columns of V (column-major), as zheev returns them.
#pragma once
#include <complex>
#include <stdexcept>
#include <string>
#include <vector>

#include <hip/hip_runtime_api.h>
#include <hipsolver/hipsolver.h>

namespace afqmc_gpu {

inline void hermitian_eigen(int n, const std::vector<std::complex<double>>& A, std::vector<double>& w,
                            std::vector<std::complex<double>>& V) {
  if (n <= 0 || A.size() != (size_t)n * n) throw std::runtime_error("hermitian_eigen: bad matrix size");
  auto hip_ok = [](hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string("hermitian_eigen: ") + what + ": " + hipGetErrorString(e));
  };
  auto sol_ok = [](hipsolverStatus_t s, const char* what) {
    if (s != HIPSOLVER_STATUS_SUCCESS) throw std::runtime_error(std::string("hermitian_eigen: ") + what + " failed");
  };
  hipsolverHandle_t h = nullptr;
  hipDoubleComplex *dA = nullptr, *work = nullptr;
  double* dW = nullptr;
  int* dinfo = nullptr;
  auto release = [&]() {
    if (work) (void)hipFree(work);
    if (dinfo) (void)hipFree(dinfo);
    if (dW) (void)hipFree(dW);
    if (dA) (void)hipFree(dA);
    if (h) (void)hipsolverDestroy(h);
  };
  try {
    sol_ok(hipsolverCreate(&h), "hipsolverCreate");
    hip_ok(hipMalloc(&dA, sizeof(hipDoubleComplex) * n * n), "hipMalloc");
    hip_ok(hipMalloc(&dW, sizeof(double) * n), "hipMalloc");
    hip_ok(hipMalloc(&dinfo, sizeof(int)), "hipMalloc");
    hip_ok(hipMemcpy(dA, A.data(), sizeof(hipDoubleComplex) * n * n, hipMemcpyHostToDevice), "hipMemcpy");
    int lwork = 0;
    sol_ok(hipsolverZheevd_bufferSize(h, HIPSOLVER_EIG_MODE_VECTOR, HIPSOLVER_FILL_MODE_UPPER, n, dA, n, dW, &lwork),
           "zheevd_bufferSize");
    hip_ok(hipMalloc(&work, sizeof(hipDoubleComplex) * (size_t)(lwork > 0 ? lwork : 1)), "hipMalloc");
    sol_ok(hipsolverZheevd(h, HIPSOLVER_EIG_MODE_VECTOR, HIPSOLVER_FILL_MODE_UPPER, n, dA, n, dW, work, lwork, dinfo),
           "zheevd");
    int info = 0;
    hip_ok(hipMemcpy(&info, dinfo, sizeof(int), hipMemcpyDeviceToHost), "hipMemcpy");
    if (info != 0) throw std::runtime_error("hermitian_eigen: zheevd info = " + std::to_string(info));
    w.resize(n);
    V.resize((size_t)n * n);
    hip_ok(hipMemcpy(w.data(), dW, sizeof(double) * n, hipMemcpyDeviceToHost), "hipMemcpy");
    hip_ok(hipMemcpy(V.data(), dA, sizeof(hipDoubleComplex) * n * n, hipMemcpyDeviceToHost), "hipMemcpy");
  } catch (...) {
    release();
    throw;
  }
  release();
}

// V diag(exp(-alpha w)) V^H (HubbardSOC::returnExpMinusAlphaK, HubbardSOC.cpp:109-118)
inline std::vector<std::complex<double>> exp_minus_alpha(int n, const std::vector<double>& w,
                                                         const std::vector<std::complex<double>>& V, double alpha) {
  std::vector<std::complex<double>> out((size_t)n * n, 0.0), VD((size_t)n * n);
  for (int k = 0; k < n; ++k) {
    const double e = std::exp(-alpha * w[k]);
    for (int i = 0; i < n; ++i) VD[i + (size_t)k * n] = V[i + (size_t)k * n] * e;
  }
  // out(i, j) = sum_k VD(i, k) conj(V(j, k)), column j at a time
  for (int j = 0; j < n; ++j)
    for (int k = 0; k < n; ++k) {
      const std::complex<double> c = std::conj(V[j + (size_t)k * n]);
      const std::complex<double>* vd = &VD[(size_t)k * n];
      std::complex<double>* o = &out[(size_t)j * n];
      for (int i = 0; i < n; ++i) o[i] += vd[i] * c;
    }
  return out;
}

}  // namespace afqmc_gpu
