// Shared helpers for the MI355X (gfx950) AFQMC walker-path kernels.
//
// Complex numbers are stored interleaved (re, im), column-major, exactly as
// the reference's TensorHao<complex<double>,2> (tensor_core.h:85-106), so host
// buffers upload without repacking.  T is the storage / GEMM type of a walker
// set (double for the FP64 path, float for the mixed path); all small dense
// work (LU, inverse, Cholesky, log-determinants, weights) runs in double.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <string>

#define AFQMC_HD __host__ __device__ __forceinline__

template <typename T> struct cplx { T re, im; };
using zd = cplx<double>;
using zf = cplx<float>;

template <typename T> AFQMC_HD cplx<T> cmk(T r, T i) { cplx<T> z; z.re = r; z.im = i; return z; }
template <typename T> AFQMC_HD cplx<T> cadd(cplx<T> a, cplx<T> b) { return cmk<T>(a.re + b.re, a.im + b.im); }
template <typename T> AFQMC_HD cplx<T> csub(cplx<T> a, cplx<T> b) { return cmk<T>(a.re - b.re, a.im - b.im); }
template <typename T> AFQMC_HD cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return cmk<T>(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
template <typename T> AFQMC_HD cplx<T> cconj(cplx<T> a) { return cmk<T>(a.re, -a.im); }
template <typename T> AFQMC_HD cplx<T> cscale(cplx<T> a, T s) { return cmk<T>(a.re * s, a.im * s); }
// complex division following the straightforward formula (inputs are well scaled here)
AFQMC_HD zd cdiv(zd a, zd b) {
  double d = b.re * b.re + b.im * b.im;
  return cmk<double>((a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d);
}
AFQMC_HD zd cinv(zd b) {
  double d = b.re * b.re + b.im * b.im;
  return cmk<double>(b.re / d, -b.im / d);
}
AFQMC_HD double cabs(zd a) { return hypot(a.re, a.im); }
// complex exp / log in the principal branch
__device__ __forceinline__ zd cexp_d(zd a) {
  double e = exp(a.re);
  double s, c;
  sincos(a.im, &s, &c);
  return cmk<double>(e * c, e * s);
}
__device__ __forceinline__ zd clog_d(zd a) { return cmk<double>(log(hypot(a.re, a.im)), atan2(a.im, a.re)); }

template <typename T> AFQMC_HD zd to_zd(cplx<T> a) { return cmk<double>((double)a.re, (double)a.im); }
template <typename T> AFQMC_HD cplx<T> from_zd(zd a) { return cmk<T>((T)a.re, (T)a.im); }

// ---------------------------------------------------------------------------
// error handling: every C entry point returns 0 or a negative code and stores
// a message retrievable with afqmc_last_error().
// ---------------------------------------------------------------------------
void afqmc_set_error(const char* fmt, ...);

#define HIP_TRY(expr)                                                              \
  do {                                                                             \
    hipError_t _e = (expr);                                                        \
    if (_e != hipSuccess) {                                                        \
      afqmc_set_error("%s:%d %s -> %s", __FILE__, __LINE__, #expr, hipGetErrorString(_e)); \
      return -2;                                                                   \
    }                                                                              \
  } while (0)

// the grouped-launch collector open on this thread (group.hpp), if any: a kernel launched
// directly while it collects would run ahead of the recorded ones
struct GroupCollector;
extern thread_local GroupCollector* g_group;

#define HIP_LAUNCH_CHECK()                                                         \
  do {                                                                             \
    if (g_group) {                                                                 \
      afqmc_set_error("%s:%d a kernel with no grouped form ran inside a grouped call", __FILE__, __LINE__); \
      return -1;                                                                   \
    }                                                                              \
    hipError_t _e = hipGetLastError();                                             \
    if (_e != hipSuccess) {                                                        \
      afqmc_set_error("%s:%d launch -> %s", __FILE__, __LINE__, hipGetErrorString(_e)); \
      return -2;                                                                   \
    }                                                                              \
  } while (0)

#define REQUIRE(cond, ...)                                                         \
  do {                                                                             \
    if (!(cond)) {                                                                 \
      afqmc_set_error(__VA_ARGS__);                                                \
      return -1;                                                                   \
    }                                                                              \
  } while (0)

// ---------------------------------------------------------------------------
// MFMA 16x16x4 wrappers.  Operand layout, checked on gfx950 with a one-wave probe and
// by the GEMM tests:
//   A (16x4):  lane l holds A[l%16][l/16]
//   B (4x16):  lane l holds B[l/16][l%16]
//   D (16x16): lane l, register r holds D[row(l/16, r)][l%16] with
//              row = 4*(l/16) + r for f32 and row = l/16 + 4*r for f64
// ---------------------------------------------------------------------------
typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

template <typename T> struct mfma_traits;
template <> struct mfma_traits<double> {
  typedef d4 acc_t;
  static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lk, int r) { return lk + 4 * r; }
};
template <> struct mfma_traits<float> {
  typedef f4 acc_t;
  static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lk, int r) { return 4 * lk + r; }
};

// wave64 reductions
template <typename T> __device__ __forceinline__ T wave_sum(T v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
