#pragma once
#include "common.hpp"

struct SmallDenseArgs {
  int N, batch;
  const void* in; long long ld_in, stride_in;   // complex T, N x N column-major per matrix
  void* out; long long ld_out, stride_out;      // complex T (may alias in)
  zd* logdet;          // gj: log det (principal branch); chol: (sum log R_kk, sum_{k<nsub} log R_kk)
  int* info;           // gj: 1 if singular; chol: 1 if not positive definite
  int write_inverse;   // gj only
  int nsub;            // chol only
  const int* alive;    // optional: skip matrix b if alive[b / alive_div] == 0
  int alive_div;
  double pivot_tol;    // chol only: info = 1 also when a pivot of the Jacobi-scaled matrix is <= this
  int layout;          // gj only: 0 the default kernel for N, 1 the 1-D lane layout, 2 the 2-D lane layout,
                       // 3 panels of four columns (N <= 112)
  void* work;          // gj with 161 <= N <= 256 only: FP64 complex working copies, N * N per matrix
                       // of the batch (the matrix does not fit a CU's registers and LDS)
};
// bytes of SmallDenseArgs::work a Gauss-Jordan call of this size needs (0 up to 160 columns)
long long small_dense_gj_work_bytes(int N, int batch);

template <typename T> int launch_gj_inverse(const SmallDenseArgs& a, hipStream_t s);
template <typename T> int launch_chol_inverse(const SmallDenseArgs& a, hipStream_t s);
