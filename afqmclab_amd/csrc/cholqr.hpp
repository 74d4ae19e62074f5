// FP64 CholeskyQR restabilisation of a batch of walker blocks (SD::stabilize, SD.cpp:42-56, and
// SD2is::stabilize, SD2is.cpp:48-76, whose Householder QR, tensor_2d_bl_cpu.cpp:263-315, this
// replaces), robust to ill-conditioned walkers.
#pragma once
#include "cp_kernels.hpp"
#include "runtime.hpp"

struct CholQrBlock {
  int nw = 0, M = 0, N = 0, nsub = 0;   // walkers, rows, columns; nsub: columns of the second log sum
  DevBuf<char>* wf = nullptr;           // FP64 complex walkers [nw][N][M]; the result ends in *wf
  DevBuf<char>* tmp = nullptr;          // spare buffer of the same size (the two may be swapped)
  zd* S = nullptr;                      // [nw][N][N] workspace
  zd* ld = nullptr;                     // [nw] (sum log R_kk, sum_{k<nsub} log R_kk)
  int *info = nullptr, *need = nullptr, *flag = nullptr;  // [nw] each
  int* alive = nullptr;
  zd* logw = nullptr;
  double m0 = 1.0, m1 = 0.0;            // logw += m0 * sum log R_kk + m1 * sum_{k<nsub} log R_kk
  double tol = 1e-13;                   // second pass where |Q^H Q - I| exceeds this
  DevScalars* sc = nullptr;             // counts of shifted and dead walkers
  bool check_all = true;                // pass 2 over every alive walker (else the caller's own check)
};

// W <- Q with W = Q R, R upper triangular with positive diagonal; logw += log det R (by m0, m1).
// Pass 1: CholeskyQR of every alive walker.  A walker whose Gram matrix W^H W breaks down or has a
// Jacobi-scaled pivot below 1e-10 (condition number of W above about 1e5, where CholeskyQR2 starts
// to lose orthogonality) is redone from its W by shifted CholeskyQR: R1 = chol(W^H W + s I) with
// s = 11 (M N + N (N + 1)) u tr(W^H W) (Fukaya, Kannan, Nakatsukasa, Zhang and Yamamoto, SIAM J.
// Sci. Comput. 42 (2020) A477), positive definite for any finite W.  Pass 2: a second CholeskyQR
// of the walkers whose Q^H Q is farther than tol from I.  Pass 3: a third for the shifted walkers
// still farther than tol (shifted CholeskyQR3, accurate up to condition numbers near 1/u).  The
// product of the passes' R is R, whose diagonal is the product of theirs, so log det R is the sum
// of the passes' log sums.  A numerically rank-deficient walker thus gets noise-level R_kk and a
// very negative log-weight, as Householder gives it; one with an exactly zero column after the
// shifted pass (Householder's R_kk = 0, log det R = -inf, weight zero), or with non-finite
// entries, breaks down in a later pass and is marked dead instead.
// With check_all false pass 2 is left to the caller, which runs cholqr_refine over the walkers its
// own check selects.
int cholqr_f64(const CholQrBlock& q, hipStream_t s);
// A further CholeskyQR of the walkers in `among` whose Q^H Q is farther than tol from I (need
// lists them afterwards); the others stay put.
int cholqr_refine(const CholQrBlock& q, const int* among, hipStream_t s);
