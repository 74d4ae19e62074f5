// Robust batched CholeskyQR in FP64 (see cholqr.hpp).
#include "cholqr.hpp"

#include <cstring>
#include <utility>

#include "gemm.hpp"
#include "small_dense.hpp"

namespace {

GemmArgs qr_gemm_base() {
  GemmArgs a;
  memset(&a, 0, sizeof(a));
  a.alpha_re = 1.0;
  a.scale_group = 1;
  return a;
}

// S = W^H W (upper tiles only, Hermitian) for the walkers in mask
int gram(const CholQrBlock& q, const int* mask, hipStream_t s) {
  GemmArgs a = qr_gemm_base();
  a.M = q.N; a.N = q.N; a.K = q.M; a.batch = q.nw;
  a.A = q.wf->p; a.lda = q.M; a.strideA = (long long)q.M * q.N;
  a.B = q.wf->p; a.ldb = q.M; a.strideB = (long long)q.M * q.N;
  a.C = q.S; a.ldc = q.N; a.strideC = (long long)q.N * q.N;
  a.alive = mask;
  a.herm = 1;
  // measured fastest for the Gram shapes: one 128x128 tile per matrix above 64 columns
  a.tile = q.N > 64 ? GEMM_TILE_128x128 : GEMM_TILE_64x64;
  return launch_cgemm<double>(OP_C, EPI_STORE, a, s);
}

// S <- R^{-1} with S = R^H R, ld = log sums, info = breakdown (or a Jacobi-scaled pivot <= pivot_tol),
// for the walkers in mask
int chol(const CholQrBlock& q, const int* mask, double pivot_tol, hipStream_t s) {
  SmallDenseArgs c;
  memset(&c, 0, sizeof(c));
  c.N = q.N; c.batch = q.nw; c.in = q.S; c.ld_in = q.N; c.stride_in = (long long)q.N * q.N;
  c.out = q.S; c.ld_out = q.N; c.stride_out = c.stride_in;
  c.logdet = q.ld; c.info = q.info; c.nsub = q.nsub; c.alive = mask; c.alive_div = 1;
  c.pivot_tol = pivot_tol;
  return launch_chol_inverse<double>(c, s);
}

// dst = src R^{-1} for the walkers in mask (R^{-1} upper triangular)
int times_rinv(const CholQrBlock& q, const void* src, void* dst, const int* mask, hipStream_t s) {
  GemmArgs g = qr_gemm_base();
  g.M = q.M; g.N = q.N; g.K = q.N; g.batch = q.nw;
  g.A = src; g.lda = q.M; g.strideA = (long long)q.M * q.N;
  g.B = q.S; g.ldb = q.N; g.strideB = (long long)q.N * q.N;
  g.C = dst; g.ldc = q.M; g.strideC = (long long)q.M * q.N;
  g.alive = mask;
  g.b_upper = 1;
  g.tile = GEMM_TILE_32x32;  // measured fastest for W R^{-1}
  return launch_cgemm<double>(OP_N, EPI_STORE, g, s);
}

// a further CholeskyQR of the walkers in `among` that are farther than tol from orthonormal:
// they are copied to tmp and multiplied back into wf, the others stay put
int refine(const CholQrBlock& q, const int* among, hipStream_t s) {
  int rc;
  if ((rc = gram(q, among, s))) return rc;
  if ((rc = launch_ortho_check(q.nw, q.N, q.S, among, q.tol, q.need, s))) return rc;
  if ((rc = chol(q, q.need, 0.0, s))) return rc;
  // an exactly zero column left after the shifted pass (Householder's R_kk = 0): the walker dies, and
  // leaves the shifted set so that the pass over it does not count the walker a second time
  if ((rc = launch_qr_kill(q.nw, q.need, q.flag, q.info, q.alive, q.sc, 0, s))) return rc;
  if ((rc = launch_select_alive(q.nw, (long long)q.M * q.N, sizeof(zd), q.need, q.wf->p, q.tmp->p, s))) return rc;
  if ((rc = times_rinv(q, q.tmp->p, q.wf->p, q.need, s))) return rc;
  return launch_logw_add_real(q.nw, q.logw, q.need, q.ld, nullptr, 0, q.m0, q.m1, s);
}

}  // namespace

int cholqr_f64(const CholQrBlock& q, hipStream_t s) {
  int rc;
  // pass 1, with the shifted Gram matrix for the walkers whose Cholesky broke down
  if ((rc = gram(q, q.alive, s))) return rc;
  // a scaled pivot below 1e-10 (condition number above about 1e5) takes the shifted pass
  if ((rc = chol(q, q.alive, 1e-10, s))) return rc;
  if ((rc = launch_qr_flag(q.nw, q.alive, q.info, q.flag, s))) return rc;
  if ((rc = gram(q, q.flag, s))) return rc;
  const double u = 1.1102230246251565e-16;  // 2^-53
  const double c = 11.0 * ((double)q.M * q.N + (double)q.N * (q.N + 1)) * u;
  if ((rc = launch_qr_shift(q.nw, q.N, q.S, q.flag, c, s))) return rc;
  if ((rc = chol(q, q.flag, 0.0, s))) return rc;
  // the shifted Gram matrix is positive definite unless the walker holds non-finite numbers
  if ((rc = launch_qr_kill(q.nw, q.flag, nullptr, q.info, q.alive, q.sc, 1, s))) return rc;
  if ((rc = times_rinv(q, q.wf->p, q.tmp->p, q.alive, s))) return rc;
  std::swap(*q.wf, *q.tmp);
  if ((rc = launch_logw_add_real(q.nw, q.logw, q.alive, q.ld, nullptr, 0, q.m0, q.m1, s))) return rc;
  // pass 2: every alive walker not yet orthonormal to tol; pass 3: shifted ones still not
  if (q.check_all && (rc = refine(q, q.alive, s))) return rc;
  return refine(q, q.flag, s);
}

int cholqr_refine(const CholQrBlock& q, const int* among, hipStream_t s) { return refine(q, among, s); }
