// Internal state of the constraint-path propagator and its walker sets, and the batched GEMM /
// dense helpers that the step, the measurement and the adjoint share.
#pragma once
#include <hip/hip_runtime.h>

#include <complex>
#include <type_traits>
#include <cstring>
#include <vector>

#include "../../include/afqmc_mi355x.h"
#include "cp_kernels.hpp"
#include "gemm.hpp"
#include "group.hpp"
#include "runtime.hpp"
#include "small_dense.hpp"

using cd = std::complex<double>;

struct afqmc_cp_s {
  afqmc_ctx ctx = nullptr;
  int device = 0;  // destroy must not touch ctx, which may already be gone
  int L = 0, N = 0, Nup = 0, Ndn = 0, nblk = 1, dense = 1, fp32 = 0;
  int split_bf16 = 0;  // the propagation GEMM in split-BF16 products (precision 2)
  int Mb[2] = {0, 0}, Nb[2] = {0, 0}, row0[2] = {0, 0}, col0[2] = {0, 0};
  int decomp = 1, force_type = 0, diag_only = 1, maxw = 0;
  double dt = 0.01, cap = 1.5, log_half_L = 0.0, dtUSum = 0.0;
  zd trial_logw{0, 0};
  zd hop_logw[3];
  DevBuf<char> phiT[2], expK[3][2], wdkT[2];
  DevBuf<zd> gamma, cdiag;
  DevBuf<double> U, mu, hx, hy, hz, cforce;
  DevBuf<DevScalars> sc;
  DevBuf<double> ubuf;
  // workspace
  DevBuf<char> ovl[2], scale[2], d2x2, qrS[2], qr64[2], qr64b[2];
  DevBuf<char> gjwork[2];  // FP64 working copies of the overlap matrices above 160 columns
  DevBuf<zd> ldet[2], qrld[2], logO, gpart[3][2], per_walker, site_acc, site_part, den;
  DevBuf<int> info[2], qrinfo[2], prefix, table, comb_num, qr_need, qr_flag;
  DevBuf<double> weight, comb_prob, force_rec;
  DevBuf<int> aux_rec, aux_in;
  int ntile[2] = {1, 1};
  int tile[3] = {GEMM_TILE_64x64, GEMM_TILE_64x64, GEMM_TILE_64x64};  // propagate, overlap, green
  std::vector<cd> h_gamma, h_cdiag;
  std::vector<cd> h_K;                                 // 2L x 2L, for energies from Green matrices
  std::vector<double> h_U, h_mu, h_hx, h_hy, h_hz;
  size_t esz() const { return fp32 ? sizeof(zf) : sizeof(zd); }
};

struct afqmc_walkers_s {
  afqmc_cp cp = nullptr;
  int device = 0;  // destroy must not touch cp, which may already be gone
  int nw = 0;
  DevBuf<char> wf[2], tmp[2];
  DevBuf<zd> logw, logw_tmp;
  DevBuf<int> alive;
  long long wsize(int b) const { return (long long)cp->Mb[b] * cp->Nb[b]; }
};

// ------------------------------------------------------------------ helpers shared by the entry points
namespace cpi {

template <typename T>
void pack_block(const double* dense, int L2, int N, int r0, int c0, int Mb, int Nb, std::vector<cplx<T>>& out,
                long long off) {
  for (int j = 0; j < Nb; ++j)
    for (int i = 0; i < Mb; ++i) {
      const double* s = dense + 2 * ((long long)(r0 + i) + (long long)(c0 + j) * L2);
      out[off + i + (long long)j * Mb] = cmk<T>((T)s[0], (T)s[1]);
    }
}

template <typename T>
int upload_matrix_blocks(afqmc_cp cp, const double* dense, int cols, DevBuf<char>* dst, bool square) {
  const int L2 = 2 * cp->L;
  for (int b = 0; b < cp->nblk; ++b) {
    const int Mb = cp->Mb[b], Nb = square ? cp->Mb[b] : cp->Nb[b];
    const int c0 = square ? cp->row0[b] : cp->col0[b];
    std::vector<cplx<T>> h((size_t)Mb * Nb);
    pack_block<T>(dense, L2, cols, cp->row0[b], c0, Mb, Nb, h, 0);
    int rc = dst[b].alloc_bytes(h.size() * sizeof(cplx<T>));
    if (rc) return rc;
    HIP_TRY(hipMemcpy(dst[b].p, h.data(), h.size() * sizeof(cplx<T>), hipMemcpyHostToDevice));
  }
  return 0;
}

template <typename T>
GemmArgs gemm_base() {
  GemmArgs a;
  memset(&a, 0, sizeof(a));
  a.alpha_re = 1.0;
  a.scale_group = 1;
  return a;
}

template <typename T>
int cp_gemm(afqmc_cp cp, int opa, int epi, const GemmArgs& a) {
  return launch_cgemm<T>(opa, epi, a, cp->ctx->stream);
}

// overlap O_b = phiT_b^H W_b for all walkers ("giant" GEMM with shared A)
template <typename T>
int overlap_gemm(afqmc_cp cp, afqmc_walkers w, DevBuf<char>* wf) {
  GroupFork fork;  // the blocks are independent: a grouped call launches them together
  for (int b = 0; b < cp->nblk; ++b) {
    fork.branch();
    GemmArgs a = gemm_base<T>();
    a.M = cp->Nb[b]; a.N = cp->Nb[b] * w->nw; a.K = cp->Mb[b]; a.batch = 1;
    a.A = cp->phiT[b].p; a.lda = cp->Mb[b]; a.strideA = 0;
    a.B = wf[b].p; a.ldb = cp->Mb[b]; a.strideB = 0;
    a.C = cp->ovl[b].p; a.ldc = cp->Nb[b]; a.strideC = 0;
    a.tile = cp->tile[1];
    int rc = cp_gemm<T>(cp, OP_C, EPI_STORE, a);
    if (rc) return rc;
  }
  return 0;
}

template <typename T>
int gj_blocks(afqmc_cp cp, afqmc_walkers w, bool write_inverse, const int* alive) {
  GroupFork fork;  // the blocks are independent: a grouped call launches them together
  for (int b = 0; b < cp->nblk; ++b) {
    fork.branch();
    SmallDenseArgs s;
    memset(&s, 0, sizeof(s));
    s.N = cp->Nb[b]; s.batch = w->nw;
    s.in = cp->ovl[b].p; s.ld_in = cp->Nb[b]; s.stride_in = (long long)cp->Nb[b] * cp->Nb[b];
    s.out = cp->ovl[b].p; s.ld_out = s.ld_in; s.stride_out = s.stride_in;
    s.logdet = cp->ldet[b].p; s.info = cp->info[b].p; s.write_inverse = write_inverse ? 1 : 0;
    s.alive = alive; s.alive_div = 1;
    s.work = cp->gjwork[b].p;
    int rc = launch_gj_inverse<T>(s, cp->ctx->stream);
    if (rc) return rc;
  }
  return 0;
}

// X_b = W_b O_b^{-1} with green row-dot epilogue
template <typename T>
int green_gemm(afqmc_cp cp, afqmc_walkers w, DevBuf<char>* wf, const int* alive, bool kinetic) {
  GroupFork fork;  // the blocks are independent: a grouped call launches them together
  for (int b = 0; b < cp->nblk; ++b) {
    fork.branch();
    GemmArgs a = gemm_base<T>();
    a.M = cp->Mb[b]; a.N = cp->Nb[b]; a.K = cp->Nb[b]; a.batch = w->nw;
    a.A = wf[b].p; a.lda = cp->Mb[b]; a.strideA = w->wsize(b);
    a.B = cp->ovl[b].p; a.ldb = cp->Nb[b]; a.strideB = (long long)cp->Nb[b] * cp->Nb[b];
    a.P0 = cp->phiT[b].p; a.P2 = kinetic ? cp->wdkT[b].p : nullptr; a.ldp = cp->Mb[b];
    a.partner_half = cp->dense ? cp->L : 0;
    a.out0 = cp->gpart[0][b].p; a.out1 = cp->dense ? cp->gpart[1][b].p : nullptr;
    a.out2 = kinetic ? cp->gpart[2][b].p : nullptr;
    a.alive = alive;
    a.tile = cp->tile[2];
    int rc = cp_gemm<T>(cp, OP_N, EPI_GREEN, a);
    if (rc) return rc;
  }
  return 0;
}

// W'_b = B_b diag(s) W_b  (all walkers, one GEMM per block)
template <typename T>
int propagate_gemm(afqmc_cp cp, afqmc_walkers w, int which, DevBuf<char>* in, DevBuf<char>* out, bool scaled) {
  GroupFork fork;  // the blocks are independent: a grouped call launches them together
  for (int b = 0; b < cp->nblk; ++b) {
    fork.branch();
    GemmArgs a = gemm_base<T>();
    a.M = cp->Mb[b]; a.N = cp->Nb[b] * w->nw; a.K = cp->Mb[b]; a.batch = 1;
    a.A = cp->expK[which][b].p; a.lda = cp->Mb[b];
    a.B = in[b].p; a.ldb = cp->Mb[b];
    a.C = out[b].p; a.ldc = cp->Mb[b];
    if (scaled) {
      a.scale = cp->scale[b].p; a.scale_group = cp->Nb[b]; a.scale_strideG = cp->Mb[b]; a.scale_strideB = 0;
    }
    a.tile = cp->tile[0];
    a.split_bf16 = std::is_same<T, float>::value ? cp->split_bf16 : 0;
    int rc = cp_gemm<T>(cp, OP_N, EPI_STORE, a);
    if (rc) return rc;
  }
  return 0;
}

// the walker step with each alive walker's aux fields written to aux_dev ([nw][L], device)
int step_record(afqmc_cp cp, afqmc_walkers w, int* aux_dev);
// cp->den[w] = the mixed-estimator weight of walker w (all walkers; dead ones are masked by the user)
int mixed_weight(afqmc_cp cp, afqmc_walkers w);
// SD::stabilize of every alive walker, enqueued without synchronising
int stabilize(afqmc_cp cp, afqmc_walkers w);

}  // namespace cpi
