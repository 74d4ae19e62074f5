// C-ABI of the phaseless walker path: RealMaterialMolecule + CholeskyReal + SD2s trial,
// SD2is walker sets, and the batched drop-ins for afqmcPhaselessProjection.cpp,
// afqmcPhaselessMeasure.cpp and their blocks.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

#include "cholqr.hpp"
#include "../../include/afqmc_mi355x.h"
#include "cp_kernels.hpp"
#include "gemm.hpp"
#include "ph_kernels.hpp"
#include "runtime.hpp"
#include "small_dense.hpp"

using cd = std::complex<double>;

namespace {
// the series terms fall as |A|^k / k!, so a term count this large means a runaway operator
constexpr int TAYLOR_MAX_ORDER = 200;

template <typename T>
__global__ void transpose_half_rot(int L, int Ns, int G, const cplx<double>* in, cplx<T>* out) {
  // in: reference layout (Ns, L, G) column-major = A[g](j, i) at j + i*Ns + g*Ns*L
  // out: A'[g] = A[g]^T, (L x Ns) column-major per g
  const long long tot = (long long)L * Ns * G;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long long)gridDim.x * blockDim.x) {
    const long long g = e / ((long long)L * Ns);
    const long long r = e - g * L * Ns;
    const int i = (int)(r % L), j = (int)(r / L);
    out[e] = from_zd<T>(in[j + (long long)i * Ns + g * (long long)Ns * L]);
  }
}
}  // namespace

struct afqmc_ph_s {
  afqmc_ctx ctx = nullptr;
  int device = 0;
  int L = 0, Nup = 0, Ndn = 0, Nmax = 0, Nmin = 0, G = 0, fp32 = 0, force_type = 0, maxw = 0;
  int Ns[2] = {0, 0};
  double dt = 0.01, cap = 1.5;
  zd sqrt_minus_dt{0, 0};
  zd trial_logw{0, 0};
  DevBuf<char> phiT[2], wdtT[2], Lvec, Ahalf[2], onebody[3];
  int onebody_log[3] = {0, 0, 0};
  zd onebody_logw[3];
  int ob_set[3] = {0, 0, 0};
  int tay_dynamic = 1, tay_order = 0, tay_base = 3;
  double tay_acc = 1e-10;
  DevBuf<double> cur_bg;
  std::vector<double> h_bg;
  DevBuf<zd> cforce;
  DevBuf<DevScalars> sc;
  DevBuf<double> ubuf;
  // workspace
  DevBuf<char> ovl[2], X[2], bgbuf, xs, Msamp, Tk[2], Dbuf, qrS, qr64, qr64b;
  DevBuf<zd> ldet[2], logO, den, tkin, ex, per_walker, bg_acc, ex_acc, frec, auxrec, slw, qrld;
  DevBuf<int> info[2], prefix, table, comb_num, active, orders, nactive, qrinfo, qr_need, qr_flag, lead;
  DevBuf<double> weight, comb_prob;
  int dchunk = 1;
  size_t esz() const { return fp32 ? sizeof(zf) : sizeof(zd); }
};

struct afqmc_ph_walkers_s {
  afqmc_ph ph = nullptr;
  int device = 0;
  int nw = 0;
  DevBuf<char> wf, tmp;
  DevBuf<zd> logw, logw_tmp;
  DevBuf<int> alive;
  long long wsize() const { return (long long)ph->L * ph->Nmax; }
};

namespace {

GemmArgs gbase() {
  GemmArgs a;
  memset(&a, 0, sizeof(a));
  a.alpha_re = 1.0;
  a.scale_group = 1;
  return a;
}

// O_s = phiT_s^H W_s (W_s = first Ns columns of each walker), then Gauss-Jordan
template <typename T>
int ph_overlap(afqmc_ph ph, afqmc_ph_walkers w, const DevBuf<char>& wf, bool inverse, const int* alive) {
  hipStream_t s = ph->ctx->stream;
  int rc;
  for (int sp = 0; sp < 2; ++sp) {
    const int Ns = ph->Ns[sp];
    GemmArgs a = gbase();
    a.M = Ns; a.N = Ns; a.K = ph->L; a.batch = w->nw;
    a.A = ph->phiT[sp].p; a.lda = ph->L; a.strideA = 0;
    a.B = wf.p; a.ldb = ph->L; a.strideB = w->wsize();
    a.C = ph->ovl[sp].p; a.ldc = Ns; a.strideC = (long long)Ns * Ns;
    a.alive = alive;
    if ((rc = launch_cgemm<T>(OP_C, EPI_STORE, a, s))) return rc;
    SmallDenseArgs g;
    memset(&g, 0, sizeof(g));
    g.N = Ns; g.batch = w->nw;
    g.in = ph->ovl[sp].p; g.ld_in = Ns; g.stride_in = (long long)Ns * Ns;
    g.out = ph->ovl[sp].p; g.ld_out = Ns; g.stride_out = g.stride_in;
    g.logdet = ph->ldet[sp].p; g.info = ph->info[sp].p; g.write_inverse = inverse ? 1 : 0;
    g.alive = alive; g.alive_div = 1;
    if ((rc = launch_gj_inverse<T>(g, s))) return rc;
  }
  return 0;
}

// X_s = W_s O_s^{-1} (L x Ns per walker): theta_s = trans(theta_s_T)
template <typename T>
int ph_theta(afqmc_ph ph, afqmc_ph_walkers w, const DevBuf<char>& wf, const int* alive) {
  for (int sp = 0; sp < 2; ++sp) {
    const int Ns = ph->Ns[sp];
    GemmArgs a = gbase();
    a.M = ph->L; a.N = Ns; a.K = Ns; a.batch = w->nw;
    a.A = wf.p; a.lda = ph->L; a.strideA = w->wsize();
    a.B = ph->ovl[sp].p; a.ldb = Ns; a.strideB = (long long)Ns * Ns;
    a.C = ph->X[sp].p; a.ldc = ph->L; a.strideC = (long long)ph->L * Ns;
    a.alive = alive;
    int rc = launch_cgemm<T>(OP_N, EPI_STORE, a, ph->ctx->stream);
    if (rc) return rc;
  }
  return 0;
}

// bg[w][g] = sum_s vec(A'_s[g]) . vec(X_s(w))  (calculateCholeskyBg, zgemv 'T')
template <typename T>
int ph_background(afqmc_ph ph, afqmc_ph_walkers w) {
  for (int sp = 0; sp < 2; ++sp) {
    const long long K = (long long)ph->L * ph->Ns[sp];
    GemmArgs a = gbase();
    a.M = ph->G; a.N = w->nw; a.K = (int)K; a.batch = 1;
    a.A = ph->Ahalf[sp].p; a.lda = K;
    a.B = ph->X[sp].p; a.ldb = K;
    a.C = ph->bgbuf.p; a.ldc = ph->G;
    a.beta_one = sp;
    a.tile = GEMM_TILE_32x32;  // fastest shape for G x nw outputs with K = L Ns (both precisions)
    int rc = launch_cgemm<T>(OP_T, EPI_STORE, a, ph->ctx->stream);
    if (rc) return rc;
  }
  return 0;
}

// Msamp[w] = sum_g xs[w][g] L_g  (real L^2 x G times complex G x nw)
template <typename T>
int ph_sample_matrix(afqmc_ph ph, afqmc_ph_walkers w) {
  GemmArgs a = gbase();
  const long long L2 = (long long)ph->L * ph->L;
  a.M = (int)L2; a.N = w->nw; a.K = ph->G; a.batch = 1;
  a.A = ph->Lvec.p; a.lda = L2;
  a.B = ph->xs.p; a.ldb = ph->G;
  a.C = ph->Msamp.p; a.ldc = L2;
  return launch_cgemm<T>(OP_NR, EPI_STORE, a, ph->ctx->stream);
}

// tmp <- sum_k A^k W / k! for the walkers in mask: `base` terms, then (dyn) more until the largest
// element of the last term is at most the accuracy (LogHop2isSD2isOperation.cpp:99-195); A per walker
// (strideA = L^2) or shared (strideA = 0).  maxo: the largest number of terms taken.
template <typename T>
int ph_taylor_pass(afqmc_ph ph, afqmc_ph_walkers w, DevBuf<char>& wf, const void* A, long long strideA,
                   const int* mask, bool dyn, int base, int* maxo) {
  hipStream_t s = ph->ctx->stream;
  const long long n = w->wsize();
  const size_t bytes = ph->esz() * n * w->nw;
  int rc;
  // tmp <- W (the accumulated sum), Tk[0] <- W (the current term)
  HIP_TRY(hipMemcpyAsync(w->tmp.p, wf.p, bytes, hipMemcpyDeviceToDevice, s));
  HIP_TRY(hipMemcpyAsync(ph->Tk[0].p, wf.p, bytes, hipMemcpyDeviceToDevice, s));
  HIP_TRY(hipMemcpyAsync(ph->active.p, mask, sizeof(int) * w->nw, hipMemcpyDeviceToDevice, s));
  int cur = 0;
  *maxo = 0;
  for (int m = 1;; ++m) {
    REQUIRE(m <= TAYLOR_MAX_ORDER, "the Taylor series of exp(A) W has not converged in %d terms (accuracy %g)",
            TAYLOR_MAX_ORDER, ph->tay_acc);
    GemmArgs a = gbase();
    a.M = ph->L; a.N = ph->Nmax; a.K = ph->L; a.batch = w->nw;
    a.A = A; a.lda = ph->L; a.strideA = strideA;
    a.B = ph->Tk[cur].p; a.ldb = ph->L; a.strideB = n;
    a.C = ph->Tk[cur ^ 1].p; a.ldc = ph->L; a.strideC = n;
    a.alpha_re = 1.0 / m;
    a.alive = ph->active.p;
    if ((rc = launch_cgemm<T>(OP_N, EPI_STORE, a, s))) return rc;
    HIP_TRY(hipMemsetAsync(ph->nactive.p, 0, sizeof(int), s));
    if ((rc = launch_taylor_accumulate<T>(w->nw, n, ph->Tk[cur ^ 1].p, w->tmp.p, ph->active.p, ph->orders.p, m, base,
                                          ph->tay_acc, dyn ? 1 : 0, ph->nactive.p, s)))
      return rc;
    cur ^= 1;
    *maxo = m;
    if (m >= base) {
      int h = 0;
      HIP_TRY(hipMemcpyAsync(&h, ph->nactive.p, sizeof(int), hipMemcpyDeviceToHost, s));
      HIP_TRY(hipStreamSynchronize(s));
      if (h == 0) break;
    }
  }
  return 0;
}

// W <- sum_k A^k W / k! (LogHop2isSD2isOperation::applyToRight) for the alive walkers; result written
// back to wf.  dynamicOrder: each walker's own order.  fixedOrder with a given order: that many
// terms.  fixedOrder with order 0: as the reference's serial run (determinantAndAddFixedOrders,
// logHop2isSD2isOperation.cpp:122-133), the first application finds the dynamic order of its first
// walker (here the first alive walker), freezes it, and every walker takes that many terms.
template <typename T>
int ph_taylor(afqmc_ph ph, afqmc_ph_walkers w, DevBuf<char>& wf, const void* A, long long strideA, const int* alive,
              int* max_order_out) {
  hipStream_t s = ph->ctx->stream;
  const long long n = w->wsize();
  int rc, maxo = 0;
  if (!ph->tay_dynamic && ph->tay_order == 0) {
    std::vector<int> al(w->nw), lead(w->nw, 0);
    HIP_TRY(hipMemcpyAsync(al.data(), alive, sizeof(int) * w->nw, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    int first = -1;
    for (int i = 0; i < w->nw && first < 0; ++i)
      if (al[i]) first = i;
    if (first < 0) {  // no walker to propagate: the order stays to be found
      if (max_order_out) *max_order_out = 0;
      return 0;
    }
    lead[first] = 1;
    HIP_TRY(hipMemcpyAsync(ph->lead.p, lead.data(), sizeof(int) * w->nw, hipMemcpyHostToDevice, s));
    if ((rc = ph_taylor_pass<T>(ph, w, wf, A, strideA, ph->lead.p, true, ph->tay_base, &maxo))) return rc;
    HIP_TRY(hipStreamSynchronize(s));  // the host copy of the mask must outlive the copy
    ph->tay_order = maxo;
  }
  const bool dyn = ph->tay_dynamic != 0;
  const int base = dyn ? ph->tay_base : ph->tay_order;
  if ((rc = ph_taylor_pass<T>(ph, w, wf, A, strideA, alive, dyn, base, &maxo))) return rc;
  // alive walkers take the sum; killed ones keep their state (the reference skips them)
  if ((rc = launch_select_alive(w->nw, n, ph->esz(), alive, w->tmp.p, wf.p, s))) return rc;
  if (max_order_out) *max_order_out = maxo;
  return 0;
}

template <typename T>
int ph_onebody(afqmc_ph ph, afqmc_ph_walkers w, int which, DevBuf<char>& wf, const int* alive) {
  REQUIRE(ph->ob_set[which], "one-body operator %d has not been set (afqmc_ph_set_onebody)", which);
  int rc;
  if (ph->onebody_log[which]) {
    if ((rc = ph_taylor<T>(ph, w, wf, ph->onebody[which].p, 0, alive, nullptr))) return rc;
  } else {
    GemmArgs a = gbase();
    a.M = ph->L; a.N = ph->Nmax * w->nw; a.K = ph->L; a.batch = 1;
    a.A = ph->onebody[which].p; a.lda = ph->L;
    a.B = wf.p; a.ldb = ph->L;
    a.C = w->tmp.p; a.ldc = ph->L;
    if ((rc = launch_cgemm<T>(OP_N, EPI_STORE, a, ph->ctx->stream))) return rc;
    if ((rc = launch_select_alive(w->nw, w->wsize(), ph->esz(), alive, w->tmp.p, wf.p, ph->ctx->stream))) return rc;
  }
  return 0;
}

LogOArgs ph_logo_args(afqmc_ph ph, afqmc_ph_walkers w, int mode, const zd* logw) {
  LogOArgs l;
  memset(&l, 0, sizeof(l));
  l.nw = w->nw; l.nblk = 2; l.combine = COMBINE_SUM; l.mode = mode;
  for (int b = 0; b < 2; ++b) { l.ldet[b] = ph->ldet[b].p; l.info[b] = ph->info[b].p; }
  l.logw = logw; l.logw_out = w->logw.p; l.alive = w->alive.p; l.trial_logw = ph->trial_logw;
  l.logO = ph->logO.p; l.weight = ph->weight.p; l.den = ph->den.p;
  return l;
}

template <typename T>
int ph_step_impl(afqmc_ph ph, afqmc_ph_walkers w, double* force_out, double* aux_out) {
  hipStream_t s = ph->ctx->stream;
  int rc;
  {
    RegionTimer t(ph->ctx, REG_OVERLAP);
    if ((rc = ph_overlap<T>(ph, w, w->wf, true, w->alive.p))) return rc;
  }
  LogOArgs l = ph_logo_args(ph, w, LOGO_STEP, w->logw.p);
  if ((rc = launch_logo(l, s))) return rc;
  if ((rc = launch_alive_scan(w->nw, w->alive.p, ph->prefix.p, ph->sc.p, s))) return rc;
  {
    RegionTimer t(ph->ctx, REG_PH_FORCE);
    if ((rc = ph_theta<T>(ph, w, w->wf, w->alive.p))) return rc;
    if ((rc = ph_background<T>(ph, w))) return rc;
  }
  PhSampleArgs a;
  memset(&a, 0, sizeof(a));
  a.nw = w->nw; a.G = ph->G; a.force_type = ph->force_type; a.fp32 = ph->fp32;
  a.bg = ph->bgbuf.p; a.cur_bg = ph->cur_bg.p; a.const_force = ph->cforce.p;
  a.sqrt_minus_dt = ph->sqrt_minus_dt; a.cap = ph->cap; a.dt = ph->dt; a.onebody_logw = ph->onebody_logw[0];
  a.alive = w->alive.p; a.prefix = ph->prefix.p; a.sc = ph->sc.p; a.ubuf = ph->ubuf.p; a.logw = w->logw.p;
  a.xs = ph->xs.p;
  if (force_out || aux_out) { a.force_out = ph->frec.p; a.aux_out = ph->auxrec.p; }
  {
    RegionTimer t(ph->ctx, REG_PH_SAMPLE);
    if ((rc = launch_ph_sample(a, s))) return rc;
    if ((rc = launch_advance_counter(ph->sc.p, 2 * ph->G, s))) return rc;
    if ((rc = ph_sample_matrix<T>(ph, w))) return rc;
  }
  {
    RegionTimer t(ph->ctx, REG_PH_TAYLOR);
    if ((rc = ph_taylor<T>(ph, w, w->wf, ph->Msamp.p, (long long)ph->L * ph->L, w->alive.p, nullptr))) return rc;
  }
  {
    RegionTimer t(ph->ctx, REG_PROPAGATE);
    if ((rc = ph_onebody<T>(ph, w, 0, w->wf, w->alive.p))) return rc;
  }
  if (force_out) HIP_TRY(hipMemcpyAsync(force_out, ph->frec.p, sizeof(zd) * w->nw * ph->G, hipMemcpyDeviceToHost, s));
  if (aux_out) HIP_TRY(hipMemcpyAsync(aux_out, ph->auxrec.p, sizeof(zd) * w->nw * ph->G, hipMemcpyDeviceToHost, s));
  if (force_out || aux_out) HIP_TRY(hipStreamSynchronize(s));
  return 0;
}

// SD2is::stabilize (SD2is.cpp:48-61) by robust CholeskyQR in FP64 (cholqr.hpp) on the L x Nmax matrix:
// logw += sum log R_kk + sum_{k<Nmin} log R_kk; tol: second pass where |Q^H Q - I| exceeds it
int ph_qr_f64(afqmc_ph ph, afqmc_ph_walkers w, DevBuf<char>& wf, DevBuf<char>& tmp, double tol) {
  CholQrBlock q;
  q.nw = w->nw; q.M = ph->L; q.N = ph->Nmax; q.nsub = ph->Nmin;
  q.wf = &wf; q.tmp = &tmp;
  q.S = reinterpret_cast<zd*>(ph->qrS.p); q.ld = ph->qrld.p;
  q.info = ph->qrinfo.p; q.need = ph->qr_need.p; q.flag = ph->qr_flag.p;
  q.alive = w->alive.p; q.logw = w->logw.p;
  q.m0 = 1.0; q.m1 = 1.0; q.tol = tol; q.sc = ph->sc.p;
  return cholqr_f64(q, ph->ctx->stream);
}

int ph_stabilize_impl(afqmc_ph ph, afqmc_ph_walkers w) {
  RegionTimer t(ph->ctx, REG_QR);
  if (!ph->fp32) return ph_qr_f64(ph, w, w->wf, w->tmp, 1e-13);
  hipStream_t s = ph->ctx->stream;
  const long long n = w->wsize() * w->nw;
  int rc;
  if ((rc = launch_convert_f32_f64(w->wf.p, ph->qr64.p, n, s))) return rc;
  // the result is rounded to FP32, so a first-pass Q orthonormal to 1e-8 is final
  if ((rc = ph_qr_f64(ph, w, ph->qr64, ph->qr64b, 1e-8))) return rc;
  return launch_convert_f64_f32(ph->qr64.p, w->wf.p, n, s);
}

template <typename T>
int ph_pop_control_impl(afqmc_ph ph, afqmc_ph_walkers w, double* weights_out, int* table_out, double* stats) {
  hipStream_t s = ph->ctx->stream;
  RegionTimer t(ph->ctx, REG_COMB);
  int rc = ph_overlap<T>(ph, w, w->wf, false, w->alive.p);
  if (rc) return rc;
  LogOArgs l = ph_logo_args(ph, w, LOGO_WEIGHT, w->logw.p);
  if ((rc = launch_logo(l, s))) return rc;
  if ((rc = launch_comb(w->nw, ph->weight.p, ph->ubuf.p, ph->sc.p, ph->table.p, ph->comb_prob.p, ph->comb_num.p, s)))
    return rc;
  void* wfp[2] = {w->wf.p, nullptr};
  long long ws[2] = {w->wsize(), 0};
  if ((rc = launch_apply_table<T>(w->nw, 1, wfp, ws, w->logw.p, w->alive.p, ph->table.p, s))) return rc;
  if (weights_out) HIP_TRY(hipMemcpyAsync(weights_out, ph->weight.p, sizeof(double) * w->nw, hipMemcpyDeviceToHost, s));
  if (table_out) HIP_TRY(hipMemcpyAsync(table_out, ph->table.p, sizeof(int) * w->nw, hipMemcpyDeviceToHost, s));
  if (stats) {
    DevScalars h;
    HIP_TRY(hipMemcpyAsync(&h, ph->sc.p, sizeof(h), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    stats[0] = h.wmin; stats[1] = h.wmax; stats[2] = h.wavg;
  }
  if (weights_out || table_out) HIP_TRY(hipStreamSynchronize(s));
  return 0;
}

// addMixedMeasurement (afqmcPhaselessMeasure.cpp:9-26) with the energy of
// RealMaterialMoleculeMeasureFixedSD2sSD2is::addEnergy (:197-284)
template <typename T>
int ph_measure_impl(afqmc_ph ph, afqmc_ph_walkers w) {
  hipStream_t s = ph->ctx->stream;
  RegionTimer t(ph->ctx, REG_MEASURE);
  int rc;
  // temp = exp(+dt/2 K) W in the walkers' spare buffer (Taylor mode works in place on a copy)
  DevBuf<char>& tmpw = ph->qr64b;  // reused as the measured walkers (size >= walker set)
  const size_t bytes = ph->esz() * w->wsize() * w->nw;
  HIP_TRY(hipMemcpyAsync(tmpw.p, w->wf.p, bytes, hipMemcpyDeviceToDevice, s));
  if ((rc = ph_onebody<T>(ph, w, 2, tmpw, w->alive.p))) return rc;
  if ((rc = launch_logw_shift(w->nw, w->logw.p, w->logw_tmp.p, ph->onebody_logw[2], s))) return rc;
  if ((rc = ph_overlap<T>(ph, w, tmpw, true, w->alive.p))) return rc;
  LogOArgs l = ph_logo_args(ph, w, LOGO_MEASURE, w->logw_tmp.p);
  if ((rc = launch_logo(l, s))) return rc;
  if ((rc = ph_theta<T>(ph, w, tmpw, w->alive.p))) return rc;
  if ((rc = ph_background<T>(ph, w))) return rc;
  // kinetic: sum_s sum_ij X_s(i,j) (phiT_s^H t)^T(i,j)
  for (int sp = 0; sp < 2; ++sp) {
    const long long n = (long long)ph->L * ph->Ns[sp];
    if ((rc = (ph->fp32 ? launch_rowdot_sum<float>(w->nw, n, ph->X[sp].p, n, ph->wdtT[sp].p, ph->tkin.p, sp, s)
                        : launch_rowdot_sum<double>(w->nw, n, ph->X[sp].p, n, ph->wdtT[sp].p, ph->tkin.p, sp, s))))
      return rc;
  }
  // exchange: D_s = A_s[g] theta_s stacked over g, (G Ns) x Ns per walker, for a chunk of walkers
  // at once: one GEMM with the shared half-rotated vectors as A and the chunk's theta matrices
  // side by side as B (N = Ns * chunk), A panels reused across the chunk (N tiles fastest)
  for (int sp = 0; sp < 2; ++sp) {
    const int Ns = ph->Ns[sp];
    for (int w0 = 0; w0 < w->nw; w0 += ph->dchunk) {
      const int nb = std::min(ph->dchunk, w->nw - w0);
      GemmArgs a = gbase();
      a.M = ph->G * Ns; a.N = Ns * nb; a.K = ph->L; a.batch = 1;
      a.A = ph->Ahalf[sp].p; a.lda = ph->L;
      a.B = (char*)ph->X[sp].p + ph->esz() * (size_t)w0 * ph->L * Ns; a.ldb = ph->L;
      a.C = ph->Dbuf.p; a.ldc = (long long)ph->G * Ns;
      a.n_fastest = 1;
      // fastest shapes for (G Ns) x (Ns chunk) outputs with K = L: 128x128 in FP64, 64x64 in FP32
      a.tile = std::is_same<T, double>::value ? GEMM_TILE_128x128 : GEMM_TILE_64x64;
      if ((rc = launch_cgemm<T>(OP_T, EPI_STORE, a, s))) return rc;
      if ((rc = launch_exchange_trace<T>(nb, ph->G, Ns, ph->Dbuf.p, ph->ex.p + (size_t)w0 * ph->G, sp, s))) return rc;
    }
  }
  PhEnergyArgs e;
  memset(&e, 0, sizeof(e));
  e.nw = w->nw; e.G = ph->G; e.fp32 = ph->fp32; e.alive = w->alive.p; e.den = ph->den.p; e.tkin = ph->tkin.p;
  e.bg = ph->bgbuf.p; e.ex = ph->ex.p; e.per_walker = ph->per_walker.p; e.bg_acc = ph->bg_acc.p;
  e.ex_acc = ph->ex_acc.p; e.sc = ph->sc.p;
  return launch_ph_energy(e, s);
}

template <typename T>
int upload_cplx(DevBuf<char>& dst, const double* src, size_t n) {
  std::vector<cplx<T>> h(n);
  for (size_t i = 0; i < n; ++i) h[i] = cmk<T>((T)src[2 * i], (T)src[2 * i + 1]);
  int rc = dst.alloc_bytes(sizeof(cplx<T>) * n);
  if (rc) return rc;
  HIP_TRY(hipMemcpy(dst.p, h.data(), sizeof(cplx<T>) * n, hipMemcpyHostToDevice));
  return 0;
}

}  // namespace

#define PH_DISPATCH(ph, fn, ...) ((ph)->fp32 ? fn<float>(__VA_ARGS__) : fn<double>(__VA_ARGS__))

extern "C" {

int afqmc_ph_create(afqmc_ctx ctx, const afqmc_ph_desc* d, afqmc_ph* out) {
  REQUIRE(ctx && d && out, "afqmc_ph_create: null argument");
  REQUIRE(d->L > 0 && d->Nup > 0 && d->Ndn > 0 && d->G > 0 && d->max_walkers > 0, "afqmc_ph_create: sizes must be positive");
  REQUIRE(d->t && d->chol && d->phiT_up && d->phiT_dn && d->half_rot_up && d->half_rot_dn,
          "afqmc_ph_create: t, chol, trial and half-rotated vectors are required");
  REQUIRE(std::max(d->Nup, d->Ndn) <= 160, "electron count per spin above 160 is not supported");
  REQUIRE(d->force_type == 0 || d->force_type == 1, "Do not know method.forceType %d", d->force_type);
  REQUIRE(d->precision == 0 || d->precision == 1, "afqmc_ph_create: precision must be 0 or 1");
  ScopedDevice sdv(ctx);
  afqmc_ph ph = new afqmc_ph_s();
  ph->ctx = ctx; ph->device = ctx->device;
  ph->L = d->L; ph->Nup = d->Nup; ph->Ndn = d->Ndn; ph->G = d->G;
  ph->Nmax = std::max(d->Nup, d->Ndn); ph->Nmin = std::min(d->Nup, d->Ndn);
  ph->Ns[0] = d->Nup; ph->Ns[1] = d->Ndn;
  ph->fp32 = d->precision ? 1 : 0; ph->force_type = d->force_type;
  ph->dt = d->dt; ph->cap = d->force_cap; ph->maxw = d->max_walkers;
  const cd sq = std::sqrt(cd(-d->dt, 0.0));
  ph->sqrt_minus_dt = cmk<double>(sq.real(), sq.imag());
  ph->trial_logw = cmk<double>(d->phiT_logw[0], d->phiT_logw[1]);
  ph->tay_dynamic = d->taylor_fixed ? 0 : 1;
  ph->tay_order = d->taylor_order;
  ph->tay_base = d->taylor_base > 0 ? d->taylor_base : 3;
  ph->tay_acc = d->taylor_accuracy > 0 ? d->taylor_accuracy : 1e-10;
  for (int k = 0; k < 3; ++k) ph->onebody_logw[k] = cmk<double>(0, 0);
  const int L = d->L, G = d->G;
  const long long L2 = (long long)L * L;
  int rc = 0;
  const size_t es = ph->esz();
  // trial, (phiT_s^H t)^T = t^T conj(phiT_s), Cholesky vectors, half-rotated vectors
  const double* phis[2] = {d->phiT_up, d->phiT_dn};
  for (int sp = 0; sp < 2 && !rc; ++sp) {
    const int Ns = ph->Ns[sp];
    rc = ph->fp32 ? upload_cplx<float>(ph->phiT[sp], phis[sp], (size_t)L * Ns)
                  : upload_cplx<double>(ph->phiT[sp], phis[sp], (size_t)L * Ns);
    std::vector<double> w2(2 * (size_t)L * Ns, 0.0);
    const cd* phi = reinterpret_cast<const cd*>(phis[sp]);
    for (int j = 0; j < Ns; ++j)
      for (int i = 0; i < L; ++i) {
        cd s = 0.0;
        for (int a = 0; a < L; ++a) s += d->t[a + (size_t)i * L] * std::conj(phi[a + (size_t)j * L]);
        w2[2 * (i + (size_t)j * L)] = s.real();
        w2[2 * (i + (size_t)j * L) + 1] = s.imag();
      }
    if (!rc) rc = ph->fp32 ? upload_cplx<float>(ph->wdtT[sp], w2.data(), (size_t)L * Ns)
                           : upload_cplx<double>(ph->wdtT[sp], w2.data(), (size_t)L * Ns);
  }
  if (!rc) {
    rc = ph->Lvec.alloc_bytes((ph->fp32 ? sizeof(float) : sizeof(double)) * L2 * G);
    if (!rc) {
      if (ph->fp32) {
        std::vector<float> h((size_t)(L2 * G));
        for (size_t i = 0; i < h.size(); ++i) h[i] = (float)d->chol[i];
        if (hipMemcpy(ph->Lvec.p, h.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice) != hipSuccess) rc = -2;
      } else if (hipMemcpy(ph->Lvec.p, d->chol, sizeof(double) * L2 * G, hipMemcpyHostToDevice) != hipSuccess) {
        rc = -2;
      }
    }
  }
  const double* hr[2] = {d->half_rot_up, d->half_rot_dn};
  for (int sp = 0; sp < 2 && !rc; ++sp) {
    const long long n = (long long)L * ph->Ns[sp] * G;
    DevBuf<zd> staging;
    rc = staging.alloc(n);
    if (!rc && hipMemcpy(staging.p, hr[sp], sizeof(zd) * n, hipMemcpyHostToDevice) != hipSuccess) rc = -2;
    if (!rc) rc = ph->Ahalf[sp].alloc_bytes(es * n);
    if (!rc) {
      unsigned grid = (unsigned)std::min<long long>((n + 255) / 256, 65536);
      if (ph->fp32) hipLaunchKernelGGL(transpose_half_rot<float>, dim3(grid), dim3(256), 0, 0, L, ph->Ns[sp], G, staging.p, (zf*)ph->Ahalf[sp].p);
      else hipLaunchKernelGGL(transpose_half_rot<double>, dim3(grid), dim3(256), 0, 0, L, ph->Ns[sp], G, staging.p, (zd*)ph->Ahalf[sp].p);
      if (hipDeviceSynchronize() != hipSuccess) rc = -2;
    }
  }
  ph->h_bg.assign(G, 0.0);
  if (d->bg) std::copy(d->bg, d->bg + G, ph->h_bg.begin());
  if (!rc) rc = ph->cur_bg.alloc(G);
  if (!rc && hipMemcpy(ph->cur_bg.p, ph->h_bg.data(), sizeof(double) * G, hipMemcpyHostToDevice) != hipSuccess) rc = -2;
  if (!rc) rc = ph->cforce.alloc(G);
  if (!rc) {
    std::vector<zd> cf(G, cmk<double>(0, 0));
    if (d->const_force)
      for (int g = 0; g < G; ++g) cf[g] = cmk<double>(d->const_force[2 * g], d->const_force[2 * g + 1]);
    if (hipMemcpy(ph->cforce.p, cf.data(), sizeof(zd) * G, hipMemcpyHostToDevice) != hipSuccess) rc = -2;
  }
  // workspace
  const int nw = d->max_walkers;
  const long long wsz = (long long)L * ph->Nmax;
  for (int sp = 0; sp < 2 && !rc; ++sp) {
    const long long Ns = ph->Ns[sp];
    rc = ph->ovl[sp].alloc_bytes(es * Ns * Ns * nw);
    if (!rc) rc = ph->X[sp].alloc_bytes(es * L * Ns * nw);
    if (!rc) rc = ph->ldet[sp].alloc(nw);
    if (!rc) rc = ph->info[sp].alloc(nw);
  }
  if (!rc) rc = ph->bgbuf.alloc_bytes(es * (size_t)G * nw);
  if (!rc) rc = ph->xs.alloc_bytes(es * (size_t)G * nw);
  if (!rc) rc = ph->Msamp.alloc_bytes(es * (size_t)L2 * nw);
  for (int k = 0; k < 2 && !rc; ++k) rc = ph->Tk[k].alloc_bytes(es * wsz * nw);
  // exchange chunk: keep the (G Ns x Ns) per-walker blocks under 4 GiB
  const long long dper = (long long)G * ph->Nmax * ph->Nmax * (long long)es;
  ph->dchunk = (int)std::max<long long>(1, std::min<long long>(nw, (4ll << 30) / std::max<long long>(dper, 1)));
  if (!rc) rc = ph->Dbuf.alloc_bytes((size_t)dper * ph->dchunk);
  if (!rc) rc = ph->qrS.alloc_bytes(sizeof(zd) * (size_t)ph->Nmax * ph->Nmax * nw);
  if (!rc) rc = ph->qrld.alloc(nw);
  if (!rc) rc = ph->qrinfo.alloc(nw);
  if (!rc) rc = ph->qr_need.alloc(nw);
  if (!rc) rc = ph->qr_flag.alloc(nw);
  if (!rc) rc = ph->lead.alloc(nw);
  if (!rc) rc = ph->qr64.alloc_bytes(sizeof(zd) * wsz * nw);
  if (!rc) rc = ph->qr64b.alloc_bytes(sizeof(zd) * wsz * nw);
  if (!rc) rc = ph->logO.alloc(nw);
  if (!rc) rc = ph->den.alloc(nw);
  if (!rc) rc = ph->tkin.alloc(nw);
  if (!rc) rc = ph->ex.alloc((size_t)G * nw);
  if (!rc) rc = ph->per_walker.alloc((size_t)nw * 4);
  if (!rc) rc = ph->bg_acc.alloc(G);
  if (!rc) rc = ph->ex_acc.alloc(G);
  if (!rc) rc = ph->frec.alloc((size_t)G * nw);
  if (!rc) rc = ph->auxrec.alloc((size_t)G * nw);
  if (!rc) rc = ph->slw.alloc(nw);
  if (!rc) rc = ph->prefix.alloc(nw);
  if (!rc) rc = ph->table.alloc(nw);
  if (!rc) rc = ph->comb_num.alloc(nw);
  if (!rc) rc = ph->comb_prob.alloc(nw);
  if (!rc) rc = ph->weight.alloc(nw);
  if (!rc) rc = ph->active.alloc(nw);
  if (!rc) rc = ph->orders.alloc(nw);
  if (!rc) rc = ph->nactive.alloc(1);
  if (!rc) rc = ph->sc.alloc(1);
  if (!rc) {
    DevScalars h;
    memset(&h, 0, sizeof(h));
    h.ET = -10.0;
    h.key = afqmc_seed_mix_host(985456376ull);
    if (hipMemcpy(ph->sc.p, &h, sizeof(h), hipMemcpyHostToDevice) != hipSuccess) rc = -2;
    if (!rc && hipMemset(ph->bg_acc.p, 0, sizeof(zd) * G) != hipSuccess) rc = -2;
    if (!rc && hipMemset(ph->ex_acc.p, 0, sizeof(zd) * G) != hipSuccess) rc = -2;
  }
  if (rc) {
    delete ph;
    if (rc == -2) afqmc_set_error("afqmc_ph_create: device allocation or copy failed");
    return rc;
  }
  *out = ph;
  return 0;
}

int afqmc_ph_destroy(afqmc_ph ph) {
  if (!ph) return 0;
  ScopedDevice sdv(ph->device);
  hipDeviceSynchronize();
  delete ph;
  return 0;
}

int afqmc_ph_set_onebody(afqmc_ph ph, int which, const double* matrix, const double* logw, int is_log) {
  REQUIRE(ph && matrix && logw, "null argument");
  REQUIRE(which >= 0 && which <= 2, "which must be 0, 1 or 2");
  ScopedDevice sdv(ph->ctx);
  HIP_TRY(hipStreamSynchronize(ph->ctx->stream));
  const size_t n = (size_t)ph->L * ph->L;
  int rc = ph->fp32 ? upload_cplx<float>(ph->onebody[which], matrix, n) : upload_cplx<double>(ph->onebody[which], matrix, n);
  if (rc) return rc;
  ph->onebody_log[which] = is_log ? 1 : 0;
  ph->onebody_logw[which] = cmk<double>(logw[0], logw[1]);
  ph->ob_set[which] = 1;
  return 0;
}

int afqmc_ph_set_background(afqmc_ph ph, const double* bg) {
  REQUIRE(ph && bg, "null argument");
  ScopedDevice sdv(ph->ctx);
  HIP_TRY(hipStreamSynchronize(ph->ctx->stream));
  std::copy(bg, bg + ph->G, ph->h_bg.begin());
  HIP_TRY(hipMemcpy(ph->cur_bg.p, bg, sizeof(double) * ph->G, hipMemcpyHostToDevice));
  return 0;
}

int afqmc_ph_walkers_create(afqmc_ph ph, int nw, afqmc_ph_walkers* out) {
  REQUIRE(ph && out, "null argument");
  REQUIRE(nw > 0 && nw <= ph->maxw, "walker count %d outside 1..max_walkers=%d", nw, ph->maxw);
  ScopedDevice sdv(ph->ctx);
  afqmc_ph_walkers w = new afqmc_ph_walkers_s();
  w->ph = ph; w->device = ph->device; w->nw = nw;
  const size_t bytes = ph->esz() * w->wsize() * nw;
  int rc = w->wf.alloc_bytes(bytes);
  if (!rc) rc = w->tmp.alloc_bytes(bytes);
  if (!rc && hipMemset(w->wf.p, 0, bytes) != hipSuccess) rc = -2;
  if (!rc) rc = w->logw.alloc(nw);
  if (!rc) rc = w->logw_tmp.alloc(nw);
  if (!rc) rc = w->alive.alloc(nw);
  if (!rc) {
    std::vector<int> one(nw, 1);
    std::vector<zd> z(nw, cmk<double>(0, 0));
    if (hipMemcpy(w->alive.p, one.data(), sizeof(int) * nw, hipMemcpyHostToDevice) != hipSuccess) rc = -2;
    if (!rc && hipMemcpy(w->logw.p, z.data(), sizeof(zd) * nw, hipMemcpyHostToDevice) != hipSuccess) rc = -2;
  }
  if (rc) {
    delete w;
    if (rc == -2) afqmc_set_error("afqmc_ph_walkers_create: device allocation failed");
    return rc;
  }
  *out = w;
  return 0;
}

int afqmc_ph_walkers_destroy(afqmc_ph_walkers w) {
  if (!w) return 0;
  ScopedDevice sdv(w->device);
  hipDeviceSynchronize();
  delete w;
  return 0;
}

int afqmc_ph_walkers_upload(afqmc_ph_walkers w, const double* wf, const double* logw, const int* alive) {
  REQUIRE(w && wf && logw, "null argument");
  afqmc_ph ph = w->ph;
  ScopedDevice sdv(ph->ctx);
  HIP_TRY(hipStreamSynchronize(ph->ctx->stream));
  const size_t n = (size_t)w->wsize() * w->nw;
  if (ph->fp32) {
    std::vector<zf> h(n);
    for (size_t i = 0; i < n; ++i) h[i] = cmk<float>((float)wf[2 * i], (float)wf[2 * i + 1]);
    HIP_TRY(hipMemcpy(w->wf.p, h.data(), sizeof(zf) * n, hipMemcpyHostToDevice));
  } else {
    HIP_TRY(hipMemcpy(w->wf.p, wf, sizeof(zd) * n, hipMemcpyHostToDevice));
  }
  HIP_TRY(hipMemcpy(w->logw.p, logw, sizeof(zd) * w->nw, hipMemcpyHostToDevice));
  std::vector<int> a(w->nw, 1);
  if (alive) std::copy(alive, alive + w->nw, a.begin());
  HIP_TRY(hipMemcpy(w->alive.p, a.data(), sizeof(int) * w->nw, hipMemcpyHostToDevice));
  return 0;
}

int afqmc_ph_walkers_download(afqmc_ph_walkers w, double* wf, double* logw, int* alive) {
  REQUIRE(w, "null argument");
  afqmc_ph ph = w->ph;
  ScopedDevice sdv(ph->ctx);
  HIP_TRY(hipStreamSynchronize(ph->ctx->stream));
  const size_t n = (size_t)w->wsize() * w->nw;
  if (wf) {
    if (ph->fp32) {
      std::vector<zf> h(n);
      HIP_TRY(hipMemcpy(h.data(), w->wf.p, sizeof(zf) * n, hipMemcpyDeviceToHost));
      for (size_t i = 0; i < n; ++i) { wf[2 * i] = h[i].re; wf[2 * i + 1] = h[i].im; }
    } else {
      HIP_TRY(hipMemcpy(wf, w->wf.p, sizeof(zd) * n, hipMemcpyDeviceToHost));
    }
  }
  if (logw) HIP_TRY(hipMemcpy(logw, w->logw.p, sizeof(zd) * w->nw, hipMemcpyDeviceToHost));
  if (alive) HIP_TRY(hipMemcpy(alive, w->alive.p, sizeof(int) * w->nw, hipMemcpyDeviceToHost));
  return 0;
}

static int ph_scalars(afqmc_ph ph, DevScalars* h, bool write) {
  HIP_TRY(hipStreamSynchronize(ph->ctx->stream));
  if (write) HIP_TRY(hipMemcpy(ph->sc.p, h, sizeof(*h), hipMemcpyHostToDevice));
  else HIP_TRY(hipMemcpy(h, ph->sc.p, sizeof(*h), hipMemcpyDeviceToHost));
  return 0;
}

int afqmc_ph_set_seed(afqmc_ph ph, uint64_t seed, uint64_t counter) {
  REQUIRE(ph, "null argument");
  ScopedDevice sdv(ph->ctx);
  DevScalars h;
  int rc = ph_scalars(ph, &h, false);
  if (rc) return rc;
  h.key = afqmc_seed_mix_host(seed);
  h.counter = counter;
  return ph_scalars(ph, &h, true);
}

int afqmc_ph_get_counter(afqmc_ph ph, uint64_t* counter, uint64_t* walker_steps) {
  REQUIRE(ph && counter, "null argument");
  ScopedDevice sdv(ph->ctx);
  DevScalars h;
  int rc = ph_scalars(ph, &h, false);
  if (rc) return rc;
  *counter = h.counter;
  if (walker_steps) *walker_steps = h.walker_steps;
  REQUIRE(!h.uniform_overrun, "the host-provided uniforms were exhausted (%llu given)", h.ulen);
  return 0;
}

int afqmc_ph_get_qr_stats(afqmc_ph ph, uint64_t* shifted, uint64_t* killed) {
  REQUIRE(ph && shifted && killed, "null argument");
  ScopedDevice sdv(ph->ctx);
  DevScalars h;
  int rc = ph_scalars(ph, &h, false);
  if (rc) return rc;
  *shifted = h.qr_shifted;
  *killed = h.qr_killed;
  return 0;
}

int afqmc_ph_set_uniforms(afqmc_ph ph, const double* u, int64_t n) {
  REQUIRE(ph && n >= 0, "invalid argument");
  ScopedDevice sdv(ph->ctx);
  HIP_TRY(hipStreamSynchronize(ph->ctx->stream));
  ph->ubuf.release();
  if (n > 0) {
    REQUIRE(u, "null uniforms");
    int rc = ph->ubuf.alloc((size_t)n);
    if (rc) return rc;
    HIP_TRY(hipMemcpy(ph->ubuf.p, u, sizeof(double) * n, hipMemcpyHostToDevice));
  }
  DevScalars h;
  int rc = ph_scalars(ph, &h, false);
  if (rc) return rc;
  h.counter = 0;
  h.ulen = (unsigned long long)n;
  h.uniform_overrun = 0;
  return ph_scalars(ph, &h, true);
}

int afqmc_ph_skip_uniforms(afqmc_ph ph, uint64_t n) {
  REQUIRE(ph, "null argument");
  ScopedDevice sdv(ph->ctx);
  DevScalars h;
  int rc = ph_scalars(ph, &h, false);
  if (rc) return rc;
  h.counter += n;
  return ph_scalars(ph, &h, true);
}

int afqmc_ph_set_ET(afqmc_ph ph, double ET) {
  REQUIRE(ph, "null argument");
  ScopedDevice sdv(ph->ctx);
  DevScalars h;
  int rc = ph_scalars(ph, &h, false);
  if (rc) return rc;
  h.ET = ET;
  return ph_scalars(ph, &h, true);
}

int afqmc_ph_get_ET(afqmc_ph ph, double* ET) {
  REQUIRE(ph && ET, "null argument");
  ScopedDevice sdv(ph->ctx);
  DevScalars h;
  int rc = ph_scalars(ph, &h, false);
  if (rc) return rc;
  *ET = h.ET;
  return 0;
}

int afqmc_ph_log_overlap(afqmc_ph ph, afqmc_ph_walkers w, double* logO) {
  REQUIRE(ph && w && logO && w->ph == ph, "invalid argument");
  ScopedDevice sdv(ph->ctx);
  int rc = ph->fp32 ? ph_overlap<float>(ph, w, w->wf, false, w->alive.p) : ph_overlap<double>(ph, w, w->wf, false, w->alive.p);
  if (rc) return rc;
  LogOArgs l = ph_logo_args(ph, w, LOGO_ONLY, w->logw.p);
  if ((rc = launch_logo(l, ph->ctx->stream))) return rc;
  HIP_TRY(hipMemcpyAsync(logO, ph->logO.p, sizeof(zd) * w->nw, hipMemcpyDeviceToHost, ph->ctx->stream));
  HIP_TRY(hipStreamSynchronize(ph->ctx->stream));
  return 0;
}

int afqmc_ph_background(afqmc_ph ph, afqmc_ph_walkers w, double* bg) {
  REQUIRE(ph && w && bg && w->ph == ph, "invalid argument");
  ScopedDevice sdv(ph->ctx);
  int rc = ph->fp32 ? ph_overlap<float>(ph, w, w->wf, true, w->alive.p) : ph_overlap<double>(ph, w, w->wf, true, w->alive.p);
  if (!rc) rc = ph->fp32 ? ph_theta<float>(ph, w, w->wf, w->alive.p) : ph_theta<double>(ph, w, w->wf, w->alive.p);
  if (!rc) rc = ph->fp32 ? ph_background<float>(ph, w) : ph_background<double>(ph, w);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(ph->ctx->stream));
  const size_t n = (size_t)ph->G * w->nw;
  if (ph->fp32) {
    std::vector<zf> h(n);
    HIP_TRY(hipMemcpy(h.data(), ph->bgbuf.p, sizeof(zf) * n, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; ++i) { bg[2 * i] = h[i].re; bg[2 * i + 1] = h[i].im; }
  } else {
    HIP_TRY(hipMemcpy(bg, ph->bgbuf.p, sizeof(zd) * n, hipMemcpyDeviceToHost));
  }
  return 0;
}

int afqmc_ph_sample_apply(afqmc_ph ph, afqmc_ph_walkers w, const double* aux, const double* force, double* sample_logw,
                          int* max_order) {
  REQUIRE(ph && w && aux && force && w->ph == ph, "invalid argument");
  ScopedDevice sdv(ph->ctx);
  hipStream_t s = ph->ctx->stream;
  const int G = ph->G;
  // logw of getTwoBodySampleFromAuxForce (choleskyReal.cpp:104-124) and xs = sqrt(-dt) aux
  std::vector<zd> lw((size_t)w->nw);
  std::vector<double> xsd(2 * (size_t)G * w->nw);
  const cd sq(ph->sqrt_minus_dt.re, ph->sqrt_minus_dt.im);
  for (int iw = 0; iw < w->nw; ++iw) {
    cd f2 = 0.0, xf = 0.0, xb = 0.0;
    for (int g = 0; g < G; ++g) {
      const size_t k = 2 * ((size_t)iw * G + g);
      const cd x(aux[k], aux[k + 1]), f(force[k], force[k + 1]);
      f2 += f * f; xf += x * f; xb += x * ph->h_bg[g];
      const cd v = sq * x;
      xsd[k] = v.real(); xsd[k + 1] = v.imag();
    }
    const cd l = 0.5 * f2 - xf - sq * xb;
    lw[iw] = cmk<double>(l.real(), l.imag());
    if (sample_logw) { sample_logw[2 * iw] = l.real(); sample_logw[2 * iw + 1] = l.imag(); }
  }
  int rc;
  if (ph->fp32) {
    std::vector<zf> h((size_t)G * w->nw);
    for (size_t i = 0; i < h.size(); ++i) h[i] = cmk<float>((float)xsd[2 * i], (float)xsd[2 * i + 1]);
    HIP_TRY(hipMemcpy(ph->xs.p, h.data(), sizeof(zf) * h.size(), hipMemcpyHostToDevice));
    if ((rc = ph_sample_matrix<float>(ph, w))) return rc;
    if ((rc = ph_taylor<float>(ph, w, w->wf, ph->Msamp.p, (long long)ph->L * ph->L, w->alive.p, max_order))) return rc;
  } else {
    HIP_TRY(hipMemcpy(ph->xs.p, xsd.data(), sizeof(double) * xsd.size(), hipMemcpyHostToDevice));
    if ((rc = ph_sample_matrix<double>(ph, w))) return rc;
    if ((rc = ph_taylor<double>(ph, w, w->wf, ph->Msamp.p, (long long)ph->L * ph->L, w->alive.p, max_order))) return rc;
  }
  DevBuf<zd> dl;
  if ((rc = dl.alloc(w->nw))) return rc;
  HIP_TRY(hipMemcpy(dl.p, lw.data(), sizeof(zd) * w->nw, hipMemcpyHostToDevice));
  if ((rc = launch_logw_vec_add(w->nw, w->logw.p, w->alive.p, dl.p, s))) return rc;
  HIP_TRY(hipStreamSynchronize(s));
  return 0;
}

int afqmc_ph_sample_matrix(afqmc_ph ph, afqmc_ph_walkers w, double* M) {
  REQUIRE(ph && w && M && w->ph == ph, "invalid argument");
  ScopedDevice sdv(ph->ctx);
  HIP_TRY(hipStreamSynchronize(ph->ctx->stream));
  const size_t n = (size_t)ph->L * ph->L * w->nw;
  if (ph->fp32) {
    std::vector<zf> h(n);
    HIP_TRY(hipMemcpy(h.data(), ph->Msamp.p, sizeof(zf) * n, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; ++i) { M[2 * i] = h[i].re; M[2 * i + 1] = h[i].im; }
  } else {
    HIP_TRY(hipMemcpy(M, ph->Msamp.p, sizeof(zd) * n, hipMemcpyDeviceToHost));
  }
  return 0;
}

int afqmc_ph_taylor_orders(afqmc_ph ph, afqmc_ph_walkers w, int* orders) {
  REQUIRE(ph && w && orders, "null argument");
  ScopedDevice sdv(ph->ctx);
  HIP_TRY(hipStreamSynchronize(ph->ctx->stream));
  HIP_TRY(hipMemcpy(orders, ph->orders.p, sizeof(int) * w->nw, hipMemcpyDeviceToHost));
  return 0;
}

int afqmc_ph_onebody(afqmc_ph ph, afqmc_ph_walkers w, int which) {
  REQUIRE(ph && w && w->ph == ph, "invalid argument");
  REQUIRE(which >= 0 && which <= 2, "which must be 0, 1 or 2");
  ScopedDevice sdv(ph->ctx);
  int rc = ph->fp32 ? ph_onebody<float>(ph, w, which, w->wf, w->alive.p) : ph_onebody<double>(ph, w, which, w->wf, w->alive.p);
  if (rc) return rc;
  if ((rc = launch_logw_const_add(w->nw, w->logw.p, w->alive.p, ph->onebody_logw[which], ph->ctx->stream))) return rc;
  HIP_TRY(hipStreamSynchronize(ph->ctx->stream));
  return 0;
}

int afqmc_ph_stabilize(afqmc_ph ph, afqmc_ph_walkers w) {
  REQUIRE(ph && w && w->ph == ph, "invalid argument");
  ScopedDevice sdv(ph->ctx);
  int rc = ph_stabilize_impl(ph, w);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(ph->ctx->stream));
  return 0;
}

int afqmc_ph_step(afqmc_ph ph, afqmc_ph_walkers w, double* force_out, double* aux_out) {
  REQUIRE(ph && w && w->ph == ph, "invalid argument");
  ScopedDevice sdv(ph->ctx);
  return PH_DISPATCH(ph, ph_step_impl, ph, w, force_out, aux_out);
}

int afqmc_ph_pop_control(afqmc_ph ph, afqmc_ph_walkers w, double* weights_out, int* table_out, double* stats) {
  REQUIRE(ph && w && w->ph == ph, "invalid argument");
  ScopedDevice sdv(ph->ctx);
  return PH_DISPATCH(ph, ph_pop_control_impl, ph, w, weights_out, table_out, stats);
}

int afqmc_ph_measure(afqmc_ph ph, afqmc_ph_walkers w) {
  REQUIRE(ph && w && w->ph == ph, "invalid argument");
  ScopedDevice sdv(ph->ctx);
  return PH_DISPATCH(ph, ph_measure_impl, ph, w);
}

int afqmc_ph_energy(afqmc_ph ph, double* out, double* bg_num, double* ex_num) {
  REQUIRE(ph && out, "null argument");
  ScopedDevice sdv(ph->ctx);
  DevScalars h;
  int rc = ph_scalars(ph, &h, false);
  if (rc) return rc;
  zd v[3] = {h.den, h.HNum, h.KNum};
  for (int k = 0; k < 3; ++k) { out[2 * k] = v[k].re; out[2 * k + 1] = v[k].im; }
  if (bg_num) HIP_TRY(hipMemcpy(bg_num, ph->bg_acc.p, sizeof(zd) * ph->G, hipMemcpyDeviceToHost));
  if (ex_num) HIP_TRY(hipMemcpy(ex_num, ph->ex_acc.p, sizeof(zd) * ph->G, hipMemcpyDeviceToHost));
  return 0;
}

int afqmc_ph_reset_measure(afqmc_ph ph) {
  REQUIRE(ph, "null argument");
  ScopedDevice sdv(ph->ctx);
  hipStream_t s = ph->ctx->stream;
  int rc = launch_reset_measure(ph->sc.p, ph->bg_acc.p, 0, s);
  if (rc) return rc;
  HIP_TRY(hipMemsetAsync(ph->bg_acc.p, 0, sizeof(zd) * ph->G, s));
  HIP_TRY(hipMemsetAsync(ph->ex_acc.p, 0, sizeof(zd) * ph->G, s));
  return 0;
}

int afqmc_ph_adjust_ET(afqmc_ph ph) {
  REQUIRE(ph, "null argument");
  ScopedDevice sdv(ph->ctx);
  int rc = launch_set_et_from_energy(ph->sc.p, ph->ctx->stream);
  if (rc) return rc;
  return afqmc_ph_reset_measure(ph);
}

}  // extern "C"
