// C-ABI of the constraint-path walker path: context, walker sets and the
// batched drop-ins for the reference's per-walker calls (see
// include/afqmc_mi355x.h for the reference function each entry replaces).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstdarg>
#include <cstring>
#include <mutex>
#include <vector>

#include "cholqr.hpp"
#include "cp_internal.hpp"

// ------------------------------------------------------------------ NiupNidn constants (host)
// solveCoshxEqExpy / solveCosxEqExpy / exp2by2Matrix, simple_fun.cpp:7-91
// Written out in real arithmetic on the branches the reference's libm (glibc) takes for
// an argument with +0 imaginary part: sqrt(-a + 0i) = +i sqrt(a) and
// acos(x + 0i) = -i acosh(x) for x > 1.  (The complex overloads of the host compiler
// used with hipcc return +i acosh(x) there, which flips gamma of hopSpin.)
static cd solve_cosh_exp(double y) {
  const double e = std::exp(y);
  cd g;
  if (e * e - 1.0 >= 0.0) {
    g = cd(std::log(e - std::sqrt(e * e - 1.0)), 0.0);
  } else {
    const double s = std::sqrt(1.0 - e * e);  // sqrt(e^2 - 1) = i s
    g = cd(std::log(std::hypot(e, s)), std::atan2(-s, e));
  }
  return std::abs(g.real()) < std::abs(g.imag()) ? cd(0, g.imag()) : cd(g.real(), 0);
}
static cd solve_cos_exp(double y) {
  const double e = std::exp(y);
  cd g = (e <= 1.0) ? cd(std::acos(e), 0.0) : cd(0.0, -std::acosh(e));
  return std::abs(g.real()) < std::abs(g.imag()) ? cd(0, g.imag()) : cd(g.real(), 0);
}
static void exp2by2(double& a, double& b, cd& c) {
  if (std::abs(c) < 1e-60) { a = std::exp(a); b = std::exp(b); c = 0.0; return; }
  double ac = std::abs(c), arg = std::arg(c);
  double d0, d1;
  cd v00, v01, v10, v11, im(0, 1);
  if (std::abs(a - b) < 1e-60) {
    double n = 1.0 / std::sqrt(2.0);
    d0 = std::exp(-ac + a); d1 = std::exp(ac + b);
    v00 = -n; v10 = n * std::exp(im * arg); v01 = n; v11 = n * std::exp(im * arg);
  } else {
    double xi = 0.5 * std::atan(2.0 * ac / (a - b));
    d0 = std::exp(0.5 * std::cos(2 * xi) * (a - b) + std::sin(2 * xi) * ac + (a + b) * 0.5);
    d1 = std::exp(-0.5 * std::cos(2 * xi) * (a - b) - std::sin(2 * xi) * ac + (a + b) * 0.5);
    v00 = std::exp(-im * arg * 0.5) * std::cos(xi);
    v10 = std::exp(im * arg * 0.5) * std::sin(xi);
    v01 = -std::exp(-im * arg * 0.5) * std::sin(xi);
    v11 = std::exp(im * arg * 0.5) * std::cos(xi);
  }
  a = (d0 * v00 * std::conj(v00) + d1 * v01 * std::conj(v01)).real();
  b = (d0 * v10 * std::conj(v10) + d1 * v11 * std::conj(v11)).real();
  c = d0 * v10 * std::conj(v00) + d1 * v11 * std::conj(v01);
}

using namespace cpi;

namespace {

template <typename T>
LogOArgs logo_args(afqmc_cp cp, afqmc_walkers w, int mode) {
  LogOArgs l;
  memset(&l, 0, sizeof(l));
  l.nw = w->nw; l.nblk = cp->nblk; l.combine = COMBINE_WRAP; l.mode = mode;
  for (int b = 0; b < cp->nblk; ++b) { l.ldet[b] = cp->ldet[b].p; l.info[b] = cp->info[b].p; }
  l.logw = w->logw.p; l.logw_out = w->logw.p; l.alive = w->alive.p; l.trial_logw = cp->trial_logw;
  l.logO = cp->logO.p; l.weight = cp->weight.p; l.den = cp->den.p;
  return l;
}

template <typename T>
int do_overlap(afqmc_cp cp, afqmc_walkers w, bool inverse) {
  int rc = overlap_gemm<T>(cp, w, w->wf);
  if (!rc) rc = gj_blocks<T>(cp, w, inverse, w->alive.p);
  return rc;
}

template <typename T>
HsSampleArgs hs_args(afqmc_cp cp, afqmc_walkers w) {
  HsSampleArgs h;
  memset(&h, 0, sizeof(h));
  h.nw = w->nw; h.L = cp->L; h.decomp = cp->decomp; h.force_type = cp->force_type; h.nblk = cp->nblk;
  h.dense = cp->dense;
  for (int b = 0; b < cp->nblk; ++b) {
    h.gd[b] = cp->gpart[0][b].p; h.go[b] = cp->dense ? cp->gpart[1][b].p : nullptr;
    h.ntile[b] = cp->ntile[b]; h.Mb[b] = cp->Mb[b];
    h.scale[b] = cp->scale[b].p;
  }
  h.gamma = cp->gamma.p; h.cdiag = cp->cdiag.p; h.const_force = cp->cforce.p;
  h.cap = cp->cap; h.dt = cp->dt; h.log_half_L = cp->log_half_L; h.dtUSum = cp->dtUSum;
  h.onebody_logw = cp->hop_logw[0];
  h.alive = w->alive.p; h.prefix = cp->prefix.p; h.sc = cp->sc.p;
  h.ubuf = cp->ubuf.p; h.logw = w->logw.p; h.diag_only = cp->diag_only; h.d2x2 = cp->d2x2.p;
  h.fp32 = cp->fp32;
  return h;
}

template <typename T>
int step_impl(afqmc_cp cp, afqmc_walkers w, double* force_out, int* aux_out, int* aux_dev = nullptr,
              const int* aux_in_dev = nullptr) {
  hipStream_t s = cp->ctx->stream;
  int rc;
  REQUIRE(!capturing() || (!force_out && !aux_out), "a captured step cannot return records");
  for (int b = 0; b < cp->nblk; ++b) capture_note_pair(&w->wf[b], &w->tmp[b]);
  {
    RegionTimer t(cp->ctx, REG_OVERLAP);
    if ((rc = overlap_gemm<T>(cp, w, w->wf))) return rc;
  }
  {
    RegionTimer t(cp->ctx, REG_GJ);
    if ((rc = gj_blocks<T>(cp, w, true, w->alive.p))) return rc;
  }
  LogOArgs l = logo_args<T>(cp, w, LOGO_STEP);
  if ((rc = launch_logo(l, s))) return rc;
  if ((rc = launch_alive_scan(w->nw, w->alive.p, cp->prefix.p, cp->sc.p, s))) return rc;
  {
    RegionTimer t(cp->ctx, REG_GREEN);
    if ((rc = green_gemm<T>(cp, w, w->wf, w->alive.p, false))) return rc;
  }
  HsSampleArgs h = hs_args<T>(cp, w);
  if (force_out || aux_out) {
    h.force_out = cp->force_rec.p;
    h.aux_out = cp->aux_rec.p;
  }
  if (aux_dev) h.aux_out = aux_dev;
  h.aux_in = aux_in_dev;
  {
    RegionTimer t(cp->ctx, REG_HS_SAMPLE);
    if ((rc = launch_hs_sample(h, s))) return rc;
    // given fields draw no uniforms: only the walker-step count advances
    if ((rc = launch_advance_counter(cp->sc.p, aux_in_dev ? 0 : cp->L, s))) return rc;
  }
  if (cp->diag_only) {
    RegionTimer t(cp->ctx, REG_PROPAGATE);
    if ((rc = propagate_gemm<T>(cp, w, 0, w->wf, w->tmp, true))) return rc;
  } else {
    {
      RegionTimer t(cp->ctx, REG_HS_APPLY);
      if ((rc = launch_hs_apply_2x2<T>(w->nw, cp->L, cp->N, w->wf[0].p, w->tmp[0].p, cp->d2x2.p, s))) return rc;
    }
    for (int b = 0; b < cp->nblk; ++b) std::swap(w->wf[b], w->tmp[b]);
    RegionTimer t(cp->ctx, REG_PROPAGATE);
    if ((rc = propagate_gemm<T>(cp, w, 0, w->wf, w->tmp, false))) return rc;
  }
  // the propagated walkers are in tmp: make them current
  for (int b = 0; b < cp->nblk; ++b) std::swap(w->wf[b], w->tmp[b]);
  if (force_out) HIP_TRY(hipMemcpyAsync(force_out, cp->force_rec.p, sizeof(double) * w->nw * cp->L, hipMemcpyDeviceToHost, s));
  if (aux_out) HIP_TRY(hipMemcpyAsync(aux_out, cp->aux_rec.p, sizeof(int) * w->nw * cp->L, hipMemcpyDeviceToHost, s));
  if (force_out || aux_out) HIP_TRY(hipStreamSynchronize(s));
  return 0;
}

CholQrBlock qr_block(afqmc_cp cp, afqmc_walkers w, DevBuf<char>* wf, DevBuf<char>* tmp, int b, double tol) {
  CholQrBlock q;
  q.nw = w->nw; q.M = cp->Mb[b]; q.N = cp->Nb[b]; q.nsub = cp->Nb[b];
  q.wf = &wf[b]; q.tmp = &tmp[b];
  q.S = reinterpret_cast<zd*>(cp->qrS[b].p); q.ld = cp->qrld[b].p;
  q.info = cp->qrinfo[b].p; q.need = cp->qr_need.p; q.flag = cp->qr_flag.p;
  q.alive = w->alive.p; q.logw = w->logw.p;
  q.m0 = 1.0; q.m1 = 0.0; q.tol = tol; q.sc = cp->sc.p;
  return q;
}

// SD::stabilize of each block by robust CholeskyQR in FP64 (cholqr.hpp): W <- Q, logw += sum log R_kk
int qr_impl_f64(afqmc_cp cp, afqmc_walkers w, DevBuf<char>* wf, DevBuf<char>* tmp, double tol) {
  for (int b = 0; b < cp->nblk; ++b) {
    int rc = cholqr_f64(qr_block(cp, w, wf, tmp, b, tol), cp->ctx->stream);
    if (rc) return rc;
  }
  return 0;
}
int stabilize_impl(afqmc_cp cp, afqmc_walkers w) {
  RegionTimer t(cp->ctx, REG_QR);
  for (int b = 0; b < cp->nblk; ++b) {
    capture_note_pair(&w->wf[b], &w->tmp[b]);
    capture_note_pair(&cp->qr64[b], &cp->qr64b[b]);
  }
  if (!cp->fp32) return qr_impl_f64(cp, w, w->wf, w->tmp, 1e-13);
  // mixed path: QR in FP64 on a widened copy of the FP32 walkers
  hipStream_t s = cp->ctx->stream;
  int rc;
  {
    GroupFork fork;
    for (int b = 0; b < cp->nblk; ++b) {
      fork.branch();
      if ((rc = launch_convert_f32_f64(w->wf[b].p, cp->qr64[b].p, w->wsize(b) * w->nw, s))) return rc;
    }
  }
  // pass 1 and the shifted walkers' further passes in FP64, then the result rounded to FP32
  for (int b = 0; b < cp->nblk; ++b) {
    CholQrBlock q = qr_block(cp, w, cp->qr64, cp->qr64b, b, 1e-8);
    q.check_all = false;
    if ((rc = cholqr_f64(q, s))) return rc;
  }
  {
    GroupFork fork;
    for (int b = 0; b < cp->nblk; ++b) {
      fork.branch();
      if ((rc = launch_convert_f64_f32(cp->qr64[b].p, w->wf[b].p, w->wsize(b) * w->nw, s))) return rc;
    }
  }
  // Pass 2 from the rounded walkers, whose orthonormality FP32 rounding already limits to about
  // 1e-7: their Gram matrix in FP32 (a quarter of the FP64 Gram's time), and the walkers farther
  // than 1e-4 from orthonormal (beyond that Gram's own rounding) take the FP64 check and pass to
  // 1e-8 on their FP64 copy and are rounded again.
  GroupFork fork;
  for (int b = 0; b < cp->nblk; ++b) {
    fork.branch();
    const int N = cp->Nb[b], M = cp->Mb[b];
    GemmArgs a;
    memset(&a, 0, sizeof(a));
    a.alpha_re = 1.0;
    a.scale_group = 1;
    a.M = N; a.N = N; a.K = M; a.batch = w->nw;
    a.A = w->wf[b].p; a.lda = M; a.strideA = (long long)M * N;
    a.B = w->wf[b].p; a.ldb = M; a.strideB = (long long)M * N;
    a.C = cp->qrS[b].p; a.ldc = N; a.strideC = (long long)N * N;
    a.alive = w->alive.p;
    a.herm = 1;
    a.tile = GEMM_TILE_64x64;
    if ((rc = launch_cgemm<float>(OP_C, EPI_STORE, a, s))) return rc;
    if ((rc = launch_ortho_check_f32(w->nw, N, reinterpret_cast<const zf*>(cp->qrS[b].p), w->alive.p, 1e-4,
                                     cp->qr_need.p, s)))
      return rc;
    const CholQrBlock q = qr_block(cp, w, cp->qr64, cp->qr64b, b, 1e-8);
    if ((rc = cholqr_refine(q, cp->qr_need.p, s))) return rc;
    if ((rc = launch_convert_f64_f32_masked(cp->qr64[b].p, w->wf[b].p, w->wsize(b), w->nw, cp->qr_need.p, s)))
      return rc;
  }
  return 0;
}

template <typename T>
int pop_control_impl(afqmc_cp cp, afqmc_walkers w, double* weights_out, int* table_out, double* stats,
                     const int* table_in = nullptr) {
  hipStream_t s = cp->ctx->stream;
  RegionTimer t(cp->ctx, REG_COMB);
  int rc = do_overlap<T>(cp, w, false);
  if (rc) return rc;
  LogOArgs l = logo_args<T>(cp, w, LOGO_WEIGHT);
  if ((rc = launch_logo(l, s))) return rc;
  if (table_in) {
    HIP_TRY(hipMemcpyAsync(cp->table.p, table_in, sizeof(int) * w->nw, hipMemcpyHostToDevice, s));
  } else if ((rc = launch_comb(w->nw, cp->weight.p, cp->ubuf.p, cp->sc.p, cp->table.p, cp->comb_prob.p,
                               cp->comb_num.p, s))) {
    return rc;
  }
  void* wfp[2] = {w->wf[0].p, cp->nblk > 1 ? w->wf[1].p : nullptr};
  long long ws[2] = {w->wsize(0), cp->nblk > 1 ? w->wsize(1) : 0};
  if ((rc = launch_apply_table<T>(w->nw, cp->nblk, wfp, ws, w->logw.p, w->alive.p, cp->table.p, s))) return rc;
  if (weights_out) HIP_TRY(hipMemcpyAsync(weights_out, cp->weight.p, sizeof(double) * w->nw, hipMemcpyDeviceToHost, s));
  if (table_out) HIP_TRY(hipMemcpyAsync(table_out, cp->table.p, sizeof(int) * w->nw, hipMemcpyDeviceToHost, s));
  if (stats) {
    DevScalars h;
    HIP_TRY(hipMemcpyAsync(&h, cp->sc.p, sizeof(h), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    stats[0] = h.wmin; stats[1] = h.wmax; stats[2] = h.wavg;
  }
  if (weights_out || table_out) HIP_TRY(hipStreamSynchronize(s));
  return 0;
}

template <typename T>
int measure_impl(afqmc_cp cp, afqmc_walkers w) {
  hipStream_t s = cp->ctx->stream;
  RegionTimer t(cp->ctx, REG_MEASURE);
  int rc;
  if ((rc = propagate_gemm<T>(cp, w, 2, w->wf, w->tmp, false))) return rc;
  if ((rc = launch_logw_shift(w->nw, w->logw.p, w->logw_tmp.p, cp->hop_logw[2], s))) return rc;
  if ((rc = overlap_gemm<T>(cp, w, w->tmp))) return rc;
  if ((rc = gj_blocks<T>(cp, w, true, w->alive.p))) return rc;
  LogOArgs l = logo_args<T>(cp, w, LOGO_MEASURE);
  l.logw = w->logw_tmp.p;
  if ((rc = launch_logo(l, s))) return rc;
  if ((rc = green_gemm<T>(cp, w, w->tmp, w->alive.p, true))) return rc;
  EnergyArgs e;
  memset(&e, 0, sizeof(e));
  e.nw = w->nw; e.L = cp->L; e.nblk = cp->nblk; e.dense = cp->dense;
  for (int b = 0; b < cp->nblk; ++b) {
    e.gd[b] = cp->gpart[0][b].p; e.go[b] = cp->dense ? cp->gpart[1][b].p : nullptr; e.gk[b] = cp->gpart[2][b].p;
    e.ntile[b] = cp->ntile[b]; e.Mb[b] = cp->Mb[b];
  }
  e.U = cp->U.p; e.mu = cp->mu.p; e.hx = cp->hx.p; e.hy = cp->hy.p; e.hz = cp->hz.p;
  e.alive = w->alive.p; e.den = cp->den.p; e.per_walker = cp->per_walker.p; e.site_acc = cp->site_acc.p; e.site_part = cp->site_part.p; e.sc = cp->sc.p;
  return launch_energy(e, s);
}

// den[w] = exp(logO) of phiT and exp(dt/2 K) W_w, the weight of addMixedMeasurement
// (afqmcConstraintPathMeasure.cpp:14-26), without accumulating the energy
template <typename T>
int mixed_weight_impl(afqmc_cp cp, afqmc_walkers w) {
  hipStream_t s = cp->ctx->stream;
  int rc;
  if ((rc = propagate_gemm<T>(cp, w, 2, w->wf, w->tmp, false))) return rc;
  if ((rc = launch_logw_shift(w->nw, w->logw.p, w->logw_tmp.p, cp->hop_logw[2], s))) return rc;
  if ((rc = overlap_gemm<T>(cp, w, w->tmp))) return rc;
  if ((rc = gj_blocks<T>(cp, w, false, w->alive.p))) return rc;
  LogOArgs l = logo_args<T>(cp, w, LOGO_MEASURE);
  l.logw = w->logw_tmp.p;
  return launch_logo(l, s);
}

}  // namespace

#define DISPATCH(cp, fn, ...) ((cp)->fp32 ? fn<float>(__VA_ARGS__) : fn<double>(__VA_ARGS__))

int cpi::step_record(afqmc_cp cp, afqmc_walkers w, int* aux_dev) {
  return DISPATCH(cp, step_impl, cp, w, nullptr, nullptr, aux_dev);
}
int cpi::mixed_weight(afqmc_cp cp, afqmc_walkers w) { return DISPATCH(cp, mixed_weight_impl, cp, w); }
int cpi::stabilize(afqmc_cp cp, afqmc_walkers w) { return stabilize_impl(cp, w); }

// ------------------------------------------------------------------ C entry points
namespace {
template <typename T>
int tune_role(afqmc_cp cp, afqmc_walkers w, int role, int reps, double* ms_row) {
  hipStream_t s = cp->ctx->stream;
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  double best = 1e300;
  int best_tile = cp->tile[role];
  int rc = 0;
  for (int t = 0; t < GEMM_TILE_COUNT && !rc; ++t) {
    ms_row[t] = -1.0;
    if (role == 2 && !gemm_tile_green(t)) continue;
    cp->tile[role] = t;
    auto run = [&]() {
      // each role reads the walkers and writes only workspace (tmp, overlaps, Green partials)
      if (role == 0) return propagate_gemm<T>(cp, w, 0, w->wf, w->tmp, cp->diag_only != 0);
      if (role == 1) return overlap_gemm<T>(cp, w, w->wf);
      return green_gemm<T>(cp, w, w->wf, w->alive.p, false);
    };
    if ((rc = run())) break;
    if (hipEventRecord(e0, s) != hipSuccess) { rc = -2; break; }
    for (int r = 0; r < reps && !rc; ++r) rc = run();
    if (rc) break;
    if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) { rc = -2; break; }
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    ms_row[t] = ms / reps;
    if (ms_row[t] < best) { best = ms_row[t]; best_tile = t; }
  }
  cp->tile[role] = best_tile;
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (rc == -2) afqmc_set_error("afqmc_cp_autotune: event failure");
  return rc;
}
}  // namespace

extern "C" {

int afqmc_cp_create(afqmc_ctx ctx, const afqmc_cp_desc* d, afqmc_cp* out) {
  NOT_CAPTURING();
  REQUIRE(ctx && d && out, "afqmc_cp_create: null argument");
  REQUIRE(d->L > 0 && d->N > 0 && d->max_walkers > 0, "afqmc_cp_create: L, N, max_walkers must be positive");
  REQUIRE(d->K && d->exp_minus_dt_k && d->exp_minus_half_dt_k && d->exp_half_dt_k && d->phiT && d->U,
          "afqmc_cp_create: K, propagators, phiT and U are required");
  REQUIRE(d->decomp_type >= 0 && d->decomp_type <= 3, "Can not find the matched decompType! %d", d->decomp_type);
  REQUIRE(d->force_type == 0 || d->force_type == 1, "Do not know method.forceType %d", d->force_type);
  REQUIRE(d->force_cap >= 0.0, "forceCap must be positive or zero!");
  REQUIRE(d->precision >= 0 && d->precision <= 2, "afqmc_cp_create: precision must be 0, 1 or 2");
  ScopedDevice sd(ctx);
  afqmc_cp cp = new afqmc_cp_s();
  cp->ctx = ctx;
  cp->device = ctx->device;
  cp->L = d->L; cp->N = d->N; cp->fp32 = d->precision ? 1 : 0; cp->split_bf16 = d->precision == 2 ? 1 : 0;
  cp->decomp = d->decomp_type; cp->force_type = d->force_type; cp->dt = d->dt; cp->cap = d->force_cap;
  cp->maxw = d->max_walkers;
  const int L = d->L;
  auto zero_or = [&](const double* p, std::vector<double>& v) { v.assign(L, 0.0); if (p) std::copy(p, p + L, v.begin()); };
  std::vector<double> U, mu, hx, hy, hz, cf;
  zero_or(d->U, U); zero_or(d->mu, mu); zero_or(d->hx, hx); zero_or(d->hy, hy); zero_or(d->hz, hz);
  zero_or(d->const_force, cf);
  bool mixing = false;
  for (int i = 0; i < L; ++i) if (hx[i] != 0.0 || hy[i] != 0.0) mixing = true;
  const bool hop = cp->decomp >= 2;
  if (d->spin_blocks) {
    if (hop || mixing) {
      delete cp;
      REQUIRE(false, "spin_blocks needs a density decomposition and hx = hy = 0 (the HS field must not mix spins)");
    }
    REQUIRE(d->Nup >= 0 && d->Nup <= d->N, "invalid Nup");
    cp->nblk = 2; cp->dense = 0;
    cp->Nup = d->Nup; cp->Ndn = d->N - d->Nup;
    cp->Mb[0] = cp->Mb[1] = L;
    cp->Nb[0] = cp->Nup; cp->Nb[1] = cp->Ndn;
    cp->row0[0] = 0; cp->row0[1] = L; cp->col0[0] = 0; cp->col0[1] = cp->Nup;
    if (cp->Nup == 0 || cp->Ndn == 0) {
      delete cp;
      REQUIRE(false, "spin_blocks needs Nup > 0 and Ndn > 0");
    }
  } else {
    cp->nblk = 1; cp->dense = 1;
    cp->Mb[0] = 2 * L; cp->Nb[0] = d->N;
  }
  for (int b = 0; b < cp->nblk; ++b) {
    if (cp->Nb[b] > 256) {
      const int nb = cp->Nb[b];
      delete cp;
      REQUIRE(false, "block column count %d exceeds the supported 256", nb);
    }
  }
  cp->diag_only = (!hop && !mixing) ? 1 : 0;
  // NiupNidn constants (NiupNidn.cpp:12-26, 201-250)
  cp->h_gamma.resize(L);
  cp->h_cdiag.resize(4 * L);
  double usum = 0.0;
  for (int i = 0; i < L; ++i) usum += U[i];
  cp->dtUSum = d->dt * usum;
  cp->log_half_L = std::log(0.5) * L;
  for (int i = 0; i < L; ++i) {
    double dtU = d->dt * U[i];
    switch (cp->decomp) {
      case 0: cp->h_gamma[i] = solve_cosh_exp(-dtU * 0.5); break;
      case 1: case 2: cp->h_gamma[i] = solve_cosh_exp(dtU * 0.5); break;
      default: cp->h_gamma[i] = solve_cos_exp(dtU * 0.5); break;
    }
    double a = -d->dt * (U[i] * 0.5 - mu[i] + hz[i] * 0.5);
    double b = -d->dt * (U[i] * 0.5 - mu[i] - hz[i] * 0.5);
    cd c = -d->dt * 0.5 * cd(hx[i], hy[i]);
    exp2by2(a, b, c);
    cp->h_cdiag[i] = a; cp->h_cdiag[L + i] = std::conj(c); cp->h_cdiag[2 * L + i] = c; cp->h_cdiag[3 * L + i] = b;
  }
  int rc = 0;
  auto up = [&](DevBuf<double>& dst, const std::vector<double>& v) {
    if (rc) return;
    rc = dst.alloc(v.size());
    if (!rc && hipMemcpy(dst.p, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) rc = -2;
  };
  up(cp->U, U); up(cp->mu, mu); up(cp->hx, hx); up(cp->hy, hy); up(cp->hz, hz); up(cp->cforce, cf);
  if (!rc) rc = cp->gamma.alloc(L);
  if (!rc) rc = cp->cdiag.alloc(4 * L);
  if (!rc && hipMemcpy(cp->gamma.p, cp->h_gamma.data(), sizeof(zd) * L, hipMemcpyHostToDevice) != hipSuccess) rc = -2;
  if (!rc && hipMemcpy(cp->cdiag.p, cp->h_cdiag.data(), sizeof(zd) * 4 * L, hipMemcpyHostToDevice) != hipSuccess) rc = -2;
  // trial, propagators, (phiT^H K)^T
  const int L2 = 2 * L;
  std::vector<cd> K((const cd*)d->K, (const cd*)d->K + (size_t)L2 * L2);
  cp->h_K = K;
  cp->h_U = U; cp->h_mu = mu; cp->h_hx = hx; cp->h_hy = hy; cp->h_hz = hz;
  std::vector<cd> phi((const cd*)d->phiT, (const cd*)d->phiT + (size_t)L2 * d->N);
  std::vector<cd> wdkT((size_t)L2 * d->N);  // (phiT^H K)^T = K^T conj(phiT): 2L x N
  for (int n = 0; n < d->N; ++n)
    for (int r = 0; r < L2; ++r) {
      cd s = 0.0;
      for (int q = 0; q < L2; ++q) s += std::conj(phi[q + (size_t)n * L2]) * K[q + (size_t)r * L2];
      wdkT[r + (size_t)n * L2] = s;
    }
  const double* mats[3] = {d->exp_minus_dt_k, d->exp_minus_half_dt_k, d->exp_half_dt_k};
  for (int k = 0; k < 3; ++k) cp->hop_logw[k] = cmk<double>(d->hop_logw[2 * k], d->hop_logw[2 * k + 1]);
  cp->trial_logw = cmk<double>(d->phiT_logw[0], d->phiT_logw[1]);
  if (!rc) {
    if (cp->fp32) {
      rc = upload_matrix_blocks<float>(cp, d->phiT, d->N, cp->phiT, false);
      for (int k = 0; k < 3 && !rc; ++k) rc = upload_matrix_blocks<float>(cp, mats[k], L2, cp->expK[k], true);
      if (!rc) rc = upload_matrix_blocks<float>(cp, (const double*)wdkT.data(), d->N, cp->wdkT, false);
    } else {
      rc = upload_matrix_blocks<double>(cp, d->phiT, d->N, cp->phiT, false);
      for (int k = 0; k < 3 && !rc; ++k) rc = upload_matrix_blocks<double>(cp, mats[k], L2, cp->expK[k], true);
      if (!rc) rc = upload_matrix_blocks<double>(cp, (const double*)wdkT.data(), d->N, cp->wdkT, false);
    }
  }
  // workspace
  const int nw = d->max_walkers;
  const size_t es = cp->esz();
  for (int b = 0; b < cp->nblk && !rc; ++b) {
    const long long Mb = cp->Mb[b], Nb = cp->Nb[b];
    cp->ntile[b] = (int)((Nb + 63) / 64);
    rc = cp->ovl[b].alloc_bytes(es * Nb * Nb * nw);
    if (!rc && small_dense_gj_work_bytes((int)Nb, nw) > 0)
      rc = cp->gjwork[b].alloc_bytes(small_dense_gj_work_bytes((int)Nb, nw));
    if (!rc) rc = cp->ldet[b].alloc(nw);
    if (!rc) rc = cp->info[b].alloc(nw);
    for (int g = 0; g < 3 && !rc; ++g) rc = cp->gpart[g][b].alloc((size_t)nw * cp->ntile[b] * Mb);
    if (!rc) rc = cp->scale[b].alloc_bytes(es * Mb * nw);
    if (!rc) rc = cp->qrS[b].alloc_bytes(sizeof(zd) * Nb * Nb * nw);
    if (!rc) rc = cp->qrld[b].alloc(nw);
    if (!rc) rc = cp->qrinfo[b].alloc(nw);
    if (!rc && cp->fp32) rc = cp->qr64[b].alloc_bytes(sizeof(zd) * Mb * Nb * nw);
    if (!rc && cp->fp32) rc = cp->qr64b[b].alloc_bytes(sizeof(zd) * Mb * Nb * nw);
  }
  if (!rc && !cp->diag_only) rc = cp->d2x2.alloc_bytes(es * 4 * L * nw);
  if (!rc) rc = cp->logO.alloc(nw);
  if (!rc) rc = cp->den.alloc(nw);
  if (!rc) rc = cp->prefix.alloc(nw);
  if (!rc) rc = cp->table.alloc(nw);
  if (!rc) rc = cp->comb_num.alloc(nw);
  if (!rc) rc = cp->qr_need.alloc(nw);
  if (!rc) rc = cp->qr_flag.alloc(nw);
  if (!rc) rc = cp->comb_prob.alloc(nw);
  if (!rc) rc = cp->weight.alloc(nw);
  if (!rc) rc = cp->per_walker.alloc((size_t)nw * 8);
  if (!rc) rc = cp->site_acc.alloc((size_t)4 * L);
  if (!rc) rc = cp->site_part.alloc((size_t)SITE_CHUNKS * 4 * L);
  if (!rc) rc = cp->force_rec.alloc((size_t)nw * L);
  if (!rc) rc = cp->aux_rec.alloc((size_t)nw * L);
  if (!rc) rc = cp->sc.alloc(1);
  if (!rc) {
    DevScalars h;
    memset(&h, 0, sizeof(h));
    h.ET = -10.0;
    h.key = afqmc_seed_mix_host(985456376ull);
    if (hipMemcpy(cp->sc.p, &h, sizeof(h), hipMemcpyHostToDevice) != hipSuccess) rc = -2;
    if (!rc && hipMemset(cp->site_acc.p, 0, sizeof(zd) * 4 * L) != hipSuccess) rc = -2;
  }
  if (rc) {
    delete cp;
    if (rc == -2) afqmc_set_error("afqmc_cp_create: device allocation or copy failed");
    return rc;
  }
  *out = cp;
  return 0;
}

int afqmc_cp_destroy(afqmc_cp cp) {
  if (!cp) return 0;
  ScopedDevice sd(cp->device);
  hipDeviceSynchronize();
  delete cp;
  return 0;
}

int afqmc_cp_get_hs_constants(afqmc_cp cp, double* out) {
  NOT_CAPTURING();
  REQUIRE(cp && out, "null argument");
  const int L = cp->L;
  for (int i = 0; i < L; ++i) { out[2 * i] = cp->h_gamma[i].real(); out[2 * i + 1] = cp->h_gamma[i].imag(); }
  for (int i = 0; i < 4 * L; ++i) {
    out[2 * (L + i)] = cp->h_cdiag[i].real();
    out[2 * (L + i) + 1] = cp->h_cdiag[i].imag();
  }
  return 0;
}

int afqmc_cp_walkers_create(afqmc_cp cp, int nw, afqmc_walkers* out) {
  NOT_CAPTURING();
  REQUIRE(cp && out, "null argument");
  REQUIRE(nw > 0 && nw <= cp->maxw, "walker count %d outside 1..max_walkers=%d", nw, cp->maxw);
  ScopedDevice sd(cp->ctx);
  afqmc_walkers w = new afqmc_walkers_s();
  w->cp = cp;
  w->device = cp->device;
  w->nw = nw;
  int rc = 0;
  for (int b = 0; b < cp->nblk && !rc; ++b) {
    rc = w->wf[b].alloc_bytes(cp->esz() * w->wsize(b) * nw);
    if (!rc) rc = w->tmp[b].alloc_bytes(cp->esz() * w->wsize(b) * nw);
    if (!rc && hipMemset(w->wf[b].p, 0, cp->esz() * w->wsize(b) * nw) != hipSuccess) rc = -2;
  }
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
    if (rc == -2) afqmc_set_error("afqmc_cp_walkers_create: device allocation failed");
    return rc;
  }
  *out = w;
  return 0;
}

int afqmc_walkers_destroy(afqmc_walkers w) {
  if (!w) return 0;
  ScopedDevice sd(w->device);
  hipDeviceSynchronize();
  delete w;
  return 0;
}

int afqmc_walkers_count(afqmc_walkers w) { return w ? w->nw : 0; }

int afqmc_walkers_upload(afqmc_walkers w, const double* wf, const double* logw, const int* alive) {
  NOT_CAPTURING();
  REQUIRE(w && wf && logw, "null argument");
  afqmc_cp cp = w->cp;
  ScopedDevice sd(cp->ctx);
  HIP_TRY(hipStreamSynchronize(cp->ctx->stream));
  const int L2 = 2 * cp->L, N = cp->N;
  const long long per = (long long)L2 * N;
  if (!cp->dense) {
    // the off-block entries must be zero for the block representation to be exact
    for (int iw = 0; iw < w->nw; ++iw)
      for (int j = 0; j < N; ++j)
        for (int i = 0; i < L2; ++i) {
          bool in_up = i < cp->L && j < cp->Nup;
          bool in_dn = i >= cp->L && j >= cp->Nup;
          if (!in_up && !in_dn) {
            const double* z = wf + 2 * (iw * per + i + (long long)j * L2);
            REQUIRE(z[0] == 0.0 && z[1] == 0.0, "walker %d is not block diagonal (row %d, column %d)", iw, i, j);
          }
        }
  }
  for (int b = 0; b < cp->nblk; ++b) {
    const long long ws = w->wsize(b);
    if (cp->fp32) {
      std::vector<zf> h(ws * w->nw);
      for (int iw = 0; iw < w->nw; ++iw)
        pack_block<float>(wf + 2 * iw * per, L2, N, cp->row0[b], cp->col0[b], cp->Mb[b], cp->Nb[b], h, iw * ws);
      HIP_TRY(hipMemcpy(w->wf[b].p, h.data(), sizeof(zf) * h.size(), hipMemcpyHostToDevice));
    } else {
      std::vector<zd> h(ws * w->nw);
      for (int iw = 0; iw < w->nw; ++iw)
        pack_block<double>(wf + 2 * iw * per, L2, N, cp->row0[b], cp->col0[b], cp->Mb[b], cp->Nb[b], h, iw * ws);
      HIP_TRY(hipMemcpy(w->wf[b].p, h.data(), sizeof(zd) * h.size(), hipMemcpyHostToDevice));
    }
  }
  HIP_TRY(hipMemcpy(w->logw.p, logw, sizeof(zd) * w->nw, hipMemcpyHostToDevice));
  std::vector<int> a(w->nw, 1);
  if (alive) std::copy(alive, alive + w->nw, a.begin());
  HIP_TRY(hipMemcpy(w->alive.p, a.data(), sizeof(int) * w->nw, hipMemcpyHostToDevice));
  return 0;
}

int afqmc_walkers_download(afqmc_walkers w, double* wf, double* logw, int* alive) {
  NOT_CAPTURING();
  REQUIRE(w, "null argument");
  afqmc_cp cp = w->cp;
  ScopedDevice sd(cp->ctx);
  HIP_TRY(hipStreamSynchronize(cp->ctx->stream));
  const int L2 = 2 * cp->L, N = cp->N;
  const long long per = (long long)L2 * N;
  if (wf) {
    std::fill(wf, wf + 2 * per * w->nw, 0.0);
    for (int b = 0; b < cp->nblk; ++b) {
      const long long ws = w->wsize(b);
      std::vector<zd> h(ws * w->nw);
      if (cp->fp32) {
        std::vector<zf> f(ws * w->nw);
        HIP_TRY(hipMemcpy(f.data(), w->wf[b].p, sizeof(zf) * f.size(), hipMemcpyDeviceToHost));
        for (size_t k = 0; k < f.size(); ++k) h[k] = cmk<double>(f[k].re, f[k].im);
      } else {
        HIP_TRY(hipMemcpy(h.data(), w->wf[b].p, sizeof(zd) * h.size(), hipMemcpyDeviceToHost));
      }
      for (int iw = 0; iw < w->nw; ++iw)
        for (int j = 0; j < cp->Nb[b]; ++j)
          for (int i = 0; i < cp->Mb[b]; ++i) {
            zd v = h[iw * ws + i + (long long)j * cp->Mb[b]];
            double* dst = wf + 2 * (iw * per + (cp->row0[b] + i) + (long long)(cp->col0[b] + j) * L2);
            dst[0] = v.re;
            dst[1] = v.im;
          }
    }
  }
  if (logw) HIP_TRY(hipMemcpy(logw, w->logw.p, sizeof(zd) * w->nw, hipMemcpyDeviceToHost));
  if (alive) HIP_TRY(hipMemcpy(alive, w->alive.p, sizeof(int) * w->nw, hipMemcpyDeviceToHost));
  return 0;
}

int afqmc_walkers_download_some(afqmc_walkers w, int n, const int* index, double* wf, double* logw, int* alive) {
  NOT_CAPTURING();
  REQUIRE(w && (index || n == 0) && n >= 0, "null argument");
  for (int k = 0; k < n; ++k)
    REQUIRE(index[k] >= 0 && index[k] < w->nw, "walker index %d out of range (%d walkers)", index[k], w->nw);
  afqmc_cp cp = w->cp;
  ScopedDevice sd(cp->ctx);
  HIP_TRY(hipStreamSynchronize(cp->ctx->stream));
  const int L2 = 2 * cp->L, N = cp->N;
  const long long per = (long long)L2 * N;
  if (wf) {
    std::fill(wf, wf + 2 * per * n, 0.0);
    for (int b = 0; b < cp->nblk; ++b) {
      const long long ws = w->wsize(b);
      std::vector<char> raw(cp->esz() * ws);
      for (int k = 0; k < n; ++k) {
        HIP_TRY(hipMemcpy(raw.data(), w->wf[b].p + cp->esz() * ws * index[k], raw.size(), hipMemcpyDeviceToHost));
        for (int j = 0; j < cp->Nb[b]; ++j)
          for (int i = 0; i < cp->Mb[b]; ++i) {
            const long long e = i + (long long)j * cp->Mb[b];
            double re, im;
            if (cp->fp32) {
              const zf v = reinterpret_cast<const zf*>(raw.data())[e];
              re = v.re; im = v.im;
            } else {
              const zd v = reinterpret_cast<const zd*>(raw.data())[e];
              re = v.re; im = v.im;
            }
            double* dst = wf + 2 * (k * per + (cp->row0[b] + i) + (long long)(cp->col0[b] + j) * L2);
            dst[0] = re;
            dst[1] = im;
          }
      }
    }
  }
  if (logw || alive) {
    std::vector<zd> lw(w->nw);
    std::vector<int> al(w->nw);
    HIP_TRY(hipMemcpy(lw.data(), w->logw.p, sizeof(zd) * w->nw, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(al.data(), w->alive.p, sizeof(int) * w->nw, hipMemcpyDeviceToHost));
    for (int k = 0; k < n; ++k) {
      if (logw) { logw[2 * k] = lw[index[k]].re; logw[2 * k + 1] = lw[index[k]].im; }
      if (alive) alive[k] = al[index[k]];
    }
  }
  return 0;
}

int afqmc_walkers_set_alive(afqmc_walkers w, const int* alive) {
  NOT_CAPTURING();
  REQUIRE(w && alive, "null argument");
  ScopedDevice sd(w->cp->ctx);
  HIP_TRY(hipStreamSynchronize(w->cp->ctx->stream));
  std::vector<int> a(alive, alive + w->nw);
  for (int& x : a) x = x ? 1 : 0;
  HIP_TRY(hipMemcpy(w->alive.p, a.data(), sizeof(int) * w->nw, hipMemcpyHostToDevice));
  return 0;
}

int afqmc_cp_walkers_from_trial(afqmc_cp cp, afqmc_walkers w, const double* logw) {
  NOT_CAPTURING();
  REQUIRE(cp && w && w->cp == cp, "invalid argument");
  ScopedDevice sd(cp->ctx);
  hipStream_t s = cp->ctx->stream;
  HIP_TRY(hipStreamSynchronize(s));
  // phiT into walker 0, then doubling device copies [0, k) -> [k, 2k)
  for (int b = 0; b < cp->nblk; ++b) {
    const size_t bytes = cp->esz() * (size_t)w->wsize(b);
    HIP_TRY(hipMemcpyAsync(w->wf[b].p, cp->phiT[b].p, bytes, hipMemcpyDeviceToDevice, s));
    for (int k = 1; k < w->nw; k *= 2) {
      const int m = std::min(k, w->nw - k);
      HIP_TRY(hipMemcpyAsync(w->wf[b].p + bytes * k, w->wf[b].p, bytes * m, hipMemcpyDeviceToDevice, s));
    }
  }
  const zd lw = cmk<double>(logw ? logw[0] : 0.0, logw ? logw[1] : 0.0);
  std::vector<zd> h(w->nw, lw);
  std::vector<int> a(w->nw, 1);
  HIP_TRY(hipMemcpyAsync(w->logw.p, h.data(), sizeof(zd) * w->nw, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(w->alive.p, a.data(), sizeof(int) * w->nw, hipMemcpyHostToDevice, s));
  HIP_TRY(hipStreamSynchronize(s));
  return 0;
}

int afqmc_cp_set_seed(afqmc_cp cp, uint64_t seed, uint64_t counter) {
  NOT_CAPTURING();
  REQUIRE(cp, "null argument");
  ScopedDevice sd(cp->ctx);
  HIP_TRY(hipStreamSynchronize(cp->ctx->stream));
  DevScalars h;
  HIP_TRY(hipMemcpy(&h, cp->sc.p, sizeof(h), hipMemcpyDeviceToHost));
  h.key = afqmc_seed_mix_host(seed);
  h.counter = counter;
  HIP_TRY(hipMemcpy(cp->sc.p, &h, sizeof(h), hipMemcpyHostToDevice));
  return 0;
}

int afqmc_cp_get_counter(afqmc_cp cp, uint64_t* counter) {
  NOT_CAPTURING();
  REQUIRE(cp && counter, "null argument");
  ScopedDevice sd(cp->ctx);
  HIP_TRY(hipStreamSynchronize(cp->ctx->stream));
  DevScalars h;
  HIP_TRY(hipMemcpy(&h, cp->sc.p, sizeof(h), hipMemcpyDeviceToHost));
  *counter = h.counter;
  REQUIRE(!h.uniform_overrun, "the host-provided uniforms were exhausted (%llu given)", h.ulen);
  return 0;
}

int afqmc_cp_get_walker_steps(afqmc_cp cp, uint64_t* walker_steps) {
  NOT_CAPTURING();
  REQUIRE(cp && walker_steps, "null argument");
  ScopedDevice sd(cp->ctx);
  HIP_TRY(hipStreamSynchronize(cp->ctx->stream));
  DevScalars h;
  HIP_TRY(hipMemcpy(&h, cp->sc.p, sizeof(h), hipMemcpyDeviceToHost));
  *walker_steps = h.walker_steps;
  return 0;
}

int afqmc_cp_get_qr_stats(afqmc_cp cp, uint64_t* shifted, uint64_t* killed) {
  NOT_CAPTURING();
  REQUIRE(cp && shifted && killed, "null argument");
  ScopedDevice sd(cp->ctx);
  HIP_TRY(hipStreamSynchronize(cp->ctx->stream));
  DevScalars h;
  HIP_TRY(hipMemcpy(&h, cp->sc.p, sizeof(h), hipMemcpyDeviceToHost));
  *shifted = h.qr_shifted;
  *killed = h.qr_killed;
  return 0;
}

int afqmc_cp_set_uniforms(afqmc_cp cp, const double* u, int64_t n) {
  NOT_CAPTURING();
  REQUIRE(cp && n >= 0, "invalid argument");
  ScopedDevice sd(cp->ctx);
  HIP_TRY(hipStreamSynchronize(cp->ctx->stream));
  cp->ubuf.release();
  if (n > 0) {
    REQUIRE(u, "null uniforms");
    int rc = cp->ubuf.alloc((size_t)n);
    if (rc) return rc;
    HIP_TRY(hipMemcpy(cp->ubuf.p, u, sizeof(double) * n, hipMemcpyHostToDevice));
  }
  // host-provided uniforms are indexed from 0
  DevScalars h;
  HIP_TRY(hipMemcpy(&h, cp->sc.p, sizeof(h), hipMemcpyDeviceToHost));
  h.counter = 0;
  h.ulen = (unsigned long long)n;
  h.uniform_overrun = 0;
  HIP_TRY(hipMemcpy(cp->sc.p, &h, sizeof(h), hipMemcpyHostToDevice));
  return 0;
}

int afqmc_cp_skip_uniforms(afqmc_cp cp, uint64_t n) {
  NOT_CAPTURING();
  REQUIRE(cp, "null argument");
  ScopedDevice sd(cp->ctx);
  HIP_TRY(hipStreamSynchronize(cp->ctx->stream));
  unsigned long long c = 0;
  HIP_TRY(hipMemcpy(&c, (char*)cp->sc.p + offsetof(DevScalars, counter), sizeof(c), hipMemcpyDeviceToHost));
  c += n;
  HIP_TRY(hipMemcpy((char*)cp->sc.p + offsetof(DevScalars, counter), &c, sizeof(c), hipMemcpyHostToDevice));
  return 0;
}

int afqmc_cp_set_ET(afqmc_cp cp, double ET) {
  NOT_CAPTURING();
  REQUIRE(cp, "null argument");
  ScopedDevice sd(cp->ctx);
  HIP_TRY(hipStreamSynchronize(cp->ctx->stream));
  HIP_TRY(hipMemcpy((char*)cp->sc.p + offsetof(DevScalars, ET), &ET, sizeof(double), hipMemcpyHostToDevice));
  return 0;
}

int afqmc_cp_get_ET(afqmc_cp cp, double* ET) {
  NOT_CAPTURING();
  REQUIRE(cp && ET, "null argument");
  ScopedDevice sd(cp->ctx);
  HIP_TRY(hipStreamSynchronize(cp->ctx->stream));
  HIP_TRY(hipMemcpy(ET, (char*)cp->sc.p + offsetof(DevScalars, ET), sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

int afqmc_cp_log_overlap(afqmc_cp cp, afqmc_walkers w, double* logO) {
  NOT_CAPTURING();
  REQUIRE(cp && w && logO && w->cp == cp, "invalid argument");
  ScopedDevice sd(cp->ctx);
  int rc = cp->fp32 ? do_overlap<float>(cp, w, false) : do_overlap<double>(cp, w, false);
  if (rc) return rc;
  LogOArgs l = cp->fp32 ? logo_args<float>(cp, w, LOGO_ONLY) : logo_args<double>(cp, w, LOGO_ONLY);
  if ((rc = launch_logo(l, cp->ctx->stream))) return rc;
  HIP_TRY(hipMemcpyAsync(logO, cp->logO.p, sizeof(zd) * w->nw, hipMemcpyDeviceToHost, cp->ctx->stream));
  HIP_TRY(hipStreamSynchronize(cp->ctx->stream));
  return 0;
}

int afqmc_cp_green_diagonal(afqmc_cp cp, afqmc_walkers w, double* gdiag, double* goff) {
  NOT_CAPTURING();
  REQUIRE(cp && w && gdiag && w->cp == cp, "invalid argument");
  ScopedDevice sd(cp->ctx);
  int rc = cp->fp32 ? do_overlap<float>(cp, w, true) : do_overlap<double>(cp, w, true);
  if (!rc) rc = cp->fp32 ? green_gemm<float>(cp, w, w->wf, w->alive.p, false)
                         : green_gemm<double>(cp, w, w->wf, w->alive.p, false);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(cp->ctx->stream));
  const int L = cp->L, L2 = 2 * L;
  for (int b = 0; b < cp->nblk; ++b) {
    const size_t n = (size_t)w->nw * cp->ntile[b] * cp->Mb[b];
    std::vector<zd> g0(n), g1;
    HIP_TRY(hipMemcpy(g0.data(), cp->gpart[0][b].p, sizeof(zd) * n, hipMemcpyDeviceToHost));
    if (cp->dense) {
      g1.resize(n);
      HIP_TRY(hipMemcpy(g1.data(), cp->gpart[1][b].p, sizeof(zd) * n, hipMemcpyDeviceToHost));
    }
    for (int iw = 0; iw < w->nw; ++iw)
      for (int r = 0; r < cp->Mb[b]; ++r) {
        zd s0 = cmk<double>(0, 0), s1 = s0;
        for (int t = 0; t < cp->ntile[b]; ++t) {
          size_t k = ((size_t)iw * cp->ntile[b] + t) * cp->Mb[b] + r;
          s0 = cadd(s0, g0[k]);
          if (cp->dense) s1 = cadd(s1, g1[k]);
        }
        const int row = cp->row0[b] + r;
        gdiag[2 * ((size_t)iw * L2 + row)] = s0.re;
        gdiag[2 * ((size_t)iw * L2 + row) + 1] = s0.im;
        if (goff) {
          // X row r pairs with trial row r^: G(r^, r) = greenOffDiagonal(r^)
          const int dst = cp->dense ? (row < L ? row + L : row - L) : row;
          goff[2 * ((size_t)iw * L2 + dst)] = cp->dense ? s1.re : 0.0;
          goff[2 * ((size_t)iw * L2 + dst) + 1] = cp->dense ? s1.im : 0.0;
        }
      }
  }
  return 0;
}

int afqmc_cp_force(afqmc_cp cp, afqmc_walkers w, double* force) {
  NOT_CAPTURING();
  REQUIRE(cp && w && force && w->cp == cp, "invalid argument");
  const int L = cp->L, L2 = 2 * L;
  std::vector<double> gd((size_t)w->nw * L2 * 2), go((size_t)w->nw * L2 * 2);
  int rc = afqmc_cp_green_diagonal(cp, w, gd.data(), go.data());
  if (rc) return rc;
  for (int iw = 0; iw < w->nw; ++iw)
    for (int i = 0; i < L; ++i) {
      auto G = [&](const std::vector<double>& v, int r) { return cd(v[2 * ((size_t)iw * L2 + r)], v[2 * ((size_t)iw * L2 + r) + 1]); };
      cd bg;
      switch (cp->decomp) {
        case 0: bg = G(gd, i) + G(gd, i + L) - 1.0; break;
        case 1: bg = G(gd, i) - G(gd, i + L); break;
        case 2: bg = G(go, i) + G(go, i + L); break;
        default: bg = G(go, i) - G(go, i + L); break;
      }
      double f = (cp->h_gamma[i] * bg).real();
      f = std::min(cp->cap, std::max(-cp->cap, f));
      force[(size_t)iw * L + i] = f;
    }
  return 0;
}

int afqmc_cp_onebody(afqmc_cp cp, afqmc_walkers w, int which) {
  REQUIRE(cp && w && w->cp == cp, "invalid argument");
  REQUIRE(which >= 0 && which <= 2, "which must be 0, 1 or 2");
  ScopedDevice sd(cp->ctx);
  int rc = cp->fp32 ? propagate_gemm<float>(cp, w, which, w->wf, w->tmp, false)
                    : propagate_gemm<double>(cp, w, which, w->wf, w->tmp, false);
  if (rc) return rc;
  // alive walkers take the product, killed ones keep their state (the reference skips them)
  for (int b = 0; b < cp->nblk; ++b) {
    if ((rc = launch_select_alive(w->nw, w->wsize(b), cp->esz(), w->alive.p, w->tmp[b].p, w->wf[b].p, cp->ctx->stream)))
      return rc;
  }
  if ((rc = launch_logw_const_add(w->nw, w->logw.p, w->alive.p, cp->hop_logw[which], cp->ctx->stream))) return rc;
  HIP_TRY(sync_unless_capturing(cp->ctx->stream));
  return 0;
}

int afqmc_cp_hs_apply(afqmc_cp cp, afqmc_walkers w, const double* diag4, const double* sample_logw) {
  NOT_CAPTURING();
  REQUIRE(cp && w && diag4 && sample_logw && w->cp == cp, "invalid argument");
  REQUIRE(cp->dense, "afqmc_cp_hs_apply needs a dense (spin_blocks = 0) set");
  ScopedDevice sd(cp->ctx);
  hipStream_t s = cp->ctx->stream;
  const int L = cp->L;
  const size_t n = (size_t)4 * L * w->nw;
  int rc = 0;
  DevBuf<char> d;
  if ((rc = d.alloc_bytes(cp->esz() * n))) return rc;
  if (cp->fp32) {
    std::vector<zf> h(n);
    for (size_t k = 0; k < n; ++k) h[k] = cmk<float>((float)diag4[2 * k], (float)diag4[2 * k + 1]);
    HIP_TRY(hipMemcpy(d.p, h.data(), sizeof(zf) * n, hipMemcpyHostToDevice));
    RegionTimer t(cp->ctx, REG_HS_APPLY);
    rc = launch_hs_apply_2x2<float>(w->nw, L, cp->N, w->wf[0].p, w->tmp[0].p, d.p, s);
  } else {
    HIP_TRY(hipMemcpy(d.p, diag4, sizeof(zd) * n, hipMemcpyHostToDevice));
    RegionTimer t(cp->ctx, REG_HS_APPLY);
    rc = launch_hs_apply_2x2<double>(w->nw, L, cp->N, w->wf[0].p, w->tmp[0].p, d.p, s);
  }
  if (rc) return rc;
  if ((rc = launch_select_alive(w->nw, w->wsize(0), cp->esz(), w->alive.p, w->tmp[0].p, w->wf[0].p, s))) return rc;
  DevBuf<zd> lw;
  if ((rc = lw.alloc(w->nw))) return rc;
  HIP_TRY(hipMemcpy(lw.p, sample_logw, sizeof(zd) * w->nw, hipMemcpyHostToDevice));
  if ((rc = launch_logw_vec_add(w->nw, w->logw.p, w->alive.p, lw.p, s))) return rc;
  HIP_TRY(hipStreamSynchronize(s));
  return 0;
}

int afqmc_cp_stabilize(afqmc_cp cp, afqmc_walkers w) {
  REQUIRE(cp && w && w->cp == cp, "invalid argument");
  ScopedDevice sd(cp->ctx);
  int rc = stabilize_impl(cp, w);
  if (rc) return rc;
  HIP_TRY(sync_unless_capturing(cp->ctx->stream));
  return 0;
}

int afqmc_comb(afqmc_ctx ctx, int nw, const double* weight, const double* uniforms, int* table, double* stats) {
  NOT_CAPTURING();
  REQUIRE(ctx && weight && uniforms && table && nw > 0, "invalid argument");
  ScopedDevice sd(ctx);
  hipStream_t s = ctx->stream;
  DevBuf<double> w, u, prob;
  DevBuf<int> t, num;
  DevBuf<DevScalars> sc;
  int rc = w.alloc(nw);
  if (!rc) rc = u.alloc(nw);
  if (!rc) rc = prob.alloc(nw);
  if (!rc) rc = t.alloc(nw);
  if (!rc) rc = num.alloc(nw);
  if (!rc) rc = sc.alloc(1);
  if (rc) return rc;
  DevScalars h;
  memset(&h, 0, sizeof(h));
  h.ulen = (unsigned long long)nw;
  HIP_TRY(hipMemcpy(sc.p, &h, sizeof(h), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(w.p, weight, sizeof(double) * nw, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(u.p, uniforms, sizeof(double) * nw, hipMemcpyHostToDevice));
  if ((rc = launch_comb(nw, w.p, u.p, sc.p, t.p, prob.p, num.p, s))) return rc;
  HIP_TRY(hipMemcpyAsync(table, t.p, sizeof(int) * nw, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(&h, sc.p, sizeof(h), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  if (stats) { stats[0] = h.wmin; stats[1] = h.wmax; stats[2] = h.wavg; }
  return 0;
}

int afqmc_cp_apply_table(afqmc_cp cp, afqmc_walkers w, const int* table) {
  NOT_CAPTURING();
  REQUIRE(cp && w && table && w->cp == cp, "invalid argument");
  for (int i = 0; i < w->nw; ++i) REQUIRE(table[i] >= 0 && table[i] < w->nw, "table entry %d out of range", i);
  ScopedDevice sd(cp->ctx);
  hipStream_t s = cp->ctx->stream;
  HIP_TRY(hipMemcpy(cp->table.p, table, sizeof(int) * w->nw, hipMemcpyHostToDevice));
  void* wfp[2] = {w->wf[0].p, cp->nblk > 1 ? w->wf[1].p : nullptr};
  long long ws[2] = {w->wsize(0), cp->nblk > 1 ? w->wsize(1) : 0};
  int rc = cp->fp32 ? launch_apply_table<float>(w->nw, cp->nblk, wfp, ws, w->logw.p, w->alive.p, cp->table.p, s)
                    : launch_apply_table<double>(w->nw, cp->nblk, wfp, ws, w->logw.p, w->alive.p, cp->table.p, s);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(s));
  return 0;
}

int afqmc_cp_step(afqmc_cp cp, afqmc_walkers w, double* force_out, int* aux_out) {
  REQUIRE(cp && w && w->cp == cp, "invalid argument");
  ScopedDevice sd(cp->ctx);
  return DISPATCH(cp, step_impl, cp, w, force_out, aux_out);
}

int afqmc_cp_pop_control(afqmc_cp cp, afqmc_walkers w, double* weights_out, int* table_out, double* stats) {
  REQUIRE(cp && w && w->cp == cp, "invalid argument");
  REQUIRE(!capturing() || (!weights_out && !table_out && !stats), "a captured comb cannot return outputs");
  ScopedDevice sd(cp->ctx);
  return DISPATCH(cp, pop_control_impl, cp, w, weights_out, table_out, stats);
}

int afqmc_cp_step_fields(afqmc_cp cp, afqmc_walkers w, const int* aux) {
  NOT_CAPTURING();
  REQUIRE(cp && w && w->cp == cp && aux, "invalid argument");
  const long long n = (long long)w->nw * cp->L;
  ScopedDevice sd(cp->ctx);
  hipStream_t s = cp->ctx->stream;
  std::vector<int> alive(w->nw);
  HIP_TRY(hipStreamSynchronize(s));
  HIP_TRY(hipMemcpy(alive.data(), w->alive.p, sizeof(int) * w->nw, hipMemcpyDeviceToHost));
  for (int iw = 0; iw < w->nw; ++iw) {
    if (!alive[iw]) continue;
    for (int i = 0; i < cp->L; ++i) {
      const int x = aux[(long long)iw * cp->L + i];
      REQUIRE(x == 1 || x == -1, "field of walker %d, site %d is %d, not +1 or -1", iw, i, x);
    }
  }
  if (cp->aux_in.n < (size_t)n) {
    int rc = cp->aux_in.alloc((size_t)n);
    if (rc) return rc;
  }
  HIP_TRY(hipMemcpyAsync(cp->aux_in.p, aux, sizeof(int) * n, hipMemcpyHostToDevice, s));
  int rc = DISPATCH(cp, step_impl, cp, w, nullptr, nullptr, nullptr, cp->aux_in.p);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(s));
  return 0;
}

int afqmc_cp_pop_control_table(afqmc_cp cp, afqmc_walkers w, const int* table, double* weights_out) {
  NOT_CAPTURING();
  REQUIRE(cp && w && w->cp == cp && table, "invalid argument");
  for (int i = 0; i < w->nw; ++i) {
    REQUIRE(table[i] >= 0 && table[i] < w->nw, "table entry %d out of range", i);
    REQUIRE(table[table[i]] == table[i], "table entry %d: parent %d does not keep its own slot", i, table[i]);
  }
  ScopedDevice sd(cp->ctx);
  int rc = DISPATCH(cp, pop_control_impl, cp, w, weights_out, nullptr, nullptr, table);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(cp->ctx->stream));
  return 0;
}

int afqmc_cp_measure(afqmc_cp cp, afqmc_walkers w) {
  REQUIRE(cp && w && w->cp == cp, "invalid argument");
  ScopedDevice sd(cp->ctx);
  return DISPATCH(cp, measure_impl, cp, w);
}

int afqmc_cp_energy(afqmc_cp cp, double* out) {
  NOT_CAPTURING();
  REQUIRE(cp && out, "null argument");
  ScopedDevice sd(cp->ctx);
  HIP_TRY(hipStreamSynchronize(cp->ctx->stream));
  DevScalars h;
  HIP_TRY(hipMemcpy(&h, cp->sc.p, sizeof(h), hipMemcpyDeviceToHost));
  zd v[7] = {h.den, h.HNum, h.KNum, h.VNum, h.RNum, h.NupTot, h.NdnTot};
  for (int k = 0; k < 7; ++k) { out[2 * k] = v[k].re; out[2 * k + 1] = v[k].im; }
  return 0;
}

int afqmc_cp_site_measure(afqmc_cp cp, double* out) {
  NOT_CAPTURING();
  REQUIRE(cp && out, "null argument");
  ScopedDevice sd(cp->ctx);
  HIP_TRY(hipStreamSynchronize(cp->ctx->stream));
  HIP_TRY(hipMemcpy(out, cp->site_acc.p, sizeof(zd) * 4 * cp->L, hipMemcpyDeviceToHost));
  return 0;
}

int afqmc_cp_reset_measure(afqmc_cp cp) {
  REQUIRE(cp, "null argument");
  ScopedDevice sd(cp->ctx);
  int rc = launch_reset_measure(cp->sc.p, cp->site_acc.p, cp->L, cp->ctx->stream);
  if (rc) return rc;
  return 0;
}

int afqmc_cp_adjust_ET(afqmc_cp cp) {
  REQUIRE(cp, "null argument");
  ScopedDevice sd(cp->ctx);
  int rc = launch_set_et_from_energy(cp->sc.p, cp->ctx->stream);
  if (rc) return rc;
  return launch_reset_measure(cp->sc.p, cp->site_acc.p, cp->L, cp->ctx->stream);
}

int afqmc_cp_set_tiles(afqmc_cp cp, const int* tiles) {
  REQUIRE(cp && tiles, "null argument");
  for (int r = 0; r < 3; ++r) REQUIRE(tiles[r] >= -1 && tiles[r] < GEMM_TILE_COUNT, "tile %d out of range", tiles[r]);
  REQUIRE(tiles[2] < 0 || gemm_tile_green(tiles[2]), "the Green GEMM takes tile shape 0, 1, 4, 5, 7, 8 or 9");
  for (int r = 0; r < 3; ++r) cp->tile[r] = tiles[r];
  return 0;
}


int afqmc_cp_autotune(afqmc_cp cp, afqmc_walkers w, int reps, int* tiles, double* ms) {
  REQUIRE(cp && w && w->cp == cp && reps > 0, "invalid argument");
  NOT_CAPTURING();
  ScopedDevice sd(cp->ctx);
  double t[3][GEMM_TILE_COUNT];
  int rc = 0;
  for (int role = 0; role < 3 && !rc; ++role)
    rc = cp->fp32 ? tune_role<float>(cp, w, role, reps, t[role]) : tune_role<double>(cp, w, role, reps, t[role]);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(cp->ctx->stream));
  if (tiles)
    for (int r = 0; r < 3; ++r) tiles[r] = cp->tile[r];
  if (ms)
    for (int r = 0; r < 3; ++r)
      for (int k = 0; k < GEMM_TILE_COUNT; ++k) ms[r * GEMM_TILE_COUNT + k] = t[r][k];
  return 0;
}

}  // extern "C"

// ------------------------------------------------------------------ groups of runs
struct afqmc_cp_group_s {
  afqmc_ctx ctx = nullptr;
  std::vector<afqmc_cp> cps;
  std::vector<afqmc_walkers> ws;
  int fp32 = 0;
  int tile[3] = {GEMM_TILE_64x64, GEMM_TILE_64x64, GEMM_TILE_64x64};
  GroupTables tables;
};

namespace {

// per_member(cp, w) for every member with the launches recorded, then launched grouped.  The
// members' own tile shapes give way to the group's for the call, so each GEMM phase is one launch.
template <class F>
int run_grouped(afqmc_cp_group g, F&& per_member) {
  REQUIRE(!grouping(), "grouped calls do not nest");
  ScopedDevice sd(g->ctx);
  GroupCollector c;
  c.tables = &g->tables;
  const size_t n = g->cps.size();
  std::vector<std::array<int, 3>> saved(n);
  for (size_t i = 0; i < n; ++i)
    for (int r = 0; r < 3; ++r) {
      saved[i][r] = g->cps[i]->tile[r];
      if (g->tile[r] >= 0) g->cps[i]->tile[r] = g->tile[r];  // -1: the member's own shape
    }
  g_group = &c;
  int rc = 0;
  for (size_t i = 0; i < n && !rc; ++i) {
    c.member = (int)i;
    c.pos = 0;
    rc = per_member(g->cps[i], g->ws[i]);
  }
  g_group = nullptr;
  for (size_t i = 0; i < n; ++i)
    for (int r = 0; r < 3; ++r) g->cps[i]->tile[r] = saved[i][r];
  if (rc) return rc;
  return group_flush_all(c, g->ctx->stream);
}

}  // namespace

extern "C" {

int afqmc_cp_group_create(int n, const afqmc_cp* cps, const afqmc_walkers* walkers, afqmc_cp_group* out) {
  REQUIRE(n >= 1 && cps && walkers && out, "invalid argument");
  for (int i = 0; i < n; ++i) {
    REQUIRE(cps[i] && walkers[i] && walkers[i]->cp == cps[i], "member %d: a propagator and one of its walker sets", i);
    REQUIRE(cps[i]->ctx == cps[0]->ctx, "member %d: every member on the first member's context", i);
    REQUIRE(cps[i]->fp32 == cps[0]->fp32, "member %d: every member in the same precision", i);
    REQUIRE(!cps[i]->split_bf16, "member %d: the split-BF16 propagation has no grouped form", i);
    for (int j = 0; j < i; ++j) REQUIRE(cps[j] != cps[i], "members %d and %d share a propagator", j, i);
  }
  afqmc_cp_group g = new afqmc_cp_group_s();
  g->ctx = cps[0]->ctx;
  g->cps.assign(cps, cps + n);
  g->ws.assign(walkers, walkers + n);
  g->fp32 = cps[0]->fp32;
  *out = g;
  return 0;
}

int afqmc_cp_group_destroy(afqmc_cp_group g) {
  if (!g) return 0;
  ScopedDevice sd(g->ctx);
  (void)hipStreamSynchronize(g->ctx->stream);
  delete g;
  return 0;
}

int afqmc_cp_group_set_tiles(afqmc_cp_group g, const int* tiles) {
  REQUIRE(g && tiles, "null argument");
  for (int r = 0; r < 3; ++r) REQUIRE(tiles[r] >= -1 && tiles[r] < GEMM_TILE_COUNT, "tile %d out of range", tiles[r]);
  REQUIRE(tiles[2] < 0 || gemm_tile_green(tiles[2]), "the Green GEMM takes tile shape 0, 1, 4, 5, 7, 8 or 9");
  for (int r = 0; r < 3; ++r) g->tile[r] = tiles[r];
  return 0;
}

int afqmc_cp_group_step(afqmc_cp_group g) {
  REQUIRE(g, "null argument");
  return run_grouped(g, [&](afqmc_cp cp, afqmc_walkers w) {
    return g->fp32 ? step_impl<float>(cp, w, nullptr, nullptr) : step_impl<double>(cp, w, nullptr, nullptr);
  });
}

int afqmc_cp_group_stabilize(afqmc_cp_group g) {
  REQUIRE(g, "null argument");
  return run_grouped(g, [&](afqmc_cp cp, afqmc_walkers w) { return stabilize_impl(cp, w); });
}

int afqmc_cp_group_pop_control(afqmc_cp_group g) {
  REQUIRE(g, "null argument");
  return run_grouped(g, [&](afqmc_cp cp, afqmc_walkers w) {
    return g->fp32 ? pop_control_impl<float>(cp, w, nullptr, nullptr, nullptr)
                   : pop_control_impl<double>(cp, w, nullptr, nullptr, nullptr);
  });
}

int afqmc_cp_group_table_count(afqmc_cp_group g) { return g ? (int)g->tables.count() : -1; }

}  // extern "C"
