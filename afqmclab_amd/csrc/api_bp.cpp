// Back-propagation of the constraint-path run (tutorials/afqmcConstraintPathBP): C entry points.
//
// The reference keeps, per walker, the aux fields of the last backPropagationStep steps and the comb
// tables of those steps, moves them with every table when it measures, and rebuilds the next
// window's backup walkers by propagating the old backups again
// (afqmcConstraintPathBackPropagation.cpp:22-97).  Here the fields of every step go into a device
// ring indexed by (step mod window, slot), the tables into a ring beside it, and the walkers are
// copied at every window start (one copy per measureSkipStep steps that a window spans, HBM being
// plentiful).  A measurement follows each walker's ancestry backwards through the tables
// (anc = table[anc], the slot a walker was copied from) while it propagates the bras, so no record is
// ever moved and no walker is propagated twice.  Window w starts at step w * measureSkipStep and is
// measured at step w * measureSkipStep + backPropagationStep, the schedule of the reference's driver
// in both of its cases (backPropagationStep >= measureSkipStep and <).
#include <algorithm>
#include <vector>

#include "cp_internal.hpp"

using namespace cpi;

struct afqmc_bp_s {
  afqmc_cp cp = nullptr;
  int device = 0;
  int nw = 0, nbp = 0, skip = 1, mgs = 1, nsnap = 1, chunk = 1, slices = 1;
  long long t = -1;          // steps recorded since afqmc_bp_begin (-1: not begun)
  long long measurements = 0;
  DevBuf<int> aux;           // [max(nbp, 1)][nw][L]
  DevBuf<int> tables;        // [max(nbp, 1)][nw]
  std::vector<char> has_table;
  std::vector<long long> snap_time;
  std::vector<DevBuf<char>> snap;  // [nsnap * nblk]
  DevBuf<int> anc;
  afqmc_walkers bra = nullptr, ket = nullptr;
  DevBuf<char> gt[2];
  DevBuf<zd> part, acc;      // acc: den, green (4L^2), dd (4L^2), pm, mp, sp (L^2 each)
  size_t L2sq() const { return 4 * (size_t)cp->L * cp->L; }
  size_t Lsq() const { return (size_t)cp->L * cp->L; }
  zd* acc_den() { return acc.p; }
  zd* acc_green() { return acc.p + 1; }
  zd* acc_dd() { return acc.p + 1 + L2sq(); }
  zd* acc_pm() { return acc.p + 1 + 2 * L2sq(); }
  zd* acc_mp() { return acc.p + 1 + 2 * L2sq() + Lsq(); }
  zd* acc_sp() { return acc.p + 1 + 2 * L2sq() + 2 * Lsq(); }
  size_t acc_count() const { return 1 + 2 * L2sq() + 3 * Lsq(); }
  ~afqmc_bp_s() {
    if (bra) afqmc_walkers_destroy(bra);
    if (ket) afqmc_walkers_destroy(ket);
  }
};

namespace {

int snapshot_if_due(afqmc_bp bp, afqmc_walkers w) {
  afqmc_cp cp = bp->cp;
  if (bp->t % bp->skip != 0) return 0;
  const int slot = (int)((bp->t / bp->skip) % bp->nsnap);
  if (bp->snap_time[slot] == bp->t) return 0;
  for (int b = 0; b < cp->nblk; ++b)
    HIP_TRY(hipMemcpyAsync(bp->snap[slot * cp->nblk + b].p, w->wf[b].p, cp->esz() * w->wsize(b) * w->nw,
                           hipMemcpyDeviceToDevice, cp->ctx->stream));
  bp->snap_time[slot] = bp->t;
  return 0;
}

template <typename T>
int bra_step(afqmc_bp bp, afqmc_walkers bra, const int* aux, const int* table) {
  afqmc_cp cp = bp->cp;
  hipStream_t s = cp->ctx->stream;
  BpFieldArgs f;
  memset(&f, 0, sizeof(f));
  f.nw = bp->nw; f.L = cp->L; f.decomp = cp->decomp; f.dense = cp->dense; f.diag_only = cp->diag_only;
  f.aux = aux; f.table = table; f.anc = bp->anc.p; f.gamma = cp->gamma.p; f.cdiag = cp->cdiag.p;
  for (int b = 0; b < cp->nblk; ++b) { f.scale[b] = cp->scale[b].p; f.Mb[b] = cp->Mb[b]; }
  f.d2x2 = cp->d2x2.p; f.fp32 = cp->fp32;
  int rc = launch_bp_field(f, s);
  if (rc) return rc;
  if (cp->diag_only) {
    if ((rc = propagate_gemm<T>(cp, bra, 0, bra->wf, bra->tmp, true))) return rc;
  } else {
    if ((rc = launch_hs_apply_2x2<T>(bp->nw, cp->L, cp->N, bra->wf[0].p, bra->tmp[0].p, cp->d2x2.p, s))) return rc;
    std::swap(bra->wf[0], bra->tmp[0]);
    if ((rc = propagate_gemm<T>(cp, bra, 0, bra->wf, bra->tmp, false))) return rc;
  }
  for (int b = 0; b < cp->nblk; ++b) std::swap(bra->wf[b], bra->tmp[b]);
  return 0;
}

// addMeasurement (afqmcConstraintPathMeasure.cpp:41-72) at step t of a window that began at t - nbp
template <typename T>
int measure_impl(afqmc_bp bp, afqmc_walkers w, int snap_slot) {
  afqmc_cp cp = bp->cp;
  hipStream_t s = cp->ctx->stream;
  afqmc_walkers bra = bp->bra, ket = bp->ket;
  const int nw = bp->nw, L = cp->L;
  int rc;
  // weights: the mixed overlaps of the current walkers
  if ((rc = mixed_weight(cp, w))) return rc;
  HIP_TRY(hipMemcpyAsync(bra->alive.p, w->alive.p, sizeof(int) * nw, hipMemcpyDeviceToDevice, s));
  HIP_TRY(hipMemcpyAsync(ket->alive.p, w->alive.p, sizeof(int) * nw, hipMemcpyDeviceToDevice, s));
  // bra = exp(-dt/2 K)^H phiT for every walker
  if ((rc = launch_iota(nw, bp->anc.p, s))) return rc;
  for (int b = 0; b < cp->nblk; ++b)
    if ((rc = launch_gather_walkers(nw, bra->wsize(b), cp->esz(), nullptr, cp->phiT[b].p, bra->tmp[b].p, s)))
      return rc;
  if ((rc = propagate_gemm<T>(cp, bra, 1, bra->tmp, bra->wf, false))) return rc;
  // the recorded steps, last first: D^H then exp(-dt K)^H, normalised every mgs steps
  const long long start = bp->t - bp->nbp;
  for (int j = bp->nbp - 1; j >= 0; --j) {
    const int slot = (int)((start + j) % bp->nbp);
    const int* aux = bp->aux.p + (size_t)slot * nw * L;
    const int* table = bp->has_table[slot] ? bp->tables.p + (size_t)slot * nw : nullptr;
    if ((rc = bra_step<T>(bp, bra, aux, table))) return rc;
    if ((bp->nbp - j) % bp->mgs == 0 && (rc = stabilize(cp, bra))) return rc;
  }
  // bra = exp(dt/2 K)^H bra, ket = exp(dt/2 K) backup[anc]
  if ((rc = propagate_gemm<T>(cp, bra, 2, bra->wf, bra->tmp, false))) return rc;
  for (int b = 0; b < cp->nblk; ++b) {
    std::swap(bra->wf[b], bra->tmp[b]);
    if ((rc = launch_gather_walkers(nw, ket->wsize(b), cp->esz(), bp->anc.p, bp->snap[snap_slot * cp->nblk + b].p,
                                    ket->tmp[b].p, s)))
      return rc;
  }
  if ((rc = propagate_gemm<T>(cp, ket, 2, ket->tmp, ket->wf, false))) return rc;
  // per block: S = bra^H ket, S^{-1}, Theta = ket S^{-1} (into bra->tmp), bra^H (into ket->tmp)
  for (int b = 0; b < cp->nblk; ++b) {
    const int Mb = cp->Mb[b], Nb = cp->Nb[b];
    GemmArgs a = gemm_base<T>();
    a.M = Nb; a.N = Nb; a.K = Mb; a.batch = nw;
    a.A = bra->wf[b].p; a.lda = Mb; a.strideA = bra->wsize(b);
    a.B = ket->wf[b].p; a.ldb = Mb; a.strideB = ket->wsize(b);
    a.C = cp->ovl[b].p; a.ldc = Nb; a.strideC = (long long)Nb * Nb;
    a.alive = w->alive.p;
    if ((rc = launch_cgemm<T>(OP_C, EPI_STORE, a, s))) return rc;
  }
  if ((rc = gj_blocks<T>(cp, w, true, w->alive.p))) return rc;
  for (int b = 0; b < cp->nblk; ++b) {
    const int Mb = cp->Mb[b], Nb = cp->Nb[b];
    GemmArgs a = gemm_base<T>();
    a.M = Mb; a.N = Nb; a.K = Nb; a.batch = nw;
    a.A = ket->wf[b].p; a.lda = Mb; a.strideA = ket->wsize(b);
    a.B = cp->ovl[b].p; a.ldb = Nb; a.strideB = (long long)Nb * Nb;
    a.C = bra->tmp[b].p; a.ldc = Mb; a.strideC = bra->wsize(b);
    a.alive = w->alive.p;
    if ((rc = launch_cgemm<T>(OP_N, EPI_STORE, a, s))) return rc;
    if ((rc = launch_conj_transpose<T>(nw, Mb, Nb, bra->wsize(b), ket->wsize(b), w->alive.p, bra->wf[b].p,
                                       ket->tmp[b].p, s)))
      return rc;
  }
  // G^T = Theta bra^H per chunk of walkers, then the observables
  if ((rc = launch_den_sum(nw, cp->den.p, w->alive.p, bp->acc_den(), s))) return rc;
  for (int c0 = 0; c0 < nw; c0 += bp->chunk) {
    const int cn = std::min(bp->chunk, nw - c0);
    WickArgs k;
    memset(&k, 0, sizeof(k));
    for (int b = 0; b < cp->nblk; ++b) {
      const int Mb = cp->Mb[b], Nb = cp->Nb[b];
      const size_t es = cp->esz();
      GemmArgs a = gemm_base<T>();
      a.M = Mb; a.N = Mb; a.K = Nb; a.batch = cn;
      a.A = bra->tmp[b].p + es * bra->wsize(b) * c0; a.lda = Mb; a.strideA = bra->wsize(b);
      a.B = ket->tmp[b].p + es * ket->wsize(b) * c0; a.ldb = Nb; a.strideB = ket->wsize(b);
      a.C = bp->gt[b].p; a.ldc = Mb; a.strideC = (long long)Mb * Mb;
      a.alive = w->alive.p + c0;
      if ((rc = launch_cgemm<T>(OP_N, EPI_STORE, a, s))) return rc;
      k.gt[b] = bp->gt[b].p; k.gstride[b] = (long long)Mb * Mb; k.ld[b] = Mb;
    }
    k.L = L; k.dense = cp->dense; k.c0 = c0; k.cn = cn; k.slices = bp->slices; k.first = c0 == 0;
    k.den = cp->den.p; k.alive = w->alive.p; k.part = bp->part.p; k.fp32 = cp->fp32;
    if ((rc = launch_wick(k, s))) return rc;
  }
  return launch_wick_reduce(L, bp->slices, bp->part.p, bp->acc_green(), bp->acc_dd(), bp->acc_pm(), bp->acc_mp(),
                            bp->acc_sp(), s);
}

}  // namespace

extern "C" {

int afqmc_bp_create(afqmc_cp cp, int nw, int bp_steps, int measure_skip, int mgs, afqmc_bp* out) {
  NOT_CAPTURING();
  REQUIRE(cp && out, "afqmc_bp_create: null argument");
  REQUIRE(nw > 0 && nw <= cp->maxw, "walker count %d outside 1..max_walkers=%d", nw, cp->maxw);
  REQUIRE(bp_steps >= 0, "backPropagationStep must be non-negative");
  REQUIRE(measure_skip > 0, "measureSkipStep ==0, code will not measure anything!");
  REQUIRE(mgs > 0, "mgsStep must be positive");
  ScopedDevice sd(cp->ctx);
  afqmc_bp bp = new afqmc_bp_s();
  bp->cp = cp;
  bp->device = cp->device;
  bp->nw = nw; bp->nbp = bp_steps; bp->skip = measure_skip; bp->mgs = mgs;
  bp->nsnap = bp_steps / measure_skip + 1;
  const int L = cp->L;
  const size_t ring = (size_t)std::max(bp_steps, 1);
  int rc = bp->aux.alloc(ring * nw * L);
  if (!rc && hipMemset(bp->aux.p, 0, sizeof(int) * ring * nw * L) != hipSuccess) rc = -2;
  if (!rc) rc = bp->tables.alloc(ring * nw);
  bp->has_table.assign(ring, 0);
  bp->snap_time.assign(bp->nsnap, -1);
  bp->snap.resize((size_t)bp->nsnap * cp->nblk);
  long long wbytes = 0;
  for (int b = 0; b < cp->nblk; ++b) wbytes += (long long)cp->Mb[b] * cp->Nb[b];
  for (int k = 0; k < bp->nsnap && !rc; ++k)
    for (int b = 0; b < cp->nblk && !rc; ++b)
      rc = bp->snap[(size_t)k * cp->nblk + b].alloc_bytes(cp->esz() * cp->Mb[b] * cp->Nb[b] * (size_t)nw);
  if (!rc) rc = bp->anc.alloc(nw);
  if (!rc) rc = afqmc_cp_walkers_create(cp, nw, &bp->bra);
  if (!rc) rc = afqmc_cp_walkers_create(cp, nw, &bp->ket);
  // G^T of a chunk of walkers: at most 512 MiB
  size_t gper = 0;
  for (int b = 0; b < cp->nblk; ++b) gper += cp->esz() * cp->Mb[b] * cp->Mb[b];
  bp->chunk = (int)std::max<size_t>(1, std::min<size_t>(nw, (512ull << 20) / gper));
  for (int b = 0; b < cp->nblk && !rc; ++b)
    rc = bp->gt[b].alloc_bytes(cp->esz() * cp->Mb[b] * cp->Mb[b] * (size_t)bp->chunk);
  const int tiles = ((L + 15) / 16) * ((L + 15) / 16);
  bp->slices = std::max(1, std::min(64, (2048 + tiles - 1) / tiles));
  if (!rc) rc = bp->part.alloc((size_t)bp->slices * WICK_Q * L * L);
  if (!rc) rc = bp->acc.alloc(bp->acc_count());
  if (!rc && hipMemset(bp->acc.p, 0, sizeof(zd) * bp->acc_count()) != hipSuccess) rc = -2;
  (void)wbytes;
  if (rc) {
    delete bp;
    if (rc == -2) afqmc_set_error("afqmc_bp_create: device allocation failed");
    return rc;
  }
  *out = bp;
  return 0;
}

int afqmc_bp_destroy(afqmc_bp bp) {
  if (!bp) return 0;
  ScopedDevice sd(bp->device);
  hipDeviceSynchronize();
  delete bp;
  return 0;
}

int afqmc_bp_begin(afqmc_bp bp) {
  NOT_CAPTURING();
  REQUIRE(bp, "null argument");
  bp->t = 0;
  std::fill(bp->has_table.begin(), bp->has_table.end(), 0);
  std::fill(bp->snap_time.begin(), bp->snap_time.end(), -1);
  return 0;
}

int afqmc_bp_step(afqmc_bp bp, afqmc_walkers w) {
  NOT_CAPTURING();
  REQUIRE(bp && w && w->cp == bp->cp && w->nw == bp->nw, "invalid argument");
  REQUIRE(bp->t >= 0, "afqmc_bp_step before afqmc_bp_begin");
  afqmc_cp cp = bp->cp;
  ScopedDevice sd(cp->ctx);
  int rc = snapshot_if_due(bp, w);
  if (rc) return rc;
  int* aux = nullptr;
  int slot = 0;
  if (bp->nbp > 0) {
    slot = (int)(bp->t % bp->nbp);
    aux = bp->aux.p + (size_t)slot * bp->nw * cp->L;
  }
  if ((rc = step_record(cp, w, aux))) return rc;
  if (bp->nbp > 0) bp->has_table[slot] = 0;
  bp->t += 1;
  return 0;
}

int afqmc_bp_pop_control(afqmc_bp bp, afqmc_walkers w, double* stats) {
  NOT_CAPTURING();
  REQUIRE(bp && w && w->cp == bp->cp && w->nw == bp->nw, "invalid argument");
  REQUIRE(bp->t >= 1, "afqmc_bp_pop_control records the table of the last afqmc_bp_step");
  afqmc_cp cp = bp->cp;
  ScopedDevice sd(cp->ctx);
  int rc = afqmc_cp_pop_control(cp, w, nullptr, nullptr, stats);
  if (rc) return rc;
  if (bp->nbp > 0) {
    const int slot = (int)((bp->t - 1) % bp->nbp);
    HIP_TRY(hipMemcpyAsync(bp->tables.p + (size_t)slot * bp->nw, cp->table.p, sizeof(int) * bp->nw,
                           hipMemcpyDeviceToDevice, cp->ctx->stream));
    bp->has_table[slot] = 1;
  }
  return 0;
}

int afqmc_bp_measure(afqmc_bp bp, afqmc_walkers w) {
  NOT_CAPTURING();
  REQUIRE(bp && w && w->cp == bp->cp && w->nw == bp->nw, "invalid argument");
  REQUIRE(bp->t >= bp->nbp, "a measurement needs %d recorded steps", bp->nbp);
  afqmc_cp cp = bp->cp;
  ScopedDevice sd(cp->ctx);
  int rc = snapshot_if_due(bp, w);
  if (rc) return rc;
  const long long start = bp->t - bp->nbp;
  const int slot = (int)((start / bp->skip) % bp->nsnap);
  REQUIRE(start % bp->skip == 0 && bp->snap_time[slot] == start,
          "no walker copy at step %lld: a window starts every measureSkipStep steps", start);
  rc = cp->fp32 ? measure_impl<float>(bp, w, slot) : measure_impl<double>(bp, w, slot);
  if (rc) return rc;
  bp->measurements += 1;
  return 0;
}

int afqmc_bp_get(afqmc_bp bp, double* scalars, double* green, double* dd, double* pm, double* mp, double* sp) {
  NOT_CAPTURING();
  REQUIRE(bp, "null argument");
  afqmc_cp cp = bp->cp;
  ScopedDevice sd(cp->ctx);
  HIP_TRY(hipStreamSynchronize(cp->ctx->stream));
  std::vector<cd> h(bp->acc_count());
  HIP_TRY(hipMemcpy(h.data(), bp->acc.p, sizeof(zd) * h.size(), hipMemcpyDeviceToHost));
  const size_t L = cp->L, L2 = 2 * L;
  const cd* G = h.data() + 1;
  const cd* D = G + L2 * L2;
  // HubbardSOCMeasureCommuteSDSD::addEnergy (:143-180) is linear in G but for the U term, which is
  // <n_i,up n_i,dn> = densityDensity(i, i+L): the sums over walkers give the same numerators
  cd Ke = 0.0, Ve = 0.0, Re = 0.0;
  for (size_t k = 0; k < L2 * L2; ++k) Ke += cp->h_K[k] * G[k];
  for (size_t i = 0; i < L; ++i) {
    Ve += cp->h_U[i] * D[i + (i + L) * L2];
    Re += (-cp->h_mu[i] + cp->h_hz[i] * 0.5) * G[i + i * L2];
    Re += (-cp->h_mu[i] - cp->h_hz[i] * 0.5) * G[(i + L) + (i + L) * L2];
    Re += cd(cp->h_hx[i] * 0.5, -cp->h_hy[i] * 0.5) * G[i + (i + L) * L2];
    Re += cd(cp->h_hx[i] * 0.5, cp->h_hy[i] * 0.5) * G[(i + L) + i * L2];
  }
  if (scalars) {
    const cd v[5] = {h[0], Ke + Ve + Re, Ke, Ve, Re};
    for (int k = 0; k < 5; ++k) { scalars[2 * k] = v[k].real(); scalars[2 * k + 1] = v[k].imag(); }
  }
  auto put = [&](double* dst, const cd* src, size_t n) {
    if (dst) std::memcpy(dst, src, sizeof(cd) * n);
  };
  put(green, G, L2 * L2);
  put(dd, D, L2 * L2);
  put(pm, D + L2 * L2, L * L);
  put(mp, D + L2 * L2 + L * L, L * L);
  put(sp, D + L2 * L2 + 2 * L * L, L * L);
  return 0;
}

int afqmc_bp_reset(afqmc_bp bp) {
  NOT_CAPTURING();
  REQUIRE(bp, "null argument");
  ScopedDevice sd(bp->cp->ctx);
  HIP_TRY(hipMemsetAsync(bp->acc.p, 0, sizeof(zd) * bp->acc_count(), bp->cp->ctx->stream));
  return 0;
}

int afqmc_bp_get_step(afqmc_bp bp, int64_t* t, int64_t* measurements) {
  REQUIRE(bp, "null argument");
  if (t) *t = bp->t;
  if (measurements) *measurements = bp->measurements;
  return 0;
}

}  // extern "C"
