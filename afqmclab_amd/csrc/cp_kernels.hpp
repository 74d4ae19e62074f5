// Walker-level kernels of the constraint-path (CP) step and the comb.
#pragma once
#include "common.hpp"

// decompType of NiupNidn (NiupNidn.cpp:201-226)
enum { DECOMP_DENSITY_CHARGE = 0, DECOMP_DENSITY_SPIN = 1, DECOMP_HOP_CHARGE = 2, DECOMP_HOP_SPIN = 3 };

// Device-resident scalars of a run: the driver loop never waits on the host.
struct DevScalars {
  double ET;
  unsigned long long key;      // uniform stream key
  unsigned long long counter;  // next uniform index
  int alive_count;
  int uniform_overrun;         // set when a kernel asked for a host-provided uniform beyond ulen
  zd den, HNum, KNum, VNum, RNum;
  zd NupTot, NdnTot, SplusTot, SminusTot;
  double wmin, wmax, wavg;     // popCheck of the last comb
  double taylor_orders;        // phaseless: total Taylor terms applied in the last step
  unsigned long long walker_steps;  // walkers propagated, summed over steps (alive walkers only)
  unsigned long long ulen;     // number of host-provided uniforms (when a uniform buffer is set)
  unsigned long long qr_shifted;  // walkers whose Gram matrix needed the shifted Cholesky pass
  unsigned long long qr_killed;   // walkers whose Gram matrix stayed singular (rank deficient)
};

__device__ __forceinline__ double afqmc_uniform_dev(unsigned long long key, unsigned long long k) {
  unsigned long long z = key + (k + 1ull) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z = z ^ (z >> 31);
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}
unsigned long long afqmc_seed_mix_host(unsigned long long seed);

// uniform number k of the run: from the host-provided list when one is set (reads stay inside
// the list; an index past its end flags uniform_overrun and yields 0.5), else from the stream
__device__ __forceinline__ double afqmc_take_uniform(const double* ubuf, DevScalars* sc, unsigned long long key,
                                                     unsigned long long k) {
  if (!ubuf) return afqmc_uniform_dev(key, k);
  if (k < sc->ulen) return ubuf[k];
  sc->uniform_overrun = 1;
  return 0.5;
}

// logO combination of per-block log-determinants
enum { COMBINE_WRAP = 0, COMBINE_SUM = 1 };
// what to do with logO
enum { LOGO_STEP = 0, LOGO_WEIGHT = 1, LOGO_MEASURE = 2, LOGO_ONLY = 3 };

struct LogOArgs {
  int nw, nblk, combine, mode;
  const zd* ldet[2];
  const int* info[2];
  const zd* logw;        // walker logw (for MEASURE: logw of the temp walker)
  zd* logw_out;          // updated in STEP / WEIGHT mode
  int* alive;
  zd trial_logw;
  zd* logO;              // out
  double* weight;        // WEIGHT mode out
  zd* den;               // MEASURE mode out: exp(logO)
};
int launch_logo(const LogOArgs& a, hipStream_t s);

int launch_alive_scan(int nw, const int* alive, int* prefix, DevScalars* sc, hipStream_t s);

struct HsSampleArgs {
  int nw, L, decomp, force_type, nblk, dense;
  const zd* gd[2]; const zd* go[2];  // green partial sums [nw][ntile][Mb]
  int ntile[2], Mb[2];
  const zd* gamma;                   // L
  const zd* cdiag;                   // 4 x L: 00, 01, 10, 11
  const double* const_force;         // L
  double cap, dt, log_half_L, dtUSum;
  zd onebody_logw;                   // expMinusDtK.logw
  const int* alive; const int* prefix;
  DevScalars* sc;
  const double* ubuf;                // host-provided uniforms (device copy) or nullptr
  zd* logw;
  void* scale[2]; int diag_only;     // complex T, [nw][Mb] per block
  void* d2x2;                        // complex T, [nw][4][L] when !diag_only
  double* force_out; int* aux_out;   // optional records
  const int* aux_in;                 // optional given fields [nw][L] (no uniforms drawn)
  int fp32;
};
int launch_hs_sample(const HsSampleArgs& a, hipStream_t s);
int launch_advance_counter(DevScalars* sc, int per_walker, hipStream_t s);

template <typename T>
int launch_hs_apply_2x2(int nw, int L, int N, const void* in, void* out, const void* d2x2, hipStream_t s);

int launch_comb(int nw, const double* weight, const double* ubuf, DevScalars* sc, int* table, double* prob,
                int* number, hipStream_t s);
template <typename T>
int launch_apply_table(int nw, int nblk, void* const* wf, const long long* wsize, zd* logw, int* alive,
                       const int* table, hipStream_t s);

struct EnergyArgs {
  int nw, L, nblk, dense;
  const zd* gd[2]; const zd* go[2]; const zd* gk[2];
  int ntile[2], Mb[2];
  const double *U, *mu, *hx, *hy, *hz;
  const int* alive;
  const zd* den;
  zd* per_walker;      // [nw][8]: den, H, K, V, R, Nup, Ndn, (unused)
  zd* site_acc;        // [4][L] += per site (Nup, Ndn, Splus, Sminus) * den
  zd* site_part;       // [SITE_CHUNKS][4][L] scratch: per walker-chunk partial sums
  DevScalars* sc;
};
constexpr int SITE_CHUNKS = 64;
int launch_energy(const EnergyArgs& a, hipStream_t s);
int launch_set_et_from_energy(DevScalars* sc, hipStream_t s);
// CholeskyQR: which walkers still need a second pass (S = Q^H Q farther than tol from I)
int launch_ortho_check(int nw, int N, const zd* S, const int* alive, double tol, int* need, hipStream_t s);
// the same over FP32 Gram matrices
int launch_ortho_check_f32(int nw, int N, const zf* S, const int* alive, double tol, int* need, hipStream_t s);
// flag[w] = alive[w] && info[w] (a Cholesky breakdown of walker w)
int launch_qr_flag(int nw, const int* alive, const int* info, int* flag, hipStream_t s);
// S_w += c * tr(S_w) * I for the flagged walkers (the shift of shifted CholeskyQR)
int launch_qr_shift(int nw, int N, zd* S, const int* flag, double c, hipStream_t s);
// flagged walkers whose Cholesky broke down (info) die and leave the flag (and flag2 when given, so a
// later pass over flag2 does not see them again); counts into sc (the flagged ones as shifted when
// count_shifted)
int launch_qr_kill(int nw, int* flag, int* flag2, const int* info, int* alive, DevScalars* sc,
                   int count_shifted, hipStream_t s);
int launch_logw_add_real(int nw, zd* logw, const int* alive, const zd* add0, const zd* add1, int use_imag_sub,
                         double mult0, double mult1, hipStream_t s);

// ------------------------------------------------------------------ back-propagation
// One backward step of the per-walker bras: anc[w] = table[anc[w]] when the step had a comb, then the
// conjugate transposes D^H of the site matrices of the ancestor's recorded fields, in the layout the
// walker step writes (scale per block for diagonal samples, d2x2 otherwise).
struct BpFieldArgs {
  int nw, L, decomp, dense, diag_only;
  const int* aux;    // [nw][L] fields of this step, by slot
  const int* table;  // comb table of this step or nullptr
  int* anc;          // [nw] ancestor slots, updated in place
  const zd* gamma; const zd* cdiag;
  void* scale[2]; int Mb[2];
  void* d2x2;
  int fp32;
};
int launch_bp_field(const BpFieldArgs& a, hipStream_t s);
int launch_iota(int n, int* out, hipStream_t s);
// dst[w] = src[idx[w]] for per-walker blocks of n complex elements; idx == nullptr: dst[w] = src[0]
int launch_gather_walkers(int nw, long long n, size_t esz, const int* idx, const void* src, void* dst, hipStream_t s);
// dst_w = src_w^H: src M x N (ld M), dst N x M (ld N), walkers with alive[w] != 0
template <typename T>
int launch_conj_transpose(int nw, int M, int N, long long sstride, long long dstride, const int* alive,
                          const void* src, void* dst, hipStream_t s);

// Accumulation of the observables of HubbardSOCMeasureObserveSDSD from per-walker Green matrices:
// G(r, c) = gt[c + r * ld] (a walker's G^T, column-major), spin blocks (dense = 0) hold the up-up and
// down-down L x L blocks and zero spin-flip blocks.  Each thread owns a site pair (i, j) and sums, over
// the walkers of the chunk with weight den[w], 11 complex quantities: G at the four spin pairs,
// <n_a n_b> at the four spin pairs, <S+_i S-_j>, <S-_i S+_j>, <D+_i D_j>.
constexpr int WICK_Q = 11;
struct WickArgs {
  int L, dense, c0, cn, slices, first;
  const void* gt[2]; long long gstride[2]; int ld[2];
  const zd* den; const int* alive;
  zd* part;  // [slices][WICK_Q][L * L] (pair index i + j L)
  int fp32;
};
int launch_wick(const WickArgs& a, hipStream_t s);
// acc_green (2L x 2L), acc_dd (2L x 2L), acc_pm, acc_mp, acc_sp (L x L) += sum over slices of part
int launch_wick_reduce(int L, int slices, const zd* part, zd* acc_green, zd* acc_dd, zd* acc_pm, zd* acc_mp,
                       zd* acc_sp, hipStream_t s);
// acc[0] += sum of den[w] over alive walkers
int launch_den_sum(int nw, const zd* den, const int* alive, zd* acc, hipStream_t s);
