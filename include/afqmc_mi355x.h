/*
 * C-ABI of the MI355X walker-propagation library (libafqmc_mi355x.so).
 *
 * Plain pointers and sizes only.  Complex numbers are interleaved (re, im)
 * doubles and matrices column-major, the layout of the reference's
 * TensorHao<std::complex<double>,2> (common/tensorHao/include/tensor_core.h:85-106),
 * so a reference SD's wf.data() can be passed as is.  Every function returns
 * 0 on success and a negative code on error; afqmc_last_error() then holds the
 * message (the reference prints and calls exit(1); a library must not exit).
 * Calls are host-synchronous unless stated otherwise.  Walker state lives in
 * HBM between calls.
 *
 * Each entry point names the reference function it replaces, batched over
 * all walkers of a set (the reference loops over walkers one at a time).
 */
#ifndef AFQMC_MI355X_H
#define AFQMC_MI355X_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct afqmc_ctx_s* afqmc_ctx;
typedef struct afqmc_cp_s* afqmc_cp;
typedef struct afqmc_walkers_s* afqmc_walkers;
typedef struct afqmc_ph_s* afqmc_ph;
typedef struct afqmc_graph_s* afqmc_graph;

/* ------------------------------------------------------------------ context */
const char* afqmc_last_error(void);
int afqmc_version(void);
/* device context and stream (all work of one context is ordered on one HIP stream) */
int afqmc_ctx_create(int device, afqmc_ctx* out);
int afqmc_ctx_destroy(afqmc_ctx ctx);
int afqmc_ctx_synchronize(afqmc_ctx ctx);
/* Per-region kernel timing with HIP events on the context stream (no reference counterpart;
 * the reference only has TimerHao wall time, time_hao.h:16-41).  Regions, in order:
 * 0 propagate (one-body GEMM with fused HS scaling), 1 overlap GEMM, 2 Gauss-Jordan,
 * 3 Green GEMM, 4 HS sampling, 5 QR, 6 comb, 7 measurement, 8 phaseless force bias,
 * 9 phaseless sample matrix, 10 phaseless Taylor, 11 HS 2x2 apply.
 * set_timing resets the totals; get_timing synchronises and returns n totals (ms) and counts. */
int afqmc_ctx_set_timing(afqmc_ctx ctx, int enable);
int afqmc_ctx_get_timing(afqmc_ctx ctx, double* total_ms, long long* count, int n);

/* HIP-graph capture of enqueue-only calls (no reference counterpart: the reference launches
 * nothing).  Between begin and end, the calls made on this thread are recorded instead of run:
 * afqmc_cp_step (without records), afqmc_cp_stabilize, afqmc_cp_pop_control (without outputs),
 * afqmc_cp_measure, afqmc_cp_onebody, afqmc_cp_reset_measure, afqmc_cp_adjust_ET.  Calls that
 * return data to the host fail while a capture is open.  begin_group also records work issued on
 * the member contexts (one stream each; e.g. independent lattices), which become parallel
 * branches of one graph.  afqmc_graph_launch enqueues the whole graph on the origin's stream
 * (asynchronous).  Walker sets and propagators used in a graph must outlive it. */
int afqmc_capture_begin(afqmc_ctx ctx);
int afqmc_capture_begin_group(afqmc_ctx origin, int n, const afqmc_ctx* members);
int afqmc_capture_end(afqmc_ctx origin, afqmc_graph* out);
int afqmc_graph_launch(afqmc_graph g);
int afqmc_graph_node_count(afqmc_graph g);
int afqmc_graph_destroy(afqmc_graph g);

/* ------------------------------------------------------------------ CP (Hubbard) */
/* Model + method description.  Mirrors HubbardSOC (HubbardSOC.cpp:39-59), the
 * three Hop propagators built by returnExpMinusAlphaK (HubbardSOC.cpp:109-118),
 * NiupNidn's parameters (NiupNidn.cpp:12-26) and the CP method keys used on the
 * hot path (afqmcConstraintPathMethod.cpp:19-46). */
typedef struct {
  int L;                 /* lattice sites; walkers are 2L x N */
  int N;                 /* particles (HubbardSOC::N) */
  int Nup;               /* spin_blocks: first Nup columns are spin up */
  int spin_blocks;       /* 1: walkers and trial are block diagonal (rows 0..L-1 x cols 0..Nup-1 and
                            rows L..2L-1 x cols Nup..N-1), K and the HS field do not mix spins;
                            stored and multiplied as two blocks.  0: dense 2L x N.  Each block
                            has at most 256 columns (161 to 256: the factorisations work on
                            copies in device memory, N^2 x 16 bytes per walker). */
  int precision;         /* 0: complex FP64 throughout; 1: complex FP32 walkers and GEMMs, FP64 LU,
                            log-determinants, QR, weights and energies; 2: as 1, with the
                            propagation GEMM in split-BF16 products (experimental) */
  int decomp_type;       /* NiupNidn decompType: 0 densityCharge, 1 densitySpin, 2 hopCharge, 3 hopSpin */
  int force_type;        /* 0 dynamicForce, 1 constForce */
  double dt;
  double force_cap;
  const double* K;                    /* 2L x 2L complex */
  const double* exp_minus_dt_k;       /* 2L x 2L complex: exp(-dt K) */
  const double* exp_minus_half_dt_k;  /* exp(-dt/2 K) */
  const double* exp_half_dt_k;        /* exp(+dt/2 K) */
  double hop_logw[6];                 /* logw (re, im) of the three Hop operators */
  const double* U;                    /* L */
  const double* mu;                   /* L */
  const double* hx;                   /* L */
  const double* hy;                   /* L */
  const double* hz;                   /* L */
  const double* phiT;                 /* 2L x N complex trial wave function */
  double phiT_logw[2];
  const double* const_force;          /* L, used when force_type == 1 (NiupNidn::readForce) */
  int max_walkers;                    /* workspace is sized for this many walkers */
} afqmc_cp_desc;

int afqmc_cp_create(afqmc_ctx ctx, const afqmc_cp_desc* desc, afqmc_cp* out);
int afqmc_cp_destroy(afqmc_cp cp);
/* gamma and constDiag of the NiupNidn built from desc (NiupNidn.cpp:201-250): 5 x L complex */
int afqmc_cp_get_hs_constants(afqmc_cp cp, double* gamma_cdiag);

/* walker sets: std::vector<SD> of the driver (afqmcConstraintPath.h:27-28) */
int afqmc_cp_walkers_create(afqmc_cp cp, int nw, afqmc_walkers* out);
int afqmc_walkers_destroy(afqmc_walkers w);
int afqmc_walkers_count(afqmc_walkers w);
/* wf: nw x (2L x N) complex, logw: nw complex, alive: nw ints (NULL = all alive) */
int afqmc_walkers_upload(afqmc_walkers w, const double* wf, const double* logw, const int* alive);
int afqmc_walkers_download(afqmc_walkers w, double* wf, double* logw, int* alive);
/* SD::write of a few walkers (SD.cpp:85-103): the n walkers index[0..n-1], wf n x (2L x N), logw n complex,
 * alive n ints (any output may be NULL) */
int afqmc_walkers_download_some(afqmc_walkers w, int n, const int* index, double* wf, double* logw, int* alive);
/* walkerIsAlive of the driver (afqmcConstraintPath.h:29): nw ints, nonzero = alive */
int afqmc_walkers_set_alive(afqmc_walkers w, const int* alive);
/* setWalkerFromPhiT / initialWalker (SDSDOperation.cpp:176-186, afqmcConstraintPathSetWalker.cpp:9-18):
 * every walker = the trial held on the device, logw (1 complex, NULL = 0), all alive */
int afqmc_cp_walkers_from_trial(afqmc_cp cp, afqmc_walkers w, const double* logw);

/* random numbers: a counter-based uniform stream on the device, or host-provided
 * uniforms in the reference's consumption order (uniformHao, random_hao.cpp:73-76) */
int afqmc_cp_set_seed(afqmc_cp cp, uint64_t seed, uint64_t counter);
int afqmc_cp_get_counter(afqmc_cp cp, uint64_t* counter);
/* walkers propagated by afqmc_cp_step, summed over steps (alive walkers only) */
int afqmc_cp_get_walker_steps(afqmc_cp cp, uint64_t* walker_steps);
int afqmc_cp_set_uniforms(afqmc_cp cp, const double* u, int64_t n); /* n = 0: back to the device stream */
/* consume n uniforms without using them (e.g. the aux sample drawn by estimateMemory,
 * afqmcConstraintPath.cpp:66) */
int afqmc_cp_skip_uniforms(afqmc_cp cp, uint64_t n);
int afqmc_cp_set_ET(afqmc_cp cp, double ET);
int afqmc_cp_get_ET(afqmc_cp cp, double* ET);

/* GEMM tile shapes of the step (no reference counterpart): tiles[0] the one-body propagator,
 * tiles[1] the overlap, tiles[2] the Green's function GEMM; 0 = 64x64, 1 = 32x64, 2 = 64x32,
 * 3 = 32x32, 4 = 128x64 output tiles with a 16-deep K tile, 5 = 64x64, 6 = 64x32, 7 = 32x64 with a
 * 32-deep K tile, 8 = 128x128 and 9 = 256x64 (8 waves), 10 = 64x128, with a 16-deep K tile (the Green
 * GEMM takes shapes 0, 1, 4, 5, 7, 8 and 9).  autotune times every shape on the walker set's
 * own buffers (reps launches each), keeps the fastest and returns the choices in tiles (3) and the
 * times in ms (3 x 11, ms per launch, -1 = not applicable).  Every shape gives bit-identical results. */
int afqmc_cp_set_tiles(afqmc_cp cp, const int* tiles);
int afqmc_cp_autotune(afqmc_cp cp, afqmc_walkers w, int reps, int* tiles, double* ms);

/* Groups of runs (no reference counterpart: the reference runs one model per process; this is for
 * callers that propagate many Hubbard instances of uneven size at once).  A group holds n members,
 * each a propagator with one of its walker sets, all on one context and in one precision (FP64 or
 * mixed; not the split-BF16 propagation), no two sharing a propagator.  group_step, group_stabilize
 * and group_pop_control do what afqmc_cp_step, afqmc_cp_stabilize and afqmc_cp_pop_control (without
 * outputs) do on every member, with the same results, but launch each kernel of a phase once over
 * all members (about ten launches per step instead of ten per member), which a graph captured from
 * group calls replays.  The group's tile shapes (set_tiles, as afqmc_cp_set_tiles; 64x64 by default)
 * stand in for the members' own during its calls; a shape of -1 keeps each member's own for that
 * GEMM (members with equal shapes share a launch, each other shape adds one).  Argument tables of the grouped launches are kept
 * on the device until group_destroy (table_count: how many), so graphs captured from a group's calls
 * must be destroyed before it.  After an error the members' walker sets are undefined. */
typedef struct afqmc_cp_group_s* afqmc_cp_group;
int afqmc_cp_group_create(int n, const afqmc_cp* cps, const afqmc_walkers* walkers, afqmc_cp_group* out);
int afqmc_cp_group_destroy(afqmc_cp_group g);
int afqmc_cp_group_set_tiles(afqmc_cp_group g, const int* tiles);
int afqmc_cp_group_step(afqmc_cp_group g);
int afqmc_cp_group_stabilize(afqmc_cp_group g);
int afqmc_cp_group_pop_control(afqmc_cp_group g);
int afqmc_cp_group_table_count(afqmc_cp_group g);

/* SDSDOperation::returnLogOverlap (SDSDOperation.cpp:65-72) for every walker: logO nw complex */
int afqmc_cp_log_overlap(afqmc_cp cp, afqmc_walkers w, double* logO);
/* SDSDOperation::returnGreenDiagonal / returnGreenOffDiagonal (SDSDOperation.cpp:91-141):
 * gdiag, goff: nw x 2L complex (goff may be NULL; zero in spin-block mode) */
int afqmc_cp_green_diagonal(afqmc_cp cp, afqmc_walkers w, double* gdiag, double* goff);
/* HubbardSOCMeasureFixedSDSD::getForce (HubbardSOCMeasureFixedSDSD.cpp:73-116): force nw x L */
int afqmc_cp_force(afqmc_cp cp, afqmc_walkers w, double* force);
/* HopSDOperation::applyToRight (hopSDOperation.cpp:14-19) with
 * which = 0: exp(-dt K), 1: exp(-dt/2 K), 2: exp(+dt/2 K); alive walkers only */
int afqmc_cp_onebody(afqmc_cp cp, afqmc_walkers w, int which);
/* NiupNidnSampleSDOperation::applyToRight (NiupNidnSDOperation.cpp:15-41) with one sample per
 * walker: diag4 = nw x 4 x L complex (diag00, diag01, diag10, diag11), sample_logw nw complex.
 * Dense (spin_blocks = 0) sets only. */
int afqmc_cp_hs_apply(afqmc_cp cp, afqmc_walkers w, const double* diag4, const double* sample_logw);
/* SD::stabilize (SD.cpp:42-45) for every alive walker: CholeskyQR in FP64, R with positive diagonal,
 * a second pass for walkers not yet orthonormal; a walker whose Gram matrix is not numerically positive
 * definite (condition number beyond about 1e8) is redone by shifted CholeskyQR3, and one whose shifted
 * Gram matrix still breaks down (rank deficient: Householder's R has a zero on its diagonal, log det R =
 * -inf, weight zero) is marked dead */
int afqmc_cp_stabilize(afqmc_cp cp, afqmc_walkers w);
/* walkers that took the shifted CholeskyQR and walkers marked dead as rank deficient, summed over all
 * stabilisations of this propagator */
int afqmc_cp_get_qr_stats(afqmc_cp cp, uint64_t* shifted, uint64_t* killed);
/* popConfiguration(1, weight) (populationControlConfiguration.cpp:24-101), one uniform per slot;
 * stats = popCheck min, max, mean (may be NULL) */
int afqmc_comb(afqmc_ctx ctx, int nw, const double* weight, const double* uniforms, int* table, double* stats);
/* afqmc_comb on device buffers, allocating nothing.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  weight and uniforms (nw doubles each), table (nw ints) and
 * stats (3 doubles, may be NULL) are device pointers.  The comb's working memory (run scalars,
 * probabilities, slot counts) is a fixed device array of about 0.8 MB per GPU, held until the
 * process ends, so nw is limited to AFQMC_COMB_NOALLOC_MAX.  The call is enqueued on the
 * context's stream and returns without synchronising.  All calls on one GPU share that array:
 * calls that overlap in time on one GPU (other streams, other contexts, other threads) race. */
#define AFQMC_COMB_NOALLOC_MAX 65536
int afqmc_comb_noalloc(afqmc_ctx ctx, int nw, const double* weight, const double* uniforms, int* table, double* stats);
/* populationControl, serial branch (populationControl.h:110-116): walker[i] = walker[table[i]] */
int afqmc_cp_apply_table(afqmc_cp cp, afqmc_walkers w, const int* table);

/* ---- driver level: the loops of afqmcConstraintPathProjection.cpp / ...Measure.cpp ----
 * These stay on the device; they do not synchronise unless an output pointer is given. */
/* projectExpMinusDtKExpMinusDtV (:37-74): force (nw x L) and aux (nw x L) records may be NULL */
int afqmc_cp_step(afqmc_cp cp, afqmc_walkers w, double* force_out, int* aux_out);
/* popControl (:103-156): weights (nw), table (nw), stats (3) may be NULL */
int afqmc_cp_pop_control(afqmc_cp cp, afqmc_walkers w, double* weights_out, int* table_out, double* stats);
/* projectExpMinusDtKExpMinusDtV (:37-74) with each walker's aux fields given (nw x L ints, +-1; rows of
 * dead walkers are ignored) instead of sampled: no uniforms are drawn, while the force bias still enters
 * the sample's logw (getTwoBodySampleFromAuxForce, NiupNidn.cpp:133-140) and the phase projection still
 * applies.  Replays the fields of another run (aux_out of afqmc_cp_step), e.g. the FP64 path's fields
 * in the mixed path on the same walkers, which isolates the precision error of the energy. */
int afqmc_cp_step_fields(afqmc_cp cp, afqmc_walkers w, const int* aux);
/* popControl (:103-156) with the comb's table given (nw ints, e.g. the table_out of another run; every
 * parent must keep its own slot) instead of drawn: the weights are formed and the log-weights shifted
 * by -logO as usual (weights_out nw, may be NULL); no uniforms are drawn. */
int afqmc_cp_pop_control_table(afqmc_cp cp, afqmc_walkers w, const int* table, double* weights_out);
/* addMixedMeasurement (afqmcConstraintPathMeasure.cpp:9-27) */
int afqmc_cp_measure(afqmc_cp cp, afqmc_walkers w);
/* accumulated den, HNum, KNum, VNum, RNum, NupTotNum, NdnTotNum: 7 complex */
int afqmc_cp_energy(afqmc_cp cp, double* out);
/* NupNum, NdnNum, SplusNum, SminusNum: 4 x L complex */
int afqmc_cp_site_measure(afqmc_cp cp, double* out);
int afqmc_cp_reset_measure(afqmc_cp cp);
/* adjustETAndResetMeasurement (afqmcConstraintPathMeasure.cpp:35-45): ET = Re(HNum/den) on device, then reset */
int afqmc_cp_adjust_ET(afqmc_cp cp);

/* Fixed-field propagation and its adjoint (no reference counterpart: the reference has no
 * gradients; this is the differentiable segment of a constraint-path run, for sensitivities of the
 * mixed-estimator energy to the Hamiltonian).  FP64 spin-block propagators with the densitySpin
 * decomposition and mu = h = 0.  Steps s = 0 .. nsteps-1 apply, per spin block (sigma = +1 up,
 * -1 down), W <- exp(-dt K) diag(exp(sigma gamma_i x) constDiag_i) W with the field
 * x_{s,w,i} = (u < 1/2 ? -1 : +1), u the uniform number (s*nw + w)*L + i of the counter-based
 * stream seeded with field_seed; no force bias, phase projection or comb; CholeskyQR every mgs
 * steps with logw += log det R.  The energy is the mixed estimator of the final walkers,
 * E = sum_w O_w E_L,w / sum_w O_w (O_w = exp(logw) times the overlaps with the trial, E_L the
 * local energy of HubbardSOCMeasureFixedSDSD.cpp:178-218), one complex (re, im).  Both calls start
 * from the walker set's current state and restore it.
 * afqmc_cp_fixed_field_adjoint: forward with a checkpoint every ckpt steps (a multiple of mgs),
 * then the reverse sweep, recomputing each segment from its checkpoint.  Outputs:
 *   bbar (2 x L x L complex): dE/dB_ij for B = exp(-dt K_b) of each spin block, summed over steps;
 *   mbar (2 x L x L complex): the explicit dE/dK_ij of each block's kinetic term;
 *   ubar (L complex): dE/dU_i, through the HS factors and the interaction term.
 * dE/dK_b follows from bbar by the derivative of the matrix exponential (done by the caller,
 * which holds K's eigendecomposition) plus mbar. */
int afqmc_cp_fixed_field_energy(afqmc_cp cp, afqmc_walkers w, int nsteps, int mgs, uint64_t field_seed,
                                double* energy);
int afqmc_cp_fixed_field_adjoint(afqmc_cp cp, afqmc_walkers w, int nsteps, int mgs, int ckpt, uint64_t field_seed,
                                 double* energy, double* bbar, double* mbar, double* ubar);

/* ------------------------------------------------------------------ back-propagation */
/* The back-propagated estimator of tutorials/afqmcConstraintPathBP: the aux fields of every step
 * (twoBodyAuxBackup, afqmcConstraintPathProjection.cpp:73) and the comb tables
 * (tablePerThreadBackup, :163-171) are kept on the device for the last bp_steps
 * (backPropagationStep) steps, the walkers are copied every measure_skip (measureSkipStep) steps,
 * and a measurement propagates phiT to the left through each walker's fields along its comb
 * ancestry, normalising every mgs (mgsStep) steps (afqmcConstraintPathMeasure.cpp:41-72;
 * afqmcConstraintPathBackPropagation.cpp:11-97).  The observables are those of
 * HubbardSOCMeasureObserveSDSD (HubbardSOCMeasureObserveSDSD.cpp:34-183), weighted by each walker's
 * mixed overlap.  The propagators must be Hermitian (exp(-alpha K) of a Hermitian K, as
 * HubbardSOC::returnExpMinusAlphaK builds them), since applyToLeft multiplies by their adjoints. */
typedef struct afqmc_bp_s* afqmc_bp;
int afqmc_bp_create(afqmc_cp cp, int nw, int bp_steps, int measure_skip, int mgs, afqmc_bp* out);
int afqmc_bp_destroy(afqmc_bp bp);
/* prepareBackPropagation (afqmcConstraintPathBackPropagation.cpp:11-20): step 0 of the recording */
int afqmc_bp_begin(afqmc_bp bp);
/* afqmc_cp_step with the aux fields recorded (and the walkers copied when a window starts) */
int afqmc_bp_step(afqmc_bp bp, afqmc_walkers w);
/* afqmc_cp_pop_control after the last afqmc_bp_step, its table recorded; stats may be NULL */
int afqmc_bp_pop_control(afqmc_bp bp, afqmc_walkers w, double* stats);
/* addMeasurement at step k * measure_skip + bp_steps of the recording, accumulating on the device */
int afqmc_bp_measure(afqmc_bp bp, afqmc_walkers w);
/* the accumulated numerators (complex, column-major), any pointer may be NULL:
 * scalars: den, HNum, KNum, VNum, RNum (5 complex); green: greenMatrixNum (2L x 2L);
 * dd: densityDensityNum (2L x 2L); pm, mp, sp: splusSminusNum, sminusSplusNum, spairSpairNum (L x L) */
int afqmc_bp_get(afqmc_bp bp, double* scalars, double* green, double* dd, double* pm, double* mp, double* sp);
/* observeMeasure.reSet after a write */
int afqmc_bp_reset(afqmc_bp bp);
/* steps recorded since afqmc_bp_begin and measurements made */
int afqmc_bp_get_step(afqmc_bp bp, int64_t* steps, int64_t* measurements);

/* ------------------------------------------------------------------ phaseless (real materials) */
/* Model, trial and method of the phaseless driver: RealMaterialMolecule (realMaterialMolecule.h:16-40,
 * t, choleskyVecs, choleskyBg), the SD2s trial, the half-rotated vectors of
 * RealMaterialMoleculeMeasureFixedSD2sSD2is::initWfDaggerCholeskyVecs (:165-186), the method keys
 * forceType / forceCap / dt (afqmcPhaselessMethod.cpp:158-185) and LogHop2isSD2isOperation's
 * parameters (logHop2isSD2isOperation.h:20-33). */
typedef struct afqmc_ph_walkers_s* afqmc_ph_walkers;
typedef struct {
  int L;                      /* orbitals; SD2is walkers are L x max(Nup, Ndn) */
  int Nup, Ndn;
  int G;                      /* Cholesky vectors (choleskyNumber) */
  int precision;              /* 0: complex FP64; 1: complex FP32 walkers and GEMMs, FP64 LU, logdets, QR,
                                 weights, sample logw and energies */
  int force_type;             /* 0 dynamicForce, 1 constForce */
  double dt;
  double force_cap;
  const double* t;            /* L x L real, one-body kinetic matrix of the energy */
  const double* chol;         /* L x L x G real, column-major per vector */
  const double* bg;           /* G real current background (NULL = zeros) */
  const double* phiT_up;      /* L x Nup complex */
  const double* phiT_dn;      /* L x Ndn complex */
  double phiT_logw[2];
  const double* half_rot_up;  /* Nup x L x G complex: phiT_up^H L_g, column-major (wfUpDaggerCholeskyVecs) */
  const double* half_rot_dn;  /* Ndn x L x G complex */
  const double* const_force;  /* G complex, used when force_type == 1 (NULL = zeros) */
  int max_walkers;
  int taylor_fixed;           /* 0 dynamicOrder, 1 fixedOrder */
  int taylor_order;           /* fixedOrder: terms (0 = the first application takes the dynamic order of its
                                 first alive walker and freezes it for every walker, as the serial
                                 reference's determinantAndAddFixedOrders, logHop2isSD2isOperation.cpp:122-133) */
  int taylor_base;            /* baseTaylorOrder (<= 0: 3) */
  double taylor_accuracy;     /* <= 0: 1e-10 */
} afqmc_ph_desc;

int afqmc_ph_create(afqmc_ctx ctx, const afqmc_ph_desc* desc, afqmc_ph* out);
int afqmc_ph_destroy(afqmc_ph ph);
/* one-body operators of the driver: which = 0 exp(-dt Kp), 1 exp(-dt/2 Kp), 2 exp(+dt/2 Kp).
 * is_log = 0: Hop2is matrix B applied as B W (hop2isSD2isOperation.cpp:14-19);
 * is_log = 1: LogHop2is matrix A applied as the Taylor series of exp(A) W (logHop2isSD2isOperation.cpp:99-195).
 * matrix L x L complex, logw 1 complex */
int afqmc_ph_set_onebody(afqmc_ph ph, int which, const double* matrix, const double* logw, int is_log);
/* RealMaterialMolecule::updateBackGround (realMaterialMolecule.cpp:103-116): G real */
int afqmc_ph_set_background(afqmc_ph ph, const double* bg);

/* std::vector<SD2is> of the driver (tutorials/afqmcPhaseless/include/afqmcPhaseless.h:29): wf nw x (L x max(Nup,Ndn)) complex */
int afqmc_ph_walkers_create(afqmc_ph ph, int nw, afqmc_ph_walkers* out);
int afqmc_ph_walkers_destroy(afqmc_ph_walkers w);
int afqmc_ph_walkers_upload(afqmc_ph_walkers w, const double* wf, const double* logw, const int* alive);
int afqmc_ph_walkers_download(afqmc_ph_walkers w, double* wf, double* logw, int* alive);

/* random numbers: as afqmc_cp_set_seed / afqmc_cp_set_uniforms; a phaseless step takes 2G uniforms
 * per alive walker (u1, u2 of gaussianHao, random_hao.cpp:78-85) */
int afqmc_ph_set_seed(afqmc_ph ph, uint64_t seed, uint64_t counter);
int afqmc_ph_get_counter(afqmc_ph ph, uint64_t* counter, uint64_t* walker_steps);
int afqmc_ph_set_uniforms(afqmc_ph ph, const double* u, int64_t n);
int afqmc_ph_skip_uniforms(afqmc_ph ph, uint64_t n);
int afqmc_ph_set_ET(afqmc_ph ph, double ET);
int afqmc_ph_get_ET(afqmc_ph ph, double* ET);

/* SD2sSD2isOperation::returnLogOverlap (SD2sSD2isOperation.cpp:84-94): logO nw complex */
int afqmc_ph_log_overlap(afqmc_ph ph, afqmc_ph_walkers w, double* logO);
/* RealMaterialMoleculeMeasureFixedSD2sSD2is::calculateCholeskyBg (:237-256): bg nw x G complex */
int afqmc_ph_background(afqmc_ph ph, afqmc_ph_walkers w, double* bg);
/* CholeskyReal::getTwoBodySampleFromAuxForce (choleskyReal.cpp:104-124) followed by
 * LogHop2isSD2isOperation::applyToRight: aux, force nw x G complex; sample_logw nw complex out (may be NULL);
 * max_order out (may be NULL); the per-walker orders by afqmc_ph_taylor_orders */
int afqmc_ph_sample_apply(afqmc_ph ph, afqmc_ph_walkers w, const double* aux, const double* force,
                          double* sample_logw, int* max_order);
/* the last sample matrices sum_g sqrt(-dt) x_g L_g: nw x L x L complex */
int afqmc_ph_sample_matrix(afqmc_ph ph, afqmc_ph_walkers w, double* M);
int afqmc_ph_taylor_orders(afqmc_ph ph, afqmc_ph_walkers w, int* orders);
/* Hop2isSD2isOperation / LogHop2isSD2isOperation applyToRight with one-body operator `which` */
int afqmc_ph_onebody(afqmc_ph ph, afqmc_ph_walkers w, int which);
/* SD2is::stabilize (SD2is.cpp:48-61) by CholeskyQR in FP64, as afqmc_cp_stabilize (second pass where
 * needed, shifted CholeskyQR3 for ill-conditioned walkers, rank-deficient walkers marked dead) */
int afqmc_ph_stabilize(afqmc_ph ph, afqmc_ph_walkers w);
int afqmc_ph_get_qr_stats(afqmc_ph ph, uint64_t* shifted, uint64_t* killed);
/* projectExpMinusDtKExpMinusDtV (afqmcPhaselessProjection.cpp:75-121): force and aux records nw x G complex
 * (may be NULL) */
int afqmc_ph_step(afqmc_ph ph, afqmc_ph_walkers w, double* force_out, double* aux_out);
/* popControl (afqmcPhaselessProjection.cpp:150-202) */
int afqmc_ph_pop_control(afqmc_ph ph, afqmc_ph_walkers w, double* weights_out, int* table_out, double* stats);
/* addMixedMeasurement (afqmcPhaselessMeasure.cpp:9-26) */
int afqmc_ph_measure(afqmc_ph ph, afqmc_ph_walkers w);
/* accumulated den, HNum, TNum (3 complex); choleskyBgNum and choleskyExNum (G complex each, may be NULL) */
int afqmc_ph_energy(afqmc_ph ph, double* out, double* bg_num, double* ex_num);
int afqmc_ph_reset_measure(afqmc_ph ph);
/* ET = Re(HNum/den), then reset (the ET half of adjustETAndBackGroundThenResetMeasurement,
 * afqmcPhaselessMeasure.cpp:36-45; the background half is afqmc_ph_set_background) */
int afqmc_ph_adjust_ET(afqmc_ph ph);

/* ------------------------------------------------------------------ testing hooks
 * One kernel family on host data, for unit tests of the device building blocks.
 * No reference counterpart: tensor_2d_bl_cpu.cpp's gmm / LUconstruct+inverse /
 * QR are what they are checked against. */
/* C_b = op(A_b) diag(scale) B_b; A M x K (op N) or K x M (op T/C), strideA = 0 shares A;
 * opa = 3: op N with A real (A then holds real doubles).  flags bit 0: the result is Hermitian
 * (op C, A = B; only tiles on or above the diagonal computed), bit 1: B upper triangular, bit 2:
 * split-BF16 products (precision 1, op N only; the tile is then a fixed 128 x 64); tile: a GEMM
 * tile shape (as afqmc_cp_set_tiles) */
int afqmc_test_zgemm(afqmc_ctx ctx, int precision, int opa, int M, int N, int K, int batch, const double* A,
                     long long strideA, const double* B, const double* scale, int scale_group, int flags, int tile,
                     double* C);
/* batched N x N inverse and log-determinant (Gauss-Jordan, partial pivoting), 1 <= N <= 256 */
int afqmc_test_gj_inverse(afqmc_ctx ctx, int precision, int N, int batch, const double* A, double* Ainv,
                          double* logdet, int* info);
/* the same with the kernel chosen: 1 lanes = rows (readlanes for the pivot row), 2 a 2-D lane layout
 * with both update operands from LDS (N <= 128), identical to 1; 3 panels of four columns with rank-4
 * updates (N <= 112), equal to 1 up to rounding (same pivots, updates summed in another order);
 * 0 the default for N and batch */
int afqmc_test_gj_inverse_layout(afqmc_ctx ctx, int precision, int N, int batch, int layout, const double* A,
                                 double* Ainv, double* logdet, int* info);
/* the same over the matrices b with alive[b] != 0 only; the others keep A (the inverse is written
 * in place) and the logdet values passed in */
int afqmc_test_gj_inverse_masked(afqmc_ctx ctx, int precision, int N, int batch, int layout, const int* alive,
                                 const double* A, double* Ainv, double* logdet, int* info);
/* device-to-device streaming copy of `bytes` (a multiple of 16), reps times: the reference kernel
 * for scaling HBM traffic counters; ms = average kernel time */
int afqmc_test_copy(afqmc_ctx ctx, long long bytes, int reps, double* ms);
/* batched Cholesky S = R^H R, returns R^{-1} and (sum log R_kk, same) */
int afqmc_test_chol_inverse(afqmc_ctx ctx, int N, int batch, const double* S, double* Rinv, double* logsum,
                            int* info);

#ifdef __cplusplus
}
#endif
#endif
