// C++ host driver of the constraint-path walker loop on top of the C-ABI (include/afqmc_mi355x.h).
//
// AfqmcConstraintPathGpu keeps the member-function structure and the afqmc_param keys of the
// reference's AfqmcConstraintPath (tutorials/afqmcConstraintPath/include/afqmcConstraintPath.h:12-58,
// afqmcConstraintPathMethod.h:15-45): run() -> initialWalker -> estimateMemory ->
// measureWithProjection, the same projectOneStep / modifyGM / popControl / addMixedMeasurement /
// adjustETAndResetMeasurement cadences, with the per-walker loops replaced by one batched
// device call each.  Walkers, weights, ET and the random counter stay in HBM.  With an output
// directory set, each block appends the reference's measurement files (den.dat, HNum.dat, ...,
// HubbardSOCMeasureFixedSDSD.cpp:118-136) and beta.dat.
#pragma once
#include <cmath>
#include <complex>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/afqmc_mi355x.h"
#include "reference_io.hpp"

namespace afqmc_gpu {

inline void check(int rc, const char* what) {
  if (rc != 0) throw std::runtime_error(std::string(what) + ": " + afqmc_last_error());
}

// AfqmcConstraintPathMethod's keys and defaults (afqmcConstraintPathMethod.cpp:19-61), read from an
// afqmc_param file as readBySearchString does: the value follows the key's token, "=" optional.
struct ConstraintPathMethod {
  double dt = 0.01;
  size_t thermalSize = 200, writeNumber = 60, measureNumberPerWrite = 2, measureSkipStep = 5;
  // the back-propagation driver's extra key (tutorials/afqmcConstraintPathBP/source/
  // afqmcConstraintPathMethod.cpp:26, default 100 at :166); read and written only when backPropagation
  bool backPropagation = false;
  size_t backPropagationStep = 100;
  int walkerSizePerThread = 300;
  std::string decompType = "densitySpin", forceType = "dynamicForce";
  double forceCap = 1.5;
  std::string initialPhiTFlag = "setFromModel", initialWalkerFlag = "sampleFromPhiT";
  size_t mgsStep = 10, popControlStep = 10;
  double ET = -10.0;
  size_t ETAdjustStep = 10, ETAdjustMaxSize = 100;
  long long seed = 985456376;

  void read(const std::string& file) {
    using refio::search_value;
    search_value(file, "dt", dt);
    search_value(file, "thermalSize", thermalSize);
    search_value(file, "writeNumber", writeNumber);
    search_value(file, "measureNumberPerWrite", measureNumberPerWrite);
    search_value(file, "measureSkipStep", measureSkipStep);
    if (backPropagation) search_value(file, "backPropagationStep", backPropagationStep);
    search_value(file, "walkerSizePerThread", walkerSizePerThread);
    search_value(file, "decompType", decompType);
    search_value(file, "forceType", forceType);
    search_value(file, "forceCap", forceCap);
    search_value(file, "initialPhiTFlag", initialPhiTFlag);
    search_value(file, "initialWalkerFlag", initialWalkerFlag);
    search_value(file, "mgsStep", mgsStep);
    search_value(file, "popControlStep", popControlStep);
    search_value(file, "ET", ET);
    search_value(file, "ETAdjustStep", ETAdjustStep);
    search_value(file, "ETAdjustMaxSize", ETAdjustMaxSize);
    search_value(file, "seed", seed);
    analysis();
  }
  // afqmcConstraintPathMethod.cpp:48-89
  void write(const std::string& file) const {
    std::ofstream f(file, std::ios::out | std::ios::trunc);
    if (!f) throw std::runtime_error("Error opening file in File!!! " + file);
    auto kv = [&](const char* k, const auto& v) { f << std::setw(36) << (std::string(k) + " ") << std::setw(26) << v << "\n"; };
    f << std::left << "\n";
    kv("dt", dt); kv("thermalSize", thermalSize); kv("writeNumber", writeNumber);
    kv("measureNumberPerWrite", measureNumberPerWrite); kv("measureSkipStep", measureSkipStep);
    if (backPropagation) kv("backPropagationStep", backPropagationStep);
    f << "\n";
    kv("walkerSizePerThread", walkerSizePerThread); f << "\n";
    kv("decompType", decompType); kv("forceType", forceType); kv("forceCap", forceCap);
    kv("initialPhiTFlag", initialPhiTFlag); kv("initialWalkerFlag", initialWalkerFlag); f << "\n";
    kv("mgsStep", mgsStep); kv("popControlStep", popControlStep); f << "\n";
    kv("ET", ET); kv("ETAdjustStep", ETAdjustStep); kv("ETAdjustMaxSize", ETAdjustMaxSize); f << "\n";
    kv("seed", seed); f << "\n";
  }
  // afqmcConstraintPathMethod.cpp:182-229, as errors instead of exit(1)
  void analysis() const {
    if (dt < 0.0) throw std::runtime_error("dt must be postive!");
    if (measureNumberPerWrite == 0 || measureSkipStep == 0) throw std::runtime_error("code will not measure anything!");
    if (walkerSizePerThread <= 0) throw std::runtime_error("walkerSizePerThread is not a positive number!");
    if (forceCap < 0.0) throw std::runtime_error("forceCap must be positive or zero!");
    if (ETAdjustMaxSize > thermalSize) throw std::runtime_error("We should not adjust ET after thermalizing!");
  }
  int decomp() const {
    if (decompType == "densityCharge") return 0;
    if (decompType == "densitySpin") return 1;
    if (decompType == "hopCharge") return 2;
    if (decompType == "hopSpin") return 3;
    throw std::runtime_error("Error!!! Do not know decompType " + decompType);
  }
  int force() const {
    if (forceType == "dynamicForce") return 0;
    if (forceType == "constForce") return 1;
    throw std::runtime_error("Error!!! Do not know forceType " + forceType);
  }
};

// The model and trial as the reference holds them after initialParameters / initialPhiT:
// K and the three Hop propagators of returnExpMinusAlphaK (2L x 2L, column-major), the on-site
// fields and the 2L x N trial.
struct HubbardInput {
  int L = 0, N = 0, Nup = 0;
  bool spinBlocks = false;
  std::vector<std::complex<double>> K, expMinusDtK, expMinusHalfDtK, expHalfDtK, phiT;
  std::vector<double> U, mu, hx, hy, hz, constForce;
};

struct BlockEnergy {
  std::complex<double> HNum, den, KNum, VNum;
};

class AfqmcConstraintPathGpu {
 public:
  AfqmcConstraintPathGpu(const ConstraintPathMethod& m, const HubbardInput& in, int precision = 0, int device = 0)
      : method(m), model(in) {
    method.analysis();
    check(afqmc_ctx_create(device, &ctx), "afqmc_ctx_create");
    afqmc_cp_desc d{};
    d.L = in.L; d.N = in.N; d.Nup = in.Nup; d.spin_blocks = in.spinBlocks ? 1 : 0;
    d.precision = precision;
    d.decomp_type = method.decomp();
    d.force_type = method.force();
    d.dt = method.dt; d.force_cap = method.forceCap;
    d.K = dp(in.K); d.exp_minus_dt_k = dp(in.expMinusDtK);
    d.exp_minus_half_dt_k = dp(in.expMinusHalfDtK); d.exp_half_dt_k = dp(in.expHalfDtK);
    d.U = in.U.data(); d.mu = in.mu.data(); d.hx = in.hx.data(); d.hy = in.hy.data(); d.hz = in.hz.data();
    d.phiT = dp(in.phiT);
    d.const_force = in.constForce.empty() ? nullptr : in.constForce.data();
    d.max_walkers = method.walkerSizePerThread;
    check(afqmc_cp_create(ctx, &d, &cp), "afqmc_cp_create");
    check(afqmc_cp_walkers_create(cp, method.walkerSizePerThread, &walker), "afqmc_cp_walkers_create");
    check(afqmc_cp_set_seed(cp, (uint64_t)method.seed, 0), "afqmc_cp_set_seed");
    check(afqmc_cp_set_ET(cp, method.ET), "afqmc_cp_set_ET");
    if (method.backPropagation) {
      const int nbp = std::abs(method.dt) < 1e-12 ? 0 : (int)method.backPropagationStep;
      check(afqmc_bp_create(cp, method.walkerSizePerThread, nbp, (int)method.measureSkipStep, (int)method.mgsStep, &bp),
            "afqmc_bp_create");
    }
  }
  ~AfqmcConstraintPathGpu() {
    afqmc_bp_destroy(bp);
    afqmc_walkers_destroy(walker);
    afqmc_cp_destroy(cp);
    afqmc_ctx_destroy(ctx);
  }
  AfqmcConstraintPathGpu(const AfqmcConstraintPathGpu&) = delete;
  AfqmcConstraintPathGpu& operator=(const AfqmcConstraintPathGpu&) = delete;

  // afqmcConstraintPath.cpp:14-30 (phiT and the model are given)
  void run() {
    initialWalker();
    estimateMemory();
    if (std::abs(method.dt) < 1e-12) measureWithoutProjection();
    else measureWithProjection();
    prepareStop();
  }
  const std::vector<BlockEnergy>& energies() const { return blocks; }
  const ConstraintPathMethod& getMethod() const { return method; }

  // Initial walkers other than copies of phiT (initialWalkerFlag readFromFile, readAllWalkers,
  // setFromModel): nw walkers of 2L x N (column-major each), their log weights and alive flags.
  void setInitialWalkers(std::vector<std::complex<double>> wf, std::vector<std::complex<double>> logw,
                         std::vector<int> alive) {
    const size_t n = (size_t)2 * model.L * model.N, nw = (size_t)method.walkerSizePerThread;
    if (wf.size() != n * nw || logw.size() != nw || alive.size() != nw)
      throw std::runtime_error("setInitialWalkers: expected " + std::to_string(nw) + " walkers of 2L x N");
    initWf = std::move(wf);
    initLogw = std::move(logw);
    initAlive = std::move(alive);
  }
  // Blocks append the reference's measurement files and beta.dat in this directory.
  void setOutputDirectory(const std::string& dir) { outDir = dir; }
  // The walkers after prepareStop (projected back by exp(dt K / 2)), as writeWalkers sees them.
  void finalWalkers(std::vector<std::complex<double>>& wf, std::vector<std::complex<double>>& logw,
                    std::vector<int>& alive) const {
    wf = finalWf;
    logw = finalLogw;
    alive = finalAlive;
  }
  uint64_t counter() const {
    uint64_t c = 0;
    check(afqmc_cp_get_counter(cp, &c), "afqmc_cp_get_counter");
    return c;
  }

  // initialWalkerFlag = sampleFromPhiT: every walker = phiT (or the walkers given by
  // setInitialWalkers), then QR and comb (initialMgsAndPopControl, afqmcConstraintPathSetWalker.cpp)
  void initialWalker() {
    if (!initWf.empty()) {
      check(afqmc_walkers_upload(walker, dp(initWf), dp(initLogw), initAlive.data()), "afqmc_walkers_upload");
    } else {
      const int nw = method.walkerSizePerThread;
      const size_t n = model.phiT.size();
      std::vector<std::complex<double>> wf(n * nw);
      for (int i = 0; i < nw; ++i) std::copy(model.phiT.begin(), model.phiT.end(), wf.begin() + n * i);
      std::vector<std::complex<double>> logw(nw, 0.0);
      check(afqmc_walkers_upload(walker, dp(wf), dp(logw), nullptr), "afqmc_walkers_upload");
    }
    modifyGM();
    popControl();
  }
  // the aux sample drawn there consumes L uniforms (afqmcConstraintPath.cpp:66)
  void estimateMemory() { check(afqmc_cp_skip_uniforms(cp, (uint64_t)model.L), "afqmc_cp_skip_uniforms"); }

  // afqmcConstraintPath.cpp:97-141
  void measureWithProjection() {
    if (bp) return measureWithProjectionBP();
    projectExpMinusHalfDtK();
    size_t mgsIndex = 0, popControlIndex = 0;
    for (size_t i = 0; i < method.thermalSize; ++i) {
      if (i < method.ETAdjustMaxSize && i % method.ETAdjustStep == 0) {
        addMixedMeasurement();
        adjustETAndResetMeasurement();
      }
      projectOneStep(mgsIndex, popControlIndex);
    }
    for (size_t i = 0; i < method.writeNumber; ++i) {
      for (size_t j = 0; j < method.measureNumberPerWrite; ++j) {
        addMixedMeasurement();
        for (size_t k = 0; k < method.measureSkipStep; ++k) projectOneStep(mgsIndex, popControlIndex);
      }
      if (!outDir.empty()) {
        const double beta = (method.thermalSize + (i + 0.5) * method.measureNumberPerWrite * method.measureSkipStep - 0.5) *
                            method.dt;
        refio::append(outDir + "/beta.dat", beta);
      }
      writeAndResetMeasurement();
    }
  }
  // the back-propagation driver's loop (tutorials/afqmcConstraintPathBP/source/afqmcConstraintPath.cpp:115-180):
  // every step after thermalisation is recorded on the device; addMeasurement fires where the
  // reference's does, with its window bookkeeping (afqmcConstraintPathBackPropagation.cpp:59-76)
  void measureWithProjectionBP() {
    const size_t nbp = method.backPropagationStep, skip = method.measureSkipStep;
    projectExpMinusHalfDtK();
    size_t mgsIndex = 0, popControlIndex = 0;
    for (size_t i = 0; i < method.thermalSize; ++i) {
      if (i < method.ETAdjustMaxSize && i % method.ETAdjustStep == 0) {
        addMixedMeasurement();
        adjustETAndResetMeasurement();
      }
      projectOneStep(mgsIndex, popControlIndex);
    }
    prepareBackPropagation();
    if (nbp >= skip)
      for (size_t i = 0; i < nbp; ++i) projectOneStep(mgsIndex, popControlIndex);
    for (size_t i = 0; i < method.writeNumber; ++i) {
      for (size_t j = 0; j < method.measureNumberPerWrite; ++j) {
        for (size_t k = 0; k < skip; ++k) {
          if (BPIndex == nbp) addMeasurement();
          projectOneStep(mgsIndex, popControlIndex);
        }
        if (nbp < skip) prepareBackPropagation();
      }
      if (!outDir.empty()) {
        const double beta = (method.thermalSize + (i + 0.5) * method.measureNumberPerWrite * skip - 0.5) * method.dt;
        refio::append(outDir + "/beta.dat", beta);
      }
      writeAndResetMeasurement();
    }
  }
  // afqmcConstraintPath.cpp:89-96 (dt = 0); with back-propagation the measurement is the
  // full-Green-matrix one of the mixed pair (a window of no steps)
  void measureWithoutProjection() {
    projectExpMinusHalfDtK();
    if (bp) {
      check(afqmc_bp_begin(bp), "afqmc_bp_begin");
      check(afqmc_bp_measure(bp, walker), "afqmc_bp_measure");
    } else {
      addMixedMeasurement();
    }
    writeAndResetMeasurement();
  }
  // afqmcConstraintPath.cpp:149-155: the adjusted ET goes back into the method, the walkers are
  // projected back by exp(dt K / 2) and kept for writeWalkers
  void prepareStop() {
    check(afqmc_cp_get_ET(cp, &method.ET), "afqmc_cp_get_ET");
    check(afqmc_cp_onebody(cp, walker, 2), "afqmc_cp_onebody");
    const size_t n = (size_t)2 * model.L * model.N, nw = (size_t)method.walkerSizePerThread;
    finalWf.assign(n * nw, 0.0);
    finalLogw.assign(nw, 0.0);
    finalAlive.assign(nw, 0);
    check(afqmc_walkers_download(walker, reinterpret_cast<double*>(finalWf.data()),
                                 reinterpret_cast<double*>(finalLogw.data()), finalAlive.data()),
          "afqmc_walkers_download");
  }

 private:
  void projectExpMinusHalfDtK() { check(afqmc_cp_onebody(cp, walker, 1), "afqmc_cp_onebody"); }
  void projectExpMinusDtKExpMinusDtV() { check(afqmc_cp_step(cp, walker, nullptr, nullptr), "afqmc_cp_step"); }
  void projectOneStep(size_t& mgsIndex, size_t& popControlIndex) {
    if (recording) check(afqmc_bp_step(bp, walker), "afqmc_bp_step");
    else projectExpMinusDtKExpMinusDtV();
    if (++mgsIndex == method.mgsStep) { modifyGM(); mgsIndex = 0; }
    if (++popControlIndex == method.popControlStep) { popControl(); popControlIndex = 0; }
    if (isBP) ++BPIndex;
  }
  void modifyGM() { check(afqmc_cp_stabilize(cp, walker), "afqmc_cp_stabilize"); }
  void popControl() {
    if (recording) check(afqmc_bp_pop_control(bp, walker, nullptr), "afqmc_bp_pop_control");
    else check(afqmc_cp_pop_control(cp, walker, nullptr, nullptr, nullptr), "afqmc_cp_pop_control");
  }
  // afqmcConstraintPathBackPropagation.cpp:11-20: the device records from the first call on
  void prepareBackPropagation() {
    if (!recording) {
      check(afqmc_bp_begin(bp), "afqmc_bp_begin");
      recording = true;
    }
    isBP = true;
    BPIndex = 0;
  }
  // afqmcConstraintPathMeasure.cpp:41-72
  void addMeasurement() {
    if (!isBP) return;
    check(afqmc_bp_measure(bp, walker), "afqmc_bp_measure");
    if (method.backPropagationStep >= method.measureSkipStep) BPIndex = method.backPropagationStep - method.measureSkipStep;
    else isBP = false;
  }
  void addMixedMeasurement() { check(afqmc_cp_measure(cp, walker), "afqmc_cp_measure"); }
  void adjustETAndResetMeasurement() { check(afqmc_cp_adjust_ET(cp), "afqmc_cp_adjust_ET"); }
  void writeAndResetMeasurement() {
    if (bp) return writeAndResetMeasurementBP();
    double e[14];
    check(afqmc_cp_energy(cp, e), "afqmc_cp_energy");
    blocks.push_back({{e[2], e[3]}, {e[0], e[1]}, {e[4], e[5]}, {e[6], e[7]}});
    if (!outDir.empty()) {
      using cd = std::complex<double>;
      const size_t L = (size_t)model.L;
      std::vector<cd> site(4 * L);
      check(afqmc_cp_site_measure(cp, reinterpret_cast<double*>(site.data())), "afqmc_cp_site_measure");
      auto c = [&](int k) { return cd(e[2 * k], e[2 * k + 1]); };
      auto part = [&](int q) { return std::vector<cd>(site.begin() + q * L, site.begin() + (q + 1) * L); };
      cd splusTot = 0.0, sminusTot = 0.0;
      for (size_t i = 0; i < L; ++i) { splusTot += site[2 * L + i]; sminusTot += site[3 * L + i]; }
      refio::append(outDir + "/den.dat", c(0));
      refio::append(outDir + "/KNum.dat", c(2));
      refio::append(outDir + "/VNum.dat", c(3));
      refio::append(outDir + "/RNum.dat", c(4));
      refio::append(outDir + "/HNum.dat", c(1));
      refio::append(outDir + "/NupNum.dat", part(0));
      refio::append(outDir + "/NdnNum.dat", part(1));
      refio::append(outDir + "/SplusNum.dat", part(2));
      refio::append(outDir + "/SminusNum.dat", part(3));
      refio::append(outDir + "/NupTotNum.dat", c(5));
      refio::append(outDir + "/NdnTotNum.dat", c(6));
      refio::append(outDir + "/SplusTotNum.dat", splusTot);
      refio::append(outDir + "/SminusTotNum.dat", sminusTot);
    }
    check(afqmc_cp_reset_measure(cp), "afqmc_cp_reset_measure");
  }

  // HubbardSOCMeasureObserveSDSD::write (HubbardSOCMeasureObserveSDSD.cpp:47-57): den, HNum, then
  // KNum, VNum, RNum, then the observable matrices column-major, one complex per line
  void writeAndResetMeasurementBP() {
    using cd = std::complex<double>;
    const size_t L = (size_t)model.L;
    std::vector<cd> sc(5), g(4 * L * L), dd(4 * L * L), pm(L * L), mp(L * L), sp(L * L);
    auto d = [](std::vector<cd>& v) { return reinterpret_cast<double*>(v.data()); };
    check(afqmc_bp_get(bp, d(sc), d(g), d(dd), d(pm), d(mp), d(sp)), "afqmc_bp_get");
    blocks.push_back({sc[1], sc[0], sc[2], sc[3]});
    if (!outDir.empty()) {
      refio::append(outDir + "/den.dat", sc[0]);
      refio::append(outDir + "/HNum.dat", sc[1]);
      refio::append(outDir + "/KNum.dat", sc[2]);
      refio::append(outDir + "/VNum.dat", sc[3]);
      refio::append(outDir + "/RNum.dat", sc[4]);
      refio::append(outDir + "/greenMatrixNum.dat", g);
      refio::append(outDir + "/densityDensityNum.dat", dd);
      refio::append(outDir + "/splusSminusNum.dat", pm);
      refio::append(outDir + "/sminusSplusNum.dat", mp);
      refio::append(outDir + "/spairSpairNum.dat", sp);
    }
    check(afqmc_bp_reset(bp), "afqmc_bp_reset");
  }

  static const double* dp(const std::vector<std::complex<double>>& v) {
    return reinterpret_cast<const double*>(v.data());
  }

  ConstraintPathMethod method;
  const HubbardInput& model;
  afqmc_ctx ctx = nullptr;
  afqmc_cp cp = nullptr;
  afqmc_walkers walker = nullptr;
  afqmc_bp bp = nullptr;
  bool recording = false, isBP = false;
  size_t BPIndex = 0;
  std::vector<BlockEnergy> blocks;
  std::vector<std::complex<double>> initWf, initLogw, finalWf, finalLogw;
  std::vector<int> initAlive, finalAlive;
  std::string outDir;
};

}  // namespace afqmc_gpu
