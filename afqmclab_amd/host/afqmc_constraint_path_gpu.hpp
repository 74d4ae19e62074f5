// C++ host driver of the constraint-path walker loop on top of the C-ABI (include/afqmc_mi355x.h).
//
// AfqmcConstraintPathGpu keeps the member-function structure and the afqmc_param keys of the
// reference's AfqmcConstraintPath (tutorials/afqmcConstraintPath/include/afqmcConstraintPath.h:12-58,
// afqmcConstraintPathMethod.h:15-45): run() -> initialWalker -> estimateMemory ->
// measureWithProjection, the same projectOneStep / modifyGM / popControl / addMixedMeasurement /
// adjustETAndResetMeasurement cadences, with the per-walker loops replaced by one batched
// device call each.  Walkers, weights, ET and the random counter stay in HBM.
#pragma once
#include <complex>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/afqmc_mi355x.h"

namespace afqmc_gpu {

inline void check(int rc, const char* what) {
  if (rc != 0) throw std::runtime_error(std::string(what) + ": " + afqmc_last_error());
}

// AfqmcConstraintPathMethod's keys (afqmcConstraintPathMethod.cpp:19-46), read from an
// afqmc_param file of "key value" lines as readBySearchString does.
struct ConstraintPathMethod {
  double dt = 0.01;
  size_t thermalSize = 200, writeNumber = 60, measureNumberPerWrite = 2, measureSkipStep = 5;
  int walkerSizePerThread = 300;
  std::string decompType = "densitySpin", forceType = "dynamicForce";
  double forceCap = 1.5;
  size_t mgsStep = 10, popControlStep = 10;
  double ET = -10.0;
  size_t ETAdjustStep = 10, ETAdjustMaxSize = 100;
  long long seed = 985456376;

  template <typename T>
  static bool find(const std::string& file, const std::string& key, T& value) {
    std::ifstream in(file);
    if (!in) throw std::runtime_error("cannot open " + file);
    std::string line;
    while (std::getline(in, line)) {
      std::istringstream ss(line);
      std::string k;
      if (ss >> k && k == key) {
        ss >> value;
        return true;
      }
    }
    return false;
  }
  void read(const std::string& file) {
    find(file, "dt", dt);
    find(file, "thermalSize", thermalSize);
    find(file, "writeNumber", writeNumber);
    find(file, "measureNumberPerWrite", measureNumberPerWrite);
    find(file, "measureSkipStep", measureSkipStep);
    find(file, "walkerSizePerThread", walkerSizePerThread);
    find(file, "decompType", decompType);
    find(file, "forceType", forceType);
    find(file, "forceCap", forceCap);
    find(file, "mgsStep", mgsStep);
    find(file, "popControlStep", popControlStep);
    find(file, "ET", ET);
    find(file, "ETAdjustStep", ETAdjustStep);
    find(file, "ETAdjustMaxSize", ETAdjustMaxSize);
    find(file, "seed", seed);
    analysis();
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
  }
  ~AfqmcConstraintPathGpu() {
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
    measureWithProjection();
  }
  const std::vector<BlockEnergy>& energies() const { return blocks; }
  uint64_t counter() const {
    uint64_t c = 0;
    check(afqmc_cp_get_counter(cp, &c), "afqmc_cp_get_counter");
    return c;
  }

  // initialWalkerFlag = sampleFromPhiT: every walker = phiT, then QR and comb (initialMgsAndPopControl)
  void initialWalker() {
    const int nw = method.walkerSizePerThread;
    const size_t n = model.phiT.size();
    std::vector<std::complex<double>> wf(n * nw);
    for (int i = 0; i < nw; ++i) std::copy(model.phiT.begin(), model.phiT.end(), wf.begin() + n * i);
    std::vector<std::complex<double>> logw(nw, 0.0);
    check(afqmc_walkers_upload(walker, dp(wf), dp(logw), nullptr), "afqmc_walkers_upload");
    modifyGM();
    popControl();
  }
  // the aux sample drawn there consumes L uniforms (afqmcConstraintPath.cpp:66)
  void estimateMemory() { check(afqmc_cp_skip_uniforms(cp, (uint64_t)model.L), "afqmc_cp_skip_uniforms"); }

  // afqmcConstraintPath.cpp:97-141
  void measureWithProjection() {
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
      writeAndResetMeasurement();
    }
  }

 private:
  void projectExpMinusHalfDtK() { check(afqmc_cp_onebody(cp, walker, 1), "afqmc_cp_onebody"); }
  void projectExpMinusDtKExpMinusDtV() { check(afqmc_cp_step(cp, walker, nullptr, nullptr), "afqmc_cp_step"); }
  void projectOneStep(size_t& mgsIndex, size_t& popControlIndex) {
    projectExpMinusDtKExpMinusDtV();
    if (++mgsIndex == method.mgsStep) { modifyGM(); mgsIndex = 0; }
    if (++popControlIndex == method.popControlStep) { popControl(); popControlIndex = 0; }
  }
  void modifyGM() { check(afqmc_cp_stabilize(cp, walker), "afqmc_cp_stabilize"); }
  void popControl() { check(afqmc_cp_pop_control(cp, walker, nullptr, nullptr, nullptr), "afqmc_cp_pop_control"); }
  void addMixedMeasurement() { check(afqmc_cp_measure(cp, walker), "afqmc_cp_measure"); }
  void adjustETAndResetMeasurement() { check(afqmc_cp_adjust_ET(cp), "afqmc_cp_adjust_ET"); }
  void writeAndResetMeasurement() {
    double e[14];
    check(afqmc_cp_energy(cp, e), "afqmc_cp_energy");
    blocks.push_back({{e[2], e[3]}, {e[0], e[1]}, {e[4], e[5]}, {e[6], e[7]}});
    check(afqmc_cp_reset_measure(cp), "afqmc_cp_reset_measure");
  }

  static const double* dp(const std::vector<std::complex<double>>& v) {
    return reinterpret_cast<const double*>(v.data());
  }

  ConstraintPathMethod method;
  const HubbardInput& model;
  afqmc_ctx ctx = nullptr;
  afqmc_cp cp = nullptr;
  afqmc_walkers walker = nullptr;
  std::vector<BlockEnergy> blocks;
};

}  // namespace afqmc_gpu
