// C++ host driver of the phaseless walker loop on top of the C-ABI (include/afqmc_mi355x.h).
//
// AfqmcPhaselessGpu keeps the member-function structure and the afqmc_param keys of the
// reference's AfqmcPhaseless (tutorials/afqmcPhaseless/include/afqmcPhaseless.h,
// afqmcPhaselessMethod.cpp): run() -> initialParameters -> initialPhiT -> initialWalker ->
// estimateMemory -> measureWithProjection -> prepareStop, the same projectOneStep / modifyGM /
// popControl / addMixedMeasurement / adjustETAndBackGroundThenResetMeasurement cadences and the
// background-adjust schedule (afqmcPhaseless.cpp:34-63, 128-205), with the per-walker loops
// replaced by one batched device call each.  The model is RealMaterialMolecule's (t, K, the real
// Cholesky vectors and the background), read from its HDF5 model_param; the one-body operators
// come from the eigendecomposition of Kp = K + sum_g bg_g L_g on the host
// (realMaterialMolecule.cpp:117-164).
#pragma once
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/afqmc_mi355x.h"
#include "hdf5_model.hpp"
#include "hermitian_eigen.hpp"
#include "reference_io.hpp"

namespace afqmc_gpu {

// AfqmcPhaselessMethod's keys and defaults (afqmcPhaselessMethod.cpp:14-36, 158-184)
struct PhaselessMethod {
  double dt = 0.01;
  size_t thermalSize = 200, writeNumber = 60, measureNumberPerWrite = 2, measureSkipStep = 5;
  int walkerSizePerThread = 300;
  std::string forceType = "dynamicForce";
  double forceCap = 1.5;
  std::string initialPhiTFlag = "readFromFile", initialWalkerFlag = "readFromFile";
  size_t mgsStep = 10, popControlStep = 10;
  double ET = -10.0;
  size_t ETAndBackGroundAdjustStep = 10, ETAndBackGroundAdjustMaxSize = 100;
  size_t ETAndBackGroundGrowthEstimateStep = 5, ETAndBackGroundGrowthEstimateMaxSize = 200;
  long long seed = 985456376;

  void read(const std::string& file) {
    using refio::search_value;
    search_value(file, "dt", dt);
    search_value(file, "thermalSize", thermalSize);
    search_value(file, "writeNumber", writeNumber);
    search_value(file, "measureNumberPerWrite", measureNumberPerWrite);
    search_value(file, "measureSkipStep", measureSkipStep);
    search_value(file, "walkerSizePerThread", walkerSizePerThread);
    search_value(file, "forceType", forceType);
    search_value(file, "forceCap", forceCap);
    search_value(file, "initialPhiTFlag", initialPhiTFlag);
    search_value(file, "initialWalkerFlag", initialWalkerFlag);
    search_value(file, "mgsStep", mgsStep);
    search_value(file, "popControlStep", popControlStep);
    search_value(file, "ET", ET);
    search_value(file, "ETAndBackGroundAdjustStep", ETAndBackGroundAdjustStep);
    search_value(file, "ETAndBackGroundAdjustMaxSize", ETAndBackGroundAdjustMaxSize);
    search_value(file, "ETAndBackGroundGrowthEstimateStep", ETAndBackGroundGrowthEstimateStep);
    search_value(file, "ETAndBackGroundGrowthEstimateMaxSize", ETAndBackGroundGrowthEstimateMaxSize);
    search_value(file, "seed", seed);
    analysis();
  }
  // afqmcPhaselessMethod.cpp:38-77
  void write(const std::string& file) const {
    std::ofstream f(file, std::ios::out | std::ios::trunc);
    if (!f) throw std::runtime_error("Error opening file in File!!! " + file);
    auto kv = [&](const char* k, const auto& v) { f << std::setw(36) << (std::string(k) + " ") << std::setw(26) << v << "\n"; };
    f << std::left << "\n";
    kv("dt", dt); kv("thermalSize", thermalSize); kv("writeNumber", writeNumber);
    kv("measureNumberPerWrite", measureNumberPerWrite); kv("measureSkipStep", measureSkipStep); f << "\n";
    kv("walkerSizePerThread", walkerSizePerThread); f << "\n";
    kv("forceType", forceType); kv("forceCap", forceCap);
    kv("initialPhiTFlag", initialPhiTFlag); kv("initialWalkerFlag", initialWalkerFlag); f << "\n";
    kv("mgsStep", mgsStep); kv("popControlStep", popControlStep); f << "\n";
    kv("ET", ET); kv("ETAndBackGroundAdjustStep", ETAndBackGroundAdjustStep);
    kv("ETAndBackGroundAdjustMaxSize", ETAndBackGroundAdjustMaxSize);
    kv("ETAndBackGroundGrowthEstimateStep", ETAndBackGroundGrowthEstimateStep);
    kv("ETAndBackGroundGrowthEstimateMaxSize", ETAndBackGroundGrowthEstimateMaxSize); f << "\n";
    kv("seed", seed); f << "\n";
  }
  // afqmcPhaselessMethod.cpp:187-239, as errors instead of exit(1)
  void analysis() const {
    if (std::abs(dt) < 1e-12) return;
    if (dt < 0.0) throw std::runtime_error("Error!!! dt must be postive!");
    if (measureNumberPerWrite == 0) throw std::runtime_error("Error!!! measureNumberPerWrite = 0, code will not measure anything!");
    if (measureSkipStep == 0) throw std::runtime_error("Error!!! measureSkipStep ==0, code will not measure anything!");
    if (walkerSizePerThread <= 0) throw std::runtime_error("Error!!! walkerSizePerThread is not a positive number!");
    if (forceCap < 0.0) throw std::runtime_error("Error!!! forceCap must be positive or zero!");
    if (ETAndBackGroundAdjustMaxSize > thermalSize)
      throw std::runtime_error("Error!!! We should not adjust ET and backGround after thermalizing!");
    if (ETAndBackGroundGrowthEstimateMaxSize > thermalSize)
      throw std::runtime_error("Error!!! We should not use growth estimator to adjust ET and backGround after thermalizing!");
  }
  int force() const {
    if (forceType == "dynamicForce") return 0;
    if (forceType == "constForce") return 1;
    throw std::runtime_error("Error!!! Do not know forceType " + forceType);
  }
};

struct PhaselessBlock {
  std::complex<double> HNum, den, TNum;
};

// SD2s (SD2s.cpp:103-123): logw, then the up and down wave functions as 2-D TensorHao
struct SD2sFile {
  std::complex<double> logw = 0.0;
  size_t L = 0, Nup = 0, Ndn = 0;
  std::vector<std::complex<double>> up, dn;
};
// SD2is (SD2is.cpp:105-127): logw, Nup, Ndn, then one L x max(Nup, Ndn) wave function
struct SD2isFile {
  std::complex<double> logw = 0.0;
  size_t L = 0, Nup = 0, Ndn = 0;
  std::vector<std::complex<double>> wf;
};

inline SD2sFile read_sd2s(const std::string& path) {
  auto f = refio::open_in(path);
  SD2sFile s;
  refio::read_value(f, s.logw, "logw");
  for (int b = 0; b < 2; ++b) {
    size_t rank = 0, rows = 0, cols = 0;
    refio::read_value(f, rank, "tensor rank");
    if (rank != 2) throw std::runtime_error("Error!!! Dimen is not consistent! in " + path);
    refio::read_value(f, rows, "rows");
    refio::read_value(f, cols, "cols");
    refio::read_array(f, rows * cols, b ? s.dn : s.up, "wave function");
    s.L = rows;
    (b ? s.Ndn : s.Nup) = cols;
  }
  return s;
}
inline void write_sd2s(const std::string& path, const SD2sFile& s) {
  auto f = refio::open_out(path);
  refio::write_value(f, s.logw);
  f << 2 << "\n" << s.L << " " << s.Nup << " \n";
  refio::write_array(f, s.up);
  f << 2 << "\n" << s.L << " " << s.Ndn << " \n";
  refio::write_array(f, s.dn);
}
inline SD2isFile read_sd2is(const std::string& path) {
  auto f = refio::open_in(path);
  SD2isFile s;
  refio::read_value(f, s.logw, "logw");
  refio::read_value(f, s.Nup, "Nup");
  refio::read_value(f, s.Ndn, "Ndn");
  size_t rank = 0, rows = 0, cols = 0;
  refio::read_value(f, rank, "tensor rank");
  if (rank != 2) throw std::runtime_error("Error!!! Dimen is not consistent! in " + path);
  refio::read_value(f, rows, "rows");
  refio::read_value(f, cols, "cols");
  refio::read_array(f, rows * cols, s.wf, "wave function");
  s.L = rows;
  return s;
}
inline void write_sd2is(const std::string& path, const SD2isFile& s) {
  auto f = refio::open_out(path);
  refio::write_value(f, s.logw);
  f << std::setw(26) << s.Nup << "\n" << std::setw(26) << s.Ndn << "\n";
  f << 2 << "\n" << s.L << " " << std::max(s.Nup, s.Ndn) << " \n";
  refio::write_array(f, s.wf);
}

class AfqmcPhaselessGpu {
 public:
  using cd = std::complex<double>;

  // dir: the run directory (afqmc_param, model_param, optional constForce_param, phiT.dat, phi.dat,
  // walkers/); outputs are written there as the reference writes them in its working directory
  AfqmcPhaselessGpu(const std::string& dir, int precision = 0, int device = 0) : dir_(dir), precision_(precision) {
    check(afqmc_ctx_create(device, &ctx), "afqmc_ctx_create");
  }
  ~AfqmcPhaselessGpu() {
    afqmc_ph_walkers_destroy(walker);
    afqmc_ph_destroy(ph);
    afqmc_ctx_destroy(ctx);
  }
  AfqmcPhaselessGpu(const AfqmcPhaselessGpu&) = delete;
  AfqmcPhaselessGpu& operator=(const AfqmcPhaselessGpu&) = delete;

  // afqmcPhaseless.cpp:14-30
  void run() {
    initialParameters();
    initialPhiT();
    initialWalker();
    estimateMemory();
    if (std::abs(method.dt) < 1e-12) measureWithoutProjection();
    else measureWithProjection();
    prepareStop();
  }
  const std::vector<PhaselessBlock>& energies() const { return blocks; }
  const PhaselessMethod& getMethod() const { return method; }
  const std::vector<double>& background() const { return model.choleskyBg; }

 private:
  static void check(int rc, const char* what) {
    if (rc != 0) throw std::runtime_error(std::string(what) + ": " + afqmc_last_error());
  }
  std::string path(const std::string& f) const { return dir_ + "/" + f; }

  // afqmcPhaseless.cpp:32-63; the device propagator is created once the trial is known
  void initialParameters() {
    method.read(path("afqmc_param"));
    model = h5model::read_model_param(path("model_param"));
    L = (int)model.L;
    G = (int)model.choleskyNumber;
    if (model.Nup == 0 || model.Ndn == 0) throw std::runtime_error("the device path needs Nup > 0 and Ndn > 0");
    adjustable = false;
    numberOfGrowthMeasure = 0;
    if (method.ETAndBackGroundAdjustMaxSize > 0) adjustable = true;
    if (method.ETAndBackGroundAdjustMaxSize < method.ETAndBackGroundGrowthEstimateMaxSize) {
      numberOfGrowthMeasure = (method.ETAndBackGroundGrowthEstimateMaxSize - method.ETAndBackGroundAdjustMaxSize) /
                              method.ETAndBackGroundGrowthEstimateStep;
      if (numberOfGrowthMeasure > 0) adjustable = true;
    }
    // readForce("constForce_param"): G complex values, zeros when the file is absent (choleskyReal.cpp:47-55)
    constForce.assign(G, 0.0);
    if (refio::file_exists(path("constForce_param"))) {
      auto f = refio::open_in(path("constForce_param"));
      refio::read_array(f, (size_t)G, constForce, "constForce");
    }
    setKpEigen();
  }

  // Kp = K + sum_g bg_g L_g and its eigendecomposition (setKp / setKpEigenValueAndVector,
  // realMaterialMolecule.cpp:152-176)
  void setKpEigen() {
    const size_t L2 = (size_t)L * L;
    std::vector<cd> kp(L2);
    for (size_t e = 0; e < L2; ++e) {
      double v = model.K[e];
      for (int g = 0; g < G; ++g) v += model.choleskyBg[g] * model.choleskyVecs[e + L2 * g];
      kp[e] = v;
    }
    Kp = kp;
    hermitian_eigen(L, kp, KpValue, KpVector);
    // Kp is real symmetric: rotate each eigenvector to real (its largest element real positive
    // before the phase is removed), as the reference's real eigensolver returns them
    for (int k = 0; k < L; ++k) {
      cd* v = KpVector.data() + (size_t)k * L;
      int imax = 0;
      for (int i = 1; i < L; ++i)
        if (std::abs(v[i]) > std::abs(v[imax])) imax = i;
      const cd ph = std::abs(v[imax]) > 0.0 ? std::conj(v[imax]) / std::abs(v[imax]) : cd(1.0);
      for (int i = 0; i < L; ++i) v[i] = cd((v[i] * ph).real(), 0.0);
    }
  }
  // exp(-alpha Kp) (returnExpMinusAlphaK, :117-132) or -alpha Kp (returnLogExpMinusAlphaK, :134-146),
  // both with logw = alpha/2 sum_g bg_g^2
  void setOneBody(int which, double alpha, bool log) {
    const size_t L2 = (size_t)L * L;
    std::vector<cd> m(L2);
    if (log) {
      for (size_t e = 0; e < L2; ++e) m[e] = -alpha * Kp[e];
    } else {
      for (int j = 0; j < L; ++j)
        for (int i = 0; i < L; ++i) {
          cd s = 0.0;
          for (int k = 0; k < L; ++k)
            s += KpVector[i + (size_t)k * L] * std::exp(-alpha * KpValue[k]) * std::conj(KpVector[j + (size_t)k * L]);
          m[i + (size_t)j * L] = s;
        }
    }
    double b2 = 0.0;
    for (double b : model.choleskyBg) b2 += b * b;
    const double lw[2] = {alpha * 0.5 * b2, 0.0};
    check(afqmc_ph_set_onebody(ph, which, reinterpret_cast<const double*>(m.data()), lw, log ? 1 : 0),
          "afqmc_ph_set_onebody");
  }
  void setOneBodyAll(bool log) {
    setOneBody(0, method.dt, log);
    setOneBody(1, method.dt * 0.5, log);
    setOneBody(2, -method.dt * 0.5, log);
    logOneBody = log;
  }

  // afqmcPhaselessSetWalker.cpp:9-33, then the device propagator with the trial's half-rotated
  // Cholesky vectors (initWfDaggerCholeskyVecs, realMaterialMoleculeMeasureFixedSD2sSD2is.cpp:165-186)
  void initialPhiT() {
    const size_t Nup = model.Nup, Ndn = model.Ndn;
    if (method.initialPhiTFlag == "setFromModel") {
      phiT.L = L; phiT.Nup = Nup; phiT.Ndn = Ndn; phiT.logw = 0.0;
      phiT.up.assign(KpVector.begin(), KpVector.begin() + (size_t)L * Nup);
      phiT.dn.assign(KpVector.begin(), KpVector.begin() + (size_t)L * Ndn);
      write_sd2s(path("phiT.dat"), phiT);
    } else if (method.initialPhiTFlag == "readFromFile") {
      phiT = read_sd2s(path("phiT.dat"));
      if ((int)phiT.L != L || phiT.Nup != Nup || phiT.Ndn != Ndn)
        throw std::runtime_error("phiT.dat does not match the model's L, Nup, Ndn");
    } else if (method.initialPhiTFlag == "setRandomly") {
      throw std::runtime_error("initialPhiTFlag setRandomly draws SPRNG numbers in the reference; not supported");
    } else {
      throw std::runtime_error("Error!!! Do not recognize initialPhiTFlag!");
    }
    std::vector<cd> hup = halfRotate(phiT.up, (int)Nup), hdn = halfRotate(phiT.dn, (int)Ndn);
    afqmc_ph_desc d{};
    d.L = L; d.Nup = (int)Nup; d.Ndn = (int)Ndn; d.G = G;
    d.precision = precision_;
    d.force_type = method.force();
    d.dt = method.dt; d.force_cap = method.forceCap;
    d.t = model.t.data(); d.chol = model.choleskyVecs.data(); d.bg = model.choleskyBg.data();
    d.phiT_up = reinterpret_cast<const double*>(phiT.up.data());
    d.phiT_dn = reinterpret_cast<const double*>(phiT.dn.data());
    d.phiT_logw[0] = phiT.logw.real(); d.phiT_logw[1] = phiT.logw.imag();
    d.half_rot_up = reinterpret_cast<const double*>(hup.data());
    d.half_rot_dn = reinterpret_cast<const double*>(hdn.data());
    d.const_force = reinterpret_cast<const double*>(constForce.data());
    d.max_walkers = method.walkerSizePerThread;
    check(afqmc_ph_create(ctx, &d, &ph), "afqmc_ph_create");
    check(afqmc_ph_walkers_create(ph, method.walkerSizePerThread, &walker), "afqmc_ph_walkers_create");
    check(afqmc_ph_set_seed(ph, (uint64_t)method.seed, 0), "afqmc_ph_set_seed");
    check(afqmc_ph_set_ET(ph, method.ET), "afqmc_ph_set_ET");
    setOneBodyAll(adjustable);
  }
  // phi^H L_g for every g: N x L x G, column-major (N fastest)
  std::vector<cd> halfRotate(const std::vector<cd>& phi, int N) const {
    const size_t L2 = (size_t)L * L;
    std::vector<cd> out((size_t)N * L * G);
    for (int g = 0; g < G; ++g) {
      const double* v = model.choleskyVecs.data() + L2 * g;
      for (int j = 0; j < L; ++j)
        for (int a = 0; a < N; ++a) {
          cd s = 0.0;
          for (int i = 0; i < L; ++i) s += std::conj(phi[i + (size_t)a * L]) * v[i + (size_t)j * L];
          out[a + (size_t)N * (j + (size_t)L * g)] = s;
        }
    }
    return out;
  }

  // afqmcPhaselessSetWalker.cpp:35-100, then initialMgsAndPopControl
  void initialWalker() {
    const int nw = method.walkerSizePerThread;
    const size_t Nmax = std::max(model.Nup, model.Ndn), per = (size_t)L * Nmax;
    std::vector<cd> wf(per * nw), logw(nw, 0.0);
    std::vector<int> alive(nw, 1);
    const std::string& flag = method.initialWalkerFlag;
    auto fill_all = [&](const std::vector<cd>& w, cd lw) {
      for (int i = 0; i < nw; ++i) {
        std::copy(w.begin(), w.end(), wf.begin() + per * i);
        logw[i] = lw;
      }
    };
    if (flag == "sampleFromPhiT") {
      fill_all(model.Nup >= model.Ndn ? phiT.up : phiT.dn, phiT.logw);
    } else if (flag == "setFromModel") {
      SD2isFile s{0.0, (size_t)L, model.Nup, model.Ndn, std::vector<cd>(KpVector.begin(), KpVector.begin() + per)};
      write_sd2is(path("phi.dat"), s);
      fill_all(s.wf, 0.0);
    } else if (flag == "readFromFile") {
      SD2isFile s = read_sd2is(path("phi.dat"));
      if ((int)s.L != L || s.Nup != model.Nup || s.Ndn != model.Ndn || s.wf.size() != per)
        throw std::runtime_error("phi.dat does not match the model's L, Nup, Ndn");
      fill_all(s.wf, s.logw);
    } else if (flag == "readAllWalkers") {
      for (int i = 0; i < nw; ++i) {
        const std::string f = path("walkers/phi_" + std::to_string(i) + ".dat");
        if (refio::file_exists(f)) {
          SD2isFile s = read_sd2is(f);
          if (s.wf.size() != per) throw std::runtime_error(f + " does not match the model's L, Nup, Ndn");
          std::copy(s.wf.begin(), s.wf.end(), wf.begin() + per * i);
          logw[i] = s.logw;
        } else {
          // the reference fills an absent walker randomly and marks it dead: its content is never used
          std::copy(phiT.up.begin(), phiT.up.begin() + std::min(per, phiT.up.size()), wf.begin() + per * i);
          alive[i] = 0;
        }
      }
    } else if (flag == "setRandomly") {
      throw std::runtime_error("initialWalkerFlag setRandomly draws SPRNG numbers in the reference; not supported");
    } else {
      throw std::runtime_error("Error!!! Do not recognize initialWalkerFlag!");
    }
    check(afqmc_ph_walkers_upload(walker, reinterpret_cast<const double*>(wf.data()),
                                  reinterpret_cast<const double*>(logw.data()), alive.data()),
          "afqmc_ph_walkers_upload");
    modifyGM();
    popControl();
  }
  // the aux sample drawn there consumes 2 G uniforms (afqmcPhaseless.cpp:89-92)
  void estimateMemory() { check(afqmc_ph_skip_uniforms(ph, 2 * (uint64_t)G), "afqmc_ph_skip_uniforms"); }

  void measureWithoutProjection() {
    projectOneBody(1);
    addMixedMeasurement();
    writeAndResetMeasurement();
  }

  // afqmcPhaseless.cpp:128-205
  void measureWithProjection() {
    projectOneBody(1);
    size_t mgsIndex = 0, popControlIndex = 0;
    const PhaselessMethod& m = method;
    for (size_t i = 0; i < m.thermalSize; ++i) {
      if (i < m.ETAndBackGroundAdjustMaxSize && i % m.ETAndBackGroundAdjustStep == 0) {
        addMixedMeasurement();
        adjustETAndBackGroundThenResetMeasurement();
      }
      if (m.ETAndBackGroundAdjustMaxSize > 0 && i == m.ETAndBackGroundAdjustMaxSize - 1 && numberOfGrowthMeasure == 0)
        setOneBodyAll(false);
      if (numberOfGrowthMeasure > 0 && i < m.ETAndBackGroundGrowthEstimateMaxSize && i >= m.ETAndBackGroundAdjustMaxSize) {
        if ((i + 1 - m.ETAndBackGroundAdjustMaxSize) % m.ETAndBackGroundGrowthEstimateStep == 0) addMixedMeasurement();
        if (i == m.ETAndBackGroundGrowthEstimateMaxSize - 1) {
          adjustETAndBackGroundThenResetMeasurement();
          setOneBodyAll(false);
        }
      }
      projectOneStep(mgsIndex, popControlIndex);
    }
    for (size_t i = 0; i < m.writeNumber; ++i) {
      for (size_t j = 0; j < m.measureNumberPerWrite; ++j) {
        addMixedMeasurement();
        for (size_t k = 0; k < m.measureSkipStep; ++k) projectOneStep(mgsIndex, popControlIndex);
      }
      const double beta = (m.thermalSize + (i + 0.5) * m.measureNumberPerWrite * m.measureSkipStep - 0.5) * m.dt;
      refio::append(path("beta.dat"), beta);
      writeAndResetMeasurement();
    }
  }

  // afqmcPhaseless.cpp:207-219: afqmc_param with the adjusted ET, the background back into
  // model_param, the walkers projected back by exp(dt Kp / 2) and written
  void prepareStop() {
    check(afqmc_ph_get_ET(ph, &method.ET), "afqmc_ph_get_ET");
    method.write(path("afqmc_param"));
    h5model::write_background(path("model_param"), model.choleskyBg);
    projectOneBody(2);
    writeWalkers();
  }
  void writeWalkers() {
    const int nw = method.walkerSizePerThread;
    const size_t Nmax = std::max(model.Nup, model.Ndn), per = (size_t)L * Nmax;
    std::vector<cd> wf(per * nw), logw(nw);
    std::vector<int> alive(nw);
    check(afqmc_ph_walkers_download(walker, reinterpret_cast<double*>(wf.data()), reinterpret_cast<double*>(logw.data()),
                                    alive.data()),
          "afqmc_ph_walkers_download");
    const std::string wdir = path("walkers");
    std::string cmd = "rm -rf '" + wdir + "' && mkdir -p '" + wdir + "'";
    if (std::system(cmd.c_str()) != 0) throw std::runtime_error("cannot recreate " + wdir);
    for (int i = 0; i < nw; ++i) {
      if (!alive[i]) continue;
      SD2isFile s{logw[i], (size_t)L, model.Nup, model.Ndn,
                  std::vector<cd>(wf.begin() + per * i, wf.begin() + per * (i + 1))};
      write_sd2is(wdir + "/phi_" + std::to_string(i) + ".dat", s);
    }
  }

  void projectOneBody(int which) { check(afqmc_ph_onebody(ph, walker, which), "afqmc_ph_onebody"); }
  void projectOneStep(size_t& mgsIndex, size_t& popControlIndex) {
    check(afqmc_ph_step(ph, walker, nullptr, nullptr), "afqmc_ph_step");
    if (++mgsIndex == method.mgsStep) { modifyGM(); mgsIndex = 0; }
    if (++popControlIndex == method.popControlStep) { popControl(); popControlIndex = 0; }
  }
  void modifyGM() { check(afqmc_ph_stabilize(ph, walker), "afqmc_ph_stabilize"); }
  void popControl() { check(afqmc_ph_pop_control(ph, walker, nullptr, nullptr, nullptr), "afqmc_ph_pop_control"); }
  void addMixedMeasurement() { check(afqmc_ph_measure(ph, walker), "afqmc_ph_measure"); }

  // afqmcPhaselessMeasure.cpp:36-47: ET = Re(HNum / den), background = Re(choleskyBgNum / den)
  void adjustETAndBackGroundThenResetMeasurement() {
    double e[6];
    std::vector<cd> bgNum(G);
    check(afqmc_ph_energy(ph, e, reinterpret_cast<double*>(bgNum.data()), nullptr), "afqmc_ph_energy");
    const cd den(e[0], e[1]);
    for (int g = 0; g < G; ++g) model.choleskyBg[g] = (bgNum[g] / den).real();
    check(afqmc_ph_adjust_ET(ph), "afqmc_ph_adjust_ET");
    check(afqmc_ph_set_background(ph, model.choleskyBg.data()), "afqmc_ph_set_background");
    setKpEigen();
  }

  // RealMaterialMoleculeMeasureFixedSD2sSD2is::write (:118-125)
  void writeAndResetMeasurement() {
    double e[6];
    std::vector<cd> bgNum(G), exNum(G);
    check(afqmc_ph_energy(ph, e, reinterpret_cast<double*>(bgNum.data()), reinterpret_cast<double*>(exNum.data())),
          "afqmc_ph_energy");
    const cd den(e[0], e[1]), HNum(e[2], e[3]), TNum(e[4], e[5]);
    blocks.push_back({HNum, den, TNum});
    refio::append(path("den.dat"), den);
    refio::append(path("TNum.dat"), TNum);
    refio::append(path("choleskyBgNum.dat"), bgNum);
    refio::append(path("choleskyExNum.dat"), exNum);
    refio::append(path("HNum.dat"), HNum);
    check(afqmc_ph_reset_measure(ph), "afqmc_ph_reset_measure");
  }

  std::string dir_;
  int precision_ = 0;
  PhaselessMethod method;
  h5model::RealMaterialMoleculeParam model;
  int L = 0, G = 0;
  bool adjustable = false, logOneBody = false;
  size_t numberOfGrowthMeasure = 0;
  std::vector<cd> constForce, Kp, KpVector;
  std::vector<double> KpValue;
  SD2sFile phiT;
  afqmc_ctx ctx = nullptr;
  afqmc_ph ph = nullptr;
  afqmc_ph_walkers walker = nullptr;
  std::vector<PhaselessBlock> blocks;
};

}  // namespace afqmc_gpu
