// afqmc_cp_run: the constraint-path driver as a C++ program on the GPU library.
//
// Usage: afqmc_cp_run [--bp] <dir> [precision]
//
// --bp runs the back-propagation driver of tutorials/afqmcConstraintPathBP on the same directory:
// afqmc_param also holds backPropagationStep, and the measurement files are those of
// HubbardSOCMeasureObserveSDSD (den, HNum, KNum, VNum, RNum, greenMatrixNum, densityDensityNum,
// splusSminusNum, sminusSplusNum, spairSpairNum).
//
// A run directory in the reference's own formats (afqmcConstraintPath.cpp:14-60,
// afqmcConstraintPathSetWalker.cpp), used when <dir>/model_param exists:
//   afqmc_param         "key value" method file (afqmcConstraintPathMethod.cpp:19-46)
//   model_param         HubbardSOC text: L, N, K, mu, hx, hy, hz, U (HubbardSOC.cpp:39-59)
//   constForce_param    optional, L doubles (NiupNidn.cpp:56-64)
//   phiT.dat, phi.dat, walkers/phi_<i>.dat   SD text files, read or written as the flags say
// K's eigendecomposition gives exp(-alpha K) (HubbardSOC.cpp:97-118) and the setFromModel trial
// (its first N eigenvectors, HubbardSOCSDOperation.cpp:19-30).  When K, hx, hy and the trial do not
// mix the spins (the trial's first columns up only, the rest down only), the walkers are held as
// two spin blocks.  Outputs as the reference writes them: the measurement files (den.dat,
// HNum.dat, ...), beta.dat, afqmc_param with the adjusted ET, walkers/phi_<i>.dat.  The
// setRandomly flags draw SPRNG numbers in the reference and are not supported.
//
// Otherwise the directory holds the set-up made by the Python side:
//   <dir>/afqmc_param, <dir>/model_dims "L N Nup spinBlocks", <dir>/*.bin raw little-endian arrays:
//   K, expMinusDtK, expMinusHalfDtK, expHalfDtK (2L x 2L complex), phiT (2L x N complex), U, mu,
//   hx, hy, hz (L doubles); it writes <dir>/HNum.dat and <dir>/den.dat.
#include <sys/stat.h>

#include <cstdio>
#include <fstream>
#include <iomanip>
#include <iostream>

#include "afqmc_constraint_path_gpu.hpp"
#include "hermitian_eigen.hpp"

namespace {

using cd = std::complex<double>;
namespace refio = afqmc_gpu::refio;

template <typename T>
std::vector<T> read_bin(const std::string& path, size_t n) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot open " + path);
  std::vector<T> v(n);
  f.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(n * sizeof(T)));
  if ((size_t)f.gcount() != n * sizeof(T)) throw std::runtime_error("short read: " + path);
  return v;
}

// number of leading up-only columns if the 2L x N matrix is [up block | down block], else -1
int spin_block_split(const std::vector<cd>& phi, int L, int N) {
  auto zero_rows = [&](int j, int r0, int r1) {
    for (int i = r0; i < r1; ++i)
      if (phi[i + (size_t)j * 2 * L] != 0.0) return false;
    return true;
  };
  int nup = 0;
  while (nup < N && zero_rows(nup, L, 2 * L)) ++nup;
  for (int j = nup; j < N; ++j)
    if (!zero_rows(j, 0, L)) return -1;
  return nup;
}

bool spin_diagonal_model(const refio::HubbardModelParam& m) {
  const size_t L = m.L;
  for (size_t j = 0; j < 2 * L; ++j)
    for (size_t i = 0; i < 2 * L; ++i)
      if ((i < L) != (j < L) && m.K[i + j * 2 * L] != 0.0) return false;
  for (size_t i = 0; i < L; ++i)
    if (m.hx[i] != 0.0 || m.hy[i] != 0.0) return false;
  return true;
}

refio::SDFile read_sd_checked(const std::string& path, size_t rows, size_t cols) {
  refio::SDFile s = refio::read_sd(path);
  if (s.rows != rows || s.cols != cols)
    throw std::runtime_error(path + ": expected a " + std::to_string(rows) + " x " + std::to_string(cols) +
                             " wave function, found " + std::to_string(s.rows) + " x " + std::to_string(s.cols));
  return s;
}

int run_reference_dir(const std::string& dir, int precision, bool back_propagation) {
  afqmc_gpu::ConstraintPathMethod method;
  method.backPropagation = back_propagation;
  method.read(dir + "/afqmc_param");
  const refio::HubbardModelParam m = refio::read_model_param(dir + "/model_param");
  const int L = (int)m.L, N = (int)m.N, L2 = 2 * L;
  std::vector<double> kval;
  std::vector<cd> kvec;
  afqmc_gpu::hermitian_eigen(L2, m.K, kval, kvec);

  afqmc_gpu::HubbardInput in;
  in.L = L; in.N = N;
  in.K = m.K;
  in.expMinusDtK = afqmc_gpu::exp_minus_alpha(L2, kval, kvec, method.dt);
  in.expMinusHalfDtK = afqmc_gpu::exp_minus_alpha(L2, kval, kvec, method.dt * 0.5);
  in.expHalfDtK = afqmc_gpu::exp_minus_alpha(L2, kval, kvec, -method.dt * 0.5);
  in.U = m.U; in.mu = m.mu; in.hx = m.hx; in.hy = m.hy; in.hz = m.hz;
  in.constForce = refio::read_force(dir + "/constForce_param", m.L);

  // the model's trial: the lowest N eigenvectors of K
  refio::SDFile fromModel;
  fromModel.rows = L2; fromModel.cols = N;
  fromModel.wf.assign(kvec.begin(), kvec.begin() + (size_t)L2 * N);

  refio::SDFile phiT;
  if (method.initialPhiTFlag == "setFromModel") {
    phiT = fromModel;
    refio::write_sd(dir + "/phiT.dat", phiT);
  } else if (method.initialPhiTFlag == "readFromFile") {
    phiT = read_sd_checked(dir + "/phiT.dat", L2, N);
  } else if (method.initialPhiTFlag == "setRandomly") {
    throw std::runtime_error("initialPhiTFlag setRandomly draws SPRNG numbers; give phiT.dat with readFromFile");
  } else {
    throw std::runtime_error("Error!!! Do not recognize initialPhiTFlag!");
  }
  in.phiT = phiT.wf;
  const int nup = spin_block_split(phiT.wf, L, N);
  in.spinBlocks = nup >= 0 && spin_diagonal_model(m);
  in.Nup = in.spinBlocks ? nup : 0;

  const int nw = method.walkerSizePerThread;
  const size_t per = (size_t)L2 * N;
  std::vector<cd> wf, logw;
  std::vector<int> alive;
  auto replicate = [&](const refio::SDFile& s) {
    wf.resize(per * nw);
    for (int i = 0; i < nw; ++i) std::copy(s.wf.begin(), s.wf.end(), wf.begin() + per * i);
    logw.assign(nw, s.logw);
    alive.assign(nw, 1);
  };
  if (method.initialWalkerFlag == "setFromModel") {
    refio::write_sd(dir + "/phi.dat", fromModel);
    replicate(fromModel);
  } else if (method.initialWalkerFlag == "readFromFile") {
    replicate(read_sd_checked(dir + "/phi.dat", L2, N));
  } else if (method.initialWalkerFlag == "readAllWalkers") {
    // a missing file leaves a dead walker (the reference fills it randomly and marks it dead)
    wf.resize(per * nw);
    logw.assign(nw, 0.0);
    alive.assign(nw, 0);
    for (int i = 0; i < nw; ++i) {
      const std::string f = dir + "/walkers/phi_" + std::to_string(i) + ".dat";
      if (refio::file_exists(f)) {
        refio::SDFile s = read_sd_checked(f, L2, N);
        std::copy(s.wf.begin(), s.wf.end(), wf.begin() + per * i);
        logw[i] = s.logw;
        alive[i] = 1;
      } else {
        std::copy(phiT.wf.begin(), phiT.wf.end(), wf.begin() + per * i);
      }
    }
  } else if (method.initialWalkerFlag == "setRandomly") {
    throw std::runtime_error("initialWalkerFlag setRandomly draws SPRNG numbers; give phi.dat with readFromFile");
  } else if (method.initialWalkerFlag != "sampleFromPhiT") {
    throw std::runtime_error("Error!!! Do not recognize initialWalkerFlag!");
  }
  // spin blocks need walkers without spin-mixing entries
  if (in.spinBlocks && !wf.empty()) {
    for (int i = 0; i < nw && in.spinBlocks; ++i) {
      std::vector<cd> one(wf.begin() + per * i, wf.begin() + per * (i + 1));
      if (spin_block_split(one, L, N) != nup) in.spinBlocks = false;
    }
    if (!in.spinBlocks) in.Nup = 0;
  }

  afqmc_gpu::AfqmcConstraintPathGpu afqmc(method, in, precision);
  if (!wf.empty()) afqmc.setInitialWalkers(wf, logw, alive);
  afqmc.setOutputDirectory(dir);
  afqmc.run();

  // prepareStop: afqmc_param with the adjusted ET, then writeWalkers
  afqmc.getMethod().write(dir + "/afqmc_param");
  std::vector<cd> fwf, flogw;
  std::vector<int> falive;
  afqmc.finalWalkers(fwf, flogw, falive);
  const std::string wdir = dir + "/walkers";
  mkdir(wdir.c_str(), 0755);
  for (int i = 0; i < nw; ++i) {
    const std::string f = wdir + "/phi_" + std::to_string(i) + ".dat";
    if (falive[i]) {
      refio::SDFile s;
      s.logw = flogw[i];
      s.rows = L2; s.cols = N;
      s.wf.assign(fwf.begin() + per * i, fwf.begin() + per * (i + 1));
      refio::write_sd(f, s);
    } else {
      std::remove(f.c_str());
    }
  }
  std::cout << "blocks " << afqmc.energies().size() << " counter " << afqmc.counter() << " spinBlocks "
            << (in.spinBlocks ? 1 : 0) << std::endl;
  return 0;
}

int run_bin_dir(const std::string& dir, int precision) {
  afqmc_gpu::ConstraintPathMethod method;
  method.read(dir + "/afqmc_param");
  afqmc_gpu::HubbardInput in;
  {
    std::ifstream f(dir + "/model_dims");
    int sb = 0;
    if (!(f >> in.L >> in.N >> in.Nup >> sb)) throw std::runtime_error("bad model_dims");
    in.spinBlocks = sb != 0;
  }
  const size_t L = in.L, L2 = 4 * L * L;
  in.K = read_bin<cd>(dir + "/K.bin", L2);
  in.expMinusDtK = read_bin<cd>(dir + "/expMinusDtK.bin", L2);
  in.expMinusHalfDtK = read_bin<cd>(dir + "/expMinusHalfDtK.bin", L2);
  in.expHalfDtK = read_bin<cd>(dir + "/expHalfDtK.bin", L2);
  in.phiT = read_bin<cd>(dir + "/phiT.bin", 2 * L * in.N);
  in.U = read_bin<double>(dir + "/U.bin", L);
  in.mu = read_bin<double>(dir + "/mu.bin", L);
  in.hx = read_bin<double>(dir + "/hx.bin", L);
  in.hy = read_bin<double>(dir + "/hy.bin", L);
  in.hz = read_bin<double>(dir + "/hz.bin", L);

  afqmc_gpu::AfqmcConstraintPathGpu afqmc(method, in, precision);
  afqmc.run();
  std::ofstream h(dir + "/HNum.dat"), d(dir + "/den.dat");
  h << std::setprecision(17);
  d << std::setprecision(17);
  for (const auto& b : afqmc.energies()) {
    h << b.HNum.real() << " " << b.HNum.imag() << "\n";
    d << b.den.real() << " " << b.den.imag() << "\n";
  }
  std::cout << "blocks " << afqmc.energies().size() << " counter " << afqmc.counter() << std::endl;
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  std::vector<std::string> args(argv + 1, argv + argc);
  bool back_propagation = false;
  if (!args.empty() && args[0] == "--bp") {
    back_propagation = true;
    args.erase(args.begin());
  }
  if (args.empty()) {
    std::cerr << "usage: afqmc_cp_run [--bp] <dir> [precision]\n";
    return 2;
  }
  const std::string dir = args[0];
  const int precision = args.size() > 1 ? std::atoi(args[1].c_str()) : 0;
  try {
    if (refio::file_exists(dir + "/model_param")) return run_reference_dir(dir, precision, back_propagation);
    if (back_propagation) throw std::runtime_error("--bp needs a reference run directory (model_param)");
    return run_bin_dir(dir, precision);
  } catch (const std::exception& e) {
    std::cerr << "afqmc_cp_run: " << e.what() << std::endl;
    return 1;
  }
}
