// afqmc_cp_run: the constraint-path driver as a C++ program on the GPU library.
//
// Usage: afqmc_cp_run <dir> [precision]
//   <dir>/afqmc_param   reference-format "key value" method file (afqmcConstraintPathMethod.cpp:19-46)
//   <dir>/model_dims    "L N Nup spinBlocks"
//   <dir>/*.bin         raw little-endian arrays: K, expMinusDtK, expMinusHalfDtK, expHalfDtK (2L x 2L
//                       complex), phiT (2L x N complex), U, mu, hx, hy, hz (L doubles)
// Writes <dir>/HNum.dat and <dir>/den.dat, one "re im" line per block, as the reference's
// writeAndResetMeasurement appends one value per block (afqmcConstraintPathMeasure.cpp:29-33).
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <iostream>

#include "afqmc_constraint_path_gpu.hpp"

namespace {

template <typename T>
std::vector<T> read_bin(const std::string& path, size_t n) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot open " + path);
  std::vector<T> v(n);
  f.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(n * sizeof(T)));
  if ((size_t)f.gcount() != n * sizeof(T)) throw std::runtime_error("short read: " + path);
  return v;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) {
    std::cerr << "usage: afqmc_cp_run <dir> [precision]\n";
    return 2;
  }
  const std::string dir = argv[1];
  const int precision = argc > 2 ? std::atoi(argv[2]) : 0;
  try {
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
    using cd = std::complex<double>;
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
  } catch (const std::exception& e) {
    std::cerr << "afqmc_cp_run: " << e.what() << std::endl;
    return 1;
  }
  return 0;
}
