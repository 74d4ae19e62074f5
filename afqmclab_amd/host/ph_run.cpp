// afqmc_ph_run: the phaseless driver (tutorials/afqmcPhaseless) as a C++ program on the GPU library.
//
// Usage:
//   afqmc_ph_run <dir> [precision]       run the phaseless loop in a reference run directory:
//       afqmc_param (afqmcPhaselessMethod keys), model_param (RealMaterialMolecule's HDF5 file:
//       L, Nup, Ndn, choleskyNumber, t, K, choleskyVecs, choleskyBg), optional constForce_param,
//       phiT.dat / phi.dat / walkers/phi_<i>.dat as initialPhiTFlag and initialWalkerFlag say.
//       Writes what the tutorial writes: den.dat, TNum.dat, choleskyBgNum.dat, choleskyExNum.dat,
//       HNum.dat and beta.dat per block, afqmc_param with the adjusted ET, the adjusted background
//       back into model_param, walkers/phi_<i>.dat (afqmcPhaseless.cpp:128-219).
//   afqmc_ph_run --write-model <dir>      write <dir>/model_param from raw arrays: <dir>/model_dims
//       "L Nup Ndn choleskyNumber" and little-endian doubles t.bin, K.bin (L x L), choleskyVecs.bin
//       (L x L x choleskyNumber), choleskyBg.bin (choleskyNumber), all column-major, as the
//       reference's molecule script writes its model_param (scripts/pyscf/molecule.py:122-143).
//   afqmc_ph_run --read-model <dir>       read <dir>/model_param and write the arrays back as
//       model_dims.read and *.read.bin (a check of the reader).
#include <cstdio>
#include <fstream>
#include <iostream>

#include "afqmc_phaseless_gpu.hpp"

namespace {

namespace h5m = afqmc_gpu::h5model;

std::vector<double> read_doubles(const std::string& path, size_t n) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot open " + path);
  std::vector<double> v(n);
  f.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(n * sizeof(double)));
  if ((size_t)f.gcount() != n * sizeof(double)) throw std::runtime_error("short read: " + path);
  return v;
}
void write_doubles(const std::string& path, const std::vector<double>& v) {
  std::ofstream f(path, std::ios::binary | std::ios::trunc);
  if (!f) throw std::runtime_error("cannot write " + path);
  f.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)(v.size() * sizeof(double)));
}

int write_model(const std::string& dir) {
  std::ifstream d(dir + "/model_dims");
  h5m::RealMaterialMoleculeParam m;
  if (!(d >> m.L >> m.Nup >> m.Ndn >> m.choleskyNumber)) throw std::runtime_error("cannot read " + dir + "/model_dims");
  const size_t L2 = m.L * m.L;
  m.t = read_doubles(dir + "/t.bin", L2);
  m.K = read_doubles(dir + "/K.bin", L2);
  m.choleskyVecs = read_doubles(dir + "/choleskyVecs.bin", L2 * m.choleskyNumber);
  m.choleskyBg = read_doubles(dir + "/choleskyBg.bin", m.choleskyNumber);
  h5m::write_model_param(dir + "/model_param", m);
  return 0;
}

int read_model(const std::string& dir) {
  h5m::RealMaterialMoleculeParam m = h5m::read_model_param(dir + "/model_param");
  std::ofstream d(dir + "/model_dims.read");
  d << m.L << " " << m.Nup << " " << m.Ndn << " " << m.choleskyNumber << "\n";
  write_doubles(dir + "/t.read.bin", m.t);
  write_doubles(dir + "/K.read.bin", m.K);
  write_doubles(dir + "/choleskyVecs.read.bin", m.choleskyVecs);
  write_doubles(dir + "/choleskyBg.read.bin", m.choleskyBg);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  try {
    if (argc >= 3 && std::string(argv[1]) == "--write-model") return write_model(argv[2]);
    if (argc >= 3 && std::string(argv[1]) == "--read-model") return read_model(argv[2]);
    if (argc < 2 || argv[1][0] == '-') {
      std::cerr << "usage: afqmc_ph_run <dir> [precision] | --write-model <dir> | --read-model <dir>\n";
      return 2;
    }
    const std::string dir = argv[1];
    const int precision = argc > 2 ? std::atoi(argv[2]) : 0;
    afqmc_gpu::AfqmcPhaselessGpu drv(dir, precision);
    drv.run();
    std::cout << std::setprecision(16);
    for (const auto& b : drv.energies()) std::cout << "block E = " << (b.HNum / b.den).real() << "\n";
    return 0;
  } catch (const std::exception& e) {
    std::cerr << e.what() << "\n";
    return 1;
  }
}
