// Text files in the formats the reference's constraint-path tutorial reads and writes, so that a run
// directory prepared for it (model_param, afqmc_param, phiT.dat, phi.dat, walkers/phi_i.dat,
// constForce_param) works unchanged and its outputs (HNum.dat, den.dat, ... beta.dat, walkers/)
// look the same.
//
//   scalars and arrays: whitespace separated; a complex number is "re im"
//     (readWriteFstream.h:12-60); written one value per line, 16-digit scientific, width 26
//   TensorHao: rank, the dimensions, then the elements column-major (tensor_core.h:278-300)
//   SD: logw, then its 2-D TensorHao wave function (SD.cpp:85-103)
//   HubbardSOC model_param: L, N, K (2L x 2L complex), mu, hx, hy, hz, U (HubbardSOC.cpp:39-74)
#pragma once
#include <complex>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <stdexcept>
#include <string>
#include <sys/stat.h>
#include <vector>

namespace afqmc_gpu {
namespace refio {

using cd = std::complex<double>;

inline bool file_exists(const std::string& path) {
  struct stat st;
  return stat(path.c_str(), &st) == 0 && S_ISREG(st.st_mode);
}

inline std::ifstream open_in(const std::string& path) {
  std::ifstream f(path);
  if (!f) throw std::runtime_error("Error opening file in File!!! " + path);
  return f;
}
inline std::ofstream open_out(const std::string& path, std::ios_base::openmode mode = std::ios::trunc) {
  std::ofstream f(path, std::ios::out | mode);
  if (!f) throw std::runtime_error("Error opening file in File!!! " + path);
  f << std::setprecision(16) << std::scientific;
  return f;
}

template <typename T>
void read_value(std::istream& in, T& v, const char* what) {
  if (!(in >> v)) throw std::runtime_error(std::string("cannot read ") + what);
}
inline void read_value(std::istream& in, cd& v, const char* what) {
  double re, im;
  if (!(in >> re >> im)) throw std::runtime_error(std::string("cannot read ") + what);
  v = cd(re, im);
}
template <typename T>
void read_array(std::istream& in, size_t n, std::vector<T>& v, const char* what) {
  v.resize(n);
  for (size_t i = 0; i < n; ++i) read_value(in, v[i], what);
}

inline void write_value(std::ostream& out, double v) { out << std::setw(26) << v << "\n"; }
inline void write_value(std::ostream& out, const cd& v) {
  out << std::setw(26) << v.real() << std::setw(26) << v.imag() << "\n";
}
template <typename T>
void write_array(std::ostream& out, const std::vector<T>& v) {
  for (const T& x : v) write_value(out, x);
}

// HubbardSOC's model_param
struct HubbardModelParam {
  size_t L = 0, N = 0;
  std::vector<cd> K;  // 2L x 2L, column-major
  std::vector<double> mu, hx, hy, hz, U;
};

inline HubbardModelParam read_model_param(const std::string& path) {
  auto f = open_in(path);
  HubbardModelParam m;
  read_value(f, m.L, "L");
  read_value(f, m.N, "N");
  read_array(f, 4 * m.L * m.L, m.K, "K");
  read_array(f, m.L, m.mu, "mu");
  read_array(f, m.L, m.hx, "hx");
  read_array(f, m.L, m.hy, "hy");
  read_array(f, m.L, m.hz, "hz");
  read_array(f, m.L, m.U, "U");
  return m;
}

inline void write_model_param(const std::string& path, const HubbardModelParam& m) {
  auto f = open_out(path);
  f << std::setw(26) << m.L << "\n" << std::setw(26) << m.N << "\n";
  write_array(f, m.K);
  write_array(f, m.mu);
  write_array(f, m.hx);
  write_array(f, m.hy);
  write_array(f, m.hz);
  write_array(f, m.U);
}

// SD: a Slater determinant with its log weight
struct SDFile {
  cd logw = 0.0;
  size_t rows = 0, cols = 0;
  std::vector<cd> wf;  // rows x cols, column-major
};

inline SDFile read_sd(const std::string& path) {
  auto f = open_in(path);
  SDFile s;
  read_value(f, s.logw, "logw");
  size_t rank = 0;
  read_value(f, rank, "tensor rank");
  if (rank != 2) throw std::runtime_error("Error!!! Dimen is not consistent! 2 " + std::to_string(rank) + " in " + path);
  read_value(f, s.rows, "rows");
  read_value(f, s.cols, "cols");
  read_array(f, s.rows * s.cols, s.wf, "wave function");
  return s;
}

inline void write_sd(const std::string& path, const SDFile& s) {
  auto f = open_out(path);
  write_value(f, s.logw);
  f << 2 << "\n" << s.rows << " " << s.cols << " \n";
  write_array(f, s.wf);
}

// append one complex value / an array of complex values (writeThreadSum with ios::app)
inline void append(const std::string& path, const cd& v) {
  auto f = open_out(path, std::ios::app);
  write_value(f, v);
}
inline void append(const std::string& path, const std::vector<cd>& v) {
  auto f = open_out(path, std::ios::app);
  write_array(f, v);
}
inline void append(const std::string& path, double v) {
  auto f = open_out(path, std::ios::app);
  write_value(f, v);
}

// constForce_param: L doubles, or zeros when the file is absent (NiupNidn.cpp:56-64)
inline std::vector<double> read_force(const std::string& path, size_t L) {
  std::vector<double> v(L, 0.0);
  if (!file_exists(path)) return v;
  auto f = open_in(path);
  read_array(f, L, v, "constForce");
  return v;
}

// readBySearchString: the value after the first whitespace token equal to `key`, with an optional
// "=" between them (readBySearchString.cpp:16-60); false when the key is absent
template <typename T>
bool search_value(const std::string& path, const std::string& key, T& value) {
  auto f = open_in(path);
  std::string tok;
  while (f >> tok) {
    if (tok != key) continue;
    std::string s;
    if (!(f >> s)) return false;
    if (s == "=" && !(f >> s)) return false;
    std::istringstream ss(s);
    T v;
    if (!(ss >> v)) throw std::runtime_error("cannot parse the value of " + key + " in " + path);
    value = v;
    return true;
  }
  return false;
}

}  // namespace refio
}  // namespace afqmc_gpu
