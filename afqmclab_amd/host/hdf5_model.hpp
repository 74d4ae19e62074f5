// The real-materials model_param of the reference in HDF5 (RealMaterialMolecule::read / write /
// writeBackGround, realMaterialMolecule.cpp:41-94): datasets L, Nup, Ndn, choleskyNumber (one
// unsigned 64-bit integer each, read as NATIVE_ULONG, readWriteHdf5.cpp:10-16), t and K (L x L
// doubles, column-major), choleskyVecs (L x L x choleskyNumber doubles, column-major, vector g
// at offset g L^2) and choleskyBg (choleskyNumber doubles), each a one-dimensional dataset whose
// size is checked against the expected count (readWriteHdf5.cpp:42-48).  Written big-endian
// (STD_U64BE, IEEE_F64BE) as writeFile does (readWriteHdf5.cpp:59-96).
//
// The HDF5 C library is opened at run time (dlopen of AFQMC_HDF5_LIB, or the library path fixed at
// build time): linking it would put its directory, which also holds an older C++ runtime, on the
// program's library search path.  hdf5.h supplies the types and constants only.
#pragma once
#include <dlfcn.h>
#include <hdf5.h>

#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#ifndef AFQMC_HDF5_LIB
#define AFQMC_HDF5_LIB "libhdf5.so"
#endif

namespace afqmc_gpu {
namespace h5model {

struct RealMaterialMoleculeParam {
  size_t L = 0, Nup = 0, Ndn = 0, choleskyNumber = 0;
  std::vector<double> t, K;             // L x L
  std::vector<double> choleskyVecs;     // L x L x choleskyNumber
  std::vector<double> choleskyBg;       // choleskyNumber
};

namespace detail {

// hdf5.h's H5F_ACC_* macros call H5check()/H5open() (a link-time dependency); their values
constexpr unsigned kAccRdonly = 0x0000u, kAccRdwr = 0x0001u, kAccTrunc = 0x0002u;
constexpr hid_t kDefault = 0, kAll = 0;  // H5P_DEFAULT, H5S_ALL

// the HDF5 entry points and predefined type ids, resolved once
struct Api {
  herr_t (*open)(void);
  hid_t (*fopen)(const char*, unsigned, hid_t);
  hid_t (*fcreate)(const char*, unsigned, hid_t, hid_t);
  herr_t (*fclose)(hid_t);
  hid_t (*dopen)(hid_t, const char*, hid_t);
  hid_t (*dcreate)(hid_t, const char*, hid_t, hid_t, hid_t, hid_t, hid_t);
  herr_t (*dclose)(hid_t);
  herr_t (*dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void*);
  herr_t (*dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void*);
  hid_t (*dget_space)(hid_t);
  hssize_t (*sget_npoints)(hid_t);
  hid_t (*screate_simple)(int, const hsize_t*, const hsize_t*);
  herr_t (*sclose)(hid_t);
  htri_t (*lexists)(hid_t, const char*, hid_t);
  herr_t (*ldelete)(hid_t, const char*, hid_t);
  hid_t native_ulong, native_double, u64be, f64be;
};

template <typename F>
void sym(void* h, const char* name, F& f) {
  f = reinterpret_cast<F>(dlsym(h, name));
  if (!f) throw std::runtime_error(std::string("HDF5: symbol ") + name + " not found");
}

inline const Api& api() {
  static Api a = [] {
    const char* env = std::getenv("AFQMC_HDF5_LIB");
    const std::string lib = env ? env : AFQMC_HDF5_LIB;
    void* h = dlopen(lib.c_str(), RTLD_NOW | RTLD_LOCAL);
    if (!h) throw std::runtime_error("HDF5 library " + lib + " cannot be loaded: " + dlerror());
    Api x{};
    sym(h, "H5open", x.open);
    sym(h, "H5Fopen", x.fopen);
    sym(h, "H5Fcreate", x.fcreate);
    sym(h, "H5Fclose", x.fclose);
    sym(h, "H5Dopen2", x.dopen);
    sym(h, "H5Dcreate2", x.dcreate);
    sym(h, "H5Dclose", x.dclose);
    sym(h, "H5Dread", x.dread);
    sym(h, "H5Dwrite", x.dwrite);
    sym(h, "H5Dget_space", x.dget_space);
    sym(h, "H5Sget_simple_extent_npoints", x.sget_npoints);
    sym(h, "H5Screate_simple", x.screate_simple);
    sym(h, "H5Sclose", x.sclose);
    sym(h, "H5Lexists", x.lexists);
    sym(h, "H5Ldelete", x.ldelete);
    if (x.open() < 0) throw std::runtime_error("HDF5: H5open failed");
    auto type_id = [&](const char* name) {
      hid_t* p = reinterpret_cast<hid_t*>(dlsym(h, name));
      if (!p) throw std::runtime_error(std::string("HDF5: symbol ") + name + " not found");
      return *p;
    };
    x.native_ulong = type_id("H5T_NATIVE_ULONG_g");
    x.native_double = type_id("H5T_NATIVE_DOUBLE_g");
    x.u64be = type_id("H5T_STD_U64BE_g");
    x.f64be = type_id("H5T_IEEE_F64BE_g");
    return x;
  }();
  return a;
}

struct Handle {
  hid_t id = -1;
  herr_t (*close)(hid_t) = nullptr;
  Handle(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
  ~Handle() {
    if (id >= 0 && close) close(id);
  }
  Handle(const Handle&) = delete;
  Handle& operator=(const Handle&) = delete;
};

inline size_t dataset_size(hid_t ds, const std::string& name) {
  Handle space(api().dget_space(ds), api().sclose);
  if (space.id < 0) throw std::runtime_error("HDF5: no dataspace for " + name);
  const hssize_t n = api().sget_npoints(space.id);
  if (n < 0) throw std::runtime_error("HDF5: bad dataspace for " + name);
  return (size_t)n;
}

// readFile(size_t L, T* data, file, name): the dataset must hold exactly n values
inline void read_array(hid_t file, const std::string& name, hid_t memtype, size_t n, void* out) {
  Handle ds(api().dopen(file, name.c_str(), kDefault), api().dclose);
  if (ds.id < 0) throw std::runtime_error("HDF5: cannot open dataset " + name);
  if (dataset_size(ds.id, name) != n)
    throw std::runtime_error("Error, dataset size is not consistent with pointer size. (" + name + ")");
  if (api().dread(ds.id, memtype, kAll, kAll, kDefault, out) < 0)
    throw std::runtime_error("HDF5: cannot read dataset " + name);
}

inline void write_array(hid_t file, const std::string& name, hid_t filetype, hid_t memtype, size_t n,
                        const void* data) {
  if (api().lexists(file, name.c_str(), kDefault) > 0) api().ldelete(file, name.c_str(), kDefault);
  const hsize_t dims[1] = {(hsize_t)n};
  Handle space(api().screate_simple(1, dims, nullptr), api().sclose);
  Handle ds(api().dcreate(file, name.c_str(), filetype, space.id, kDefault, kDefault, kDefault),
            api().dclose);
  if (ds.id < 0) throw std::runtime_error("HDF5: cannot create dataset " + name);
  if (api().dwrite(ds.id, memtype, kAll, kAll, kDefault, data) < 0)
    throw std::runtime_error("HDF5: cannot write dataset " + name);
}

}  // namespace detail

inline RealMaterialMoleculeParam read_model_param(const std::string& path) {
  const detail::Api& a = detail::api();
  detail::Handle f(a.fopen(path.c_str(), detail::kAccRdonly, detail::kDefault), a.fclose);
  if (f.id < 0) throw std::runtime_error("HDF5: cannot open " + path);
  RealMaterialMoleculeParam m;
  unsigned long v = 0;
  detail::read_array(f.id, "L", a.native_ulong, 1, &v); m.L = v;
  detail::read_array(f.id, "Nup", a.native_ulong, 1, &v); m.Nup = v;
  detail::read_array(f.id, "Ndn", a.native_ulong, 1, &v); m.Ndn = v;
  detail::read_array(f.id, "choleskyNumber", a.native_ulong, 1, &v); m.choleskyNumber = v;
  const size_t L2 = m.L * m.L;
  m.t.resize(L2);
  m.K.resize(L2);
  m.choleskyVecs.resize(L2 * m.choleskyNumber);
  m.choleskyBg.resize(m.choleskyNumber);
  detail::read_array(f.id, "t", a.native_double, L2, m.t.data());
  detail::read_array(f.id, "K", a.native_double, L2, m.K.data());
  detail::read_array(f.id, "choleskyVecs", a.native_double, m.choleskyVecs.size(), m.choleskyVecs.data());
  detail::read_array(f.id, "choleskyBg", a.native_double, m.choleskyBg.size(), m.choleskyBg.data());
  return m;
}

inline void write_model_param(const std::string& path, const RealMaterialMoleculeParam& m) {
  const detail::Api& a = detail::api();
  detail::Handle f(a.fcreate(path.c_str(), detail::kAccTrunc, detail::kDefault, detail::kDefault), a.fclose);
  if (f.id < 0) throw std::runtime_error("HDF5: cannot create " + path);
  const unsigned long dims[4] = {m.L, m.Nup, m.Ndn, m.choleskyNumber};
  const char* names[4] = {"L", "Nup", "Ndn", "choleskyNumber"};
  for (int i = 0; i < 4; ++i) detail::write_array(f.id, names[i], a.u64be, a.native_ulong, 1, &dims[i]);
  detail::write_array(f.id, "t", a.f64be, a.native_double, m.t.size(), m.t.data());
  detail::write_array(f.id, "K", a.f64be, a.native_double, m.K.size(), m.K.data());
  detail::write_array(f.id, "choleskyVecs", a.f64be, a.native_double, m.choleskyVecs.size(),
                      m.choleskyVecs.data());
  detail::write_array(f.id, "choleskyBg", a.f64be, a.native_double, m.choleskyBg.size(),
                      m.choleskyBg.data());
}

// RealMaterialMolecule::writeBackGround (realMaterialMolecule.cpp:97-102): rewrites choleskyBg in place
inline void write_background(const std::string& path, const std::vector<double>& bg) {
  const detail::Api& a = detail::api();
  detail::Handle f(a.fopen(path.c_str(), detail::kAccRdwr, detail::kDefault), a.fclose);
  if (f.id < 0) throw std::runtime_error("HDF5: cannot open " + path);
  detail::write_array(f.id, "choleskyBg", a.f64be, a.native_double, bg.size(), bg.data());
}

}  // namespace h5model
}  // namespace afqmc_gpu
