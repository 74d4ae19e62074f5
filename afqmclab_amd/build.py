"""Build libafqmc_mi355x.so in-tree with hipcc for gfx950.

Usage: python -m afqmclab_amd.build   (or __graft_entry__.build()).
Every source under afqmclab_amd/csrc is compiled with hipcc (in parallel) and
linked into afqmclab_amd/libafqmc_mi355x.so, which is what the Python
bindings, the tests and bench.py load.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libafqmc_mi355x.so")
HOST = os.path.join(HERE, "host")
CP_RUN = os.path.join(HERE, "afqmc_cp_run")
PH_RUN = os.path.join(HERE, "afqmc_ph_run")
# the HDF5 C library for the phaseless driver's model_param (the reference's RealMaterialMolecule file)
HDF5_PREFIX = os.environ.get("AFQMC_HDF5_PREFIX", "/opt/conda")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("AFQMC_OFFLOAD_ARCH", "gfx950")
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # keep hipcc from emitting packed FP32 math (v_pk_*_f32)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def sources():
    out = []
    for f in sorted(os.listdir(CSRC)):
        if f.endswith(".hip") or f.endswith(".cpp"):
            out.append(os.path.join(CSRC, f))
    return out


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [HIPCC, *FLAGS, "-x", "hip", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def build(verbose=True, jobs=None):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    # C++ host driver on the C-ABI (plain g++: host code only, linked against the library)
    # (hipSOLVER for the one-time eigendecomposition of K when it reads a reference run directory)
    rocm = os.path.dirname(os.path.dirname(HIPCC))
    cmd = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-Wall", "-D__HIP_PLATFORM_AMD__",
           "-I", os.path.join(HERE, "..", "include"), "-I", os.path.join(rocm, "include"),
           os.path.join(HOST, "cp_run.cpp"), "-o", CP_RUN + ".tmp", "-L", HERE, "-lafqmc_mi355x",
           "-L", os.path.join(rocm, "lib"), "-lhipsolver", "-lamdhip64",
           "-Wl,-rpath,$ORIGIN", "-Wl,-rpath," + os.path.join(rocm, "lib")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"host driver build failed:\n{r.stderr}")
    os.replace(CP_RUN + ".tmp", CP_RUN)
    built = [CP_RUN]
    # phaseless driver: opens the HDF5 C library at run time for model_param; skipped (with a note)
    # when the headers and library are absent
    h5inc, h5lib = os.path.join(HDF5_PREFIX, "include"), os.path.join(HDF5_PREFIX, "lib")
    if os.path.exists(os.path.join(h5inc, "hdf5.h")) and os.path.exists(os.path.join(h5lib, "libhdf5.so")):
        cmd = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-Wall", "-D__HIP_PLATFORM_AMD__",
               "-I", os.path.join(HERE, "..", "include"), "-I", os.path.join(rocm, "include"), "-isystem", h5inc,
               f'-DAFQMC_HDF5_LIB="{os.path.realpath(os.path.join(h5lib, "libhdf5.so"))}"',
               os.path.join(HOST, "ph_run.cpp"), "-o", PH_RUN + ".tmp", "-L", HERE, "-lafqmc_mi355x",
               "-L", os.path.join(rocm, "lib"), "-lhipsolver", "-lamdhip64", "-ldl",
               "-Wl,-rpath,$ORIGIN", "-Wl,-rpath," + os.path.join(rocm, "lib")]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"phaseless host driver build failed:\n{r.stderr}")
        os.replace(PH_RUN + ".tmp", PH_RUN)
        built.append(PH_RUN)
    elif verbose:
        print(f"HDF5 not found under {HDF5_PREFIX}: afqmc_ph_run (HDF5 model_param) not built")
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources and {', '.join(built)}")
    return LIB


def clean():
    shutil.rmtree(OBJ, ignore_errors=True)


if __name__ == "__main__":
    build()
    sys.exit(0)
