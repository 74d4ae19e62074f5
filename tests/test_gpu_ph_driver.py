"""The phaseless C++ host driver (afqmclab_amd/host, built as afqmclab_amd/afqmc_ph_run) on a
reference run directory: afqmc_param, the HDF5 model_param of RealMaterialMolecule
(realMaterialMolecule.cpp:41-60) and the trial / walker files.  Per-block HNum/den, the adjusted
ET and the background written back into model_param agree with the oracle's phaseless driver on
the same uniform stream (1e-9 relative)."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle.drivers import AfqmcPhaseless as OraclePh, PhaselessMethod as OracleMethod  # noqa: E402
from oracle.models import synthetic_molecule  # noqa: E402
from oracle.rng import UniformStream  # noqa: E402

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "afqmclab_amd", "afqmc_ph_run")

needs_exe = pytest.mark.skipif(not os.path.exists(EXE), reason="afqmc_ph_run not built (no HDF5 library at build time)")


def write_model(d, om):
    """model_param through the driver's converter (the reference's layout, big-endian datasets)."""
    L, G = om.L, om.choleskyNumber
    with open(os.path.join(d, "model_dims"), "w") as f:
        f.write(f"{L} {om.Nup} {om.Ndn} {G}\n")
    for name, a in (("t", om.t), ("K", om.K), ("choleskyVecs", om.choleskyVecs), ("choleskyBg", om.choleskyBg)):
        np.asfortranarray(np.asarray(a, dtype=np.float64)).ravel(order="F").astype("<f8").tofile(
            os.path.join(d, name + ".bin"))
    r = subprocess.run([EXE, "--write-model", d], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr


def read_model_bg(d):
    r = subprocess.run([EXE, "--read-model", d], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    return np.fromfile(os.path.join(d, "choleskyBg.read.bin"))


def write_param(d, m):
    """afqmc_param as PhaselessMethod::write lays it out (afqmcPhaselessMethod.cpp:55-90): the key and a
    space left-aligned in 36 columns, the value in 26."""
    with open(os.path.join(d, "afqmc_param"), "w") as f:
        for k, v in m.items():
            f.write(f"{k + ' ':<36s}{str(v):<26s}\n")


def write_sd2s(path, up, dn, logw=0j):
    """SD2s::write (SD2s.cpp:114-123): logw, then the up and down TensorHao."""
    with open(path, "w") as f:
        f.write(f"{logw.real:26.16e}{logw.imag:26.16e}\n")
        for w in (up, dn):
            f.write(f"2\n{w.shape[0]} {w.shape[1]} \n")
            for z in np.asarray(w).ravel(order="F"):
                f.write(f"{z.real:26.16e}{z.imag:26.16e}\n")


def read_dat(path):
    v = np.loadtxt(path, ndmin=2)
    return v[:, 0] + 1j * v[:, 1]


@needs_exe
@pytest.mark.parametrize("schedule", ["fixed_background", "adjust_then_growth"])
def test_ph_run_matches_oracle(tmp_path, schedule):
    om = synthetic_molecule(L=10, Nup=3, Ndn=2, G=12, seed=7)
    m = dict(dt=0.02, thermalSize=20, writeNumber=3, measureNumberPerWrite=2, measureSkipStep=3,
             walkerSizePerThread=12, forceType="dynamicForce", forceCap=1.5, mgsStep=5, popControlStep=10,
             ET=-10.0, ETAndBackGroundAdjustStep=10, ETAndBackGroundAdjustMaxSize=0,
             ETAndBackGroundGrowthEstimateStep=5, ETAndBackGroundGrowthEstimateMaxSize=0, seed=77)
    if schedule == "adjust_then_growth":
        m.update(ETAndBackGroundAdjustStep=5, ETAndBackGroundAdjustMaxSize=10,
                 ETAndBackGroundGrowthEstimateStep=2, ETAndBackGroundGrowthEstimateMaxSize=16)
    ref = OraclePh(OracleMethod(**m), om, UniformStream(77))
    d = str(tmp_path)
    write_model(d, om)
    write_sd2s(os.path.join(d, "phiT.dat"), ref.phiT.wfUp, ref.phiT.wfDn)
    write_param(d, dict(m, initialPhiTFlag="readFromFile", initialWalkerFlag="sampleFromPhiT"))
    print(f"[ph driver] {schedule}", flush=True)
    r = subprocess.run([EXE, d], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    ref.run()
    H, D = read_dat(os.path.join(d, "HNum.dat")), read_dat(os.path.join(d, "den.dat"))
    assert len(H) == len(ref.energies) == m["writeNumber"]
    for h, dd, e in zip(H, D, ref.energies):
        assert abs(h - e[0]) < 1e-9 * abs(e[0]) and abs(dd - e[1]) < 1e-9 * abs(e[1])
    T = read_dat(os.path.join(d, "TNum.dat"))
    assert np.allclose(T, [e[2] for e in ref.energies], rtol=1e-9, atol=0)
    bgnum = np.loadtxt(os.path.join(d, "choleskyBgNum.dat"), ndmin=2)
    assert bgnum.shape == (m["writeNumber"] * om.choleskyNumber, 2)
    assert len(np.loadtxt(os.path.join(d, "beta.dat"), ndmin=1)) == m["writeNumber"]
    # the adjusted background back in model_param, the adjusted ET back in afqmc_param
    assert np.allclose(read_model_bg(d), om.choleskyBg, rtol=1e-9, atol=1e-12)
    et = None
    for line in open(os.path.join(d, "afqmc_param")):
        parts = line.split()
        if parts and parts[0] == "ET":
            et = float(parts[1])
    # written as PhaselessMethod::write writes it: the stream's default six significant digits
    assert et is not None and et == float(f"{ref.method.ET:g}")
    # walkers written for the alive ones, in the SD2is format
    files = os.listdir(os.path.join(d, "walkers"))
    assert len(files) == sum(ref.walkerIsAlive)
    with open(os.path.join(d, "walkers", sorted(files)[0])) as f:
        lines = f.read().split("\n")
    assert int(lines[1]) == 3 and int(lines[2]) == 2


@needs_exe
def test_ph_run_trial_from_model_and_restart(tmp_path):
    """initialPhiTFlag setFromModel writes phiT.dat as the lowest Kp eigenvectors (the oracle's
    trial up to column signs) and gives the oracle's energies; a second run from the written
    walkers (readAllWalkers) and phiT.dat runs."""
    om = synthetic_molecule(L=10, Nup=3, Ndn=2, G=12, seed=9)
    m = dict(dt=0.02, thermalSize=10, writeNumber=2, measureNumberPerWrite=2, measureSkipStep=3,
             walkerSizePerThread=10, forceType="dynamicForce", forceCap=1.5, mgsStep=5, popControlStep=10,
             ET=-10.0, ETAndBackGroundAdjustMaxSize=0, ETAndBackGroundGrowthEstimateMaxSize=0, seed=5)
    d = str(tmp_path)
    write_model(d, om)
    write_param(d, dict(m, initialPhiTFlag="setFromModel", initialWalkerFlag="sampleFromPhiT"))
    r = subprocess.run([EXE, d], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    ref = OraclePh(OracleMethod(**m), om, UniformStream(5))
    with open(os.path.join(d, "phiT.dat")) as f:
        vals = f.read().split()
    # SD2s::write (SD2s.cpp:114-123): logw (2 numbers), then rank 2, L, Nup and the up elements
    assert (int(vals[2]), int(vals[3]), int(vals[4])) == (2, 10, 3)
    up = np.array([float(vals[5 + 2 * k]) + 1j * float(vals[6 + 2 * k]) for k in range(30)]).reshape(3, 10).T
    for j in range(3):
        c = np.vdot(ref.phiT.wfUp[:, j], up[:, j])
        assert abs(abs(c) - 1) < 1e-10
    ref.run()
    H, D = read_dat(os.path.join(d, "HNum.dat")), read_dat(os.path.join(d, "den.dat"))
    for h, dd, e in zip(H, D, ref.energies):
        assert abs(h / dd - e[0] / e[1]) < 1e-9 * abs(e[0] / e[1])
    for f in ("HNum.dat", "den.dat", "TNum.dat", "choleskyBgNum.dat", "choleskyExNum.dat", "beta.dat"):
        os.remove(os.path.join(d, f))
    write_param(d, dict(m, initialPhiTFlag="readFromFile", initialWalkerFlag="readAllWalkers"))
    r = subprocess.run([EXE, d], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    E = read_dat(os.path.join(d, "HNum.dat")) / read_dat(os.path.join(d, "den.dat"))
    assert np.all(np.isfinite(E)) and len(E) == m["writeNumber"]
