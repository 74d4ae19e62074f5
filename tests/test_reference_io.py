"""Reference file formats (afqmclab_amd/reference_io.py): the NumDenErrorAnalysis statistics
against the reference's own known answers (serialErrorAnalysisTest.cpp), and the model_param /
SD / afqmc_param readers and writers on files laid out as the reference writes them."""
import json
import os

import numpy as np
import pytest

from afqmclab_amd import reference_io as rio
from afqmclab_amd.model import HubbardModel, spin_block_trial

KAT = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")))["serialErrorAnalysis"]


def test_divisors_and_block_errors_match_reference_known_answers():
    assert rio.get_divisors(500) == KAT["divisors_of_500"]
    assert rio.get_divisors(10) == KAT["block_sizes"]
    assert rio.get_divisors(36) == [1, 2, 3, 4, 6, 9, 12, 18, 36]  # a perfect square once
    a = np.array(KAT["array"])
    rows = rio.num_den_error_analysis(a + 0j, np.ones_like(a) + 0j)
    assert [r[0] for r in rows] == KAT["block_sizes"]
    for (b, mr, mi, e), m_ex, e_ex in zip(rows, KAT["mean_block"], KAT["err_block"]):
        assert abs(mr - m_ex) < 1e-13 and mi == 0.0 and abs(e - e_ex) < 1e-13


def test_num_den_ratio_of_block_sums_and_skip(tmp_path):
    rng = np.random.default_rng(4)
    num = rng.normal(size=12) + 1j * rng.normal(size=12)
    den = 1.0 + rng.uniform(size=12) + 0j
    rows = rio.num_den_error_analysis(num, den, skip=2)
    assert [r[0] for r in rows] == [1, 2, 5, 10]
    b5 = [num[2 + 5 * i:7 + 5 * i].sum() / den[2 + 5 * i:7 + 5 * i].sum() for i in range(2)]
    assert abs(rows[2][1] - np.mean(b5).real) < 1e-14
    rio.write_num_den_analysis(tmp_path / "out.dat", rows)
    back = np.loadtxt(tmp_path / "out.dat")
    assert back.shape == (4, 4) and np.allclose(back[:, 3], [r[3] for r in rows], rtol=1e-15)


def test_model_param_and_sd_round_trip_in_reference_layout(tmp_path):
    model = HubbardModel.square(3, 4, 8, 4.0, 1.0, -0.2)
    model.mu[:] = np.linspace(-0.5, 0.5, model.L)
    model.hz[:] = 0.1
    p = tmp_path / "model_param"
    rio.write_model_param(p, model)
    lines = open(p).read().splitlines()
    # L, N, 4L^2 complex K entries, then mu, hx, hy, hz, U: one value per line, width 26
    assert len(lines) == 2 + 4 * model.L ** 2 + 5 * model.L
    assert all(len(x) == 52 for x in lines[2:2 + 4 * model.L ** 2]) and len(lines[-1]) == 26
    back = rio.read_model_param(p)
    assert back.L == model.L and back.N == model.N
    assert np.array_equal(back.K, model.K) and np.array_equal(back.mu, model.mu) and np.array_equal(back.hz, model.hz)
    assert np.array_equal(back.U, model.U)
    phi = spin_block_trial(model, 4, 4)
    rio.write_sd(tmp_path / "phiT.dat", 0.25 - 1.5j, phi)
    head = open(tmp_path / "phiT.dat").read().splitlines()[:3]
    assert head[1] == "2" and head[2].split() == [str(2 * model.L), "8"]
    logw, wf = rio.read_sd(tmp_path / "phiT.dat")
    assert logw == 0.25 - 1.5j and np.array_equal(wf, phi)


def test_afqmc_param_written_and_searched_as_the_reference_does(tmp_path):
    p = tmp_path / "afqmc_param"
    rio.write_afqmc_param(p, dict(dt=0.01, walkerSizePerThread=16, decompType="densitySpin",
                                  initialPhiTFlag="readFromFile", ET=-12.5, seed=7))
    assert rio.search_value(p, "dt", float) == 0.01
    assert rio.search_value(p, "initialPhiTFlag") == "readFromFile"
    assert rio.search_value(p, "ET", float) == -12.5 and rio.search_value(p, "thermalSize") is None
    with open(p, "a") as f:
        f.write("forceCap = 2.5\n")
    assert rio.search_value(p, "forceCap", float) == 2.5


PH_EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "afqmclab_amd", "afqmc_ph_run")


@pytest.mark.skipif(not os.path.exists(PH_EXE), reason="afqmc_ph_run not built (no HDF5 library at build time)")
def test_real_materials_model_param_hdf5_roundtrip(tmp_path):
    """RealMaterialMolecule's HDF5 model_param (realMaterialMolecule.cpp:41-94) written from raw
    arrays and read back by the phaseless driver's reader: every dataset bit for bit; a dataset of
    the wrong size is an error, as readFile's size check makes it (readWriteHdf5.cpp:42-48)."""
    import subprocess
    rng = np.random.default_rng(3)
    L, Nup, Ndn, G = 7, 3, 2, 5
    arrays = {"t": rng.normal(size=(L, L)), "K": rng.normal(size=(L, L)), "choleskyVecs": rng.normal(size=(L, L, G)),
              "choleskyBg": rng.normal(size=G)}
    (tmp_path / "model_dims").write_text(f"{L} {Nup} {Ndn} {G}\n")
    for k, a in arrays.items():
        np.asfortranarray(a).ravel(order="F").astype("<f8").tofile(str(tmp_path / (k + ".bin")))
    r = subprocess.run([PH_EXE, "--write-model", str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert (tmp_path / "model_param").read_bytes()[:8] == b"\x89HDF\r\n\x1a\n"
    r = subprocess.run([PH_EXE, "--read-model", str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert (tmp_path / "model_dims.read").read_text().split() == [str(L), str(Nup), str(Ndn), str(G)]
    for k, a in arrays.items():
        assert np.array_equal(np.fromfile(str(tmp_path / (k + ".read.bin"))), np.asfortranarray(a).ravel(order="F"))
    # a model_dims that promises more Cholesky vectors than the file holds
    (tmp_path / "model_dims").write_text(f"{L} {Nup} {Ndn} {G + 1}\n")
    np.zeros((L, L, G + 1)).ravel().tofile(str(tmp_path / "choleskyVecs.bin"))
    np.zeros(G + 1).tofile(str(tmp_path / "choleskyBg.bin"))
    sub = tmp_path / "bad"
    sub.mkdir()
    r = subprocess.run([PH_EXE, "--write-model", str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0
    import shutil
    shutil.copy(str(tmp_path / "model_param"), str(sub / "model_param"))
    # hand the reader a file whose choleskyNumber does not match its vectors
    (sub / "model_dims").write_text(f"{L} {Nup} {Ndn} {G}\n")
    for k, a in arrays.items():
        np.asfortranarray(a).ravel(order="F").astype("<f8").tofile(str(sub / (k + ".bin")))
    r = subprocess.run([PH_EXE, "--read-model", str(sub)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0  # a consistent file reads
    r = subprocess.run([PH_EXE, "--read-model", str(tmp_path / "missing")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "HDF5" in r.stderr
