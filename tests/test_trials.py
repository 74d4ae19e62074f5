"""Trial generation on the host: the UHF and GHF mean fields converge, are self-consistent, and
their energies equal the Hamiltonian's expectation value in the resulting determinant, evaluated
independently by the oracle's Wick contractions (HubbardSOCMeasureFixedSDSD with bra = ket).
CPU only."""
import numpy as np
import pytest

from afqmclab_amd.model import (HubbardModel, ScfNotConverged, ghf_trial, order_parameter_seed, uhf_trial,
                                zafm_densities)
from oracle.blocks import SD, SDSDOperation
from oracle.models import HubbardSOC, HubbardSOCMeasureFixedSDSD


def wick_energy(model, phi):
    """<phi|H|phi>/<phi|phi> of HubbardSOC by the oracle's mixed estimator with the trial as walker."""
    om = HubbardSOC(model.L, model.N, model.K, mu=model.mu, hx=model.hx, hy=model.hy, hz=model.hz, U=model.U)
    left = SD(wf=phi)
    meas = HubbardSOCMeasureFixedSDSD(om, left)
    meas.addMeasurement(SDSDOperation(left, SD(wf=phi.copy())), 1.0)
    return meas.returnEnergy()


@pytest.mark.parametrize("case", [(4, 4, 8, 8, 4.0), (8, 8, 28, 28, 6.0), (6, 6, 16, 16, 4.0), (10, 10, 45, 45, 5.0)])
def test_uhf_is_self_consistent_and_variational(case):
    nx, ny, nup, ndn, U = case
    model = HubbardModel.square(nx, ny, nup + ndn, U)
    L = model.L
    phi, info = uhf_trial(model, nup, ndn, return_info=True)
    assert info["max_density_change"] < 1e-8
    # the occupied orbitals are orthonormal, spin blocks exact zeros
    assert np.allclose(phi.conj().T @ phi, np.eye(nup + ndn), atol=1e-12)
    assert np.all(phi[:L, nup:] == 0) and np.all(phi[L:, :nup] == 0)
    # one more mean-field iteration from the state's own densities returns the same state
    nu = np.sum(np.abs(phi[:L, :nup]) ** 2, axis=1)
    nd = np.sum(np.abs(phi[L:, nup:]) ** 2, axis=1)
    ku, kd = model.spin_block_hopping()
    wu, vu = np.linalg.eigh(ku + np.diag(model.U * nd))
    wd, vd = np.linalg.eigh(kd + np.diag(model.U * nu))
    assert np.max(np.abs(np.sum(np.abs(vu[:, :nup]) ** 2, axis=1) - nu)) < 1e-7
    assert np.max(np.abs(np.sum(np.abs(vd[:, :ndn]) ** 2, axis=1) - nd)) < 1e-7
    # the mean-field energy is <H> of the determinant
    e = wick_energy(model, phi)
    assert abs(e.imag) < 1e-10
    assert abs(e.real - info["energy"]) < 1e-7 * abs(e.real)


def test_uhf_at_16x16_u8_converges():
    """The 16 x 16, U = 8, 112 + 112 trial of the benchmark workload."""
    model = HubbardModel.square(16, 16, 224, 8.0)
    phi, info = uhf_trial(model, 112, 112, return_info=True)
    assert info["max_density_change"] < 1e-8
    assert abs(wick_energy(model, phi).real - info["energy"]) < 1e-7 * abs(info["energy"])


def test_uhf_reports_non_convergence():
    model = HubbardModel.square(8, 8, 56, 6.0)
    with pytest.raises(ScfNotConverged):
        uhf_trial(model, 28, 28, iters=3)


def test_uhf_antiferromagnet_below_paramagnet_at_half_filling():
    """At half filling and U = 4 the mean field orders antiferromagnetically, below the energy of
    the paramagnetic free-electron determinant."""
    from afqmclab_amd.model import spin_block_trial
    model = HubbardModel.square(4, 4, 16, 4.0)
    _, info = uhf_trial(model, 8, 8, return_info=True)
    para = wick_energy(model, spin_block_trial(model, 8, 8)).real
    assert info["energy"] < para - 0.5
    nu, nd = info["n_up"], info["n_dn"]
    sign = np.array([(-1) ** ((i % 4) + (i // 4)) for i in range(16)])
    assert np.all(sign * (nu - nd) > 0.3)  # a staggered moment on every site


def test_order_parameter_seeds():
    model = HubbardModel.square(4, 4, 12, 4.0)
    d, s = order_parameter_seed(model, "zafm")
    L = model.L
    assert np.isclose(np.sum(d).real, 12 / 16 * 16) and np.all(s == 0)
    assert d[0] == 12 / 16 and d[L] == 0 and d[1] == 0 and d[1 + L] == 12 / 16
    d, s = order_parameter_seed(model, "xafm")
    assert np.all(d == 0) and s[0] == s[L] == 0.375 and s[1] == s[1 + L] == -0.375
    d, s = order_parameter_seed(model, "yafm")
    assert s[0] == 0.375j and s[L] == -0.375j
    d, s = order_parameter_seed(model, "random", rng=np.random.default_rng(0))
    assert np.allclose(s[L:], np.conj(s[:L]))
    with pytest.raises(ValueError):
        order_parameter_seed(model, "nope")


def test_ghf_spin_rotation_and_wick_energy():
    """The Hubbard model is spin-rotation invariant: GHF from z, x and y antiferromagnetic seeds
    reaches the same energy as the UHF, and each energy is <H> of its determinant (the x and y
    states have spin-mixing orbitals, so the spin-flip Wick terms enter)."""
    model = HubbardModel.square(4, 4, 16, 4.0)
    _, uinfo = uhf_trial(model, 8, 8, return_info=True)
    for kind in ("zafm", "xafm", "yafm"):
        d, s = order_parameter_seed(model, kind)
        state, info = ghf_trial(model, d, s)
        assert np.allclose(state.conj().T @ state, np.eye(16), atol=1e-12)
        assert abs(info["energy"] - uinfo["energy"]) < 1e-7 * abs(uinfo["energy"]), kind
        e = wick_energy(model, state)
        assert abs(e.imag) < 1e-9 and abs(e.real - info["energy"]) < 1e-7 * abs(e.real), kind
        if kind != "zafm":
            L = model.L
            assert np.max(np.abs(state[:L])) > 0.1 and np.max(np.abs(state[L:])) > 0.1


def test_ghf_with_fields_matches_wick_energy():
    """Chemical potential and Zeeman fields hx, hy, hz enter H0 (setH0, ghfDetail.cpp:9-25); the
    converged GHF energy must equal <H> with the same fields."""
    model = HubbardModel.square(4, 4, 14, 3.0)
    L = model.L
    rng = np.random.default_rng(3)
    model.mu = rng.uniform(-0.2, 0.2, L)
    model.hx = rng.uniform(-0.3, 0.3, L)
    model.hy = rng.uniform(-0.3, 0.3, L)
    model.hz = rng.uniform(-0.3, 0.3, L)
    d, s = order_parameter_seed(model, "random", rng=rng)
    state, info = ghf_trial(model, d, s, converge_type="orderParameter", tol=1e-11, max_iter=500, relax=0.5)
    e = wick_energy(model, state)
    assert abs(e.imag) < 1e-9
    assert abs(e.real - info["energy"]) < 1e-7 * abs(e.real)


def test_ghf_u0_is_the_free_electron_energy():
    model = HubbardModel.square(4, 4, 10, 0.0)
    d, s = order_parameter_seed(model, "zafm")
    _, info = ghf_trial(model, d, s)
    w = np.linalg.eigvalsh(model.K)
    assert abs(info["energy"] - np.sum(w[:10])) < 1e-10


def test_ghf_reports_non_convergence():
    model = HubbardModel.square(4, 4, 14, 4.0)
    d, s = order_parameter_seed(model, "random", rng=np.random.default_rng(1))
    with pytest.raises(ScfNotConverged):
        ghf_trial(model, d, s, max_iter=2)


def test_zafm_densities_fill():
    model = HubbardModel.square(6, 6, 28, 2.0)
    nu, nd = zafm_densities(model, 14, 14)
    assert np.isclose(nu.sum(), 14) and np.isclose(nd.sum(), 14)
