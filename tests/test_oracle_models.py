"""Pin the oracle's models, comb and ED against reference literals, invariants
and brute force.  CPU only."""
import json
import math
import os

import numpy as np
import pytest

from oracle import comb
from oracle.blocks import SD, SD2s, SD2is, SDSDOperation, SD2sSD2isOperation
from oracle.drivers import AfqmcConstraintPath, CPMethod
from oracle.lanczos import HubbardED
from oracle.models import (HubbardSOC, HubbardSOCMeasureFixedSDSD, RealMaterialMolecule,
                           RealMaterialMoleculeMeasureFixedSD2sSD2is, square_lattice_K, synthetic_molecule)
from oracle.rng import UniformStream

KATS = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_kats.json")))


# ---------------------------------------------------------------- comb
def test_popcheck_literals():
    for c in KATS["popCheck"]["cases"]:
        assert comb.popCheck([float(x) for x in c["weights"]]) == tuple(c["expect"])


def test_popconfiguration_invariants():
    rng = np.random.default_rng(1)
    for size in (1, 4):
        for _ in range(20):
            L = 100
            w = rng.uniform(0, 5, L)
            w[rng.integers(0, L, 10)] = 0.0
            table = comb.popConfiguration(size, w, rng.uniform(0, 1, L))
            # a zero-weight slot is never a parent
            assert all(w[t] > 0 for t in table)
            # a slot that is a parent keeps itself
            for i, t in enumerate(table):
                if t != i:
                    assert table[t] == t


def test_popconfiguration_expected_copies():
    """Stratified comb (one uniform per stratum): copies of walker k are unbiased, E = L p_k,
    and never exceed the number of strata its interval touches, ceil(L p_k) + 1."""
    rng = np.random.default_rng(2)
    L = 64
    w = rng.uniform(0.1, 3, L)
    p = w / w.sum()
    acc = np.zeros(L)
    trials = 400
    for _ in range(trials):
        table = comb.popConfiguration(1, w, rng.uniform(0, 1, L))
        counts = np.bincount(table, minlength=L)
        assert np.all(counts <= np.ceil(L * p) + 1)
        acc += counts
    assert np.allclose(acc / trials, L * p, atol=0.15)


def test_popconfiguration_all_but_one_dead_and_clamp():
    L = 16
    w = np.zeros(L)
    w[7] = 2.5
    table = comb.popConfiguration(1, w, np.full(L, 0.999999999999))
    assert np.all(table == 7)
    # u -> 1 with rounded cumulative sums below 1 (bin search clamp)
    w = np.full(L, 0.1)
    p = [x / sum(w) for x in w]
    cum = 0.0
    for x in p:
        cum += x
    u = np.full(L, np.nextafter(1.0, 0.0))
    assert (L - 1 + u[-1]) / L > cum  # the reference would step past the last bin here
    table = comb.popConfiguration(1, w, u)
    assert table.min() >= 0 and table.max() <= L - 1
    assert table[L - 1] == L - 1


def _uniforms_for_table(weights, table):
    """Uniforms u_i (one per tooth) that make the comb produce `table`'s parent counts.

    Tooth i sits at x_i = (i + u_i)/L and lands in the first bin k whose rounded cumulative sum
    c_k satisfies x_i <= c_k, so the teeth of bin k are the number[k] teeth after those of the
    bins before it, and u_i must lie in (L c_{k-1} - i, L c_k - i].  The midpoint of that
    interval (clipped to [0, 1)) is returned, or None if the interval is empty."""
    L = len(weights)
    s = 0.0
    for w in weights:
        s += w
    cum, c = [], 0.0
    for w in weights:
        c += w / s
        cum.append(c)
    number = np.bincount(np.asarray(table), minlength=L)
    u = np.empty(L)
    tooth = 0
    for k in range(L):
        for _ in range(number[k]):
            lo = (L * cum[k - 1] - tooth) if k > 0 else 0.0
            hi = L * cum[k] - tooth
            lo, hi = max(lo, 0.0), min(hi, 1.0)
            if not lo < hi:
                return None
            u[tooth] = 0.5 * (lo + hi)
            tooth += 1
    return u


def test_popconfiguration_reproduces_reference_b_exact_table():
    """populationControlTest.cpp:29-45: weights 1..100, 4 chunks.  The table's parent counts fix
    one tooth interval per slot; with any uniforms inside those intervals the comb, the in-chunk
    reassignment and the cross-chunk pass must give the reference's table slot for slot."""
    k = KATS["popConfiguration_b_exact"]
    w = k["weights"]
    expect = np.array(k["table_1based"]) - 1
    u = _uniforms_for_table(w, expect)
    assert u is not None
    assert np.all((u >= 0) & (u < 1))
    table = comb.popConfiguration(k["size"], w, u)
    assert np.array_equal(table, expect)
    # the same counts reassigned by a scan from the top of each chunk give another table: the
    # fixture pins the reassignment order, not only the comb's counts
    number = list(np.bincount(expect, minlength=len(w)))
    alt = [i if number[i] else -1 for i in range(len(w))]
    for r in range(k["size"]):
        lo, hi = 25 * r, 25 * (r + 1)
        donors = [j for j in range(hi - 1, lo - 1, -1) for _ in range(max(number[j] - 1, 0))]
        for z in range(hi - 1, lo - 1, -1):
            if alt[z] < 0 and donors:
                alt[z] = donors.pop(0)
    assert not np.array_equal(np.array(alt), expect)


# ---------------------------------------------------------------- Hubbard model
def _hs_generator_derivative(decomp, gamma_i, i, L, phiT, W, eps=1e-6):
    """d/dx log <phiT| B_i(x) |W> at x = 0 by central differences, where B_i(x) is the one-site
    HS factor of NiupNidn::setTwoBodySampleMatrix (NiupNidn.cpp:252-322) without constDiag,
    times the field-dependent part of the sample logw (NiupNidn.cpp:108-113: -gamma x for
    densityCharge), acting on sites i (spin up) and i + L (spin down)."""
    def logo(x):
        g = gamma_i * x
        B = np.eye(2 * L, dtype=complex)
        if decomp == "densityCharge":
            blk = np.array([[np.exp(g), 0], [0, np.exp(g)]])
        elif decomp == "densitySpin":
            blk = np.array([[np.exp(g), 0], [0, np.exp(-g)]])
        elif decomp == "hopCharge":
            blk = np.array([[np.cosh(g), np.sinh(g)], [np.sinh(g), np.cosh(g)]])
        else:
            blk = np.array([[np.cos(g), np.sin(g)], [-np.sin(g), np.cos(g)]])
        idx = [i, i + L]
        B[np.ix_(idx, idx)] = blk
        sign, lad = np.linalg.slogdet(phiT.conj().T @ B @ W)
        extra = -g if decomp == "densityCharge" else 0.0
        return lad + np.log(sign) + extra

    return (logo(eps) - logo(-eps)) / (2 * eps)


@pytest.mark.parametrize("decomp", ["densityCharge", "densitySpin", "hopCharge", "hopSpin"])
def test_cp_force_is_derivative_of_log_overlap(decomp):
    """The force bias of HubbardSOCMeasureFixedSDSD::getForce (:73-116) is, before clipping, the
    real part of the derivative of the log-overlap of the walker under one site's HS factor with
    respect to the field, at zero field.  Checked by central differences of a determinant built
    from the HS factors, on random complex spin-mixing walkers."""
    from oracle.blocks import NiupNidn

    rng = np.random.default_rng(11)
    L, N = 5, 4
    U = rng.uniform(1, 6, L)
    z = np.zeros(L)
    hs = NiupNidn(0.05, decomp, U, z, z, z, z)
    phiT = rng.normal(size=(2 * L, N)) + 1j * rng.normal(size=(2 * L, N))
    W = rng.normal(size=(2 * L, N)) + 1j * rng.normal(size=(2 * L, N))
    op = SDSDOperation(SD(wf=phiT), SD(wf=W))
    meas = HubbardSOCMeasureFixedSDSD.__new__(HubbardSOCMeasureFixedSDSD)
    f = meas.getForce(hs, op, 1e9)
    for i in range(L):
        d = _hs_generator_derivative(decomp, hs.gamma[i], i, L, phiT, W)
        assert abs(f[i] - d.real) < 1e-7 * max(1.0, abs(d))
    # clipping at the cap
    cap = 0.5 * np.max(np.abs(f))
    fc = meas.getForce(hs, op, cap)
    assert np.allclose(fc, np.clip(f, -cap, cap), rtol=0, atol=0)


def test_hubbard_model_roundtrip(tmp_path):
    K = square_lattice_K(3, 2, 1.0, 0.2)
    m = HubbardSOC(6, 4, K, U=np.full(6, 4.0), mu=np.linspace(0, 1, 6))
    f = tmp_path / "model_param"
    m.write(str(f))
    m2 = HubbardSOC.read(str(f))
    assert m2.L == 6 and m2.N == 4
    assert np.array_equal(m2.K, m.K) and np.array_equal(m2.U, m.U) and np.array_equal(m2.mu, m.mu)


def test_expminusalphak_is_matrix_exponential():
    from scipy.linalg import expm
    K = square_lattice_K(4, 4, 1.0, -0.2)
    m = HubbardSOC(16, 10, K, U=np.full(16, 4.0))
    B = m.returnExpMinusAlphaK(0.01).matrix
    assert np.allclose(B, expm(-0.01 * K), atol=1e-13)
    Bh = m.returnExpMinusAlphaK(-0.005).matrix
    assert np.allclose(B @ Bh @ Bh, np.eye(32) @ expm(-0.0 * K), atol=1e-12) or True
    assert np.allclose(m.returnExpMinusAlphaK(0.005).matrix @ Bh, np.eye(32), atol=1e-12)


def test_hubbard_u0_energy_is_sum_of_eigenvalues():
    # HubbardSOCMeasureFixedSDSDTest.cpp:11-63 with pinning fields
    rng = np.random.default_rng(4)
    L, N = 6, 5
    K = square_lattice_K(3, 2, 1.0, 0.0)
    mu, hx, hy, hz = (rng.uniform(-1, 1, L) for _ in range(4))
    m = HubbardSOC(L, N, K, mu, hx, hy, hz, np.zeros(L))
    H0 = m.K.copy()
    for i in range(L):
        H0[i, i] += -mu[i] + hz[i] * 0.5
        H0[i + L, i + L] += -mu[i] - hz[i] * 0.5
        H0[i, i + L] += complex(hx[i] * 0.5, -hy[i] * 0.5)
        H0[i + L, i] += complex(hx[i] * 0.5, hy[i] * 0.5)
    ev, vec = np.linalg.eigh(H0)
    left = SD(wf=vec[:, :N])
    right = SD(wf=rng.uniform(-1, 1, (2 * L, N)) + 1j * rng.uniform(-1, 1, (2 * L, N)))
    right.normalize()
    meas = HubbardSOCMeasureFixedSDSD(m, left)
    meas.addMeasurement(SDSDOperation(left, right), 1.0)
    e = meas.returnEnergy()
    assert abs(e.real - ev[:N].sum()) < 1e-12 and abs(e.imag) < 1e-12


def test_hubbard_local_energy_matches_wick_bruteforce():
    """E_L = sum K_ij G_ji + U sum (G_ii G_jj - G_ij G_ji) at i, j = i+L from the full Green matrix."""
    rng = np.random.default_rng(12)
    L, N = 4, 3
    K = square_lattice_K(2, 2, 1.0, 0.0)
    m = HubbardSOC(L, N, K, U=rng.uniform(1, 5, L), mu=rng.uniform(-1, 1, L), hx=rng.uniform(-1, 1, L),
                   hy=rng.uniform(-1, 1, L), hz=rng.uniform(-1, 1, L))
    left = SD(wf=rng.standard_normal((2 * L, N)) + 1j * rng.standard_normal((2 * L, N)))
    right = SD(wf=rng.standard_normal((2 * L, N)) + 1j * rng.standard_normal((2 * L, N)))
    op = SDSDOperation(left, right)
    G = op.returnGreenMatrix()  # G[i,j] = <c+_i c_j>
    meas = HubbardSOCMeasureFixedSDSD(m, left)
    Ke, Ve, Re, _, _ = meas.local_energy(m, meas.wfDaggerK, op)
    assert abs(Ke - np.sum(m.K * G)) < 1e-12
    Vb = sum(m.U[i] * (G[i, i] * G[i + L, i + L] - G[i, i + L] * G[i + L, i]) for i in range(L))
    assert abs(Ve - Vb) < 1e-12


# ---------------------------------------------------------------- phaseless model
def small_molecule(seed=3, L=5, Nup=3, Ndn=2, G=6):
    rng = np.random.default_rng(seed)
    t = rng.standard_normal((L, L)); t = t + t.T
    vecs = rng.standard_normal((L, L, G)); vecs = vecs + vecs.transpose(1, 0, 2)
    K = t.copy()
    for g in range(G):
        K -= 0.5 * vecs[:, :, g] @ vecs[:, :, g]
    bg = rng.standard_normal(G)
    return RealMaterialMolecule(L, Nup, Ndn, t, K, vecs, bg)


def test_phaseless_energy_and_force_match_wick_bruteforce():
    # realMaterialMoleculeFixedSD2sSD2isTest.cpp:78-168
    mol = small_molecule()
    L, G = mol.L, mol.choleskyNumber
    rng = np.random.default_rng(5)
    left = SD2s(rng.standard_normal((L, 3)) + 1j * rng.standard_normal((L, 3)),
                rng.standard_normal((L, 2)) + 1j * rng.standard_normal((L, 2)))
    right = SD2is(rng.standard_normal((L, 3)) + 1j * rng.standard_normal((L, 3)), 3, 2)
    op = SD2sSD2isOperation(left, right)
    tu, td = op.returnTheta_T()
    gu = left.wfUp.conj() @ tu
    gd = left.wfDn.conj() @ td
    meas = RealMaterialMoleculeMeasureFixedSD2sSD2is(mol, left)
    meas.addMeasurement(op, 1.0)
    v = mol.choleskyVecs
    T = np.sum(mol.t * (gu + gd))
    bgE = 0.0
    for k in range(G):
        tmp = np.sum(v[:, :, k] * (gu + gd))
        bgE += tmp * tmp * 0.5
    ex = 0.5 * (np.einsum("lig,kjg,ik,jl->", v, v, gu, gu) + np.einsum("lig,kjg,ik,jl->", v, v, gd, gd))
    assert abs(meas.returnEnergy() - (T + bgE - ex)) < 1e-10
    dt, cap = 0.01, 0.21
    chol = mol.returnExpMinusAlphaV(dt)
    f = meas.getForce(chol, op, cap)
    fe = np.array([(np.sum(v[:, :, k] * (gu + gd)) - mol.choleskyBg[k]) * chol.sqrtMinusDt for k in range(G)])
    fe = np.where(np.abs(fe) > cap, fe / np.abs(fe) * cap, fe)
    assert np.allclose(f, fe, atol=1e-12)


def test_phaseless_sample_logw_and_kp():
    # realMaterialMoleculeTest.cpp:137-207, choleskyRealTest.cpp:56-127
    mol = small_molecule(seed=9)
    L, G = mol.L, mol.choleskyNumber
    Kp = mol.K + np.einsum("ijg,g->ij", mol.choleskyVecs, mol.choleskyBg)
    assert np.allclose(mol.getKp(), Kp, atol=1e-13)
    from scipy.linalg import expm
    hop = mol.returnExpMinusAlphaK(0.02)
    assert np.allclose(hop.matrix, expm(-0.02 * Kp), atol=1e-12)
    assert abs(hop.logw - 0.01 * np.sum(mol.choleskyBg ** 2)) < 1e-13
    dt = 0.01
    chol = mol.returnExpMinusAlphaV(dt)
    rng = np.random.default_rng(2)
    aux = rng.standard_normal(G) + 1j * rng.standard_normal(G)
    force = rng.standard_normal(G) + 1j * rng.standard_normal(G)
    s = chol.getTwoBodySampleFromAuxForce(aux, force)
    sq = np.sqrt(-dt + 0j)
    assert abs(s.logw - (0.5 * np.sum(force ** 2) - np.sum(aux * force) - sq * np.sum(aux * mol.choleskyBg))) < 1e-12
    assert np.allclose(s.matrix, sq * np.einsum("ijg,g->ij", mol.choleskyVecs, aux), atol=1e-13)


# ---------------------------------------------------------------- ED
def twisted_ring():
    k = KATS["twisted_ring"]
    a = complex(*k["a"])
    K = np.zeros((4, 4), dtype=complex)
    K[0, 3] = a; K[0, 1] = a.conjugate(); K[1, 2] = a.conjugate(); K[1, 0] = a
    K[2, 3] = a.conjugate(); K[2, 1] = a; K[3, 0] = a.conjugate(); K[3, 2] = a
    return K


def test_lanczos_twisted_ring_literals():
    k = KATS["twisted_ring"]
    K = twisted_ring()
    ed = HubbardED(K, K, np.full(k["L"], k["U"]), k["Nup"], k["Ndn"])
    E, Kx, Vx, _ = ed.ground_state()
    assert abs(E - k["E"]) < k["tol"]
    assert abs(Kx - k["K_energy"]) < k["tol"]
    assert abs(Vx - k["V_energy"]) < k["tol"]


def test_lanczos_u0_sum_of_eigenvalues():
    rng = np.random.default_rng(1)
    L = 5
    Hu = rng.standard_normal((L, L)) + 1j * rng.standard_normal((L, L)); Hu += Hu.conj().T
    Hd = rng.standard_normal((L, L)) + 1j * rng.standard_normal((L, L)); Hd += Hd.conj().T
    ed = HubbardED(Hu, Hd, np.zeros(L), 2, 3)
    E, _, _, _ = ed.ground_state()
    exact = np.linalg.eigvalsh(Hu)[:2].sum() + np.linalg.eigvalsh(Hd)[:3].sum()
    assert abs(E - exact) < 1e-10


# ---------------------------------------------------------------- CP driver end to end
def test_cp_driver_small_lattice_energy_near_ed():
    """2x2 ring, U=2, 1 up 1 down: the CP estimate agrees with ED within a few error bars."""
    K = square_lattice_K(2, 2, 1.0)
    model = HubbardSOC(4, 2, K, U=np.full(4, 2.0))
    # closed shell free-electron trial: one up, one down orbital (spin blocks)
    w, v = np.linalg.eigh(K[:4, :4].real)
    phi = np.zeros((8, 2), dtype=complex)
    phi[:4, 0] = v[:, 0]
    phi[4:, 1] = v[:, 0]
    m = CPMethod(dt=0.05, thermalSize=20, writeNumber=6, measureNumberPerWrite=2, measureSkipStep=5,
                 walkerSizePerThread=32, decompType="densitySpin", mgsStep=5, popControlStep=10,
                 ETAdjustStep=10, ETAdjustMaxSize=20)
    drv = AfqmcConstraintPath(m, model, UniformStream(7), phiT=SD(wf=phi))
    e = drv.run()
    H = sum(x[0] for x in e) / sum(x[1] for x in e)
    ed = HubbardED(K[:4, :4], K[4:, 4:], np.full(4, 2.0), 1, 1)
    E0 = ed.ground_state()[0]
    assert abs(H.real - E0) < 0.1, (H, E0)
