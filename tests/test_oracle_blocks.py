"""Pin the CPU oracle against the reference's own known answers and against
closed forms / brute force.  CPU only."""
import cmath
import itertools
import json
import math
import os

import numpy as np
import pytest

from oracle import linalg, mathfun
from oracle.blocks import (SD, SD2is, Hop, LogHop2is, HopSDOperation, LogHop2isSD2isOperation,
                           NiupNidn, NiupNidnSample, SDSDOperation, niupnidn_apply_right,
                           niupnidn_apply_left)
from oracle.rng import UniformStream, gaussians_from_uniforms, uniforms, seed_mix


KATS = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_kats.json")))


def col_major(vals, rows, cols):
    """Literal initialiser lists of the reference fill TensorHao in storage (column-major) order."""
    a = np.array([complex(*v) if isinstance(v, (tuple, list)) else v for v in vals], dtype=np.complex128)
    return a.reshape(rows, cols, order="F")


# ---------------------------------------------------------------- math helpers
def test_solve_cosh_and_cos_identities():
    # simple_fun_test.cpp:8-33
    for y in [-5.0, -3.3, -1.2, 0.0, 1.1, 2.0, 3.0, 4.0]:
        g = mathfun.solveCoshxEqExpy(y)
        assert abs(cmath.cosh(g).real - math.exp(y)) < 1e-10
        assert abs(cmath.cosh(g).imag) < 1e-10
        g = mathfun.solveCosxEqExpy(y)
        assert abs(cmath.cos(g).real - math.exp(y)) < 1e-10
        assert abs(cmath.cos(g).imag) < 1e-10


def test_exp2by2_literals():
    kat = KATS["exp2by2Matrix"]
    for c in kat["cases"]:
        a, b, z = mathfun.exp2by2Matrix(c["a"], c["b"], complex(*c["c"]))
        assert abs(a - c["expect_a"]) < c["tol"]
        assert abs(b - c["expect_b"]) < c["tol"]
        assert abs(z - complex(*c["expect_c"])) < c["tol"]
    a, b, c = mathfun.exp2by2Matrix(1.0, 2.0, 0j)
    assert a == math.exp(1.0) and b == math.exp(2.0) and c == 0


def test_exp2by2_matches_matrix_exponential():
    from scipy.linalg import expm
    rng = np.random.default_rng(3)
    for _ in range(20):
        a, b = rng.uniform(-2, 2, 2)
        c = complex(*rng.uniform(-2, 2, 2))
        a2, b2, c2 = mathfun.exp2by2Matrix(a, b, c)
        e = expm(np.array([[a, np.conj(c)], [c, b]]))
        assert np.allclose([[a2, np.conj(c2)], [c2, b2]], e, rtol=1e-12, atol=1e-12)


# ---------------------------------------------------------------- walker QR
WF43 = KATS["walker_4x3"]


def test_sd_normalize_literal():
    kat = KATS["SD_normalize"]
    sd = SD(wf=col_major(WF43, 4, 3))
    logw, ratio = sd.normalize_ratio()
    assert logw == pytest.approx(complex(kat["logw"], 0.0), rel=1e-14)
    assert ratio == pytest.approx(kat["ratio"], rel=1e-14)
    assert np.allclose(sd.wf.conj().T @ sd.wf, np.eye(3), atol=1e-14)


def test_sd2is_normalize_literal():
    kat = KATS["SD2is_normalize"]
    w = SD2is(col_major(WF43, 4, 3), kat["Nup"], kat["Ndn"])
    logw, ratio = w.normalize_ratio()
    assert logw == pytest.approx(complex(kat["logw"], 0.0), rel=1e-14)
    assert ratio == pytest.approx(kat["ratio"], rel=1e-14)


def test_qr_sign_rule_and_invariants():
    rng = np.random.default_rng(11)
    for _ in range(10):
        a = rng.standard_normal((12, 5)) + 1j * rng.standard_normal((12, 5))
        q, det, dl = linalg.qr_matrix(a)
        r = q.conj().T @ a
        assert det > 0
        assert np.allclose(np.triu(r), r, atol=1e-12)  # upper triangular
        assert abs(np.prod(np.diag(r)).real - det) < 1e-10 * det
        assert np.isclose(abs(np.linalg.det(a.conj().T @ a)) ** 0.5, det, rtol=1e-10)


# ---------------------------------------------------------------- Taylor exp
TAYLOR = KATS["LogHop2isSD2isOperation"]


@pytest.mark.parametrize("case", TAYLOR["cases"], ids=lambda c: f'{c["flag"]}-{c["order"]}')
def test_taylor_literals(case):
    k = TAYLOR
    L, nup, ndn = k["L"], k["Nup"], k["Ndn"]
    A = k["A_scale"] * col_major(k["A"], L, L)
    op = LogHop2is(A, complex(*k["A_logw"]))
    w = SD2is(col_major(k["W"], L, max(nup, ndn)), nup, ndn, complex(*k["W_logw"]))
    t = LogHop2isSD2isOperation(case["flag"], case["order"], case["accuracy"])
    r = t.applyToRight(op, w)
    assert t.currentOrder == case["expect_order"]
    assert np.max(np.abs(r.wf - col_major(k["right"], L, 2))) < case["tol"]
    assert r.logw == complex(*k["right_logw"])
    l = t.applyToLeft(op, w)
    assert t.currentOrder == case["expect_order"]
    assert np.max(np.abs(l.wf - col_major(k["left"], L, 2))) < case["tol"]
    assert l.logw == complex(*k["left_logw"])


# ---------------------------------------------------------------- one-body / HS apply
def test_hop_apply_logw_and_product():
    # hopSDOperationTest.cpp:30-50
    rng = np.random.default_rng(5)
    m = rng.uniform(-1, 1, (10, 10)) + 1j * rng.uniform(-1, 1, (10, 10))
    w = rng.uniform(-1, 1, (10, 5)) + 1j * rng.uniform(-1, 1, (10, 5))
    sd = SD(wf=w, logw=1.6)
    r = HopSDOperation().applyToRight(Hop(m, complex(1.2, 1.5)), sd)
    assert np.allclose(r.wf, m @ w, atol=1e-13) and r.logw == complex(2.8, 1.5)
    l = HopSDOperation().applyToLeft(Hop(m, complex(1.2, 1.5)), sd)
    assert np.allclose(l.wf, m.conj().T @ w, atol=1e-13) and l.logw == complex(2.8, -1.5)


def test_niupnidn_apply_equals_dense_matrix():
    # NiupNidnSDOperationTest.cpp:15-72: the 2x2 site blocks as one dense matrix
    rng = np.random.default_rng(6)
    h, N = 5, 6
    s = NiupNidnSample(h)
    for name in ("diag00", "diag01", "diag10", "diag11"):
        setattr(s, name, rng.uniform(-1, 1, h) + 1j * rng.uniform(-1, 1, h))
    s.logw = complex(1.2, 1.5)
    m = np.zeros((2 * h, 2 * h), dtype=complex)
    for i in range(h):
        m[i, i], m[i, i + h] = s.diag00[i], s.diag01[i]
        m[i + h, i], m[i + h, i + h] = s.diag10[i], s.diag11[i]
    w = rng.uniform(-1, 1, (2 * h, N)) + 1j * rng.uniform(-1, 1, (2 * h, N))
    wf, lw = niupnidn_apply_right(s, w, 1.6)
    assert np.allclose(wf, m @ w, atol=1e-14) and lw == complex(2.8, 1.5)
    wf, lw = niupnidn_apply_left(s, w, 1.6)
    assert np.allclose(wf, m.conj().T @ w, atol=1e-14) and lw == complex(2.8, -1.5)


@pytest.mark.parametrize("decomp", ["densityCharge", "densitySpin", "hopCharge", "hopSpin"])
def test_niupnidn_sample_is_hs_identity(decomp):
    """sum over x of the 2x2 sample operators (weights 1/2) reproduces exp(-dt U n_up n_dn + ...)
    for the density decompositions on every occupation state of one site."""
    dt, U = 0.05, 4.0
    z = np.zeros(1)
    hs = NiupNidn(dt, decomp, np.array([U]), z, z, z, z)
    # probabilities: logProb closed form (NiupNidnTest.cpp:53-68)
    f = np.array([0.3])
    for x in (-1, 1):
        lp = hs.logProbOfAuxFromForce(np.array([x]), f)
        assert abs(lp - math.log(math.exp(x * f[0]) / (math.exp(f[0]) + math.exp(-f[0])))) < 1e-14
    if decomp not in ("densityCharge", "densitySpin"):
        return
    # density decompositions are diagonal: per spin-occupation (n_up, n_dn)
    for nu in (0, 1):
        for nd in (0, 1):
            tot = 0j
            for x in (-1, 1):
                s = hs.getTwoBodySampleFromAux(np.array([x]))
                val = (s.diag00[0] if nu else 1.0) * (s.diag11[0] if nd else 1.0)
                # the densityCharge constant factor is in logw beyond log(1/2)
                tot += 0.5 * val * cmath.exp(s.logw - math.log(0.5))
            assert abs(tot - cmath.exp(-dt * U * nu * nd)) < 1e-12


# ---------------------------------------------------------------- overlap / Green
def brute_det(m):
    n = m.shape[0]
    tot = 0j
    for p in itertools.permutations(range(n)):
        inv = sum(1 for i in range(n) for j in range(i + 1, n) if p[i] > p[j])
        term = (-1) ** inv
        for i in range(n):
            term *= m[i, p[i]]
        tot += term
    return tot


def test_logoverlap_matches_bruteforce_determinant():
    rng = np.random.default_rng(8)
    L, N = 8, 4
    wl = rng.uniform(-1, 1, (L, N)) + 1j * rng.uniform(-1, 1, (L, N))
    wr = rng.uniform(-1, 1, (L, N)) + 1j * rng.uniform(-1, 1, (L, N))
    left = SD(wf=wl, logw=complex(2.0, 3.0))
    right = SD(wf=wr, logw=complex(1.0, 5.0))
    op = SDSDOperation(left, right)
    lo = op.returnLogOverlap()
    det = brute_det(wl.conj().T @ wr)
    # SDSDOperationTest.cpp:83-92: logO = logdet + conj(2+3i) + (1+5i)
    assert abs(lo.real - (math.log(abs(det)) + 3.0)) < 1e-12
    assert abs(cmath.exp(1j * lo.imag) - det / abs(det) * cmath.exp(2j)) < 1e-12
    assert -math.pi < (lo - complex(3.0, 2.0)).imag <= math.pi


def test_green_matches_inverse_formula():
    # SDSDOperationTest.cpp:94-119
    rng = np.random.default_rng(9)
    L, N = 10, 5
    wl = rng.uniform(-1, 1, (L, N)) + 1j * rng.uniform(-1, 1, (L, N))
    wr = rng.uniform(-1, 1, (L, N)) + 1j * rng.uniform(-1, 1, (L, N))
    op = SDSDOperation(SD(wf=wl), SD(wf=wr))
    inv = np.linalg.inv(wl.conj().T @ wr)
    green = (wr @ inv @ wl.conj().T).T
    assert np.allclose(op.returnGreenMatrix(), green, atol=1e-12)
    assert np.allclose(op.returnGreenDiagonal(), np.diag(green), atol=1e-12)
    off = np.concatenate([[green[i, i + L // 2] for i in range(L // 2)],
                          [green[i + L // 2, i] for i in range(L // 2)]])
    assert np.allclose(op.returnGreenOffDiagonal(), off, atol=1e-12)
    # idempotence of the (transposed) Green's matrix: G^T G^T = G^T
    g = op.returnGreenMatrix()
    assert np.allclose(g @ g, g, atol=1e-10)


def test_logdet_singular_and_pivot_sign():
    lu = linalg.lu_construct(np.zeros((3, 3)))
    assert linalg.log_determinant(*lu) == complex(-1e300, 0)
    p = np.array([[0, 1], [1, 0]], dtype=complex)  # one row swap, det = -1
    lo = linalg.log_determinant(*linalg.lu_construct(p))
    assert abs(cmath.exp(lo) - (-1)) < 1e-15


# ---------------------------------------------------------------- RNG stream
def test_uniform_stream_properties():
    u = uniforms(seed_mix(985456376), 0, 200000)
    assert u.min() >= 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 3e-3 and abs(u.var() - 1 / 12) < 2e-3
    s = UniformStream(985456376)
    a = s.take(10)
    b = s.take(5)
    assert np.array_equal(np.concatenate([a, b]), u[:15])
    g = gaussians_from_uniforms(u[:200000])
    assert abs(g.mean()) < 1e-2 and abs(g.var() - 1) < 2e-2
    # 2^53 scaling: values are multiples of 2^-53
    assert np.all(np.floor(u * 2.0 ** 53) == u * 2.0 ** 53)
