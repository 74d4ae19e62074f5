"""Walkers and walker operations of the reference, restated in numpy.

Test infrastructure only (see oracle/__init__.py).  Class and method names
mirror afqmc/blocks so the oracle reads like the reference.
"""
import cmath
import math

import numpy as np

from . import linalg
from .mathfun import exp2by2Matrix, solveCoshxEqExpy, solveCosxEqExpy

C128 = np.complex128


# --------------------------------------------------------------------------
# Walkers
# --------------------------------------------------------------------------
class SD:
    """Single determinant {logw, wf (L x N)} (afqmc/blocks/walker/SD/source/SD.cpp)."""

    def __init__(self, L=0, N=0, wf=None, logw=0j):
        self.wf = np.zeros((L, N), dtype=C128, order="F") if wf is None else np.array(wf, dtype=C128, order="F")
        self.logw = complex(logw)

    def copy(self):
        return SD(wf=self.wf.copy(), logw=self.logw)

    def getL(self):
        return self.wf.shape[0]

    def getN(self):
        return self.wf.shape[1]

    def addLogw(self, x):
        self.logw += complex(x)

    def stabilize(self):
        """logw += log(QRMatrix(wf)) (SD.cpp:42-45)."""
        q, det, _ = linalg.qr_matrix(self.wf)
        self.wf = q
        self.logw += math.log(det)

    def stabilize_ratio(self):
        """SD::stabilize(double &ratio) (SD.cpp:47-56)."""
        q, det, det_list = linalg.qr_matrix(self.wf)
        self.wf = q
        self.logw += math.log(det)
        a = np.abs(det_list)
        return a.min() / a.max()

    def normalize(self):
        """stabilize, return logw and reset it to 0 (SD.cpp:58-72)."""
        self.stabilize()
        lw = self.logw
        self.logw = 0j
        return lw

    def normalize_ratio(self):
        r = self.stabilize_ratio()
        lw = self.logw
        self.logw = 0j
        return lw, r


class SD2s:
    """Separate up/dn determinants (afqmc/blocks/walker/SD2s)."""

    def __init__(self, wfUp, wfDn, logw=0j):
        self.wfUp = np.array(wfUp, dtype=C128, order="F")
        self.wfDn = np.array(wfDn, dtype=C128, order="F")
        self.logw = complex(logw)

    def getL(self):
        return self.wfUp.shape[0]

    def getNup(self):
        return self.wfUp.shape[1]

    def getNdn(self):
        return self.wfDn.shape[1]


class SD2is:
    """One L x max(Nup,Ndn) matrix shared by both spins (SD2is.cpp:13-16).

    Up = first Nup columns, down = first Ndn columns."""

    def __init__(self, wf, Nup, Ndn, logw=0j):
        self.wf = np.array(wf, dtype=C128, order="F")
        self.Nup = int(Nup)
        self.Ndn = int(Ndn)
        self.logw = complex(logw)
        if self.wf.shape[1] != max(self.Nup, self.Ndn):
            raise ValueError("SD2is wf must have max(Nup, Ndn) columns")

    def copy(self):
        return SD2is(self.wf.copy(), self.Nup, self.Ndn, self.logw)

    def getL(self):
        return self.wf.shape[0]

    def addLogw(self, x):
        self.logw += complex(x)

    def stabilize(self):
        """SD2is::stabilize (SD2is.cpp:48-61)."""
        if self.Nup == self.Ndn:
            q, det, _ = linalg.qr_matrix(self.wf)
            self.wf = q
            self.logw += 2.0 * math.log(det)
        else:
            q, det, det_list = linalg.qr_matrix(self.wf)
            self.wf = q
            self.logw += math.log(det)
            nmin = min(self.Nup, self.Ndn)
            norm = complex(1.0, 0.0)
            for i in range(nmin):
                norm *= det_list[i]
            self.logw += cmath.log(norm)

    def stabilize_ratio(self):
        """SD2is::stabilize(double &ratio) (SD2is.cpp:63-76)."""
        q, det, det_list = linalg.qr_matrix(self.wf)
        self.wf = q
        self.logw += math.log(det)
        nmin = min(self.Nup, self.Ndn)
        norm = complex(1.0, 0.0)
        for i in range(nmin):
            norm *= det_list[i]
        self.logw += cmath.log(norm)
        a = np.abs(det_list)
        return a.min() / a.max()

    def normalize_ratio(self):
        r = self.stabilize_ratio()
        lw = self.logw
        self.logw = 0j
        return lw, r


# --------------------------------------------------------------------------
# One-body operators
# --------------------------------------------------------------------------
class Hop:
    """{logw, matrix (L x L)} (afqmc/blocks/oneBodyOperator/hop)."""

    def __init__(self, matrix, logw=0j):
        self.matrix = np.array(matrix, dtype=C128, order="F")
        self.logw = complex(logw)


Hop2is = Hop
LogHop2is = Hop  # matrix holds the log; the operator is exp(matrix)


def hop_apply_right(op: Hop, wf: np.ndarray, logw: complex):
    """wf' = B wf, logw' = B.logw + logw (hopSDOperation.cpp:14-19)."""
    return op.matrix @ wf, op.logw + logw


def hop_apply_left(op: Hop, wf: np.ndarray, logw: complex):
    """wf' = B^H wf, logw' = conj(B.logw) + logw (hopSDOperation.cpp:21-26)."""
    return op.matrix.conj().T @ wf, op.logw.conjugate() + logw


class HopSDOperation:
    def applyToRight(self, op, walker, walkerNew=None):
        wf, lw = hop_apply_right(op, walker.wf, walker.logw)
        return _like(walker, wf, lw)

    def applyToLeft(self, op, walker, walkerNew=None):
        wf, lw = hop_apply_left(op, walker.wf, walker.logw)
        return _like(walker, wf, lw)


Hop2isSD2isOperation = HopSDOperation


def _like(walker, wf, logw):
    if isinstance(walker, SD2is):
        return SD2is(wf, walker.Nup, walker.Ndn, logw)
    return SD(wf=wf, logw=logw)


class LogHop2isSD2isOperation:
    """Taylor series exp(A) W (logHop2isSD2isOperation.cpp:12-195).

    flag "fixedOrder" with taylorOrder > 0: exactly taylorOrder terms;
    "fixedOrder" with taylorOrder == 0: first call determines the order
    dynamically and freezes it; "dynamicOrder": baseTaylorOrder terms, then
    at least one more, until max |T_k| <= accuracy.  Pinned by the literal
    values of logHop2isSD2isOperationTest.cpp:25-94."""

    def __init__(self, flag="dynamicOrder", taylorOrder=0, accuracy=1e-10, baseTaylorOrder=3):
        if flag not in ("fixedOrder", "dynamicOrder"):
            raise ValueError("Do not know the input flag for LogHop2isSD2isOperation: " + flag)
        self.flag = flag
        self.taylorOrder = taylorOrder
        self.accuracy = accuracy
        self.baseTaylorOrder = baseTaylorOrder
        self.currentOrder = 0
        self.orders = []

    def _apply(self, A, wf):
        new = wf.copy()
        old = wf.copy()
        k = 0

        def add_fixed(maxOrder):
            nonlocal new, old, k
            while k < maxOrder:
                t = (A @ old) * (1.0 / (k + 1.0))
                new = new + t
                old = t
                k += 1

        def add_dynamic():
            nonlocal new, old, k
            while True:
                t = (A @ old) * (1.0 / (k + 1.0))
                new = new + t
                old = t
                k += 1
                if not np.any(np.abs(old) > self.accuracy):
                    break

        if self.flag == "fixedOrder":
            if self.taylorOrder == 0:
                add_fixed(self.baseTaylorOrder)
                add_dynamic()
                self.taylorOrder = k
            else:
                add_fixed(self.taylorOrder)
        else:
            add_fixed(self.baseTaylorOrder)
            add_dynamic()
        self.currentOrder = k
        self.orders.append(k)
        return new

    def applyToRight(self, op, walker):
        wf = self._apply(op.matrix, walker.wf)
        return _like(walker, wf, op.logw + walker.logw)

    def applyToLeft(self, op, walker):
        wf = self._apply(op.matrix.conj().T, walker.wf)
        return _like(walker, wf, op.logw.conjugate() + walker.logw)


# --------------------------------------------------------------------------
# Walker-walker operations
# --------------------------------------------------------------------------
class SDSDOperation:
    """<L|R> for SD x SD (SDSDOperation.cpp:28-186)."""

    def __init__(self, walkerLeft: SD, walkerRight: SD):
        if walkerLeft.wf.shape != walkerRight.wf.shape:
            raise ValueError("Find Inconsistency between walkerLeft and walkerRight!")
        self.walkerLeft = walkerLeft
        self.walkerRight = walkerRight
        self._lu = None
        self._theta = None

    def returnLUOverlap(self):
        if self._lu is None:
            ovlp = self.walkerLeft.wf.conj().T @ self.walkerRight.wf
            self._lu = linalg.lu_construct(ovlp)
        return self._lu

    def returnLogOverlap(self):
        lu = self.returnLUOverlap()
        return self.walkerLeft.logw.conjugate() + self.walkerRight.logw + linalg.log_determinant(*lu)

    def returnTheta_T(self):
        """theta_T = O^{-T} R^T (zgetrs 'T', SDSDOperation.cpp:166-173)."""
        if self._theta is None:
            lu, ipiv, info = self.returnLUOverlap()
            self._theta = linalg.solve_lineq(lu, ipiv, self.walkerRight.wf.T, "T")
        return self._theta

    def returnGreenMatrix(self):
        return self.walkerLeft.wf.conj() @ self.returnTheta_T()

    def returnGreenDiagonal(self):
        th = self.returnTheta_T()
        return np.einsum("ij,ji->i", self.walkerLeft.wf.conj(), th)

    def returnGreenOffDiagonal(self):
        """(i, i+L/2) for i < L/2 and (i+L/2, i) (SDSDOperation.cpp:115-141)."""
        th = self.returnTheta_T()
        wl = self.walkerLeft.wf.conj()
        L = wl.shape[0]
        h = L // 2
        if L != 2 * h:
            raise ValueError("Green Matrix rank size is odd number!")
        out = np.empty(L, dtype=C128)
        out[:h] = np.einsum("ij,ji->i", wl[:h], th[:, h:])
        out[h:] = np.einsum("ij,ji->i", wl[h:], th[:, :h])
        return out


class SD2sSD2isOperation:
    """<SD2s|SD2is> (SD2sSD2isOperation.cpp:28-230)."""

    def __init__(self, walkerLeft: SD2s, walkerRight: SD2is):
        if (walkerLeft.getL() != walkerRight.getL() or walkerLeft.getNup() != walkerRight.Nup
                or walkerLeft.getNdn() != walkerRight.Ndn):
            raise ValueError("Find Inconsistency between walkerLeft and walkerRight!")
        self.walkerLeft = walkerLeft
        self.walkerRight = walkerRight
        self.wfRightUp = walkerRight.wf[:, :walkerRight.Nup]
        self.wfRightDn = walkerRight.wf[:, :walkerRight.Ndn]
        self._lu = None
        self._theta = None

    def returnLUOverlap(self):
        if self._lu is None:
            ou = self.walkerLeft.wfUp.conj().T @ self.wfRightUp
            od = self.walkerLeft.wfDn.conj().T @ self.wfRightDn
            self._lu = (linalg.lu_construct(ou), linalg.lu_construct(od))
        return self._lu

    def returnLogOverlap(self):
        lu_up, lu_dn = self.returnLUOverlap()
        return (self.walkerLeft.logw.conjugate() + self.walkerRight.logw
                + linalg.log_determinant(*lu_up) + linalg.log_determinant(*lu_dn))

    def returnTheta_T(self):
        if self._theta is None:
            (lu, ip, _), (lud, ipd, _) = self.returnLUOverlap()
            tu = linalg.solve_lineq(lu, ip, self.wfRightUp.T, "T")
            td = linalg.solve_lineq(lud, ipd, self.wfRightDn.T, "T")
            self._theta = (tu, td)
        return self._theta

    def returnGreenDiagonalUp(self):
        tu, _ = self.returnTheta_T()
        return np.einsum("ij,ji->i", self.walkerLeft.wfUp.conj(), tu)

    def returnGreenDiagonalDn(self):
        _, td = self.returnTheta_T()
        return np.einsum("ij,ji->i", self.walkerLeft.wfDn.conj(), td)


# --------------------------------------------------------------------------
# Discrete HS transform of U n_up n_dn
# --------------------------------------------------------------------------
class NiupNidnSample:
    """4 x L complex 2x2 site blocks + logw (NiupNidnSample.h)."""

    def __init__(self, L):
        self.diag00 = np.zeros(L, dtype=C128)
        self.diag01 = np.zeros(L, dtype=C128)
        self.diag10 = np.zeros(L, dtype=C128)
        self.diag11 = np.zeros(L, dtype=C128)
        self.logw = 0j


class NiupNidn:
    """afqmc/blocks/twoBodyOperator/NiupNidn/source/NiupNidn.cpp."""

    def __init__(self, dt, decompType, U, mu, hx, hy, hz):
        U = np.asarray(U, dtype=np.float64)
        self.L = U.size
        self.decompType = decompType
        s = 0.0
        for x in U:
            s += x
        self.dtUSum = dt * s
        self.dtU = dt * U
        self.gamma = np.zeros(self.L, dtype=C128)
        for i in range(self.L):
            if decompType == "densityCharge":
                self.gamma[i] = solveCoshxEqExpy(-self.dtU[i] * 0.5)
            elif decompType in ("densitySpin", "hopCharge"):
                self.gamma[i] = solveCoshxEqExpy(self.dtU[i] * 0.5)
            elif decompType == "hopSpin":
                self.gamma[i] = solveCosxEqExpy(self.dtU[i] * 0.5)
            else:
                raise ValueError("Can not find the matched decompType! " + decompType)
        self.constDiag00 = np.zeros(self.L, dtype=C128)
        self.constDiag01 = np.zeros(self.L, dtype=C128)
        self.constDiag10 = np.zeros(self.L, dtype=C128)
        self.constDiag11 = np.zeros(self.L, dtype=C128)
        for i in range(self.L):
            a = -dt * (U[i] * 0.5 - mu[i] + hz[i] * 0.5)
            b = -dt * (U[i] * 0.5 - mu[i] - hz[i] * 0.5)
            c = -dt * 0.5 * complex(hx[i], hy[i])
            a, b, c = exp2by2Matrix(a, b, c)
            self.constDiag00[i] = a
            self.constDiag01[i] = c.conjugate()
            self.constDiag10[i] = c
            self.constDiag11[i] = b

    def sampleAuxFromForce(self, force, u):
        """x_i = -1 if u_i < e^{-f}/(e^{-f}+e^{f}) else +1 (NiupNidn.cpp:66-84)."""
        aux = np.empty(self.L, dtype=np.int32)
        for i in range(self.L):
            em = math.exp(-force[i])
            ep = math.exp(force[i])
            prob = em / (em + ep)
            aux[i] = -1 if u[i] < prob else 1
        return aux

    def logProbOfAuxFromForce(self, aux, force):
        """sum_i x_i f_i - log(e^{f_i}+e^{-f_i}) (NiupNidn.cpp:86-98)."""
        s = 0.0
        for i in range(self.L):
            s += aux[i] * force[i] - math.log(math.exp(force[i]) + math.exp(-force[i]))
        return s

    def getTwoBodySampleFromAux(self, aux):
        """NiupNidn.cpp:100-131 and 252-322."""
        s = NiupNidnSample(self.L)
        for i in range(self.L):
            g = aux[i] * 1.0 * self.gamma[i]
            if self.decompType == "densityCharge":
                e = cmath.exp(g)
                s.diag00[i] = e * self.constDiag00[i]
                s.diag10[i] = e * self.constDiag10[i]
                s.diag01[i] = e * self.constDiag01[i]
                s.diag11[i] = e * self.constDiag11[i]
            elif self.decompType == "densitySpin":
                e = cmath.exp(g)
                em = 1.0 / e
                s.diag00[i] = e * self.constDiag00[i]
                s.diag10[i] = em * self.constDiag10[i]
                s.diag01[i] = e * self.constDiag01[i]
                s.diag11[i] = em * self.constDiag11[i]
            elif self.decompType == "hopCharge":
                ch, sh = cmath.cosh(g), cmath.sinh(g)
                s.diag00[i] = ch * self.constDiag00[i] + sh * self.constDiag10[i]
                s.diag10[i] = sh * self.constDiag00[i] + ch * self.constDiag10[i]
                s.diag01[i] = ch * self.constDiag01[i] + sh * self.constDiag11[i]
                s.diag11[i] = sh * self.constDiag01[i] + ch * self.constDiag11[i]
            elif self.decompType == "hopSpin":
                co, si = cmath.cos(g), cmath.sin(g)
                s.diag00[i] = co * self.constDiag00[i] + si * self.constDiag10[i]
                s.diag10[i] = -si * self.constDiag00[i] + co * self.constDiag10[i]
                s.diag01[i] = co * self.constDiag01[i] + si * self.constDiag11[i]
                s.diag11[i] = -si * self.constDiag01[i] + co * self.constDiag11[i]
        if self.decompType == "densityCharge":
            gs = _kahan_sum([aux[i] * 1.0 * self.gamma[i] for i in range(self.L)])
            s.logw = math.log(0.5) * self.L + self.dtUSum * 0.5 - gs
        else:
            s.logw = complex(math.log(0.5) * self.L)
        return s

    def getTwoBodySampleFromAuxForce(self, aux, force):
        s = self.getTwoBodySampleFromAux(aux)
        s.logw = s.logw - self.logProbOfAuxFromForce(aux, force)
        return s


def _kahan_sum(values):
    """KahanData<complex> accumulation (common/erroranalysisHao/include/kahan.h)."""
    s = 0j
    c = 0j
    for v in values:
        y = v - c
        t = s + y
        c = (t - s) - y
        s = t
    return s


def niupnidn_apply_right(sample: NiupNidnSample, wf, logw):
    """Row-pair 2x2 scaling (NiupNidnSDOperation.cpp:15-41)."""
    h = sample.diag00.size
    if wf.shape[0] != 2 * h:
        raise ValueError("NiupNidnSample size is not consistent with walker!")
    up = wf[:h]
    dn = wf[h:]
    new = np.empty_like(wf)
    new[:h] = sample.diag00[:, None] * up + sample.diag01[:, None] * dn
    new[h:] = sample.diag10[:, None] * up + sample.diag11[:, None] * dn
    return new, sample.logw + logw


def niupnidn_apply_left(sample: NiupNidnSample, wf, logw):
    """NiupNidnSDOperation.cpp:43-68."""
    h = sample.diag00.size
    d00 = sample.diag00.conj()
    d10 = sample.diag01.conj()
    d01 = sample.diag10.conj()
    d11 = sample.diag11.conj()
    up = wf[:h]
    dn = wf[h:]
    new = np.empty_like(wf)
    new[:h] = d00[:, None] * up + d01[:, None] * dn
    new[h:] = d10[:, None] * up + d11[:, None] * dn
    return new, complex(sample.logw).conjugate() + logw


# --------------------------------------------------------------------------
# Continuous HS over real Cholesky vectors
# --------------------------------------------------------------------------
class CholeskyReal:
    """afqmc/blocks/twoBodyOperator/choleskyReal/source/choleskyReal.cpp.

    choleskyVecs has shape (L, L, G) (column-major L x L per vector);
    choleskyBg is a *reference* to the model's background array, as the
    reference keeps a pointer (choleskyReal.h:19)."""

    def __init__(self, dt, choleskyVecs, choleskyBg_ref):
        self.dt = dt
        self.sqrtMinusDt = cmath.sqrt(complex(-dt, 0.0))
        self.choleskyNumber = choleskyVecs.shape[2]
        self.vecs = choleskyVecs
        self.bg_holder = choleskyBg_ref

    @property
    def choleskyBg(self):
        return self.bg_holder.choleskyBg

    def sampleAuxFromForce(self, force, gauss):
        """x_g = gaussianHao() + f_g (choleskyReal.cpp:57-65)."""
        return np.asarray(gauss, dtype=np.float64) + np.asarray(force, dtype=C128)

    def logProbOfAuxFromForce(self, aux, force):
        t = aux - force
        s = 0j
        for v in t:
            s += v * v
        return -0.5 * math.log(2.0 * 3.14159265358979324) * self.choleskyNumber - 0.5 * s

    def sample_matrix(self, aux):
        """M = sum_g x_g sqrt(-dt) L_g (zgemv over (L^2 x G), choleskyReal.cpp:160-169)."""
        L = self.vecs.shape[0]
        v = self.vecs.reshape(L * L, self.choleskyNumber, order="F")
        m = (v.astype(C128) * self.sqrtMinusDt) @ aux
        return np.array(m.reshape(L, L, order="F"), order="F")

    def getTwoBodySampleFromAuxForce(self, aux, force):
        """logw = 0.5 sum f^2 - sum x f - sqrt(-dt) sum x bg (choleskyReal.cpp:104-124)."""
        f2 = 0j
        xf = 0j
        xb = 0j
        bg = self.choleskyBg
        for i in range(self.choleskyNumber):
            f2 += force[i] * force[i]
            xf += aux[i] * force[i]
            xb += aux[i] * bg[i]
        logw = 0.5 * f2 - xf - self.sqrtMinusDt * xb
        return LogHop2is(self.sample_matrix(aux), logw)

    def getTwoBodySampleFromAux(self, aux):
        a2 = 0j
        xb = 0j
        bg = self.choleskyBg
        for i in range(self.choleskyNumber):
            a2 += aux[i] * aux[i]
            xb += aux[i] * bg[i]
        logw = -0.5 * math.log(2.0 * 3.14159265358979324) * self.choleskyNumber - 0.5 * a2 - self.sqrtMinusDt * xb
        return LogHop2is(self.sample_matrix(aux), logw)
