"""Spin-block form of the constraint-path walker loop (test infrastructure and CPU baseline only).

When K, the trial and the HS field do not mix spins (no hx / hy, a density decomposition), the
reference's dense 2L x N walker W = [[W_up, 0], [0, W_dn]] stays block diagonal under every
operation of its step (afqmcConstraintPathProjection.cpp:37-74):

* exp(-dt K) is block diagonal (HubbardSOC.cpp:109-118);
* the HS factor is diagonal (diag00 on the up rows, diag11 on the down rows,
  NiupNidnSDOperation.cpp:15-41, with diag01 = diag10 = 0);
* the overlap phiT^H W is block diagonal, so zgetrf's pivots never leave a block and the
  log-determinant (tensor_2d_common_fun.cpp:75-105) is the sum of the blocks' log-moduli plus
  the principal log of the product of the blocks' phases;
* theta and the Green's function are block diagonal (G(i, i+L) = 0);
* Householder QR (tensor_2d_bl_cpu.cpp:263-315) works column by column within each block, and
  the reference's sign rule applies to the product of both blocks' diagonals.

This module runs the oracle's per-walker step (oracle.drivers.AfqmcConstraintPath) on the two
blocks: the same numbers up to rounding at four times fewer flops, the representation the
device path executes.  tests/test_oracle_spin_blocks.py checks it against the dense oracle on
the same uniforms (fields, log-weights, walkers, comb tables and block energies).
"""
import cmath
import math

import numpy as np

from . import comb, linalg
from .blocks import SD
from .drivers import AfqmcConstraintPath


class BlockSD:
    """A spin-block walker: up (L x Nup) and down (L x Ndn) blocks and logw."""

    def __init__(self, up, dn, logw=0j):
        self.up = np.asfortranarray(np.array(up, dtype=np.complex128))
        self.dn = np.asfortranarray(np.array(dn, dtype=np.complex128))
        self.logw = complex(logw)

    @classmethod
    def from_dense(cls, sd: SD, nup):
        L = sd.wf.shape[0] // 2
        return cls(sd.wf[:L, :nup], sd.wf[L:, nup:], sd.logw)

    def to_dense(self) -> SD:
        L, nup, ndn = self.up.shape[0], self.up.shape[1], self.dn.shape[1]
        wf = np.zeros((2 * L, nup + ndn), dtype=np.complex128, order="F")
        wf[:L, :nup] = self.up
        wf[L:, nup:] = self.dn
        return SD(wf=wf, logw=self.logw)

    def copy(self):
        return BlockSD(self.up.copy(), self.dn.copy(), self.logw)

    def addLogw(self, x):
        self.logw += complex(x)

    def stabilize(self):
        """SD::stabilize (SD.cpp:42-45) of the block-diagonal walker: QR of each block, the
        reference's sign rule on the product of all R_ii (column 0 of the up block negated)."""
        qu, du, _ = _qr_unsigned(self.up)
        qd, dd, _ = _qr_unsigned(self.dn)
        det = du * dd
        if det.real < 0:
            det = -det
            qu[:, 0] = -qu[:, 0]
        self.up, self.dn = qu, qd
        self.logw += math.log(det.real)


def _qr_unsigned(a):
    """zgeqrf + zungqr of one block without the sign rule: (Q, prod R_ii, R_ii)."""
    from scipy.linalg import lapack
    qr, tau, work, info = lapack.zgeqrf(np.asfortranarray(a))
    if info != 0:
        raise ValueError("QR run is not successful")
    d = np.array([qr[i, i] for i in range(a.shape[1])])
    det = complex(1.0, 0.0)
    for x in d:
        det *= x
    q, work, info = lapack.zungqr(qr, tau)
    if info != 0:
        raise ValueError("Zungqr run is not successful")
    return np.array(q[:, :a.shape[1]], order="F"), det, d


def _block_logdet(lu, ipiv, info):
    """(sum log|u_ii|, product of the +-u_ii/|u_ii| phases) of one block's zgetrf."""
    if info > 0:
        return -1e300, complex(1.0, 0.0)
    lognorm = 0.0
    phase = complex(1.0, 0.0)
    for i in range(lu.shape[0]):
        u = complex(lu[i, i])
        au = abs(u)
        lognorm += math.log(au)
        phase *= (-u / au) if ipiv[i] != i else (u / au)
    return lognorm, phase


class BlockSDSDOperation:
    """SDSDOperation (SDSDOperation.cpp) of a block trial and a block walker."""

    def __init__(self, left: BlockSD, right: BlockSD):
        self.left, self.right = left, right
        self._lu = [linalg.lu_construct(left.up.conj().T @ right.up), linalg.lu_construct(left.dn.conj().T @ right.dn)]
        self._theta = None

    def returnLogOverlap(self):
        (n0, p0), (n1, p1) = _block_logdet(*self._lu[0]), _block_logdet(*self._lu[1])
        if n0 <= -1e300 or n1 <= -1e300:
            return complex(-1e300, 0.0)
        return self.left.logw.conjugate() + self.right.logw + complex(n0 + n1, 0.0) + cmath.log(p0 * p1)

    def returnTheta_T(self):
        if self._theta is None:
            self._theta = [linalg.solve_lineq(lu[0], lu[1], w.T, "T")
                           for lu, w in zip(self._lu, (self.right.up, self.right.dn))]
        return self._theta

    def returnGreenDiagonal(self):
        tu, td = self.returnTheta_T()
        return np.concatenate([np.einsum("ij,ji->i", self.left.up.conj(), tu),
                               np.einsum("ij,ji->i", self.left.dn.conj(), td)])


class AfqmcConstraintPathSpinBlocks(AfqmcConstraintPath):
    """oracle.drivers.AfqmcConstraintPath on spin-block walkers (densityCharge / densitySpin with
    hx = hy = 0).  Same method keys, step order, cadences and uniform consumption."""

    def __init__(self, method, model, rng, phiT: SD, nup, constForce=None):
        if method.decompType not in ("densityCharge", "densitySpin"):
            raise ValueError("spin blocks need a density decomposition")
        if np.any(model.hx != 0) or np.any(model.hy != 0):
            raise ValueError("spin blocks need hx = hy = 0")
        super().__init__(method, model, rng, phiT, constForce)
        L = model.L
        self.nup = nup
        self.blockPhiT = BlockSD.from_dense(self.phiT, nup)
        for name in ("expMinusDtK", "expMinusHalfDtK", "expHalfDtK"):
            op = getattr(self, name)
            setattr(self, name + "Blocks", (op.matrix[:L, :L].copy(), op.matrix[L:, L:].copy(), op.logw))
        wdk = self.phiT.wf.conj().T @ model.K  # initWfDaggerK (HubbardSOCMeasureFixedSDSD.cpp:160-167)
        self.wdkBlocks = (wdk[:nup, :L], wdk[nup:, L:])
        self.reSetMeasure()

    def _kinetic(self, blocks, w: BlockSD):
        bu, bd, lw = blocks
        return BlockSD(bu @ w.up, bd @ w.dn, lw + w.logw)

    def initialWalker(self):
        n = self.method.walkerSizePerThread
        self.walker = [self.blockPhiT.copy() for _ in range(n)]
        self.walkerIsAlive = [True] * n
        self.modifyGM()
        self.popControl()

    def projectExpMinusHalfDtK(self):
        for i, w in enumerate(self.walker):
            if self.walkerIsAlive[i]:
                self.walker[i] = self._kinetic(self.expMinusHalfDtKBlocks, w)

    def step_walker(self, i):
        """afqmcConstraintPathProjection.cpp:37-74 on the blocks."""
        m = self.method
        w = self.walker[i]
        op = BlockSDSDOperation(self.blockPhiT, w)
        phase = op.returnLogOverlap().imag
        if math.cos(phase) <= 0.0:
            self.walkerIsAlive[i] = False
            return
        w.addLogw(math.log(math.cos(phase)) - 1j * phase)
        L = self.model.L
        hs = self.expMinusDtV
        if m.forceType == "constForce":
            force = self.constForce
        elif m.forceType == "dynamicForce":
            g = op.returnGreenDiagonal()
            bg = (g[:L] + g[L:] - 1.0) if hs.decompType == "densityCharge" else (g[:L] - g[L:])
            force = np.clip((hs.gamma * bg).real, -m.forceCap, m.forceCap)
        else:
            raise ValueError("Do not know method.forceType " + m.forceType)
        u = self.rng.take(L)
        aux = hs.sampleAuxFromForce(force, u)
        sample = hs.getTwoBodySampleFromAuxForce(aux, force)
        up = sample.diag00[:, None] * w.up
        dn = sample.diag11[:, None] * w.dn
        lw = sample.logw + w.logw + m.dt * m.ET
        self.walker[i] = self._kinetic(self.expMinusDtKBlocks, BlockSD(up, dn, lw))
        self.records.append((i, force, aux))

    def popControl(self):
        """afqmcConstraintPathProjection.cpp:103-156 (serial branch) on the blocks."""
        n = len(self.walker)
        weight = np.zeros(n)
        for i, w in enumerate(self.walker):
            if self.walkerIsAlive[i]:
                lo = BlockSDSDOperation(self.blockPhiT, w).returnLogOverlap()
                o = cmath.exp(lo).real
                if o <= 0.0:
                    weight[i] = 0.0
                    self.walkerIsAlive[i] = False
                else:
                    weight[i] = o
                    w.addLogw(-lo)
        self.walkerIsAlive = [True] * n
        table = comb.popConfiguration(1, weight, self.rng.take(n))
        comb.populationControl(self.walker, table)
        self.last_weight, self.last_table = weight, table
        return weight, table

    def reSetMeasure(self):
        self.den = self.HNum = self.KNum = self.VNum = self.RNum = 0j

    def addMixedMeasurement(self):
        """afqmcConstraintPathMeasure.cpp:9-27 with HubbardSOCMeasureFixedSDSD::addEnergy
        (:178-218) on the blocks (the spin-flip terms vanish: G(i, i+L) = 0)."""
        L = self.model.L
        mdl = self.model
        for i, w in enumerate(self.walker):
            if not self.walkerIsAlive[i]:
                continue
            t = self._kinetic(self.expHalfDtKBlocks, w)
            op = BlockSDSDOperation(self.blockPhiT, t)
            den = cmath.exp(op.returnLogOverlap())
            tu, td = op.returnTheta_T()
            g = op.returnGreenDiagonal()
            ke = np.sum(self.wdkBlocks[0] * tu) + np.sum(self.wdkBlocks[1] * td)
            ve = np.sum(mdl.U * g[:L] * g[L:])
            re = np.sum((-mdl.mu + 0.5 * mdl.hz) * g[:L]) + np.sum((-mdl.mu - 0.5 * mdl.hz) * g[L:])
            self.den += den
            self.KNum += ke * den
            self.VNum += ve * den
            self.RNum += re * den
            self.HNum += (ke + ve + re) * den

    def measureWithProjection(self):
        """afqmcConstraintPath.cpp:99-150 with this class's measurement accumulators."""
        m = self.method
        self.projectExpMinusHalfDtK()
        idx = [0, 0]
        for i in range(m.thermalSize):
            if i < m.ETAdjustMaxSize and i % m.ETAdjustStep == 0:
                self.addMixedMeasurement()
                m.ET = (self.HNum / self.den).real
                self.reSetMeasure()
            self.projectOneStep(idx)
        for i in range(m.writeNumber):
            for j in range(m.measureNumberPerWrite):
                self.addMixedMeasurement()
                for k in range(m.measureSkipStep):
                    self.projectOneStep(idx)
            self.energies.append((self.HNum, self.den, self.KNum, self.VNum, self.RNum))
            self.reSetMeasure()
