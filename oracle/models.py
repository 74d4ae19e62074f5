"""Models and fixed-trial mixed measurements (test infrastructure only).

HubbardSOC follows afqmc/model/HubbardRealSpaceSOC/source/HubbardSOC.cpp and
HubbardSOCMeasureFixedSDSD.cpp; RealMaterialMolecule follows
afqmc/model/realMaterialMolecule/source/realMaterialMolecule.cpp and
realMaterialMoleculeMeasureFixedSD2sSD2is.cpp.
"""
import numpy as np

from . import linalg
from .blocks import SD, SD2s, SD2is, Hop, LogHop2is, NiupNidn, CholeskyReal, C128


# --------------------------------------------------------------------------
# Lattice hopping (scripts/supercubic/setHoping.py, lattClass.py:65-73)
# --------------------------------------------------------------------------
def square_lattice_K(nx, ny, t1=1.0, t2=0.0, spin_blocks=True):
    """2Ls x 2Ls spin-orbit-form K of a periodic nx x ny lattice.

    Site index = x + y*nx; nearest-neighbour hop -t1 to x+-1 and y+-1 and, if
    t2 != 0, next-nearest -t2 to (x+-1, y+-1), accumulated with += as the
    reference's generator does (generateInputFile.py:26-33); up and down
    blocks are identical."""
    Ls = nx * ny
    k = np.zeros((Ls, Ls), dtype=C128)

    def idx(x, y):
        return (x % nx) + (y % ny) * nx

    for y in range(ny):
        for x in range(nx):
            i = idx(x, y)
            for d in (-1, 1):
                k[i, idx(x + d, y)] += -t1
            for d in (-1, 1):
                k[i, idx(x, y + d)] += -t1
            if t2 != 0.0:
                for dx in (-1, 1):
                    for dy in (-1, 1):
                        k[i, idx(x + dx, y + dy)] += -t2
    K = np.zeros((2 * Ls, 2 * Ls), dtype=C128, order="F")
    K[:Ls, :Ls] = k
    K[Ls:, Ls:] = k
    return K


class HubbardSOC:
    """Hubbard model with spin-orbit-form K (HubbardSOC.cpp)."""

    def __init__(self, L, N, K, mu=None, hx=None, hy=None, hz=None, U=None):
        self.L = int(L)
        self.N = int(N)
        self.K = np.array(K, dtype=C128, order="F")
        z = np.zeros(self.L)
        self.mu = z.copy() if mu is None else np.asarray(mu, dtype=np.float64)
        self.hx = z.copy() if hx is None else np.asarray(hx, dtype=np.float64)
        self.hy = z.copy() if hy is None else np.asarray(hy, dtype=np.float64)
        self.hz = z.copy() if hz is None else np.asarray(hz, dtype=np.float64)
        self.U = z.copy() if U is None else np.asarray(U, dtype=np.float64)
        self._eig = None

    @classmethod
    def read(cls, filename):
        """model_param text format (HubbardSOC.cpp:39-59)."""
        with open(filename) as f:
            toks = f.read().split()
        p = 0
        L = int(toks[p]); p += 1
        N = int(toks[p]); p += 1
        n2 = (2 * L) ** 2
        vals = np.array(toks[p:p + 2 * n2], dtype=np.float64); p += 2 * n2
        K = (vals[0::2] + 1j * vals[1::2]).reshape(2 * L, 2 * L, order="F")
        arrs = []
        for _ in range(5):
            arrs.append(np.array(toks[p:p + L], dtype=np.float64)); p += L
        return cls(L, N, K, *arrs)

    def write(self, filename):
        with open(filename, "w") as f:
            f.write("{:16d} \n".format(self.L))
            f.write("{:16d} \n".format(self.N))
            for v in self.K.reshape(-1, order="F"):
                f.write("{:26.18e} {:26.18e} \n".format(v.real, v.imag))
            for arr in (self.mu, self.hx, self.hy, self.hz, self.U):
                for v in arr:
                    f.write("{:26.18e} \n".format(v))

    def setKEigenValueAndVector(self):
        if self._eig is None:
            if not np.allclose(self.K, self.K.conj().T, atol=1e-12):
                raise ValueError("ERROR!!! Matrix is not Hermition!")
            self._eig = linalg.eigen_hermitian(self.K)
        return self._eig

    def returnExpMinusAlphaK(self, alpha):
        """V exp(-alpha Lambda) V^H (HubbardSOC.cpp:109-118)."""
        w, v = self.setKEigenValueAndVector()
        m = v @ (np.exp(-alpha * w)[:, None] * v.conj().T)
        return Hop(np.asfortranarray(m), 0j)

    def returnExpMinusAlphaV(self, alpha, decompType):
        return NiupNidn(alpha, decompType, self.U, self.mu, self.hx, self.hy, self.hz)


def fillWalkerFromModel(model: HubbardSOC) -> SD:
    """Lowest N eigenvectors of K (HubbardSOCSDOperation.cpp:19-30)."""
    w, v = model.setKEigenValueAndVector()
    return SD(wf=np.array(v[:, :model.N], order="F"), logw=0j)


class HubbardSOCMeasureFixedSDSD:
    """Mixed estimator and force bias with a fixed trial (HubbardSOCMeasureFixedSDSD.cpp)."""

    def __init__(self, model: HubbardSOC, walkerLeft: SD):
        if 2 * model.L != walkerLeft.getL():
            raise ValueError("Model L does not consistent with walker L!")
        if model.N != walkerLeft.getN():
            raise ValueError("Model N does not consistent with walker N!")
        self.model = model
        self.walkerLeft = walkerLeft
        self.wfDaggerK = walkerLeft.wf.conj().T @ model.K  # N x 2L (initWfDaggerK, :160-167)
        self.reSet()

    def reSet(self):
        self.den = 0j
        self.KNum = 0j
        self.VNum = 0j
        self.RNum = 0j
        self.HNum = 0j
        L = self.model.L
        self.NupNum = np.zeros(L, dtype=C128)
        self.NdnNum = np.zeros(L, dtype=C128)
        self.SplusNum = np.zeros(L, dtype=C128)
        self.SminusNum = np.zeros(L, dtype=C128)
        self.NupTotNum = 0j
        self.NdnTotNum = 0j
        self.SplusTotNum = 0j
        self.SminusTotNum = 0j

    def returnEnergy(self):
        return self.HNum / self.den

    @staticmethod
    def local_energy(model, wfDaggerK, op):
        """(K, V, R) local energies of one walker (addEnergy, :178-218)."""
        L = model.L
        th = op.returnTheta_T()
        gd = op.returnGreenDiagonal()
        go = op.returnGreenOffDiagonal()
        Ke = np.sum(wfDaggerK * th)
        Ve = np.sum(model.U * (gd[:L] * gd[L:] - go[:L] * go[L:]))
        Re = (np.sum((-model.mu + model.hz * 0.5) * gd[:L]) + np.sum((-model.mu - model.hz * 0.5) * gd[L:])
              + np.sum((model.hx * 0.5 - 0.5j * model.hy) * go[:L]) + np.sum((model.hx * 0.5 + 0.5j * model.hy) * go[L:]))
        return Ke, Ve, Re, gd, go

    def addMeasurement(self, op, denIncrement):
        if op.walkerLeft is not self.walkerLeft:
            raise ValueError("HubbardSOCMeasureFixedSDSD only accept sdsdOperation with fixed SD!")
        L = self.model.L
        self.den += denIncrement
        Ke, Ve, Re, gd, go = self.local_energy(self.model, self.wfDaggerK, op)
        self.KNum += Ke * denIncrement
        self.VNum += Ve * denIncrement
        self.RNum += Re * denIncrement
        self.HNum += (Ke + Ve + Re) * denIncrement
        self.NupNum += gd[:L] * denIncrement
        self.NdnNum += gd[L:] * denIncrement
        self.NupTotNum += gd[:L].sum() * denIncrement
        self.NdnTotNum += gd[L:].sum() * denIncrement
        self.SplusNum += go[:L] * denIncrement
        self.SminusNum += go[L:] * denIncrement
        self.SplusTotNum += go[:L].sum() * denIncrement
        self.SminusTotNum += go[L:].sum() * denIncrement

    def getForce(self, niupNidn: NiupNidn, op, cap):
        """f_i = clip(Re(gamma_i * bg_i), +-cap) (:73-116)."""
        h = niupNidn.L
        t = niupNidn.decompType
        if t == "densityCharge":
            g = op.returnGreenDiagonal()
            bg = g[:h] + g[h:] - 1.0
        elif t == "densitySpin":
            g = op.returnGreenDiagonal()
            bg = g[:h] - g[h:]
        elif t == "hopCharge":
            g = op.returnGreenOffDiagonal()
            bg = g[:h] + g[h:]
        elif t == "hopSpin":
            g = op.returnGreenOffDiagonal()
            bg = g[:h] - g[h:]
        else:
            raise ValueError("Can not find the matched decompType! " + t)
        f = (niupNidn.gamma * bg).real
        return np.clip(f, -cap, cap)


class HubbardSOCMeasureObserveSDSD:
    """Observables of a general bra/ket pair from its full Green matrix G(i,j) = <c+_i c_j>
    (HubbardSOCMeasureObserveSDSD.cpp:34-183 over HubbardSOCMeasureCommuteSDSD.cpp:45-180):
    den, H/K/V/R, greenMatrixNum (2L x 2L), densityDensityNum (2L x 2L, <n_a n_b>),
    splusSminusNum, sminusSplusNum, spairSpairNum (L x L), each summed with weight denIncrement.
    The back-propagation driver measures through it.  Pinned by a Fock-space brute force of every
    operator (tests/test_oracle_bp.py)."""

    def __init__(self, model: HubbardSOC):
        self.model = model
        self.reSet()

    def reSet(self):
        L = self.model.L
        self.den = 0j
        self.HNum = 0j
        self.KNum = 0j
        self.VNum = 0j
        self.RNum = 0j
        self.greenMatrixNum = np.zeros((2 * L, 2 * L), dtype=C128)
        self.densityDensityNum = np.zeros((2 * L, 2 * L), dtype=C128)
        self.splusSminusNum = np.zeros((L, L), dtype=C128)
        self.sminusSplusNum = np.zeros((L, L), dtype=C128)
        self.spairSpairNum = np.zeros((L, L), dtype=C128)

    def returnEnergy(self):
        return self.HNum / self.den

    def getForce(self, niupNidn: NiupNidn, op, cap):
        """HubbardSOCMeasureCommuteSDSD.cpp:63-104: the same background as the fixed-trial force."""
        return HubbardSOCMeasureFixedSDSD.getForce(self, niupNidn, op, cap)

    @staticmethod
    def energy_terms(model, G):
        """(K, V, R) of addEnergy (HubbardSOCMeasureCommuteSDSD.cpp:143-180)."""
        L = model.L
        d = np.diag(G)
        Ke = np.sum(model.K * G)
        Ve = np.sum(model.U * (d[:L] * d[L:] - np.diag(G[:L, L:]) * np.diag(G[L:, :L])))
        Re = (np.sum((-model.mu + model.hz * 0.5) * d[:L]) + np.sum((-model.mu - model.hz * 0.5) * d[L:])
              + np.sum((model.hx * 0.5 - 0.5j * model.hy) * np.diag(G[:L, L:]))
              + np.sum((model.hx * 0.5 + 0.5j * model.hy) * np.diag(G[L:, :L])))
        return Ke, Ve, Re

    @staticmethod
    def correlators(G):
        """densityDensity, splusSminus, sminusSplus, spairSpair of one Green matrix
        (HubbardSOCMeasureObserveSDSD.cpp:77-183), written as loops over (i, j) as the reference."""
        L2 = G.shape[0]
        L = L2 // 2
        dd = np.empty((L2, L2), dtype=C128)
        for j in range(L2):
            for i in range(L2):
                dd[i, j] = G[i, i] if i == j else G[i, i] * G[j, j] - G[i, j] * G[j, i]
        pm = np.empty((L, L), dtype=C128)
        mp = np.empty((L, L), dtype=C128)
        sp = np.empty((L, L), dtype=C128)
        for j in range(L):
            for i in range(L):
                det = G[i, i] * G[i + L, i + L] - G[i, i + L] * G[i + L, i]
                if i == j:
                    pm[i, j] = G[i, i] - det
                    mp[i, j] = G[i + L, i + L] - det
                else:
                    pm[i, j] = -G[i, j] * G[j + L, i + L] + G[i, i + L] * G[j + L, j]
                    mp[i, j] = -G[i + L, j + L] * G[j, i] + G[i + L, i] * G[j, j + L]
                sp[i, j] = G[i, j] * G[i + L, j + L] - G[i, j + L] * G[i + L, j]
        return dd, pm, mp, sp

    def addMeasurement(self, op, denIncrement):
        G = op.returnGreenMatrix()
        self.den += denIncrement
        Ke, Ve, Re = self.energy_terms(self.model, G)
        self.HNum += (Ke + Ve + Re) * denIncrement
        self.KNum += Ke * denIncrement
        self.VNum += Ve * denIncrement
        self.RNum += Re * denIncrement
        dd, pm, mp, sp = self.correlators(G)
        self.greenMatrixNum += G * denIncrement
        self.densityDensityNum += dd * denIncrement
        self.splusSminusNum += pm * denIncrement
        self.sminusSplusNum += mp * denIncrement
        self.spairSpairNum += sp * denIncrement

    def snapshot(self):
        """The accumulators as write() puts them on disk (HubbardSOCMeasureObserveSDSD.cpp:47-57)."""
        return dict(den=self.den, HNum=self.HNum, KNum=self.KNum, VNum=self.VNum, RNum=self.RNum,
                    greenMatrixNum=self.greenMatrixNum.copy(), densityDensityNum=self.densityDensityNum.copy(),
                    splusSminusNum=self.splusSminusNum.copy(), sminusSplusNum=self.sminusSplusNum.copy(),
                    spairSpairNum=self.spairSpairNum.copy())


# --------------------------------------------------------------------------
# Real-materials molecule (phaseless)
# --------------------------------------------------------------------------
class RealMaterialMolecule:
    """realMaterialMolecule.cpp.  choleskyVecs shape (L, L, G)."""

    def __init__(self, L, Nup, Ndn, t, K, choleskyVecs, choleskyBg):
        self.L = int(L)
        self.Nup = int(Nup)
        self.Ndn = int(Ndn)
        self.t = np.asarray(t, dtype=np.float64)
        self.K = np.asarray(K, dtype=np.float64)
        self.choleskyVecs = np.asarray(choleskyVecs, dtype=np.float64)
        self.choleskyNumber = self.choleskyVecs.shape[2]
        self.choleskyBg = np.array(choleskyBg, dtype=np.float64)
        self._kp = None
        self._eig = None

    def updateBackGround(self, bg):
        bg = np.asarray(bg, dtype=np.float64)
        if bg.size != self.choleskyNumber:
            raise ValueError("Background size is not choleskyNumber!")
        self.choleskyBg = bg.copy()
        self._kp = None
        self._eig = None

    def getKp(self):
        """Kp = K + sum_g bg_g L_g (setKp, :152-164)."""
        if self._kp is None:
            L = self.L
            v = self.choleskyVecs.reshape(L * L, self.choleskyNumber, order="F")
            self._kp = self.K + (v @ self.choleskyBg).reshape(L, L, order="F")
        return self._kp

    def _eigen(self):
        if self._eig is None:
            self._eig = linalg.eigen_symmetric(self.getKp())
        return self._eig

    def returnExpMinusAlphaK(self, alpha):
        """Hop2is, logw = alpha/2 sum bg^2 (:117-132)."""
        w, v = self._eigen()
        m = v @ (np.exp(-alpha * w)[:, None] * v.T)
        bg2 = 0.0
        for b in self.choleskyBg:
            bg2 += b * b
        return Hop(np.asfortranarray(m.astype(C128)), complex(alpha * 0.5 * bg2))

    def returnLogExpMinusAlphaK(self, alpha):
        """LogHop2is with matrix -alpha Kp (:134-146)."""
        bg2 = 0.0
        for b in self.choleskyBg:
            bg2 += b * b
        return LogHop2is(np.asfortranarray((-alpha * self.getKp()).astype(C128)), complex(alpha * 0.5 * bg2))

    def returnExpMinusAlphaV(self, alpha):
        return CholeskyReal(alpha, self.choleskyVecs, self)


def synthetic_molecule(L=200, Nup=40, Ndn=40, G=1600, seed=12345):
    """Synthetic real-materials Hamiltonian used by the phaseless workload.

    h = (A+A^T)/2 /sqrt(L) + diag(linspace(-2, 2, L)), A ~ N(0,1);
    L_g = S_g/||S_g||_F * exp(-g/L), S_g = (B_g+B_g^T)/2, B ~ N(0,1);
    t = h, K = h - 1/2 sum_g L_g L_g, choleskyBg = 0
    (the scripts/pyscf/molecule.py:122-143 convention)."""
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((L, L))
    h = (A + A.T) * 0.5 / np.sqrt(L) + np.diag(np.linspace(-2.0, 2.0, L))
    vecs = np.empty((L, L, G), dtype=np.float64, order="F")
    for g in range(G):
        B = rng.standard_normal((L, L))
        S = (B + B.T) * 0.5
        vecs[:, :, g] = S / np.linalg.norm(S) * np.exp(-g / L)
    K = h.copy()
    for g in range(G):
        K -= 0.5 * vecs[:, :, g] @ vecs[:, :, g]
    return RealMaterialMolecule(L, Nup, Ndn, h, K, vecs, np.zeros(G))


def fillSD2sFromModel(model: RealMaterialMolecule) -> SD2s:
    """Lowest Nup / Ndn eigenvectors of Kp (realMaterialMoleculeSD2sOperation.cpp)."""
    w, v = model._eigen()
    return SD2s(v[:, :model.Nup].astype(C128), v[:, :model.Ndn].astype(C128), 0j)


def sd2is_from_phiT(phiT: SD2s) -> SD2is:
    """setWalkerFromPhiT for SD2is (SD2sSD2isOperation.cpp:217-230)."""
    wf = phiT.wfUp if phiT.getNup() >= phiT.getNdn() else phiT.wfDn
    return SD2is(wf.copy(), phiT.getNup(), phiT.getNdn(), phiT.logw)


class RealMaterialMoleculeMeasureFixedSD2sSD2is:
    """Mixed estimator / force bias (realMaterialMoleculeMeasureFixedSD2sSD2is.cpp)."""

    def __init__(self, model: RealMaterialMolecule, walkerLeft: SD2s):
        self.model = model
        self.walkerLeft = walkerLeft
        L = model.L
        G = model.choleskyNumber
        self.wfUpDaggerT = walkerLeft.wfUp.conj().T @ model.t
        self.wfDnDaggerT = walkerLeft.wfDn.conj().T @ model.t
        # half-rotated vectors A_s[g] = phiT_s^H L_g  (Ns x L per g), :165-186
        vecs = model.choleskyVecs
        self.upDaggerVecs = np.einsum("ia,abg->ibg", walkerLeft.wfUp.conj().T, vecs)
        self.dnDaggerVecs = np.einsum("ia,abg->ibg", walkerLeft.wfDn.conj().T, vecs)
        nu = walkerLeft.getNup()
        nd = walkerLeft.getNdn()
        # (Ns*L x G) column-major flattenings used by the zgemv 'T' of calculateCholeskyBg
        self.leftUp = self.upDaggerVecs.reshape(nu * L, G, order="F")
        self.leftDn = self.dnDaggerVecs.reshape(nd * L, G, order="F")
        self.reSet()

    def reSet(self):
        G = self.model.choleskyNumber
        self.den = 0j
        self.TNum = 0j
        self.HNum = 0j
        self.choleskyBgNum = np.zeros(G, dtype=C128)
        self.choleskyExNum = np.zeros(G, dtype=C128)

    def returnEnergy(self):
        return self.HNum / self.den

    def returnCholeskyBg(self):
        return (self.choleskyBgNum / self.den).real

    def calculateCholeskyBg(self, op):
        tu, td = op.returnTheta_T()
        return self.leftUp.T @ tu.reshape(-1, order="F") + self.leftDn.T @ td.reshape(-1, order="F")

    def calculateTenergy(self, op):
        tu, td = op.returnTheta_T()
        return np.sum(self.wfUpDaggerT * tu) + np.sum(self.wfDnDaggerT * td)

    def calculateCholeskyEx(self, op):
        tu, td = op.returnTheta_T()
        thu = tu.T
        thd = td.T
        G = self.model.choleskyNumber
        ex = np.empty(G, dtype=C128)
        for k in range(G):
            du = self.upDaggerVecs[:, :, k] @ thu
            dd = self.dnDaggerVecs[:, :, k] @ thd
            ex[k] = np.sum(du.T * du) + np.sum(dd.T * dd)
        return ex

    def local_energy(self, op):
        T = self.calculateTenergy(op)
        bg = self.calculateCholeskyBg(op)
        ex = self.calculateCholeskyEx(op)
        H = 0.5 * np.sum(bg * bg - ex) + T
        return H, T, bg, ex

    def addMeasurement(self, op, denIncrement):
        self.den += denIncrement
        H, T, bg, ex = self.local_energy(op)
        self.TNum += T * denIncrement
        self.choleskyBgNum += bg * denIncrement
        self.choleskyExNum += ex * denIncrement
        self.HNum += H * denIncrement

    def getForce(self, cholesky: CholeskyReal, op, cap):
        """f_g = sqrt(-dt)(bg_g - currentBg_g), rescaled to |f_g| <= cap (:93-116)."""
        bg = self.calculateCholeskyBg(op)
        f = (bg - self.model.choleskyBg) * cholesky.sqrtMinusDt
        a = np.abs(f)
        return np.where(a > cap, f * cap / np.where(a > 0, a, 1.0), f)
