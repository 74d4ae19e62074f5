"""Phaseless AFQMC on the GPU: the real-materials model set-up, device handles and the driver loop.

``AfqmcPhaseless`` mirrors tutorials/afqmcPhaseless (afqmcPhaseless.cpp:14-205,
afqmcPhaselessProjection.cpp:11-214, afqmcPhaselessMeasure.cpp:9-45): the same
step order, cadences, background-adjust schedule and random-number consumption
order, with the per-walker loop replaced by batched device calls.
"""
import ctypes as C
import sys
from dataclasses import dataclass

import numpy as np

from ._lib import PhDesc, check, dptr, iptr, lib
from .cp import FORCE, Context  # noqa: F401


class MoleculeModel:
    """RealMaterialMolecule (realMaterialMolecule.h:16-40): L orbitals, Nup/Ndn electrons, the one-body
    matrices t (energy) and K (propagation), G real Cholesky vectors L_g (L x L each) and the
    background choleskyBg.  Kp = K + sum_g bg_g L_g (setKp, realMaterialMolecule.cpp:152-164)."""

    def __init__(self, L, Nup, Ndn, t, K, chol, bg=None):
        self.L, self.Nup, self.Ndn = int(L), int(Nup), int(Ndn)
        self.t = np.asfortranarray(np.asarray(t, dtype=np.float64))
        self.K = np.asfortranarray(np.asarray(K, dtype=np.float64))
        self.chol = np.asfortranarray(np.asarray(chol, dtype=np.float64))
        if self.chol.shape[:2] != (self.L, self.L):
            raise ValueError("Cholesky vectors must be L x L x G")
        self.G = self.chol.shape[2]
        self.bg = np.zeros(self.G) if bg is None else np.array(bg, dtype=np.float64)
        self._eig = None

    def updateBackGround(self, bg):
        bg = np.asarray(bg, dtype=np.float64)
        if bg.size != self.G:
            raise ValueError("Background size is not choleskyNumber!")
        self.bg = bg.copy()
        self._eig = None

    def Kp(self):
        v = self.chol.reshape(self.L * self.L, self.G, order="F")
        return self.K + (v @ self.bg).reshape(self.L, self.L, order="F")

    def eig(self):
        if self._eig is None:
            self._eig = np.linalg.eigh(self.Kp())
        return self._eig

    def bg_logw(self, alpha):
        return complex(alpha * 0.5 * float(np.dot(self.bg, self.bg)))

    def exp_minus_alpha_k(self, alpha):
        """Hop2is exp(-alpha Kp) with logw alpha/2 sum bg^2 (returnExpMinusAlphaK, realMaterialMolecule.cpp:117-132)."""
        w, v = self.eig()
        return (v * np.exp(-alpha * w)) @ v.T, self.bg_logw(alpha)

    def log_exp_minus_alpha_k(self, alpha):
        """LogHop2is -alpha Kp (returnLogExpMinusAlphaK, realMaterialMolecule.cpp:134-146)."""
        return -alpha * self.Kp(), self.bg_logw(alpha)

    def trial(self):
        """SD2s of the lowest Nup / Ndn orbitals of Kp (fillWalkerFromModel, realMaterialMoleculeSD2sOperation.cpp)."""
        w, v = self.eig()
        return v[:, :self.Nup].astype(np.complex128), v[:, :self.Ndn].astype(np.complex128)

    def half_rotated(self, phi):
        """phi^H L_g for every g as an (N, L, G) column-major array (initWfDaggerCholeskyVecs,
        realMaterialMoleculeMeasureFixedSD2sSD2is.cpp:165-186)."""
        phi = np.asarray(phi, dtype=np.complex128)
        a = phi.conj().T @ self.chol.reshape(self.L, self.L * self.G, order="F")
        return np.asfortranarray(a.reshape(phi.shape[1], self.L, self.G, order="F"))


def synthetic_molecule(L=200, Nup=40, Ndn=40, G=1600, seed=12345):
    """Synthetic real-materials Hamiltonian: h = (A+A^T)/2/sqrt(L) + diag(linspace(-2, 2, L)), A ~ N(0,1);
    L_g = S_g/||S_g||_F exp(-g/L), S_g = (B_g+B_g^T)/2, B ~ N(0,1); t = h, K = h - 1/2 sum_g L_g L_g, bg = 0
    (the one-body shift of the reference's molecule set-up, scripts/pyscf/molecule.py:122-143)."""
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((L, L))
    h = (A + A.T) * 0.5 / np.sqrt(L) + np.diag(np.linspace(-2.0, 2.0, L))
    vecs = np.empty((L, L, G), dtype=np.float64, order="F")
    K = h.copy()
    for g in range(G):
        B = rng.standard_normal((L, L))
        S = (B + B.T) * 0.5
        vecs[:, :, g] = S / np.linalg.norm(S) * np.exp(-g / L)
    for g in range(G):
        K -= 0.5 * vecs[:, :, g] @ vecs[:, :, g]
    return MoleculeModel(L, Nup, Ndn, h, K, vecs)


@dataclass
class PhaselessMethod:
    """afqmc_param keys of the phaseless driver (afqmcPhaselessMethod.cpp:158-185)."""
    dt: float = 0.01
    thermalSize: int = 200
    writeNumber: int = 60
    measureNumberPerWrite: int = 2
    measureSkipStep: int = 5
    walkerSizePerThread: int = 300
    forceType: str = "dynamicForce"
    forceCap: float = 1.5
    mgsStep: int = 10
    popControlStep: int = 10
    ET: float = -10.0
    ETAndBackGroundAdjustStep: int = 10
    ETAndBackGroundAdjustMaxSize: int = 0
    ETAndBackGroundGrowthEstimateStep: int = 5
    ETAndBackGroundGrowthEstimateMaxSize: int = 0
    seed: int = 985456376

    def analysis(self):
        """afqmcPhaselessMethod.cpp:187-239."""
        if self.dt < 0.0:
            raise ValueError("dt must be postive!")
        if self.measureNumberPerWrite == 0 or self.measureSkipStep == 0:
            raise ValueError("code will not measure anything!")
        if self.walkerSizePerThread <= 0:
            raise ValueError("walkerSizePerThread is not a positive number!")
        if self.forceCap < 0.0:
            raise ValueError("forceCap must be positive or zero!")
        if self.forceType not in FORCE:
            raise ValueError("Do not know method.forceType " + self.forceType)


class PhaselessPropagator:
    """Device-side model, trial, one-body operators, Cholesky vectors and workspace (afqmc_ph)."""

    def __init__(self, ctx: Context, model: MoleculeModel, method: PhaselessMethod, phiT=None, precision=0,
                 max_walkers=None, const_force=None, phiT_logw=0j, taylor="dynamicOrder", taylor_order=0,
                 taylor_base=3, taylor_accuracy=1e-10):
        if taylor not in ("dynamicOrder", "fixedOrder"):
            raise ValueError("Do not know the input flag for LogHop2isSD2isOperation: " + taylor)
        self.ctx = ctx
        self.model = model
        self.method = method
        self.precision = int(precision)
        up, dn = model.trial() if phiT is None else phiT
        self.phiT = (np.asfortranarray(np.asarray(up, dtype=np.complex128)),
                     np.asfortranarray(np.asarray(dn, dtype=np.complex128)))
        if self.phiT[0].shape != (model.L, model.Nup) or self.phiT[1].shape != (model.L, model.Ndn):
            raise ValueError("trial must be (L x Nup, L x Ndn)")
        self.phiT_logw = complex(phiT_logw)
        keep = []

        def ptr(a):
            a = np.asfortranarray(a)
            keep.append(a)
            return dptr(a)

        d = PhDesc()
        d.L, d.Nup, d.Ndn, d.G = model.L, model.Nup, model.Ndn, model.G
        d.precision = self.precision
        d.force_type = FORCE[method.forceType]
        d.dt, d.force_cap = method.dt, method.forceCap
        d.t = ptr(model.t)
        d.chol = ptr(model.chol)
        d.bg = ptr(model.bg)
        d.phiT_up, d.phiT_dn = ptr(self.phiT[0]), ptr(self.phiT[1])
        d.phiT_logw[0], d.phiT_logw[1] = self.phiT_logw.real, self.phiT_logw.imag
        d.half_rot_up = ptr(model.half_rotated(self.phiT[0]))
        d.half_rot_dn = ptr(model.half_rotated(self.phiT[1]))
        cf = np.zeros(model.G, dtype=np.complex128) if const_force is None else np.asarray(const_force, np.complex128)
        d.const_force = ptr(cf)
        d.max_walkers = int(max_walkers or method.walkerSizePerThread)
        d.taylor_fixed = int(taylor == "fixedOrder")
        d.taylor_order = int(taylor_order)
        d.taylor_base = int(taylor_base)
        d.taylor_accuracy = float(taylor_accuracy)
        h = C.c_void_p()
        check(lib().afqmc_ph_create(ctx.h, C.byref(d), C.byref(h)))
        del keep
        self.h = h
        self.max_walkers = d.max_walkers

    def close(self):
        if getattr(self, "h", None):
            lib().afqmc_ph_destroy(self.h)
            self.h = None

    def __del__(self):
        # no native calls while the interpreter shuts down (the runtime may already be tearing
        # down; the process exit reclaims the memory)
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    # ---- operators, background, random numbers, ET ------------------------
    def set_onebody(self, which, matrix, logw, is_log):
        m = np.asfortranarray(np.asarray(matrix, dtype=np.complex128))
        lw = np.array([complex(logw).real, complex(logw).imag])
        check(lib().afqmc_ph_set_onebody(self.h, int(which), dptr(m), dptr(lw), int(bool(is_log))))

    def set_background(self, bg):
        b = np.ascontiguousarray(bg, dtype=np.float64)
        check(lib().afqmc_ph_set_background(self.h, dptr(b)))

    def set_seed(self, seed, counter=0):
        check(lib().afqmc_ph_set_seed(self.h, C.c_uint64(seed & 0xFFFFFFFFFFFFFFFF), C.c_uint64(counter)))

    def counter(self):
        c, s = C.c_uint64(), C.c_uint64()
        check(lib().afqmc_ph_get_counter(self.h, C.byref(c), C.byref(s)))
        return c.value

    def walker_steps(self):
        c, s = C.c_uint64(), C.c_uint64()
        check(lib().afqmc_ph_get_counter(self.h, C.byref(c), C.byref(s)))
        return s.value

    def set_uniforms(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        check(lib().afqmc_ph_set_uniforms(self.h, dptr(u) if u.size else None, u.size))

    def skip_uniforms(self, n):
        check(lib().afqmc_ph_skip_uniforms(self.h, C.c_uint64(int(n))))

    def set_ET(self, et):
        check(lib().afqmc_ph_set_ET(self.h, float(et)))

    def get_ET(self):
        v = np.zeros(1)
        check(lib().afqmc_ph_get_ET(self.h, dptr(v)))
        return float(v[0])

    def walkers(self, nw):
        return WalkerSet2is(self, nw)

    # ---- batched drop-ins -------------------------------------------------
    def log_overlap(self, w):
        out = np.zeros(w.nw, dtype=np.complex128)
        check(lib().afqmc_ph_log_overlap(self.h, w.h, dptr(out)))
        return out

    def background(self, w):
        out = np.zeros((w.nw, self.model.G), dtype=np.complex128)
        check(lib().afqmc_ph_background(self.h, w.h, dptr(out)))
        return out

    def sample_apply(self, w, aux, force):
        a = np.ascontiguousarray(aux, dtype=np.complex128)
        f = np.ascontiguousarray(force, dtype=np.complex128)
        slw = np.zeros(w.nw, dtype=np.complex128)
        mo = C.c_int()
        check(lib().afqmc_ph_sample_apply(self.h, w.h, dptr(a), dptr(f), dptr(slw), C.byref(mo)))
        return slw, mo.value

    def sample_matrix(self, w):
        L = self.model.L
        out = np.zeros((w.nw, L, L), dtype=np.complex128)
        check(lib().afqmc_ph_sample_matrix(self.h, w.h, dptr(out)))
        return np.transpose(out, (0, 2, 1))

    def taylor_orders(self, w):
        o = np.zeros(w.nw, dtype=np.int32)
        check(lib().afqmc_ph_taylor_orders(self.h, w.h, iptr(o)))
        return o

    def onebody(self, w, which=0):
        check(lib().afqmc_ph_onebody(self.h, w.h, int(which)))

    def qr_stats(self):
        """(walkers redone by shifted CholeskyQR, walkers marked dead as rank deficient)."""
        a, b = C.c_uint64(), C.c_uint64()
        check(lib().afqmc_ph_get_qr_stats(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def stabilize(self, w):
        check(lib().afqmc_ph_stabilize(self.h, w.h))

    def step(self, w, record=False):
        if record:
            f = np.zeros((w.nw, self.model.G), dtype=np.complex128)
            a = np.zeros((w.nw, self.model.G), dtype=np.complex128)
            check(lib().afqmc_ph_step(self.h, w.h, dptr(f), dptr(a)))
            return f, a
        check(lib().afqmc_ph_step(self.h, w.h, None, None))
        return None

    def pop_control(self, w, record=False):
        if record:
            wt = np.zeros(w.nw)
            tb = np.zeros(w.nw, dtype=np.int32)
            st = np.zeros(3)
            check(lib().afqmc_ph_pop_control(self.h, w.h, dptr(wt), iptr(tb), dptr(st)))
            return wt, tb, st
        check(lib().afqmc_ph_pop_control(self.h, w.h, None, None, None))
        return None

    def measure(self, w):
        check(lib().afqmc_ph_measure(self.h, w.h))

    def energy(self):
        v = np.zeros(3, dtype=np.complex128)
        bg = np.zeros(self.model.G, dtype=np.complex128)
        ex = np.zeros(self.model.G, dtype=np.complex128)
        check(lib().afqmc_ph_energy(self.h, dptr(v), dptr(bg), dptr(ex)))
        return {"den": v[0], "HNum": v[1], "TNum": v[2], "choleskyBgNum": bg, "choleskyExNum": ex}

    def reset_measure(self):
        check(lib().afqmc_ph_reset_measure(self.h))

    def adjust_ET(self):
        check(lib().afqmc_ph_adjust_ET(self.h))


class WalkerSet2is:
    """Device-resident std::vector<SD2is>: nw walkers of L x max(Nup, Ndn)."""

    def __init__(self, prop: PhaselessPropagator, nw):
        h = C.c_void_p()
        check(lib().afqmc_ph_walkers_create(prop.h, int(nw), C.byref(h)))
        self.h = h
        self.nw = int(nw)
        self.prop = prop
        self.ncol = max(prop.model.Nup, prop.model.Ndn)

    def close(self):
        if getattr(self, "h", None):
            lib().afqmc_ph_walkers_destroy(self.h)
            self.h = None

    def __del__(self):
        # no native calls while the interpreter shuts down (the runtime may already be tearing
        # down; the process exit reclaims the memory)
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    def upload(self, wfs, logw=None, alive=None):
        buf = np.ascontiguousarray(np.transpose(np.asarray(wfs, dtype=np.complex128), (0, 2, 1)))
        lw = np.zeros(self.nw, dtype=np.complex128) if logw is None else np.ascontiguousarray(logw, dtype=np.complex128)
        al = None if alive is None else np.ascontiguousarray(alive, dtype=np.int32)
        check(lib().afqmc_ph_walkers_upload(self.h, dptr(buf), dptr(lw), iptr(al)))

    def download(self):
        L = self.prop.model.L
        buf = np.zeros((self.nw, self.ncol, L), dtype=np.complex128)
        lw = np.zeros(self.nw, dtype=np.complex128)
        al = np.zeros(self.nw, dtype=np.int32)
        check(lib().afqmc_ph_walkers_download(self.h, dptr(buf), dptr(lw), iptr(al)))
        return np.transpose(buf, (0, 2, 1)), lw, al

    def set_from_trial(self, phiT, logw=0j):
        """setWalkerFromPhiT (SD2sSD2isOperation.cpp:217-230): the larger spin block of the trial."""
        up, dn = phiT
        wf = up if up.shape[1] >= dn.shape[1] else dn
        self.upload([wf] * self.nw, np.full(self.nw, complex(logw)))


class AfqmcPhaseless:
    """Device-resident phaseless driver with the reference's step loop.

    run() follows afqmcPhaseless.cpp:14-30: initial walkers from the trial ->
    QR + comb -> one aux sample skipped (estimateMemory, 2G uniforms) ->
    measureWithProjection with the ET / background adjust schedule of
    afqmcPhaseless.cpp:135-205.  While the background is adjustable,
    exp(-dt Kp) is applied as the Taylor series of the LogHop2is built from the
    initial background (afqmcPhaseless.cpp:34-63); once adjusting ends it is
    rebuilt as a Hop2is of the current background."""

    def __init__(self, ctx, model: MoleculeModel, method: PhaselessMethod, phiT=None, precision=0, uniforms=None,
                 seed=None, **taylor):
        method.analysis()
        self.ctx = ctx
        self.model = model
        self.method = method
        m = method
        self.adjustable = False
        self.nGrowth = 0
        if m.ETAndBackGroundAdjustMaxSize > 0:
            self.adjustable = True
        if m.ETAndBackGroundAdjustMaxSize < m.ETAndBackGroundGrowthEstimateMaxSize:
            self.nGrowth = (m.ETAndBackGroundGrowthEstimateMaxSize - m.ETAndBackGroundAdjustMaxSize) \
                // m.ETAndBackGroundGrowthEstimateStep
            if self.nGrowth > 0:
                self.adjustable = True
        self.prop = PhaselessPropagator(ctx, model, method, phiT=phiT, precision=precision,
                                        max_walkers=m.walkerSizePerThread, **taylor)
        if self.adjustable:
            for which, alpha in ((0, m.dt), (1, 0.5 * m.dt), (2, -0.5 * m.dt)):
                A, lw = model.log_exp_minus_alpha_k(alpha)
                self.prop.set_onebody(which, A, lw, True)
        else:
            self._set_exp_k()
        if uniforms is not None:
            self.prop.set_uniforms(uniforms)
        else:
            self.prop.set_seed(m.seed if seed is None else seed, 0)
        self.prop.set_ET(m.ET)
        self.walker = self.prop.walkers(m.walkerSizePerThread)
        self.energies = []

    def _set_exp_k(self):
        m = self.method
        for which, alpha in ((0, m.dt), (1, 0.5 * m.dt), (2, -0.5 * m.dt)):
            B, lw = self.model.exp_minus_alpha_k(alpha)
            self.prop.set_onebody(which, B, lw, False)

    def initialWalker(self):
        self.walker.set_from_trial(self.prop.phiT, self.prop.phiT_logw)
        self.prop.stabilize(self.walker)
        self.prop.pop_control(self.walker)

    def estimateMemory(self):
        self.prop.skip_uniforms(2 * self.model.G)

    def projectOneStep(self, idx):
        self.prop.step(self.walker)
        idx[0] += 1
        if idx[0] == self.method.mgsStep:
            self.prop.stabilize(self.walker)
            idx[0] = 0
        idx[1] += 1
        if idx[1] == self.method.popControlStep:
            self.prop.pop_control(self.walker)
            idx[1] = 0

    def adjustETAndBackGround(self):
        """afqmcPhaselessMeasure.cpp:36-45: ET = Re(E), background = Re(bgNum / den), reset."""
        e = self.prop.energy()
        bg = (e["choleskyBgNum"] / e["den"]).real
        self.prop.adjust_ET()
        self.model.updateBackGround(bg)
        self.prop.set_background(bg)

    def measureWithProjection(self):
        m = self.method
        self.prop.onebody(self.walker, 1)
        idx = [0, 0]
        for i in range(m.thermalSize):
            if i < m.ETAndBackGroundAdjustMaxSize and i % m.ETAndBackGroundAdjustStep == 0:
                self.prop.measure(self.walker)
                self.adjustETAndBackGround()
            if i == m.ETAndBackGroundAdjustMaxSize - 1 and self.nGrowth == 0:
                self._set_exp_k()
            if self.nGrowth > 0 and m.ETAndBackGroundAdjustMaxSize <= i < m.ETAndBackGroundGrowthEstimateMaxSize:
                if (i + 1 - m.ETAndBackGroundAdjustMaxSize) % m.ETAndBackGroundGrowthEstimateStep == 0:
                    self.prop.measure(self.walker)
                if i == m.ETAndBackGroundGrowthEstimateMaxSize - 1:
                    self.adjustETAndBackGround()
                    self._set_exp_k()
            self.projectOneStep(idx)
        for i in range(m.writeNumber):
            for j in range(m.measureNumberPerWrite):
                self.prop.measure(self.walker)
                for k in range(m.measureSkipStep):
                    self.projectOneStep(idx)
            e = self.prop.energy()
            self.energies.append((e["HNum"], e["den"], e["TNum"]))
            self.prop.reset_measure()

    def run(self):
        self.initialWalker()
        self.estimateMemory()
        self.measureWithProjection()
        return self.energies

    def energy(self):
        return sum(e[0] for e in self.energies) / sum(e[1] for e in self.energies)
