"""Pin the oracle's back-propagation driver and its observables.  CPU only.

* every observable of HubbardSOCMeasureObserveSDSD against the same operator built in Fock space
  (Jordan-Wigner matrices on 2L modes) between two Slater determinants;
* backPropagationStep = 0 reduces the estimator to the mixed one: the write-out equals the
  constraint-path driver's mixed energy on the same random stream;
* U = 0 with a trial that is not the ground state: the back-propagated Green matrix converges to the
  ground-state projector of exp(-dt K), the mixed one does not.
"""
import itertools

import numpy as np
import pytest

from oracle.blocks import SD, SDSDOperation
from oracle.drivers import AfqmcConstraintPath, AfqmcConstraintPathBP, BPMethod, CPMethod
from oracle.models import HubbardSOC, HubbardSOCMeasureObserveSDSD, square_lattice_K
from oracle.rng import UniformStream


# ---------------------------------------------------------------- Fock space
def annihilators(M):
    """c_a on the 2^M occupation basis (bit a = mode a), Jordan-Wigner sign over modes below a."""
    dim = 1 << M
    ops = []
    for a in range(M):
        c = np.zeros((dim, dim))
        for s in range(dim):
            if (s >> a) & 1:
                sign = (-1) ** bin(s & ((1 << a) - 1)).count("1")
                c[s ^ (1 << a), s] = sign
        ops.append(c)
    return ops


def slater_state(wf, cdag):
    """prod_k (sum_a wf[a, k] c+_a) |0>."""
    v = np.zeros(cdag[0].shape[0], dtype=complex)
    v[0] = 1.0
    for k in range(wf.shape[1]):
        v = sum(wf[a, k] * (cdag[a] @ v) for a in range(wf.shape[0]))
    return v


def test_observables_match_fock_space():
    rng = np.random.default_rng(21)
    L, N = 3, 3
    M = 2 * L
    K = rng.standard_normal((M, M)) + 1j * rng.standard_normal((M, M))
    K = K + K.conj().T
    m = HubbardSOC(L, N, K, mu=rng.uniform(-1, 1, L), hx=rng.uniform(-1, 1, L), hy=rng.uniform(-1, 1, L),
                   hz=rng.uniform(-1, 1, L), U=rng.uniform(1, 4, L))
    wl = rng.standard_normal((M, N)) + 1j * rng.standard_normal((M, N))
    wr = rng.standard_normal((M, N)) + 1j * rng.standard_normal((M, N))
    meas = HubbardSOCMeasureObserveSDSD(m)
    w = complex(0.7, -0.2)
    meas.addMeasurement(SDSDOperation(SD(wf=wl, logw=0.3), SD(wf=wr, logw=1.1j)), w)

    c = annihilators(M)
    cd = [x.T for x in c]
    bra, ket = slater_state(wl, cd), slater_state(wr, cd)
    ovl = bra.conj() @ ket

    def ev(op):
        return (bra.conj() @ (op @ ket)) / ovl

    n = [cd[a] @ c[a] for a in range(M)]
    G = np.array([[ev(cd[i] @ c[j]) for j in range(M)] for i in range(M)])
    assert np.abs(meas.greenMatrixNum - w * G).max() < 1e-12
    dd = np.array([[ev(n[a] @ n[b]) for b in range(M)] for a in range(M)])
    assert np.abs(meas.densityDensityNum - w * dd).max() < 1e-12
    sp = [cd[i] @ c[i + L] for i in range(L)]   # S+_i = c+_{i up} c_{i dn}
    sm = [cd[i + L] @ c[i] for i in range(L)]
    pair = [c[i + L] @ c[i] for i in range(L)]  # c_{i dn} c_{i up}
    pm = np.array([[ev(sp[i] @ sm[j]) for j in range(L)] for i in range(L)])
    mp = np.array([[ev(sm[i] @ sp[j]) for j in range(L)] for i in range(L)])
    pp = np.array([[ev(pair[i].conj().T @ pair[j]) for j in range(L)] for i in range(L)])
    assert np.abs(meas.splusSminusNum - w * pm).max() < 1e-12
    assert np.abs(meas.sminusSplusNum - w * mp).max() < 1e-12
    assert np.abs(meas.spairSpairNum - w * pp).max() < 1e-12

    H = sum(K[a, b] * cd[a] @ c[b] for a in range(M) for b in range(M))
    H = H + sum(m.U[i] * n[i] @ n[i + L] for i in range(L))
    for i in range(L):
        H = H + (-m.mu[i] + 0.5 * m.hz[i]) * n[i] + (-m.mu[i] - 0.5 * m.hz[i]) * n[i + L]
        H = H + complex(0.5 * m.hx[i], -0.5 * m.hy[i]) * sp[i] + complex(0.5 * m.hx[i], 0.5 * m.hy[i]) * sm[i]
    assert abs(meas.HNum - w * ev(H)) < 1e-11
    assert abs(meas.den - w) == 0
    assert abs(meas.returnEnergy() - ev(H)) < 1e-11


# ---------------------------------------------------------------- driver
def hubbard(nx=2, ny=2, N=4, U=4.0, seed=5):
    L = nx * ny
    rng = np.random.default_rng(seed)
    K = square_lattice_K(nx, ny, 1.0, 0.0)
    return HubbardSOC(L, N, K, mu=np.zeros(L), hx=np.zeros(L), hy=np.zeros(L),
                      hz=rng.uniform(-0.1, 0.1, L), U=np.full(L, U))


def test_zero_steps_is_the_mixed_estimator():
    model = hubbard()
    kw = dict(dt=0.02, thermalSize=10, writeNumber=2, measureNumberPerWrite=2, measureSkipStep=3,
              walkerSizePerThread=6, decompType="densitySpin", forceType="dynamicForce", mgsStep=2,
              popControlStep=3, ET=-3.0, ETAdjustStep=5, ETAdjustMaxSize=10)
    cp = AfqmcConstraintPath(CPMethod(**kw), model, UniformStream(77))
    mixed = cp.run()
    bp = AfqmcConstraintPathBP(BPMethod(backPropagationStep=0, **kw), model, UniformStream(77))
    writes = bp.run()
    assert len(writes) == len(mixed) == 2
    for w, (HNum, den, KNum, VNum, RNum) in zip(writes, mixed):
        assert abs(w["den"] - den) <= 1e-12 * abs(den)
        for a, b in ((w["HNum"], HNum), (w["KNum"], KNum), (w["VNum"], VNum), (w["RNum"], RNum)):
            assert abs(a - b) <= 1e-10 * abs(den)
        assert abs(np.trace(w["greenMatrixNum"]) - model.N * w["den"]) <= 1e-10 * abs(den)


def green(left, right):
    """<c+_i c_j> = [R (L^H R)^-1 L^H]_{ji}."""
    return (right @ np.linalg.solve(left.conj().T @ right, left.conj().T)).T


@pytest.mark.parametrize("nbp,skip", [(80, 10), (12, 20), (7, 7)])
def test_u0_back_propagation_closed_form(nbp, skip):
    """U = 0, mu = h = 0: a step is exp(-dt K) exactly, the walkers converge to the lowest N orbitals
    of K, and the back-propagated bra is exp(-nbp dt K) phiT, so the estimate is the Green matrix of
    that pair in closed form; a long window reaches the ground-state projector, which the mixed
    estimate with a trial far from the ground state misses."""
    rng = np.random.default_rng(9)
    L, N = 3, 2
    M = 2 * L
    q, _ = np.linalg.qr(rng.standard_normal((M, M)) + 1j * rng.standard_normal((M, M)))
    lam = np.array([-2.0, -1.8, 1.0, 1.3, 1.6, 2.0])
    K = (q * lam) @ q.conj().T
    model = HubbardSOC(L, N, K, mu=np.zeros(L), hx=np.zeros(L), hy=np.zeros(L), hz=np.zeros(L), U=np.zeros(L))
    trial, _ = np.linalg.qr(q[:, :N] + 0.6 * q[:, N:N + N])
    dt = 0.1
    meth = BPMethod(dt=dt, thermalSize=150, writeNumber=1, measureNumberPerWrite=2, measureSkipStep=skip,
                    walkerSizePerThread=3, decompType="densitySpin", forceType="dynamicForce", mgsStep=5,
                    popControlStep=7, ET=-3.0, ETAdjustStep=10, ETAdjustMaxSize=0, backPropagationStep=nbp)
    bp = AfqmcConstraintPathBP(meth, model, UniformStream(3), phiT=SD(wf=trial))
    w = bp.run()[0]
    Gbp = w["greenMatrixNum"] / w["den"]
    bra = (q * np.exp(-nbp * dt * lam)) @ q.conj().T @ trial
    assert np.abs(Gbp - green(bra, q[:, :N])).max() < 1e-9
    G0 = green(q[:, :N], q[:, :N])
    assert np.abs(green(trial, q[:, :N]) - G0).max() > 0.05
    if nbp >= 80:
        assert np.abs(Gbp - G0).max() < 1e-8
        assert abs(w["HNum"] / w["den"] - lam[:N].sum()) < 1e-8


class AncestryCheck(AfqmcConstraintPathBP):
    """The same run, with a second estimate built another way: every step's aux fields and comb
    table are logged per slot, the walkers are copied at each window start, and a measurement
    follows each walker's ancestry back through the tables (a slot's parent is table[slot]) instead
    of moving records and re-propagating backups."""

    def __init__(self, *a, **k):
        super().__init__(*a, **k)
        self.t = None
        self.aux_log, self.table_log, self.snaps = {}, {}, {}
        self.alt = HubbardSOCMeasureObserveSDSD(self.model)
        self.deaths = 0

    def prepareBackPropagation(self):
        super().prepareBackPropagation()
        if self.t is None:
            self.t = 0

    def step_walker(self, i):
        before = len(self.records)
        super().step_walker(i)
        if len(self.records) > before:
            self.cur_aux[i] = self.records[-1][2].copy()
        else:
            self.deaths += 1

    def popControl(self):
        weight, table = super().popControl()
        self.cur_table = table.copy()
        return weight, table

    def projectOneStep(self, idx):
        if self.t is not None and self.t % self.method.measureSkipStep == 0:
            self.snaps[self.t] = [w.copy() for w in self.walker]
        self.cur_aux, self.cur_table = [None] * len(self.walker), None
        super().projectOneStep(idx)
        if self.t is not None:
            self.aux_log[self.t], self.table_log[self.t] = self.cur_aux, self.cur_table
            self.t += 1

    def addMeasurement(self):
        m = self.method
        T, nbp = self.t, m.backPropagationStep
        if self.isBP:
            if T not in self.snaps:
                self.snaps[T] = [w.copy() for w in self.walker]
            s = T - nbp
            for i, w in enumerate(self.walker):
                if not self.walkerIsAlive[i]:
                    continue
                anc = i
                hist = [None] * nbp
                for j in range(nbp - 1, -1, -1):
                    if self.table_log[s + j] is not None:
                        anc = int(self.table_log[s + j][anc])
                    hist[j] = self.aux_log[s + j][anc]
                    assert hist[j] is not None, "an alive walker's ancestry reached an unrecorded step"
                t = self.op1.applyToRight(self.expHalfDtK, w)
                overlap = np.exp(SDSDOperation(self.phiT, t).returnLogOverlap())
                bra = self.back_propagated_bra(hist)
                ket = self.op1.applyToRight(self.expHalfDtK, self.snaps[s][anc])
                self.alt.addMeasurement(SDSDOperation(bra, ket), overlap)
        super().addMeasurement()


@pytest.mark.parametrize("nbp,skip", [(6, 4), (4, 6), (5, 5), (0, 3)])
def test_table_replay_equals_ancestry(nbp, skip):
    rng = np.random.default_rng(31)
    L, N = 4, 3
    K = square_lattice_K(2, 2, 1.0, 0.0)
    model = HubbardSOC(L, N, K, mu=rng.uniform(-0.3, 0.3, L), hx=rng.uniform(-0.5, 0.5, L),
                       hy=rng.uniform(-0.5, 0.5, L), hz=rng.uniform(-0.5, 0.5, L), U=np.full(L, 6.0))
    meth = BPMethod(dt=0.08, thermalSize=6, writeNumber=2, measureNumberPerWrite=2, measureSkipStep=skip,
                    walkerSizePerThread=10, decompType="densityCharge", forceType="dynamicForce", mgsStep=3,
                    popControlStep=2, ET=-3.0, ETAdjustStep=3, ETAdjustMaxSize=6, backPropagationStep=nbp)
    run = AncestryCheck(meth, model, UniformStream(123))
    writes = run.run()
    assert run.deaths > 0
    for key in ("den", "HNum", "greenMatrixNum", "densityDensityNum", "splusSminusNum", "sminusSplusNum",
                "spairSpairNum"):
        total = sum(w[key] for w in writes)
        ref = getattr(run.alt, key)
        assert np.max(np.abs(total - ref)) <= 1e-10 * np.max(np.abs(ref)), key
