"""Constraint-path and phaseless step loops (test infrastructure only).

Restates tutorials/afqmcConstraintPath/source/afqmcConstraintPath*.cpp and
tutorials/afqmcPhaseless/source/afqmcPhaseless*.cpp with the random stream
injected (``rng`` has ``take(n)``).  The step order and the random-number
consumption order are the reference's:

* initial QR + comb: Nw comb uniforms;
* estimateMemory: one aux sample (Ls uniforms, or 2G for phaseless);
* each step: per alive walker in index order Ls uniforms (CP) or 2G
  (phaseless, Box-Muller pairs); killed walkers consume none;
* each comb: Nw uniforms.
"""
import cmath
import math
from dataclasses import dataclass

import numpy as np

from . import comb
from .blocks import (SD, SD2is, HopSDOperation, LogHop2isSD2isOperation, SDSDOperation,
                     SD2sSD2isOperation, niupnidn_apply_left, niupnidn_apply_right)
from .models import (HubbardSOCMeasureFixedSDSD, HubbardSOCMeasureObserveSDSD, RealMaterialMoleculeMeasureFixedSD2sSD2is,
                     fillWalkerFromModel, fillSD2sFromModel, sd2is_from_phiT)
from .rng import gaussians_from_uniforms


@dataclass
class CPMethod:
    """afqmc_param keys of the constraint-path driver (afqmcConstraintPathMethod.cpp:158-180)."""
    dt: float = 0.01
    thermalSize: int = 200
    writeNumber: int = 60
    measureNumberPerWrite: int = 2
    measureSkipStep: int = 5
    walkerSizePerThread: int = 300
    decompType: str = "densitySpin"
    forceType: str = "dynamicForce"
    forceCap: float = 1.5
    mgsStep: int = 10
    popControlStep: int = 10
    ET: float = -10.0
    ETAdjustStep: int = 10
    ETAdjustMaxSize: int = 100
    seed: int = 985456376

    def analysis(self):
        if self.dt < 0:
            raise ValueError("dt must be postive!")
        if self.measureNumberPerWrite == 0 or self.measureSkipStep == 0:
            raise ValueError("measureNumberPerWrite/measureSkipStep must be positive")
        if self.walkerSizePerThread <= 0:
            raise ValueError("walkerSizePerThread is not a positive number!")
        if self.forceCap < 0:
            raise ValueError("forceCap must be positive or zero!")
        if self.ETAdjustMaxSize > self.thermalSize:
            raise ValueError("We should not adjust ET after thermalizing!")


class AfqmcConstraintPath:
    """Serial CP-AFQMC driver (afqmcConstraintPath.cpp, ...Projection.cpp, ...Measure.cpp)."""

    def __init__(self, method: CPMethod, model, rng, phiT: SD = None, constForce=None):
        method.analysis()
        self.method = method
        self.model = model
        self.rng = rng
        self.expMinusDtK = model.returnExpMinusAlphaK(method.dt)
        self.expMinusHalfDtK = model.returnExpMinusAlphaK(method.dt * 0.5)
        self.expHalfDtK = model.returnExpMinusAlphaK(-method.dt * 0.5)
        self.expMinusDtV = model.returnExpMinusAlphaV(method.dt, method.decompType)
        self.constForce = np.zeros(model.L) if constForce is None else np.asarray(constForce, dtype=np.float64)
        self.phiT = fillWalkerFromModel(model) if phiT is None else phiT
        self.op1 = HopSDOperation()
        self.mixedMeasure = HubbardSOCMeasureFixedSDSD(model, self.phiT)
        self.walker = []
        self.walkerIsAlive = []
        self.energies = []
        self.records = []

    # -- set up ---------------------------------------------------------------
    def initialWalker(self):
        """sampleFromPhiT, then QR + comb (afqmcConstraintPathSetWalker.cpp:35-97,122-126)."""
        n = self.method.walkerSizePerThread
        self.walker = [self.phiT.copy() for _ in range(n)]
        self.walkerIsAlive = [True] * n
        self.modifyGM()
        self.popControl()

    def estimateMemory(self):
        """Consumes one aux sample (afqmcConstraintPath.cpp:66)."""
        self.rng.take(self.model.L)

    # -- propagation ----------------------------------------------------------
    def projectExpMinusHalfDtK(self):
        for i, w in enumerate(self.walker):
            if self.walkerIsAlive[i]:
                self.walker[i] = self.op1.applyToRight(self.expMinusHalfDtK, w)

    def projectExpHalfDtK(self):
        for i, w in enumerate(self.walker):
            if self.walkerIsAlive[i]:
                self.walker[i] = self.op1.applyToRight(self.expHalfDtK, w)

    def step_walker(self, i):
        """One walker of projectExpMinusDtKExpMinusDtV (afqmcConstraintPathProjection.cpp:37-74)."""
        m = self.method
        w = self.walker[i]
        op = SDSDOperation(self.phiT, w)
        phase = op.returnLogOverlap().imag
        if math.cos(phase) <= 0.0:
            self.walkerIsAlive[i] = False
            return
        w.addLogw(math.log(math.cos(phase)) - 1j * phase)
        if m.forceType == "constForce":
            force = self.constForce
        elif m.forceType == "dynamicForce":
            force = self.mixedMeasure.getForce(self.expMinusDtV, op, m.forceCap)
        else:
            raise ValueError("Do not know method.forceType " + m.forceType)
        u = self.rng.take(self.model.L)
        aux = self.expMinusDtV.sampleAuxFromForce(force, u)
        sample = self.expMinusDtV.getTwoBodySampleFromAuxForce(aux, force)
        wf, lw = niupnidn_apply_right(sample, w.wf, w.logw)
        lw += m.dt * m.ET
        wf2 = self.expMinusDtK.matrix @ wf
        self.walker[i] = SD(wf=wf2, logw=self.expMinusDtK.logw + lw)
        self.records.append((i, force, aux))

    def projectExpMinusDtKExpMinusDtV(self):
        for i in range(len(self.walker)):
            if self.walkerIsAlive[i]:
                self.step_walker(i)

    def modifyGM(self):
        for i, w in enumerate(self.walker):
            if self.walkerIsAlive[i]:
                w.stabilize()

    def popControl(self):
        """afqmcConstraintPathProjection.cpp:103-156 (serial branch)."""
        n = len(self.walker)
        weight = np.zeros(n)
        for i, w in enumerate(self.walker):
            if self.walkerIsAlive[i]:
                lo = SDSDOperation(self.phiT, w).returnLogOverlap()
                o = cmath.exp(lo).real
                if o <= 0.0:
                    weight[i] = 0.0
                    self.walkerIsAlive[i] = False
                else:
                    weight[i] = o
                    w.addLogw(-lo)
        self.walkerIsAlive = [True] * n
        u = self.rng.take(n)
        table = comb.popConfiguration(1, weight, u)
        comb.populationControl(self.walker, table)
        self.last_weight = weight
        self.last_table = table
        return weight, table

    def projectOneStep(self, idx):
        self.projectExpMinusDtKExpMinusDtV()
        idx[0] += 1
        if idx[0] == self.method.mgsStep:
            self.modifyGM()
            idx[0] = 0
        idx[1] += 1
        if idx[1] == self.method.popControlStep:
            self.popControl()
            idx[1] = 0

    # -- measurement ----------------------------------------------------------
    def addMixedMeasurement(self):
        """afqmcConstraintPathMeasure.cpp:9-27."""
        for i, w in enumerate(self.walker):
            if self.walkerIsAlive[i]:
                t = self.op1.applyToRight(self.expHalfDtK, w)
                op = SDSDOperation(self.phiT, t)
                ov = cmath.exp(op.returnLogOverlap())
                self.mixedMeasure.addMeasurement(op, ov)

    def measureWithProjection(self):
        """afqmcConstraintPath.cpp:99-150."""
        m = self.method
        self.projectExpMinusHalfDtK()
        idx = [0, 0]
        for i in range(m.thermalSize):
            if i < m.ETAdjustMaxSize and i % m.ETAdjustStep == 0:
                self.addMixedMeasurement()
                m.ET = self.mixedMeasure.returnEnergy().real
                self.mixedMeasure.reSet()
            self.projectOneStep(idx)
        for i in range(m.writeNumber):
            for j in range(m.measureNumberPerWrite):
                self.addMixedMeasurement()
                for k in range(m.measureSkipStep):
                    self.projectOneStep(idx)
            mm = self.mixedMeasure
            self.energies.append((mm.HNum, mm.den, mm.KNum, mm.VNum, mm.RNum))
            mm.reSet()

    def run(self):
        self.initialWalker()
        self.estimateMemory()
        self.measureWithProjection()
        return self.energies


# --------------------------------------------------------------------------
# Constraint path with back-propagation
# --------------------------------------------------------------------------
@dataclass
class BPMethod(CPMethod):
    """afqmc_param keys of the back-propagation driver: the constraint-path keys and
    backPropagationStep (tutorials/afqmcConstraintPathBP/source/afqmcConstraintPathMethod.cpp:19-46,
    default 100 at :166)."""
    backPropagationStep: int = 100


class AfqmcConstraintPathBP(AfqmcConstraintPath):
    """Serial back-propagation driver (tutorials/afqmcConstraintPathBP/source/*.cpp).

    The step is the constraint-path one; while isBP, each alive walker's aux field of the step is
    recorded in twoBodyAuxBackup[i][BPIndex] and each comb table in tablePerThreadBackup[BPIndex]
    (afqmcConstraintPathProjection.cpp:73-77,163-171).  A measurement replays the recorded tables on
    (aux history, backup walker) pairs, rebuilds the next window's backup by propagating the first
    measureSkipStep recorded fields (afqmcConstraintPathBackPropagation.cpp:22-97), and measures
    <bra_i| O |ket_i> with bra_i = phiT propagated to the left through walker i's recorded fields
    and ket_i = its backup walker, weighted by the walker's mixed overlap
    (afqmcConstraintPathMeasure.cpp:41-72).  Observables are HubbardSOCMeasureObserveSDSD's."""

    def __init__(self, method: BPMethod, model, rng, phiT: SD = None, constForce=None):
        super().__init__(method, model, rng, phiT, constForce)
        n, nbp = method.walkerSizePerThread, method.backPropagationStep
        self.observeMeasure = HubbardSOCMeasureObserveSDSD(model)
        self.isBP = False
        self.BPIndex = 0
        self.walkerBackup = [None] * n
        self.walkerBPMeasure = [None] * n
        self.twoBodyAuxBackup = [[None] * nbp for _ in range(n)]
        self.twoBodyAuxBPMeasure = [[None] * nbp for _ in range(n)]
        self.tablePerThreadBackup = [None] * nbp
        self.writes = []
        self.betas = []

    # -- recording (afqmcConstraintPathProjection.cpp) --------------------------
    def step_walker(self, i):
        before = len(self.records)
        super().step_walker(i)
        if self.isBP and len(self.records) > before:
            self.twoBodyAuxBackup[i][self.BPIndex] = self.records[-1][2].copy()

    def projectExpMinusDtKExpMinusDtV(self):
        super().projectExpMinusDtKExpMinusDtV()
        if self.isBP:
            self.tablePerThreadBackup[self.BPIndex] = None

    def popControl(self):
        weight, table = super().popControl()
        if self.isBP:
            self.tablePerThreadBackup[self.BPIndex] = table.copy()
        return weight, table

    def projectOneStep(self, idx):
        super().projectOneStep(idx)
        if self.isBP:
            self.BPIndex += 1

    # -- afqmcConstraintPathBackPropagation.cpp -------------------------------
    def prepareBackPropagation(self):
        self.isBP = True
        self.BPIndex = 0
        self.walkerBackup = [w.copy() for w in self.walker]

    def _apply_table_to_records(self, table, nrec):
        """populationControl over AfqmcBackPropagationPop: slot i takes slot table[i]'s first
        nrec aux fields and its backup walker (afqmcBackPropagationPop.cpp:26-34)."""
        for i in range(len(table)):
            j = int(table[i])
            if i != j:
                for k in range(nrec):
                    self.twoBodyAuxBPMeasure[i][k] = self.twoBodyAuxBPMeasure[j][k]
                self.walkerBPMeasure[i] = self.walkerBPMeasure[j].copy()

    def prepareBackPropagationMeasAndNextBackup(self):
        m = self.method
        nbp, skip = m.backPropagationStep, m.measureSkipStep
        self.walkerBPMeasure, self.walkerBackup = self.walkerBackup, self.walkerBPMeasure
        if nbp >= skip:
            self.twoBodyAuxBPMeasure = [list(h) for h in self.twoBodyAuxBackup]
        else:
            self.twoBodyAuxBPMeasure, self.twoBodyAuxBackup = self.twoBodyAuxBackup, self.twoBodyAuxBPMeasure
        for i in range(nbp):
            if self.tablePerThreadBackup[i] is not None:
                self._apply_table_to_records(self.tablePerThreadBackup[i], i + 1)
            if i + 1 == skip:
                self.generateWalkerBackupFromWalkerBPMeasure()
        if nbp >= skip:
            self.BPIndex = nbp - skip
            for i in range(nbp - skip):
                self.tablePerThreadBackup[i] = self.tablePerThreadBackup[i + skip]
                for j in range(len(self.walker)):
                    self.twoBodyAuxBackup[j][i] = self.twoBodyAuxBackup[j][i + skip]
        else:
            self.isBP = False

    def generateWalkerBackupFromWalkerBPMeasure(self):
        m = self.method
        for i in range(len(self.walker)):
            w = self.walkerBPMeasure[i].copy()
            for j in range(m.measureSkipStep):
                aux = self.twoBodyAuxBPMeasure[i][j]
                if aux is not None and aux.size == self.model.L:
                    sample = self.expMinusDtV.getTwoBodySampleFromAux(aux)
                    wf, lw = niupnidn_apply_right(sample, w.wf, w.logw)
                    w = SD(wf=self.expMinusDtK.matrix @ wf, logw=self.expMinusDtK.logw + lw)
                if j % m.mgsStep == 0:
                    w.normalize()
            self.walkerBackup[i] = w

    # -- afqmcConstraintPathMeasure.cpp ---------------------------------------
    def addMixedMeasurement(self):
        for i, w in enumerate(self.walker):
            if self.walkerIsAlive[i]:
                t = self.op1.applyToRight(self.expHalfDtK, w)
                op = SDSDOperation(self.phiT, t)
                self.observeMeasure.addMeasurement(op, cmath.exp(op.returnLogOverlap()))

    def back_propagated_bra(self, history):
        """phiT propagated to the left through one recorded history (afqmcConstraintPathMeasure.cpp:56-65)."""
        m = self.method
        nbp = m.backPropagationStep
        bra = self.op1.applyToLeft(self.expMinusHalfDtK, self.phiT)
        for j in range(nbp, 0, -1):
            aux = history[j - 1]
            if aux is None or aux.size != self.model.L:
                raise ValueError("Aux size does not consistent with L!")
            sample = self.expMinusDtV.getTwoBodySampleFromAux(aux)
            wf, lw = niupnidn_apply_left(sample, bra.wf, bra.logw)
            bra = self.op1.applyToLeft(self.expMinusDtK, SD(wf=wf, logw=lw))
            if (nbp - j + 1) % m.mgsStep == 0:
                bra.normalize()
        return self.op1.applyToLeft(self.expHalfDtK, bra)

    def addMeasurement(self):
        if not self.isBP:
            return
        self.prepareBackPropagationMeasAndNextBackup()
        for i, w in enumerate(self.walker):
            if self.walkerIsAlive[i]:
                t = self.op1.applyToRight(self.expHalfDtK, w)
                overlap = cmath.exp(SDSDOperation(self.phiT, t).returnLogOverlap())
                bra = self.back_propagated_bra(self.twoBodyAuxBPMeasure[i])
                ket = self.op1.applyToRight(self.expHalfDtK, self.walkerBPMeasure[i])
                self.observeMeasure.addMeasurement(SDSDOperation(bra, ket), overlap)

    def measureWithProjection(self):
        """afqmcConstraintPath.cpp:115-180 of the back-propagation tutorial."""
        m = self.method
        nbp, skip = m.backPropagationStep, m.measureSkipStep
        additional = nbp if nbp >= skip else 0
        self.projectExpMinusHalfDtK()
        idx = [0, 0]
        for i in range(m.thermalSize):
            if i < m.ETAdjustMaxSize and i % m.ETAdjustStep == 0:
                self.addMixedMeasurement()
                m.ET = self.observeMeasure.returnEnergy().real
                self.observeMeasure.reSet()
            self.projectOneStep(idx)
        self.prepareBackPropagation()
        if nbp >= skip:
            for _ in range(nbp):
                self.projectOneStep(idx)
        for i in range(m.writeNumber):
            for j in range(m.measureNumberPerWrite):
                for k in range(skip):
                    if self.BPIndex == nbp:
                        self.addMeasurement()
                    self.projectOneStep(idx)
                if nbp < skip:
                    self.prepareBackPropagation()
            self.betas.append((m.thermalSize + (i + 0.5) * m.measureNumberPerWrite * skip - 0.5) * m.dt)
            self.writes.append(self.observeMeasure.snapshot())
            self.observeMeasure.reSet()
        return additional

    def run(self):
        self.initialWalker()
        self.estimateMemory()
        self.measureWithProjection()
        return self.writes


# --------------------------------------------------------------------------
# Phaseless
# --------------------------------------------------------------------------
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


class AfqmcPhaseless:
    """Serial phaseless driver (afqmcPhaseless.cpp, afqmcPhaselessProjection.cpp).

    The background-adjust schedule follows afqmcPhaseless.cpp:34-63,
    135-205: while adjustable, propagation by exp(-dt Kp) uses the Taylor
    operator of the initial background, and exp(-dt Kp) is rebuilt once the
    adjustment ends."""

    def __init__(self, method: PhaselessMethod, model, rng, taylor=None):
        self.method = method
        self.model = model
        self.rng = rng
        m = method
        self.adjustable = False
        self.nGrowth = 0
        if m.ETAndBackGroundAdjustMaxSize > 0:
            self.adjustable = True
        if m.ETAndBackGroundAdjustMaxSize < m.ETAndBackGroundGrowthEstimateMaxSize:
            self.nGrowth = (m.ETAndBackGroundGrowthEstimateMaxSize - m.ETAndBackGroundAdjustMaxSize) // m.ETAndBackGroundGrowthEstimateStep
            if self.nGrowth > 0:
                self.adjustable = True
        self.expMinusDtK = self.expMinusHalfDtK = self.expHalfDtK = None
        self.logExpMinusDtK = self.logExpMinusHalfDtK = self.logExpHalfDtK = None
        if self.adjustable:
            self.logExpMinusDtK = model.returnLogExpMinusAlphaK(m.dt)
            self.logExpMinusHalfDtK = model.returnLogExpMinusAlphaK(m.dt * 0.5)
            self.logExpHalfDtK = model.returnLogExpMinusAlphaK(-m.dt * 0.5)
        else:
            self._set_exp_k()
        self.expMinusDtV = model.returnExpMinusAlphaV(m.dt)
        self.constForce = np.zeros(model.choleskyNumber, dtype=np.complex128)
        self.phiT = fillSD2sFromModel(model)
        self.mixedMeasure = RealMaterialMoleculeMeasureFixedSD2sSD2is(model, self.phiT)
        self.taylor = LogHop2isSD2isOperation() if taylor is None else taylor
        self.walker = []
        self.walkerIsAlive = []
        self.energies = []

    def _set_exp_k(self):
        m = self.method
        self.expMinusDtK = self.model.returnExpMinusAlphaK(m.dt)
        self.expMinusHalfDtK = self.model.returnExpMinusAlphaK(m.dt * 0.5)
        self.expHalfDtK = self.model.returnExpMinusAlphaK(-m.dt * 0.5)

    def _apply_k(self, exp_op, log_op, w):
        if exp_op is not None:
            return HopSDOperation().applyToRight(exp_op, w)
        return self.taylor.applyToRight(log_op, w)

    def initialWalker(self):
        n = self.method.walkerSizePerThread
        self.walker = [sd2is_from_phiT(self.phiT) for _ in range(n)]
        self.walkerIsAlive = [True] * n
        self.modifyGM()
        self.popControl()

    def estimateMemory(self):
        self.rng.take(2 * self.model.choleskyNumber)

    def step_walker(self, i):
        """afqmcPhaselessProjection.cpp:75-121."""
        m = self.method
        G = self.model.choleskyNumber
        w = self.walker[i]
        op = SD2sSD2isOperation(self.phiT, w)
        phase = op.returnLogOverlap().imag
        if math.cos(phase) <= 0.0:
            self.walkerIsAlive[i] = False
            return
        w.addLogw(math.log(math.cos(phase)) - 1j * phase)
        if m.forceType == "constForce":
            force = self.constForce
        else:
            force = self.mixedMeasure.getForce(self.expMinusDtV, op, m.forceCap)
        gauss = gaussians_from_uniforms(self.rng.take(2 * G))
        aux = self.expMinusDtV.sampleAuxFromForce(force, gauss)
        sample = self.expMinusDtV.getTwoBodySampleFromAuxForce(aux, force)
        auxForce = 0j
        for k in range(G):
            auxForce += aux[k] * force[k]
        phase = auxForce.imag
        if math.cos(phase) <= 0.0:
            self.walkerIsAlive[i] = False
            return
        w.addLogw(math.log(math.cos(phase)))
        t = self.taylor.applyToRight(sample, w)
        t.addLogw(m.dt * m.ET)
        self.walker[i] = self._apply_k(self.expMinusDtK, self.logExpMinusDtK, t)

    def projectExpMinusDtKExpMinusDtV(self):
        for i in range(len(self.walker)):
            if self.walkerIsAlive[i]:
                self.step_walker(i)

    def projectExpMinusHalfDtK(self):
        for i, w in enumerate(self.walker):
            if self.walkerIsAlive[i]:
                self.walker[i] = self._apply_k(self.expMinusHalfDtK, self.logExpMinusHalfDtK, w)

    def modifyGM(self):
        for i, w in enumerate(self.walker):
            if self.walkerIsAlive[i]:
                w.stabilize()

    def popControl(self):
        n = len(self.walker)
        weight = np.zeros(n)
        for i, w in enumerate(self.walker):
            if self.walkerIsAlive[i]:
                lo = SD2sSD2isOperation(self.phiT, w).returnLogOverlap()
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
        return weight, table

    def projectOneStep(self, idx):
        self.projectExpMinusDtKExpMinusDtV()
        idx[0] += 1
        if idx[0] == self.method.mgsStep:
            self.modifyGM()
            idx[0] = 0
        idx[1] += 1
        if idx[1] == self.method.popControlStep:
            self.popControl()
            idx[1] = 0

    def addMixedMeasurement(self):
        """afqmcPhaselessMeasure.cpp:9-26."""
        for i, w in enumerate(self.walker):
            if self.walkerIsAlive[i]:
                t = self._apply_k(self.expHalfDtK, self.logExpHalfDtK, w)
                op = SD2sSD2isOperation(self.phiT, t)
                ov = cmath.exp(op.returnLogOverlap())
                self.mixedMeasure.addMeasurement(op, ov)

    def adjustETAndBackGround(self):
        mm = self.mixedMeasure
        self.method.ET = mm.returnEnergy().real
        self.model.updateBackGround(mm.returnCholeskyBg())
        mm.reSet()

    def measureWithProjection(self):
        """afqmcPhaseless.cpp:135-205."""
        m = self.method
        self.projectExpMinusHalfDtK()
        idx = [0, 0]
        for i in range(m.thermalSize):
            if i < m.ETAndBackGroundAdjustMaxSize and i % m.ETAndBackGroundAdjustStep == 0:
                self.addMixedMeasurement()
                self.adjustETAndBackGround()
            if i == m.ETAndBackGroundAdjustMaxSize - 1 and self.nGrowth == 0:
                self._set_exp_k()
            if self.nGrowth > 0 and m.ETAndBackGroundAdjustMaxSize <= i < m.ETAndBackGroundGrowthEstimateMaxSize:
                if (i + 1 - m.ETAndBackGroundAdjustMaxSize) % m.ETAndBackGroundGrowthEstimateStep == 0:
                    self.addMixedMeasurement()
                if i == m.ETAndBackGroundGrowthEstimateMaxSize - 1:
                    self.adjustETAndBackGround()
                    self._set_exp_k()
            self.projectOneStep(idx)
        for i in range(m.writeNumber):
            for j in range(m.measureNumberPerWrite):
                self.addMixedMeasurement()
                for k in range(m.measureSkipStep):
                    self.projectOneStep(idx)
            mm = self.mixedMeasure
            self.energies.append((mm.HNum, mm.den, mm.TNum))
            mm.reSet()

    def run(self):
        self.initialWalker()
        self.estimateMemory()
        self.measureWithProjection()
        return self.energies
