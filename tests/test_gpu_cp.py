"""Parity of the CP walker path on the GPU against the CPU oracle, through the C-ABI.

FP64 bar: log-overlaps, weights, forces, Green's functions and energies to
relative 1e-10; HS fields and comb tables exact.  Walker matrices are compared
through invariants where the QR differs by column signs (CholeskyQR gives
R_kk > 0, LAPACK's Householder may not): Q_dev^H Q_ref must be diagonal with
unit-modulus entries."""
import cmath
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from afqmclab_amd.cp import AfqmcConstraintPath, CPMethod as DevMethod, CPPropagator  # noqa: E402
from afqmclab_amd.model import HubbardModel, free_electron_trial, spin_block_trial  # noqa: E402
from oracle import comb as ocomb  # noqa: E402
from oracle.blocks import SD, SDSDOperation, niupnidn_apply_right, NiupNidnSample  # noqa: E402
from oracle.drivers import AfqmcConstraintPath as OracleCP, CPMethod as OracleMethod  # noqa: E402
from oracle.models import HubbardSOC, HubbardSOCMeasureFixedSDSD  # noqa: E402
from oracle.rng import ListStream, UniformStream  # noqa: E402
from tests.gpu_util import crand, ctx  # noqa: E402

RTOL = 1e-10


def make_case(mode, nx=4, ny=4, nup=5, ndn=5, U=4.0, t2=0.0, decomp="densitySpin", precision=0, nw=12, seed=1):
    model = HubbardModel.square(nx, ny, nup + ndn, U, 1.0, t2)
    omodel = HubbardSOC(model.L, model.N, model.K, U=model.U)
    if mode == "dense":
        phi = free_electron_trial(model)
        kw = dict(spin_blocks=False)
    else:
        phi = spin_block_trial(model, nup, ndn)
        kw = dict(nup=nup, spin_blocks=True)
    method = DevMethod(dt=0.05, decompType=decomp, walkerSizePerThread=nw)
    prop = CPPropagator(ctx(), model, method, phi, precision=precision, max_walkers=nw, **kw)
    rng = np.random.default_rng(seed)
    L = model.L
    wfs = []
    for i in range(nw):
        if mode == "dense":
            w = phi + 0.3 * crand(rng, 2 * L, model.N)
        else:
            w = np.zeros_like(phi)
            w[:L, :nup] = phi[:L, :nup] + 0.3 * crand(rng, L, nup)
            w[L:, nup:] = phi[L:, nup:] + 0.3 * crand(rng, L, ndn)
        wfs.append(w)
    logw = 0.1 * crand(rng, nw)
    ws = prop.walkers(nw)
    ws.upload(wfs, logw)
    return model, omodel, phi, prop, ws, wfs, logw, method


def close_log(a, b, rtol=RTOL):
    """log-overlaps: real parts relative, phases via exp(i Im)."""
    assert abs(a.real - b.real) <= rtol * max(1.0, abs(b.real)), (a, b)
    assert abs(cmath.exp(1j * (a.imag - b.imag)) - 1.0) <= rtol * 10, (a, b)


@pytest.mark.parametrize("mode", ["dense", "blocks"])
def test_log_overlap_green_force(mode):
    model, om, phi, prop, ws, wfs, logw, method = make_case(mode)
    logO = prop.log_overlap(ws)
    gd, go = prop.green_diagonal(ws)
    f = prop.force(ws)
    left = SD(wf=phi)
    hs = om.returnExpMinusAlphaV(method.dt, method.decompType)
    meas = HubbardSOCMeasureFixedSDSD(om, left)
    for i in range(ws.nw):
        op = SDSDOperation(left, SD(wf=wfs[i], logw=logw[i]))
        close_log(logO[i], op.returnLogOverlap())
        assert np.max(np.abs(gd[i] - op.returnGreenDiagonal())) < RTOL * 10
        if mode == "dense":
            assert np.max(np.abs(go[i] - op.returnGreenOffDiagonal())) < RTOL * 10
        assert np.max(np.abs(f[i] - meas.getForce(hs, op, method.forceCap))) < RTOL * 10


@pytest.mark.parametrize("decomp", ["densityCharge", "densitySpin", "hopCharge", "hopSpin"])
def test_hs_constants_match_oracle(decomp):
    model, om, phi, prop, ws, *_ = make_case("dense", decomp=decomp)
    g, cd = prop.hs_constants()
    hs = om.returnExpMinusAlphaV(0.05, decomp)
    assert np.allclose(g, hs.gamma, rtol=1e-15, atol=1e-15)
    assert np.allclose(cd, np.stack([hs.constDiag00, hs.constDiag01, hs.constDiag10, hs.constDiag11]), rtol=1e-15)


@pytest.mark.parametrize("mode", ["dense", "blocks"])
@pytest.mark.parametrize("which", [0, 1, 2])
def test_onebody_apply(mode, which):
    model, om, phi, prop, ws, wfs, logw, method = make_case(mode)
    prop.onebody(ws, which)
    out, lw, al = ws.download()
    alpha = [method.dt, method.dt * 0.5, -method.dt * 0.5][which]
    B = om.returnExpMinusAlphaK(alpha).matrix
    for i in range(ws.nw):
        assert np.max(np.abs(out[i] - B @ wfs[i])) < 1e-13
        assert lw[i] == logw[i]


def test_hs_apply_2x2():
    # the bandwidth-bound NiupNidnSampleSDOperation::applyToRight with general 2x2 blocks
    model, om, phi, prop, ws, wfs, logw, method = make_case("dense")
    rng = np.random.default_rng(5)
    L = model.L
    d4 = crand(rng, ws.nw, 4, L)
    slw = crand(rng, ws.nw)
    prop.hs_apply(ws, d4, slw)
    out, lw, _ = ws.download()
    for i in range(ws.nw):
        s = NiupNidnSample(L)
        s.diag00, s.diag01, s.diag10, s.diag11 = d4[i]
        s.logw = slw[i]
        ref, rlw = niupnidn_apply_right(s, wfs[i], logw[i])
        assert np.max(np.abs(out[i] - ref)) < 1e-14
        assert abs(lw[i] - rlw) < 1e-14


@pytest.mark.parametrize("mode", ["dense", "blocks"])
def test_stabilize(mode):
    model, om, phi, prop, ws, wfs, logw, method = make_case(mode)
    prop.stabilize(ws)
    out, lw, _ = ws.download()
    for i in range(ws.nw):
        ref = SD(wf=wfs[i], logw=logw[i])
        ref.stabilize()
        assert abs(lw[i] - ref.logw) < RTOL * max(1, abs(ref.logw))
        D = out[i].conj().T @ ref.wf
        assert np.max(np.abs(np.abs(np.diag(D)) - 1)) < 1e-10
        assert np.max(np.abs(D - np.diag(np.diag(D)))) < 1e-10


@pytest.mark.parametrize("mode", ["dense", "blocks"])
def test_stabilize_second_pass(mode):
    """Walkers with two nearly parallel columns (condition number about 1e4) leave the first
    CholeskyQR pass about 1e-8 from orthonormal, so they take the second pass; the
    well-conditioned walkers beside them skip it.  Both kinds must match the oracle's QR."""
    model, om, phi, prop, ws, wfs, logw, method = make_case(mode)
    L = model.L
    rng = np.random.default_rng(9)
    wfs2 = []
    for i, w in enumerate(wfs):
        if i % 2:
            w = w.copy()
            rows = slice(0, 2 * L) if mode == "dense" else slice(0, L)
            w[rows, 1] = w[rows, 0] + 1e-4 * crand(rng, w[rows, 0].shape[0])
        wfs2.append(w)
    ws.upload(wfs2, logw)
    prop.stabilize(ws)
    out, lw, _ = ws.download()
    for i in range(ws.nw):
        ref = SD(wf=wfs2[i], logw=logw[i])
        ref.stabilize()
        assert abs(lw[i] - ref.logw) < RTOL * max(1, abs(ref.logw)), i
        q = out[i]
        assert np.max(np.abs(q.conj().T @ q - np.eye(model.N))) < 1e-12, i
        D = q.conj().T @ ref.wf
        assert np.max(np.abs(np.abs(np.diag(D)) - 1)) < 1e-10, i
        assert np.max(np.abs(D - np.diag(np.diag(D)))) < 1e-10, i


@pytest.mark.parametrize("mode,decomp", [("dense", "densitySpin"), ("blocks", "densitySpin"),
                                         ("dense", "densityCharge"), ("blocks", "densityCharge"),
                                         ("dense", "hopCharge"), ("dense", "hopSpin")])
def test_step_matches_oracle(mode, decomp):
    """One projectExpMinusDtKExpMinusDtV over all walkers with the same injected uniforms."""
    model, om, phi, prop, ws, wfs, logw, method = make_case(mode, decomp=decomp, nw=10)
    rng = np.random.default_rng(9)
    # kill one walker up front: it must consume no uniforms and stay untouched
    alive = np.ones(ws.nw, dtype=np.int32)
    alive[3] = 0
    ws.upload(wfs, logw, alive)
    u = rng.uniform(0, 1, model.L * ws.nw)
    prop.set_uniforms(u)
    prop.set_ET(-7.5)
    f_dev, aux_dev = prop.step(ws, record=True)
    out, lw, al = ws.download()
    om_method = OracleMethod(dt=method.dt, decompType=decomp, walkerSizePerThread=ws.nw, ET=-7.5,
                             thermalSize=0, ETAdjustMaxSize=0)
    drv = OracleCP(om_method, om, ListStream(u), phiT=SD(wf=phi))
    drv.walker = [SD(wf=wfs[i], logw=logw[i]) for i in range(ws.nw)]
    drv.walkerIsAlive = [bool(a) for a in alive]
    drv.projectExpMinusDtKExpMinusDtV()
    recs = {r[0]: r for r in drv.records}
    for i in range(ws.nw):
        assert bool(al[i]) == drv.walkerIsAlive[i]
        if not drv.walkerIsAlive[i]:
            continue
        _, f_ref, aux_ref = recs[i]
        assert np.array_equal(aux_dev[i], aux_ref)
        assert np.max(np.abs(f_dev[i] - f_ref)) < RTOL
        ref = drv.walker[i]
        close_log(lw[i], ref.logw)
        assert np.max(np.abs(out[i] - ref.wf)) < 1e-12 * max(1, np.max(np.abs(ref.wf)))


@pytest.mark.parametrize("mode", ["dense", "blocks"])
def test_step_split_bf16(mode):
    """Precision 2 (the propagation GEMM in split-BF16 products, all else as the mixed path): the
    HS fields of one step equal the oracle's, and the propagated walkers agree to 1e-4 of their
    largest element, the mixed path's bar (the split products carry up to about 2^-16 each)."""
    model, om, phi, prop, ws, wfs, logw, method = make_case(mode, nw=10, precision=2)
    rng = np.random.default_rng(9)
    u = rng.uniform(0, 1, model.L * ws.nw)
    prop.set_uniforms(u)
    prop.set_ET(-7.5)
    f_dev, aux_dev = prop.step(ws, record=True)
    out, lw, al = ws.download()
    om_method = OracleMethod(dt=method.dt, decompType=method.decompType, walkerSizePerThread=ws.nw, ET=-7.5,
                             thermalSize=0, ETAdjustMaxSize=0)
    drv = OracleCP(om_method, om, ListStream(u), phiT=SD(wf=phi))
    drv.walker = [SD(wf=wfs[i], logw=logw[i]) for i in range(ws.nw)]
    drv.walkerIsAlive = [True] * ws.nw
    drv.projectExpMinusDtKExpMinusDtV()
    recs = {r[0]: r for r in drv.records}
    for i in range(ws.nw):
        _, f_ref, aux_ref = recs[i]
        assert np.array_equal(aux_dev[i], aux_ref)
        ref = drv.walker[i]
        assert np.max(np.abs(out[i] - ref.wf)) < 1e-4 * np.max(np.abs(ref.wf))
        close_log(lw[i], ref.logw, rtol=1e-5)


@pytest.mark.parametrize("mode", ["dense", "blocks"])
def test_pop_control_matches_oracle(mode):
    model, om, phi, prop, ws, wfs, logw, method = make_case(mode, nw=16)
    rng = np.random.default_rng(4)
    # one walker with a phase beyond pi/2 -> non-positive weight -> replaced
    logw = logw.copy()
    logw[2] += 1j * math.pi
    ws.upload(wfs, logw)
    u = rng.uniform(0, 1, ws.nw)
    prop.set_uniforms(u)
    wt, tb, st = prop.pop_control(ws, record=True)
    om_method = OracleMethod(dt=method.dt, walkerSizePerThread=ws.nw, thermalSize=0, ETAdjustMaxSize=0)
    drv = OracleCP(om_method, om, ListStream(u), phiT=SD(wf=phi))
    drv.walker = [SD(wf=wfs[i], logw=logw[i]) for i in range(ws.nw)]
    drv.walkerIsAlive = [True] * ws.nw
    wref, tref = drv.popControl()
    assert np.allclose(wt, wref, rtol=RTOL, atol=0)
    assert wt[2] == 0.0
    assert np.array_equal(tb, tref)
    assert st == pytest.approx(ocomb.popCheck(list(wref)), rel=1e-12)
    out, lw, al = ws.download()
    assert np.all(al == 1)
    for i in range(ws.nw):
        close_log(lw[i], drv.walker[i].logw)
        assert np.max(np.abs(out[i] - drv.walker[i].wf)) < 1e-13


@pytest.mark.parametrize("mode", ["dense", "blocks"])
def test_measure_matches_oracle(mode):
    model, om, phi, prop, ws, wfs, logw, method = make_case(mode, U=3.0, t2=-0.2)
    # keep the overlaps O(1)
    logw = np.zeros(ws.nw, dtype=complex)
    ws.upload(wfs, logw)
    prop.reset_measure()
    prop.measure(ws)
    e = prop.energy()
    om_method = OracleMethod(dt=method.dt, walkerSizePerThread=ws.nw, thermalSize=0, ETAdjustMaxSize=0)
    drv = OracleCP(om_method, om, UniformStream(1), phiT=SD(wf=phi))
    drv.walker = [SD(wf=wfs[i], logw=logw[i]) for i in range(ws.nw)]
    drv.walkerIsAlive = [True] * ws.nw
    drv.addMixedMeasurement()
    mm = drv.mixedMeasure
    for key, ref in [("den", mm.den), ("HNum", mm.HNum), ("KNum", mm.KNum), ("VNum", mm.VNum), ("RNum", mm.RNum)]:
        assert abs(e[key] - ref) < RTOL * max(1.0, abs(ref)), (key, e[key], ref)
    sites = prop.site_measure()
    assert np.allclose(sites[0], mm.NupNum, rtol=RTOL, atol=1e-12)
    assert np.allclose(sites[1], mm.NdnNum, rtol=RTOL, atol=1e-12)
    if mode == "dense":
        assert np.allclose(sites[2], mm.SplusNum, rtol=RTOL, atol=1e-12)
        assert np.allclose(sites[3], mm.SminusNum, rtol=RTOL, atol=1e-12)


@pytest.mark.parametrize("mode", ["dense", "blocks"])
def test_driver_trajectory_matches_oracle(mode):
    """A whole small CP run (QR, comb, ET adjust, measurement) on the same uniform stream:
    per-block HNum/den agree to 1e-9 relative and the uniform counters agree exactly."""
    nx, ny, nup, ndn, U = 4, 4, 5, 5, 4.0
    model = HubbardModel.square(nx, ny, nup + ndn, U)
    om = HubbardSOC(model.L, model.N, model.K, U=model.U)
    phi = free_electron_trial(model) if mode == "dense" else spin_block_trial(model, nup, ndn)
    kw = dict(spin_blocks=False) if mode == "dense" else dict(nup=nup, spin_blocks=True)
    m = dict(dt=0.01, thermalSize=20, writeNumber=4, measureNumberPerWrite=2, measureSkipStep=5,
             walkerSizePerThread=16, decompType="densitySpin", mgsStep=10, popControlStep=10,
             ETAdjustStep=10, ETAdjustMaxSize=20, seed=2024)
    dev = AfqmcConstraintPath(ctx(), model, DevMethod(**m), phi, **kw)
    dev.run()
    ref = OracleCP(OracleMethod(**m), om, UniformStream(2024), phiT=SD(wf=phi))
    ref.run()
    assert dev.prop.counter() == ref.rng.counter
    assert len(dev.energies) == len(ref.energies)
    for a, b in zip(dev.energies, ref.energies):
        for x, y in zip(a[:2], b[:2]):
            assert abs(x - y) < 1e-9 * abs(y), (x, y)


@pytest.mark.parametrize("mode", ["dense", "blocks"])
@pytest.mark.parametrize("precision", [0, 1])
def test_step_fields_and_table_replay_reproduce_the_run(mode, precision):
    """afqmc_cp_step_fields with the fields a step recorded, and afqmc_cp_pop_control_table with
    the table a comb drew, reproduce that step and comb bit for bit on identical walkers, and
    draw no uniforms."""
    model, om, phi, prop, ws, wfs, logw, method = make_case(mode, nw=12, precision=precision)
    prop2 = CPPropagator(ctx(), model, method, phi, precision=precision, max_walkers=12,
                         **(dict(spin_blocks=False) if mode == "dense" else dict(nup=5, spin_blocks=True)))
    ws2 = prop2.walkers(12)
    ws2.upload(wfs, logw)
    for p in (prop, prop2):
        p.set_seed(3, 0)
        p.set_ET(-7.0)
    for _ in range(3):
        f, aux = prop.step(ws, record=True)
        c0 = prop2.counter()
        al = ws2.alive()
        aux = np.where(al[:, None] != 0, aux, 0)  # rows of dead walkers are not read
        prop2.step_fields(ws2, aux)
        assert prop2.counter() == c0
        a, b = ws.download(), ws2.download()
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
    wt, tb, st = prop.pop_control(ws, record=True)
    wt2 = prop2.pop_control_table(ws2, tb)
    assert np.array_equal(wt, wt2)
    a, b = ws.download(), ws2.download()
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    with pytest.raises(Exception):
        prop2.step_fields(ws2, np.zeros((12, model.L), dtype=np.int32))
    bad = np.arange(12, dtype=np.int32)
    bad[0], bad[1] = 1, 2  # slot 1 is a parent but does not keep itself
    with pytest.raises(Exception):
        prop2.pop_control_table(ws2, bad)


@pytest.mark.parametrize("mode", ["dense", "blocks"])
@pytest.mark.parametrize("precision", [0, 1])
def test_stabilize_ill_conditioned_and_rank_deficient(mode, precision):
    """Walkers beyond CholeskyQR: FP64 walkers with two columns 1e-10 apart (condition number
    about 1e10, the Gram matrix is not positive definite in FP64) take the shifted CholeskyQR3
    path, with Q orthonormal and log det R equal to the oracle's Householder QR; a walker with a
    zero column (Householder's R_kk = 0, log det R = -inf) is marked dead and counted; a walker
    with two equal columns is numerically rank deficient: where the shifted pass leaves its
    column at the rounding level it dies as well, otherwise it stays alive with an orthonormal Q
    and a log-weight lowered to the noise level of its R_kk, as Householder lowers it; the
    well-conditioned walkers beside them match the oracle as before."""
    model, om, phi, prop, ws, wfs, logw, method = make_case(mode, precision=precision)
    rows = slice(0, 2 * model.L) if mode == "dense" else slice(0, model.L)
    rng = np.random.default_rng(17)
    dup, zero = 5, 4
    ill = {1, 7, 10} if precision == 0 else set()
    wfs2 = []
    for i, w in enumerate(wfs):
        w = w.copy()
        n = w[rows, 0].shape[0]
        if i in ill:
            w[rows, 1] = w[rows, 0] + 1e-10 * crand(rng, n)
        if i == dup:
            w[rows, 2] = w[rows, 0]
        if i == zero:
            w[rows, 2] = 0.0
        wfs2.append(w)
    ws.upload(wfs2, logw)
    prop.stabilize(ws)
    out, lw, al = ws.download()
    shifted, killed = prop.qr_stats()
    assert al[zero] == 0
    assert killed == 1 + (al[dup] == 0)
    assert shifted == len(ill) + 2
    for i in range(ws.nw):
        if i == zero or (i == dup and al[dup] == 0):
            continue
        assert al[i] == 1, i
        q = out[i]
        if precision == 0:
            assert np.max(np.abs(q.conj().T @ q - np.eye(model.N))) < 1e-12, i
        if i == dup:
            good = SD(wf=wfs[i], logw=logw[i])
            good.stabilize()
            assert lw[i].real < good.logw.real - 20.0
            continue
        ref = SD(wf=wfs2[i], logw=logw[i])
        ref.stabilize()
        tol = (1e-5 if i in ill else RTOL) if precision == 0 else 1e-5
        assert abs(lw[i].real - ref.logw.real) < tol * max(1, abs(ref.logw.real)), i
        D = q.conj().T @ ref.wf
        assert np.max(np.abs(np.abs(np.diag(D)) - 1)) < (1e-5 if (i in ill or precision) else 1e-10), i


@pytest.mark.parametrize("lattice", [(3, 5), (24, 24), (16, 16)])
@pytest.mark.parametrize("precision", [0, 1])
def test_hs_apply_2x2_shapes(lattice, precision):
    """The spin-mixing HS kernel on an odd site count (8-byte FP32 accesses), on more sites than a
    workgroup has threads (several passes per column) and on the 16 x 16 lattice (several columns
    per pass), against NiupNidnSampleSDOperation::applyToRight of the oracle."""
    nx, ny = lattice
    L = nx * ny
    N = min(2 * L - 2, 24)
    model = HubbardModel.square(nx, ny, N, 4.0)
    phi = free_electron_trial(model)
    method = DevMethod(dt=0.05, walkerSizePerThread=5)
    prop = CPPropagator(ctx(), model, method, phi, spin_blocks=False, precision=precision, max_walkers=5)
    rng = np.random.default_rng(L)
    wfs = [phi + 0.3 * crand(rng, 2 * L, N) for _ in range(5)]
    logw = 0.1 * crand(rng, 5)
    ws = prop.walkers(5)
    ws.upload(wfs, logw)
    d4 = crand(rng, 5, 4, L)
    slw = crand(rng, 5)
    prop.hs_apply(ws, d4, slw)
    out, lw, _ = ws.download()
    tol = 1e-14 if precision == 0 else 2e-6
    for i in range(5):
        s = NiupNidnSample(L)
        s.diag00, s.diag01, s.diag10, s.diag11 = d4[i]
        s.logw = slw[i]
        ref, rlw = niupnidn_apply_right(s, wfs[i], logw[i])
        assert np.max(np.abs(out[i] - ref)) < tol * max(1.0, np.max(np.abs(ref)))
        assert abs(lw[i] - rlw) < 1e-14


@pytest.mark.parametrize("precision", [0, 1])
def test_dense_wide_walkers_step_stabilize_measure(precision):
    """Dense walkers of 192 columns (12 x 12 lattice, 96 + 96 electrons), beyond the
    register-resident Gauss-Jordan and Cholesky kernels: one step with injected uniforms, a QR and
    a mixed measurement against the oracle.  FP64 to 1e-10; the mixed path (FP32 walkers, FP64
    factorisations) to 1e-4 relative on log-weights and energies, fields exact."""
    model, om, phi, prop, ws, wfs, logw, method = make_case("dense", nx=12, ny=12, nup=96, ndn=96, nw=4,
                                                            precision=precision)
    assert model.N == 192
    tol = RTOL if precision == 0 else 1e-4
    rng = np.random.default_rng(31)
    # walkers near the trial, so that the constraint keeps all of them through the step
    wfs = [phi + 0.03 * crand(rng, 2 * model.L, model.N) for _ in range(ws.nw)]
    ws.upload(wfs, logw)
    u = rng.uniform(0, 1, model.L * ws.nw)
    prop.set_uniforms(u)
    prop.set_ET(-150.0)
    f_dev, aux_dev = prop.step(ws, record=True)
    om_method = OracleMethod(dt=method.dt, decompType="densitySpin", walkerSizePerThread=ws.nw, ET=-150.0,
                             thermalSize=0, ETAdjustMaxSize=0)
    drv = OracleCP(om_method, om, ListStream(u), phiT=SD(wf=phi))
    drv.walker = [SD(wf=wfs[i], logw=logw[i]) for i in range(ws.nw)]
    drv.walkerIsAlive = [True] * ws.nw
    drv.projectExpMinusDtKExpMinusDtV()
    out, lw, al = ws.download()
    recs = {r[0]: r for r in drv.records}
    for i in range(ws.nw):
        assert bool(al[i]) == drv.walkerIsAlive[i]
        if not drv.walkerIsAlive[i]:
            continue
        assert np.array_equal(aux_dev[i], recs[i][2])
        close_log(lw[i], drv.walker[i].logw, tol)
    # QR of the stepped walkers (Cholesky of 192 columns)
    prop.stabilize(ws)
    out2, lw2, _ = ws.download()
    for i in range(ws.nw):
        if not drv.walkerIsAlive[i]:
            continue
        ref = SD(wf=out[i], logw=lw[i])
        ref.stabilize()
        close_log(lw2[i], ref.logw, tol)
        q = out2[i]
        assert np.max(np.abs(q.conj().T @ q - np.eye(model.N))) < (1e-12 if precision == 0 else 1e-5)
    # mixed measurement on walkers with O(1) weights
    ws.upload(wfs, np.zeros(ws.nw, dtype=complex))
    prop.reset_measure()
    prop.measure(ws)
    e = prop.energy()
    drv2 = OracleCP(om_method, om, UniformStream(1), phiT=SD(wf=phi))
    drv2.walker = [SD(wf=wfs[i], logw=0j) for i in range(ws.nw)]
    drv2.walkerIsAlive = [True] * ws.nw
    drv2.addMixedMeasurement()
    mm = drv2.mixedMeasure
    for key, ref in [("den", mm.den), ("HNum", mm.HNum), ("KNum", mm.KNum), ("VNum", mm.VNum)]:
        assert abs(e[key] - ref) < tol * max(1.0, abs(ref)), (key, e[key], ref)
