"""Parity at the full sizes of the benchmark workloads, through the C-ABI, against the CPU oracle
on sampled walkers:

* 16 x 16 Hubbard, U = 8, 112 + 112, UHF trial, 2048 spin-block walkers: one step with
  host-provided uniforms (fields, forces, walkers, log-weights), one QR and one measurement;
* the phaseless synthetic molecule, L = 200, 40 + 40, G = 1600, 1024 walkers: one step, one
  QR + comb over all walkers, and one measurement whose exchange GEMM runs in walker chunks
  (the per-walker blocks of all 1024 walkers exceed the 4 GiB chunk buffer), with the measured
  walkers placed on both sides of chunk boundaries;
* 24 heterogeneous lattices (6 x 6 to 12 x 12) x 256 walkers: one graph step of all instances
  against the oracle, and graph replays of 10 steps + QR + comb against eager calls; the same with
  grouped launches (one launch per kernel and phase over all instances) against the oracle and,
  bit for bit, against per-instance calls.

A walker's uniforms start after those of the earlier walkers that pass the overlap-phase check
(killed walkers draw none, afqmcConstraintPathProjection.cpp:47), so each sampled walker is
replayed by the oracle from its own offset in the host-provided list.  FP64 bar: fields exact,
forces, walkers, log-weights and energies to 1e-10 relative.  Mixed precision: 1e-4 relative,
and a field may differ from the oracle's only where its uniform lies within 1e-5 of the
field's probability threshold."""
import cmath
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from afqmclab_amd.cp import Context, CPGroup, CPMethod, CPPropagator  # noqa: E402
from afqmclab_amd.model import HubbardModel, heterogeneous_hubbard_instances, uhf_trial  # noqa: E402
from afqmclab_amd.phaseless import MoleculeModel, PhaselessMethod, PhaselessPropagator  # noqa: E402
from oracle import comb as ocomb  # noqa: E402
from oracle.blocks import SD, SD2is, SD2sSD2isOperation, SDSDOperation  # noqa: E402
from oracle.drivers import AfqmcConstraintPath as OracleCP, CPMethod as OracleMethod  # noqa: E402
from oracle.drivers import AfqmcPhaseless as OraclePh, PhaselessMethod as OraclePhMethod  # noqa: E402
from oracle.models import HubbardSOC, RealMaterialMoleculeMeasureFixedSD2sSD2is, synthetic_molecule  # noqa: E402
from oracle.rng import ListStream, gaussians_from_uniforms  # noqa: E402
from tests.gpu_util import ctx  # noqa: E402

_CACHE = {}


def tol_of(precision):
    return 1e-10 if precision == 0 else 1e-4


def close_log(a, b, rtol):
    assert abs(a.real - b.real) <= rtol * max(1.0, abs(b.real)), (a, b)
    assert abs(cmath.exp(1j * (a.imag - b.imag)) - 1.0) <= rtol * 10, (a, b)


def uniform_offsets(alive, logO, per):
    """Start of each walker's uniforms: per x the earlier walkers alive and past the phase check."""
    passes = ((np.asarray(alive) != 0) & (np.cos(np.asarray(logO).imag) > 0.0)).astype(np.int64)
    return per * (np.cumsum(passes) - passes)


def check_fields(aux_dev, aux_ref, u, force, precision):
    diff = np.nonzero(np.asarray(aux_dev) != np.asarray(aux_ref))[0]
    if precision == 0:
        assert diff.size == 0, diff
    for i in diff:
        p = math.exp(-force[i]) / (math.exp(-force[i]) + math.exp(force[i]))
        assert abs(u[i] - p) < 1e-5, (i, u[i], p)


# ------------------------------------------------------------------ 16 x 16, U = 8, 2048 walkers
def c2_workload():
    if "c2" not in _CACHE:
        model = HubbardModel.square(16, 16, 224, 8.0)
        _CACHE["c2"] = (model, uhf_trial(model, 112, 112))
    return _CACHE["c2"]


@pytest.mark.parametrize("precision", [0, 1])
def test_hubbard16x16_step_qr_measure_match_oracle(precision):
    model, phi = c2_workload()
    L, nup, nw, dt, ET = model.L, 112, 2048, 0.005, -150.0
    tol = tol_of(precision)
    method = CPMethod(dt=dt, walkerSizePerThread=nw, decompType="densitySpin", forceType="dynamicForce",
                      forceCap=1.5)
    prop = CPPropagator(ctx(), model, method, phi, nup=nup, spin_blocks=True, precision=precision, max_walkers=nw)
    ws = prop.walkers(nw)
    ws.set_from_trial()
    prop.set_seed(11, 0)
    prop.set_ET(ET)
    for _ in range(3):  # distinct walkers
        prop.step(ws)
    prop.stabilize(ws)
    sample = np.array([0, 1, 2, 700, 1023, 1024, 1500, 2047])
    wf0, lw0, al0 = ws.download_some(sample)
    logO = prop.log_overlap(ws)
    alive = ws.alive()
    u = np.random.default_rng(5).uniform(0, 1, L * nw)
    prop.set_uniforms(u)
    print(f"[c2] precision={precision}: step of {nw} walkers with host uniforms", flush=True)
    f_dev, aux_dev = prop.step(ws, record=True)
    wf1, lw1, al1 = ws.download_some(sample)
    off = uniform_offsets(alive, logO, L)
    om = HubbardSOC(L, model.N, model.K, U=model.U)
    om_method = OracleMethod(dt=dt, decompType="densitySpin", walkerSizePerThread=len(sample), ET=ET,
                             thermalSize=0, ETAdjustMaxSize=0)
    drv = OracleCP(om_method, om, ListStream(u), phiT=SD(wf=phi))
    drv.walker = [SD(wf=wf0[k], logw=lw0[k]) for k in range(len(sample))]
    drv.walkerIsAlive = [bool(a) for a in al0]
    stepped = 0
    for k, i in enumerate(sample):
        if not al0[k]:
            assert al1[k] == 0
            continue
        drv.rng.counter = int(off[i])
        nrec = len(drv.records)
        drv.step_walker(k)
        assert bool(al1[k]) == drv.walkerIsAlive[k], i
        if len(drv.records) == nrec:
            continue
        _, f_ref, aux_ref = drv.records[-1]
        assert np.max(np.abs(f_dev[i] - f_ref)) < tol, i
        check_fields(aux_dev[i], aux_ref, u[off[i]:off[i] + L], f_ref, precision)
        ref = drv.walker[k]
        assert np.max(np.abs(wf1[k] - ref.wf)) < tol * np.max(np.abs(ref.wf)), i
        close_log(lw1[k], ref.logw, tol)
        stepped += 1
    assert stepped >= 6
    # QR of every walker: the sampled ones keep the oracle's invariants
    prop.stabilize(ws)
    wf2, lw2, al2 = ws.download_some(sample)
    for k in range(len(sample)):
        if not al1[k]:
            continue
        ref = SD(wf=wf1[k], logw=lw1[k])
        ref.stabilize()
        close_log(lw2[k], ref.logw, tol)
        D = wf2[k].conj().T @ ref.wf
        assert np.max(np.abs(np.abs(np.diag(D)) - 1)) < (1e-10 if precision == 0 else 1e-5)
    # one measurement of the sampled walkers (the others masked), kernels at full size
    mask = np.zeros(nw, dtype=np.int32)
    mask[sample] = al2
    ws.set_alive(mask)
    prop.reset_measure()
    prop.measure(ws)
    e = prop.energy()
    drv.walker = [SD(wf=wf2[k], logw=lw2[k]) for k in range(len(sample))]
    drv.walkerIsAlive = [bool(a) for a in al2]
    drv.mixedMeasure.reSet()
    drv.addMixedMeasurement()
    mm = drv.mixedMeasure
    for key, r in [("den", mm.den), ("HNum", mm.HNum), ("KNum", mm.KNum), ("VNum", mm.VNum)]:
        assert abs(e[key] - r) < tol * max(1.0, abs(r)), (key, e[key], r)
    ws.close()
    prop.close()


# ------------------------------------------------------------------ phaseless L = 200, G = 1600, 1024 walkers
def c3_workload():
    if "c3" not in _CACHE:
        om = synthetic_molecule(L=200, Nup=40, Ndn=40, G=1600, seed=12345)
        dm = MoleculeModel(200, 40, 40, om.t, om.K, om.choleskyVecs, om.choleskyBg)
        _CACHE["c3"] = (om, dm)
    return _CACHE["c3"]


@pytest.mark.parametrize("precision", [0, 1])
def test_phaseless_l200_step_comb_measure_match_oracle(precision):
    om, dm = c3_workload()
    L, N, G, nw, dt, ET = 200, 40, 1600, 1024, 0.005, -150.0
    tol = tol_of(precision)
    ref = OraclePh(OraclePhMethod(dt=dt, walkerSizePerThread=nw, ET=ET), om, ListStream(np.zeros(0)))
    method = PhaselessMethod(dt=dt, walkerSizePerThread=nw, ET=ET)
    prop = PhaselessPropagator(ctx(), dm, method, phiT=(ref.phiT.wfUp, ref.phiT.wfDn), precision=precision,
                               max_walkers=nw)
    B, lwB = dm.exp_minus_alpha_k(dt)
    prop.set_onebody(0, B, lwB, False)
    B2, lwB2 = dm.exp_minus_alpha_k(-0.5 * dt)
    prop.set_onebody(2, B2, lwB2, False)
    prop.set_ET(ET)
    ws = prop.walkers(nw)
    ws.set_from_trial(prop.phiT)
    prop.set_seed(21, 0)
    for _ in range(2):  # distinct walkers
        prop.step(ws)
    wf0, lw0, al0 = ws.download()
    logO = prop.log_overlap(ws)
    rng = np.random.default_rng(8)
    u = rng.uniform(0, 1, 2 * G * nw)
    prop.set_uniforms(u)
    print(f"[c3] precision={precision}: step of {nw} walkers with host uniforms", flush=True)
    f_dev, aux_dev = prop.step(ws, record=True)
    wf1, lw1, al1 = ws.download()
    off = uniform_offsets(al0, logO, 2 * G)
    ref.walker = [SD2is(wf0[i], N, N, lw0[i]) for i in range(nw)]
    ref.walkerIsAlive = [bool(a) for a in al0]
    ref.rng = ListStream(u)
    meas = RealMaterialMoleculeMeasureFixedSD2sSD2is(om, ref.phiT)
    hs = om.returnExpMinusAlphaV(dt)
    stepped = 0
    for i in (0, 1, 333, 777, 1023):
        if not al0[i]:
            continue
        op = SD2sSD2isOperation(ref.phiT, SD2is(wf0[i], N, N, lw0[i]))
        c0 = int(off[i])
        ref.rng.counter = c0
        ref.step_walker(i)
        assert bool(al1[i]) == ref.walkerIsAlive[i], i
        if ref.rng.counter == c0:  # killed by the overlap-phase check
            continue
        f_ref = meas.getForce(hs, op, method.forceCap)
        assert np.max(np.abs(f_dev[i] - f_ref)) < tol * max(1.0, np.max(np.abs(f_ref))), i
        gauss = gaussians_from_uniforms(u[c0:c0 + 2 * G])
        assert np.max(np.abs(aux_dev[i] - (gauss + f_dev[i]))) < 1e-12
        if ref.walkerIsAlive[i]:
            r = ref.walker[i]
            assert np.max(np.abs(wf1[i] - r.wf)) < tol * max(1.0, np.max(np.abs(r.wf))), i
            close_log(lw1[i], r.logw, tol)
            stepped += 1
    assert stepped >= 3
    # QR and comb over all walkers (the comb needs every weight: the oracle overlaps all 1024)
    prop.stabilize(ws)
    u2 = rng.uniform(0, 1, nw)
    prop.set_uniforms(u2)
    wt, tb, st = prop.pop_control(ws, record=True)
    ref.walker = [SD2is(wf1[i], N, N, lw1[i]) for i in range(nw)]
    ref.walkerIsAlive = [bool(a) for a in al1]
    ref.modifyGM()
    ref.rng = ListStream(u2)
    wref, tref = ref.popControl()
    assert np.allclose(wt, wref, rtol=tol, atol=0)
    assert np.array_equal(tb, ocomb.popConfiguration(1, wt, u2))
    if precision == 0:
        assert np.array_equal(tb, tref)
    # one measurement of walkers on both sides of the exchange chunk boundaries
    es = 16 if precision == 0 else 8
    dchunk = max(1, min(nw, (4 << 30) // (G * N * N * es)))
    assert dchunk < nw
    msample = sorted({0, dchunk - 1, dchunk, 2 * dchunk + 3, nw - 1})
    wf2, lw2, al2 = ws.download()
    mask = np.zeros(nw, dtype=np.int32)
    mask[msample] = 1
    ws.upload(list(wf2), lw2, mask)
    print(f"[c3] precision={precision}: measurement, exchange chunks of {dchunk} walkers", flush=True)
    prop.reset_measure()
    prop.measure(ws)
    e = prop.energy()
    ref.walker = [SD2is(wf2[i], N, N, lw2[i]) for i in range(nw)]
    ref.walkerIsAlive = [bool(a) for a in mask]
    ref.mixedMeasure.reSet()
    ref.addMixedMeasurement()
    mm = ref.mixedMeasure
    for key, r in [("den", mm.den), ("HNum", mm.HNum), ("TNum", mm.TNum)]:
        assert abs(e[key] - r) < tol * max(1.0, abs(r)), (key, e[key], r)
    assert np.max(np.abs(e["choleskyBgNum"] - mm.choleskyBgNum)) < tol * max(1.0, np.max(np.abs(mm.choleskyBgNum)))
    assert np.max(np.abs(e["choleskyExNum"] - mm.choleskyExNum)) < tol * max(1.0, np.max(np.abs(mm.choleskyExNum)))
    ws.close()
    prop.close()


# ------------------------------------------------------------------ 24 heterogeneous lattices x 256 walkers
def hetero_sets(precision, nw, ctxs):
    out = []
    for c, (name, model, phi, nup) in zip(ctxs, heterogeneous_hubbard_instances()):
        m = CPMethod(dt=0.01, walkerSizePerThread=nw, decompType="densitySpin", ET=-1.0)
        prop = CPPropagator(c, model, m, phi, nup=nup, spin_blocks=True, precision=precision, max_walkers=nw)
        prop.set_ET(-1.0)
        ws = prop.walkers(nw)
        ws.set_from_trial()
        out.append((name, model, phi, prop, ws))
    return out


@pytest.mark.parametrize("precision", [0, 1])
def test_heterogeneous_batch_graph_step_matches_oracle(precision):
    nw = 256
    tol = tol_of(precision)
    origin = Context(0)
    ctxs = [Context(0) for _ in range(24)]
    sets = hetero_sets(precision, nw, ctxs)
    rng = np.random.default_rng(13)
    us = []
    for name, model, phi, prop, ws in sets:
        u = rng.uniform(0, 1, model.L * nw)
        prop.set_uniforms(u)
        us.append(u)
    origin.capture_begin(members=ctxs)
    for name, model, phi, prop, ws in sets:
        prop.step(ws)
    g = origin.capture_end()
    print(f"[c4] precision={precision}: one graph step of 24 instances ({g.nodes()} nodes)", flush=True)
    g.launch()
    origin.synchronize()
    for c in ctxs:
        c.synchronize()
    for j, (name, model, phi, prop, ws) in enumerate(sets):
        k = (37 * j + 5) % nw
        wf, lw, al = ws.download_some([k])
        om = HubbardSOC(model.L, model.N, model.K, U=model.U)
        drv = OracleCP(OracleMethod(dt=0.01, decompType="densitySpin", walkerSizePerThread=1, ET=-1.0,
                                    thermalSize=0, ETAdjustMaxSize=0), om, ListStream(us[j]), phiT=SD(wf=phi))
        drv.walker = [SD(wf=phi.copy())]
        drv.walkerIsAlive = [True]
        drv.rng.counter = k * model.L  # every walker starts at the trial and passes the phase check
        drv.step_walker(0)
        assert al[0] == 1 and drv.walkerIsAlive[0], name
        ref = drv.walker[0]
        assert np.max(np.abs(wf[0] - ref.wf)) < tol * np.max(np.abs(ref.wf)), name
        close_log(lw[0], ref.logw, tol)
    g.close()
    for *_, prop, ws in sets:
        ws.close()
        prop.close()


def test_heterogeneous_batch_graph_replays_equal_eager_calls():
    """10 steps + QR + comb of all 24 instances as one graph (one branch per lattice), replayed
    twice with eager work between the replays, against the same calls made eagerly on identical
    propagators: log-overlaps and walkers identical."""
    nw = 256
    origin = Context(0)
    ctxs = [Context(0) for _ in range(24)]
    gsets = hetero_sets(1, nw, ctxs)
    esets = hetero_sets(1, nw, [ctx()] * 24)
    for j, (a, b) in enumerate(zip(gsets, esets)):
        a[3].set_seed(100 + j, 0)
        b[3].set_seed(100 + j, 0)

    def block(sets):
        for *_, prop, ws in sets:
            for _ in range(10):
                prop.step(ws)
            prop.stabilize(ws)
            prop.pop_control(ws)

    origin.capture_begin(members=ctxs)
    block(gsets)
    g = origin.capture_end()
    for rep in range(2):
        g.launch()
        origin.synchronize()
        for c in ctxs:
            c.synchronize()
        block(esets)
        for a, b in zip(gsets, esets):
            la, lb = a[3].log_overlap(a[4]), b[3].log_overlap(b[4])
            assert np.array_equal(la, lb), (rep, a[0])
    for a, b in zip(gsets, esets):
        wa, lwa, ala = a[4].download_some([0, 100, 255])
        wb, lwb, alb = b[4].download_some([0, 100, 255])
        assert np.array_equal(wa, wb) and np.array_equal(lwa, lwb) and np.array_equal(ala, alb)
    g.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_grouped_step_matches_oracle(precision):
    """One grouped step of the 24 instances (every kernel of a phase launched once over all of
    them) against the oracle's step of one sampled walker per instance."""
    nw = 256
    tol = tol_of(precision)
    c = Context(0)
    sets = hetero_sets(precision, nw, [c] * 24)
    rng = np.random.default_rng(29)
    us = []
    for name, model, phi, prop, ws in sets:
        u = rng.uniform(0, 1, model.L * nw)
        prop.set_uniforms(u)
        us.append(u)
    grp = CPGroup([(prop, ws) for *_, prop, ws in sets])
    print(f"[c4 grouped] precision={precision}: one grouped step", flush=True)
    grp.step()
    c.synchronize()
    for j, (name, model, phi, prop, ws) in enumerate(sets):
        k = (53 * j + 11) % nw
        wf, lw, al = ws.download_some([k])
        om = HubbardSOC(model.L, model.N, model.K, U=model.U)
        drv = OracleCP(OracleMethod(dt=0.01, decompType="densitySpin", walkerSizePerThread=1, ET=-1.0,
                                    thermalSize=0, ETAdjustMaxSize=0), om, ListStream(us[j]), phiT=SD(wf=phi))
        drv.walker = [SD(wf=phi.copy())]
        drv.walkerIsAlive = [True]
        drv.rng.counter = k * model.L
        drv.step_walker(0)
        assert al[0] == 1 and drv.walkerIsAlive[0], name
        ref = drv.walker[0]
        assert np.max(np.abs(wf[0] - ref.wf)) < tol * np.max(np.abs(ref.wf)), name
        close_log(lw[0], ref.logw, tol)
    grp.close()
    for *_, prop, ws in sets:
        ws.close()
        prop.close()


@pytest.mark.parametrize("own_tiles", [False, True])
@pytest.mark.parametrize("precision", [0, 1])
def test_grouped_graph_equals_per_instance_calls(precision, own_tiles):
    """Blocks of 10 grouped steps + grouped QR (+ grouped comb every other block), captured as
    graphs and replayed with eager per-instance work between the replays, against the same
    per-instance calls made eagerly on identical propagators: log-overlaps, walkers, weights and
    alive flags identical.  With own_tiles the members keep different GEMM tile shapes of their
    own (group tiles -1: one grouped launch per shape in use).  The block without the comb stays
    within 400 graph nodes (600 with four shapes per GEMM)."""
    nw = 256
    c = Context(0)
    gsets = hetero_sets(precision, nw, [c] * 24)
    esets = hetero_sets(precision, nw, [ctx()] * 24)
    green_shapes = [0, 1, 4, 5]
    for j, (a, b) in enumerate(zip(gsets, esets)):
        a[3].set_seed(300 + j, 0)
        b[3].set_seed(300 + j, 0)
        if own_tiles:
            for s_ in (a, b):
                s_[3].set_tiles(j % 4, (j + 1) % 4, green_shapes[j % 4])
    grp = CPGroup([(prop, ws) for *_, prop, ws in gsets], tiles=(-1, -1, -1) if own_tiles else None)

    def eager_block(sets, comb):
        for *_, prop, ws in sets:
            for _ in range(10):
                prop.step(ws)
            prop.stabilize(ws)
            if comb:
                prop.pop_control(ws)

    def group_block(comb):
        for _ in range(10):
            grp.step()
        grp.stabilize()
        if comb:
            grp.pop_control()

    # eager grouped calls first (they also create the argument tables), then graphs of them
    group_block(True)
    eager_block(esets, True)
    graphs = {}
    for comb in (False, True):
        c.capture_begin()
        group_block(comb)
        graphs[comb] = c.capture_end()
    nodes = graphs[False].nodes()
    print(f"[c4 grouped] precision={precision} own_tiles={own_tiles}: {nodes} graph nodes per 10-step "
          f"block, {grp.table_count()} argument tables", flush=True)
    assert nodes <= (600 if own_tiles else 400)
    for rep in range(4):
        comb = rep % 2 == 1
        graphs[comb].launch()
        c.synchronize()
        eager_block(esets, comb)
        for a, b in zip(gsets, esets):
            la, lb = a[3].log_overlap(a[4]), b[3].log_overlap(b[4])
            assert np.array_equal(la, lb), (rep, a[0])
    for a, b in zip(gsets, esets):
        wa, lwa, ala = a[4].download_some([0, 100, 255])
        wb, lwb, alb = b[4].download_some([0, 100, 255])
        assert np.array_equal(wa, wb) and np.array_equal(lwa, lwb) and np.array_equal(ala, alb), a[0]
    for g in graphs.values():
        g.close()
    grp.close()
    for s in (gsets, esets):
        for *_, prop, ws in s:
            ws.close()
            prop.close()
