"""Back-propagation on the GPU against the oracle's back-propagation driver, through the C-ABI.

The oracle moves recorded fields and backups with every comb table and propagates backups again, as
the reference does; the device keeps a ring of fields and tables, copies the walkers at window
starts and follows each walker's ancestry.  Both run the same step loop on the same uniform stream,
so the numerators of every write (den, HNum, KNum, VNum, RNum and the five observable matrices)
agree to 1e-9 relative in FP64.  A U = 0 run at a larger size is checked against the closed form."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from afqmclab_amd.cp import AfqmcConstraintPathBP, BPMethod as DevBP  # noqa: E402
from afqmclab_amd.model import HubbardModel, free_electron_trial, spin_block_trial  # noqa: E402
from oracle.blocks import SD  # noqa: E402
from oracle.drivers import AfqmcConstraintPathBP as OracleBP, BPMethod as OracleBPMethod  # noqa: E402
from oracle.models import HubbardSOC  # noqa: E402
from oracle.rng import UniformStream  # noqa: E402
from tests.gpu_util import ctx  # noqa: E402

KEYS = ("den", "HNum", "KNum", "VNum", "RNum", "greenMatrixNum", "densityDensityNum", "splusSminusNum",
        "sminusSplusNum", "spairSpairNum")


def compare_writes(dev, ref, rtol):
    assert len(dev) == len(ref)
    for a, b in zip(dev, ref):
        for k in KEYS:
            x, y = np.asarray(a[k]), np.asarray(b[k])
            scale = max(np.max(np.abs(y)), abs(b["den"]) * 1e-3)
            err = np.max(np.abs(x - y))
            assert err <= rtol * scale, (k, err, scale)


def run_pair(model, phi, m, kw, seed, precision=0):
    dev = AfqmcConstraintPathBP(ctx(), model, DevBP(**m), phi, precision=precision, **kw)
    dev.run()
    om = HubbardSOC(model.L, model.N, model.K, mu=model.mu, hx=model.hx, hy=model.hy, hz=model.hz, U=model.U)
    ref = OracleBP(OracleBPMethod(**m), om, UniformStream(seed), phiT=SD(wf=phi))
    ref.run()
    assert dev.prop.counter() == ref.rng.counter
    return dev, ref


def mixing_model(seed=31, U=6.0):
    rng = np.random.default_rng(seed)
    base = HubbardModel.square(2, 2, 3, U)
    L = base.L
    return HubbardModel(L=L, N=3, K=base.K, U=np.full(L, U), mu=rng.uniform(-0.3, 0.3, L),
                        hx=rng.uniform(-0.5, 0.5, L), hy=rng.uniform(-0.5, 0.5, L), hz=rng.uniform(-0.5, 0.5, L))


@pytest.mark.parametrize("nbp,skip", [(6, 4), (4, 6), (5, 5), (0, 3)])
def test_bp_spin_mixing_matches_oracle(nbp, skip):
    """Dense 2L x N walkers with hx, hy != 0 (2x2 site matrices), densityCharge (complex fields, so
    walkers die and the comb copies), comb every 2 steps."""
    model = mixing_model()
    phi = free_electron_trial(model)
    seed = 123
    m = dict(dt=0.08, thermalSize=6, writeNumber=2, measureNumberPerWrite=2, measureSkipStep=skip,
             walkerSizePerThread=10, decompType="densityCharge", forceType="dynamicForce", mgsStep=3,
             popControlStep=2, ET=-3.0, ETAdjustStep=3, ETAdjustMaxSize=6, backPropagationStep=nbp, seed=seed)
    dev, ref = run_pair(model, phi, m, dict(spin_blocks=False), seed)
    compare_writes(dev.writes, ref.writes, 1e-9)
    assert dev.betas == pytest.approx(ref.betas)


@pytest.mark.parametrize("decomp", ["hopCharge", "hopSpin"])
def test_bp_hop_decompositions_match_oracle(decomp):
    """The hop decompositions mix the spins at every site with non-Hermitian 2x2 site matrices, so
    the bras need their true adjoints."""
    model = mixing_model(seed=8, U=3.0)
    phi = free_electron_trial(model)
    seed = 77
    m = dict(dt=0.05, thermalSize=6, writeNumber=1, measureNumberPerWrite=3, measureSkipStep=3,
             walkerSizePerThread=8, decompType=decomp, forceType="dynamicForce", mgsStep=2,
             popControlStep=3, ET=-3.0, ETAdjustStep=3, ETAdjustMaxSize=6, backPropagationStep=5, seed=seed)
    dev, ref = run_pair(model, phi, m, dict(spin_blocks=False), seed)
    compare_writes(dev.writes, ref.writes, 1e-9)


@pytest.mark.parametrize("mode,nup,ndn", [("dense", 5, 5), ("blocks", 5, 5), ("blocks", 6, 3)])
def test_bp_hubbard_matches_oracle(mode, nup, ndn):
    """4x4 Hubbard, densitySpin, diagonal site matrices: dense walkers and spin blocks (also with
    Nup != Ndn)."""
    nx, ny, U = 4, 4, 4.0
    model = HubbardModel.square(nx, ny, nup + ndn, U)
    phi = free_electron_trial(model) if mode == "dense" else spin_block_trial(model, nup, ndn)
    kw = dict(spin_blocks=False) if mode == "dense" else dict(nup=nup, spin_blocks=True)
    seed = 2024
    m = dict(dt=0.05, thermalSize=10, writeNumber=2, measureNumberPerWrite=2, measureSkipStep=4,
             walkerSizePerThread=12, decompType="densitySpin", mgsStep=4, popControlStep=5,
             ETAdjustStep=5, ETAdjustMaxSize=10, backPropagationStep=10, seed=seed)
    dev, ref = run_pair(model, phi, m, kw, seed)
    compare_writes(dev.writes, ref.writes, 1e-9)


@pytest.mark.parametrize("precision", [1, 2])
def test_bp_mixed_precision_close_to_oracle(precision):
    """FP32 walkers and GEMMs (precision 2: the propagation GEMMs of kets and bras in split-BF16
    products): the same run to the mixed-precision bar (energies 1e-4 relative)."""
    nx, ny, nup, ndn, U = 4, 4, 5, 5, 4.0
    model = HubbardModel.square(nx, ny, nup + ndn, U)
    phi = spin_block_trial(model, nup, ndn)
    seed = 7
    m = dict(dt=0.05, thermalSize=10, writeNumber=1, measureNumberPerWrite=2, measureSkipStep=4,
             walkerSizePerThread=12, decompType="densitySpin", mgsStep=4, popControlStep=5,
             ETAdjustStep=5, ETAdjustMaxSize=10, backPropagationStep=8, seed=seed)
    dev, ref = run_pair(model, phi, m, dict(nup=nup, spin_blocks=True), seed, precision=precision)
    e_dev = dev.writes[0]["HNum"] / dev.writes[0]["den"]
    e_ref = ref.writes[0]["HNum"] / ref.writes[0]["den"]
    assert abs(e_dev - e_ref) < 1e-4 * abs(e_ref)
    g_dev = dev.writes[0]["greenMatrixNum"] / dev.writes[0]["den"]
    g_ref = ref.writes[0]["greenMatrixNum"] / ref.writes[0]["den"]
    assert np.max(np.abs(g_dev - g_ref)) < 1e-4


def green(left, right):
    return (right @ np.linalg.solve(left.conj().T @ right, left.conj().T)).T


@pytest.mark.parametrize("nbp,skip", [(30, 10), (6, 8)])
def test_bp_u0_closed_form_larger(nbp, skip):
    """U = 0, 2L = 128 dense orbitals, N = 24, 64 walkers: the back-propagated Green matrix is that of
    exp(-nbp dt K) phiT against the lowest N orbitals of K, and the Wick correlators are its
    products."""
    rng = np.random.default_rng(5)
    L, N, nw = 64, 24, 64
    M = 2 * L
    q, _ = np.linalg.qr(rng.standard_normal((M, M)) + 1j * rng.standard_normal((M, M)))
    lam = np.concatenate([np.linspace(-3.0, -2.0, N), np.linspace(1.0, 3.0, M - N)])
    K = (q * lam) @ q.conj().T
    K = 0.5 * (K + K.conj().T)
    model = HubbardModel(L=L, N=N, K=K, U=np.zeros(L))
    trial, _ = np.linalg.qr(q[:, :N] + 0.5 * q[:, N:2 * N])
    dt = 0.1
    m = DevBP(dt=dt, thermalSize=150, writeNumber=1, measureNumberPerWrite=2, measureSkipStep=skip,
              walkerSizePerThread=nw, decompType="densitySpin", mgsStep=5, popControlStep=7, ET=-3.0,
              ETAdjustStep=10, ETAdjustMaxSize=0, backPropagationStep=nbp, seed=11)
    dev = AfqmcConstraintPathBP(ctx(), model, m, trial, spin_blocks=False)
    dev.run()
    w = dev.writes[0]
    ev, vec = np.linalg.eigh(model.K)
    bra = (vec * np.exp(-nbp * dt * ev)) @ vec.conj().T @ trial
    G = green(bra, vec[:, :N])
    assert np.max(np.abs(w["greenMatrixNum"] / w["den"] - G)) < 1e-9
    d = np.diag(G)
    dd = np.outer(d, d) - G * G.T
    np.fill_diagonal(dd, d)
    assert np.max(np.abs(w["densityDensityNum"] / w["den"] - dd)) < 1e-9
    sp = G[:L, :L] * G[L:, L:] - G[:L, L:] * G[L:, :L]
    assert np.max(np.abs(w["spairSpairNum"] / w["den"] - sp)) < 1e-9


def test_bp_errors_are_reported():
    """Calls out of schedule return an error with a message instead of touching the device."""
    from afqmclab_amd._lib import AfqmcError
    from afqmclab_amd.cp import BackPropagation, CPPropagator, CPMethod
    model = HubbardModel.square(2, 2, 4, 4.0)
    phi = spin_block_trial(model, 2, 2)
    prop = CPPropagator(ctx(), model, CPMethod(dt=0.05, walkerSizePerThread=4), phi, nup=2, spin_blocks=True)
    w = prop.walkers(4)
    w.set_from_trial(phi)
    bp = BackPropagation(prop, 4, 6, 3, 2)
    with pytest.raises(AfqmcError, match="before afqmc_bp_begin"):
        bp.step(w)
    bp.begin()
    with pytest.raises(AfqmcError, match="table of the last"):
        bp.pop_control(w)
    bp.step(w)
    with pytest.raises(AfqmcError, match="recorded steps"):
        bp.measure(w)
    for _ in range(6):
        bp.step(w)
    # step 7: the window [1, 7) does not start at a multiple of measureSkipStep
    with pytest.raises(AfqmcError, match="no walker copy"):
        bp.measure(w)
    bp.step(w)
    bp.step(w)
    bp.measure(w)  # window [3, 9)
    assert bp.steps() == (9, 1)
    assert np.isfinite(bp.get()["den"])
    w5 = prop.walkers(3)
    with pytest.raises(AfqmcError):
        bp.step(w5)
