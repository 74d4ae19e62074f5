"""Differentiation of the fixed-field constraint-path segment on the GPU.

The device forward energy equals the CPU oracle's (oracle/fixed_field.py) to 1e-10 relative;
the device adjoint gradient (dE/dt per nearest / next-nearest bond and dE/dU) agrees with the
oracle's FP64 adjoint gradient to 1e-9 and with central finite differences (h = 1e-5) of the
oracle's energy on a small lattice and of the device energy on the 8x8 workload, to 1e-6
relative; the torch.autograd.Function harness returns the same gradient."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from afqmclab_amd.adjoint import FixedFieldEnergy, FixedFieldHubbard  # noqa: E402
from oracle import fixed_field  # noqa: E402
from tests.gpu_util import ctx  # noqa: E402

H = 1e-5


def small():
    return FixedFieldHubbard(ctx(), 4, 4, 5, 5, t1=1.0, t2=-0.25, U=4.0, dt=0.05, nw=6, nsteps=20, mgs=5, ckpt=10,
                             field_seed=77)


def oracle_energy(eng, t, U):
    L = eng.L
    return fixed_field.energy(eng.hopping(t), U, eng.phi[:L, :eng.nup], eng.phi[L:, eng.nup:], eng.nw, eng.nsteps,
                              eng.mgs, eng.dt, eng.seed)


def oracle_grad(eng, t, U):
    L = eng.L
    E, kbar, ubar = fixed_field.energy_and_grad(eng.hopping(t), U, eng.phi[:L, :eng.nup], eng.phi[L:, eng.nup:],
                                                eng.nw, eng.nsteps, eng.mgs, eng.dt, eng.seed)
    gt = np.array([-(kbar[i, j] + kbar[j, i]).real for i, j, _ in eng.bonds])
    return E, gt, float(np.sum(ubar).real)


def close(a, b, rtol):
    scale = max(1.0, np.max(np.abs(b)))
    assert np.max(np.abs(np.asarray(a) - np.asarray(b))) <= rtol * scale, (a, b)


def test_fixed_field_energy_matches_oracle():
    eng = small()
    E = eng.energy(eng.t0, eng.U0)
    ref = oracle_energy(eng, eng.t0, eng.U0)
    assert abs(E - ref) <= 1e-10 * abs(ref), (E, ref)
    # a perturbed Hamiltonian too (the trial stays fixed)
    t = eng.t0 * (1.0 + 0.05 * np.cos(np.arange(eng.t0.size)))
    E2 = eng.energy(t, 3.3)
    ref2 = oracle_energy(eng, t, 3.3)
    assert abs(E2 - ref2) <= 1e-10 * abs(ref2)


def test_adjoint_gradient_matches_oracle_finite_differences():
    eng = small()
    E, gt, gU = eng.energy_and_grad(eng.t0, eng.U0)
    assert abs(E - oracle_energy(eng, eng.t0, eng.U0)) <= 1e-10 * abs(E)
    fd = np.zeros_like(gt)
    for b in range(gt.size):
        tp, tm = eng.t0.copy(), eng.t0.copy()
        tp[b] += H
        tm[b] -= H
        fd[b] = (oracle_energy(eng, tp, eng.U0) - oracle_energy(eng, tm, eng.U0)).real / (2 * H)
    fdU = (oracle_energy(eng, eng.t0, eng.U0 + H) - oracle_energy(eng, eng.t0, eng.U0 - H)).real / (2 * H)
    close(gt, fd, 1e-6)
    assert abs(gU - fdU) <= 1e-6 * max(1.0, abs(fdU)), (gU, fdU)
    # the FP64 reference gradient (the oracle's own adjoint)
    Eo, gto, gUo = oracle_grad(eng, eng.t0, eng.U0)
    close(gt, gto, 1e-9)
    assert abs(gU - gUo) <= 1e-9 * max(1.0, abs(gUo))


def test_adjoint_gradient_8x8_matches_device_finite_differences():
    """8x8 periodic, t = 1, t' = -0.25, U = 6, 28 + 28, 512 walkers, 400 steps, checkpoints every
    50; finite differences of the device energy for U and a sample of bonds."""
    eng = FixedFieldHubbard(ctx(), 8, 8, 28, 28, t1=1.0, t2=-0.25, U=6.0, dt=0.01, nw=512, nsteps=400, mgs=10,
                            ckpt=50, field_seed=2025)
    E, gt, gU = eng.energy_and_grad(eng.t0, eng.U0)
    print(f"[adjoint] E = {E.real:.10f}, |dE/dt| = {np.linalg.norm(gt):.6f}, dE/dU = {gU:.8f}", flush=True)
    fdU = (eng.energy(eng.t0, eng.U0 + H) - eng.energy(eng.t0, eng.U0 - H)).real / (2 * H)
    assert abs(gU - fdU) <= 1e-6 * max(1.0, abs(fdU)), (gU, fdU)
    for b in (0, 37, 128, 200):
        tp, tm = eng.t0.copy(), eng.t0.copy()
        tp[b] += H
        tm[b] -= H
        fd = (eng.energy(tp, eng.U0) - eng.energy(tm, eng.U0)).real / (2 * H)
        assert abs(gt[b] - fd) <= 1e-6 * max(1.0, abs(fd)), (b, gt[b], fd)


def test_torch_autograd_harness():
    torch = pytest.importorskip("torch")
    eng = small()
    t = torch.tensor(eng.t0, dtype=torch.float64, requires_grad=True)
    U = torch.tensor(eng.U0, dtype=torch.float64, requires_grad=True)
    E = FixedFieldEnergy.apply(t, U, eng)
    E.backward()
    _, gt, gU = eng.energy_and_grad(eng.t0, eng.U0)
    assert np.allclose(t.grad.numpy(), gt, rtol=0, atol=0)
    assert float(U.grad) == gU
