"""Gradients of the constraint-path mixed-estimator energy for a fixed auxiliary-field sequence.

``FixedFieldHubbard`` runs the fixed-field segment on the GPU (afqmc_cp_fixed_field_energy /
afqmc_cp_fixed_field_adjoint: forward propagation, checkpointed reverse sweep) for a periodic
Hubbard lattice whose hopping amplitudes are given per bond (nearest and next-nearest
neighbours) and a uniform U, with the trial held fixed.  The device adjoint returns dE/dB for
B = exp(-dt K), the explicit dE/dK of the kinetic term and dE/dU_i; the host turns dE/dB into
dE/dK with the derivative of the matrix exponential on K's eigenbasis (Daleckii-Krein), then maps
dE/dK onto the bonds.  ``FixedFieldEnergy`` puts this behind torch.autograd.Function (forward:
the energy; backward: the adjoint run), a thin harness for gradient checks.
"""
import ctypes as C

import numpy as np

from ._lib import check, dptr, lib
from .cp import CPMethod, CPPropagator
from .model import HubbardModel, spin_block_trial


def lattice_bonds(nx, ny, nnn=True):
    """Bonds (i, j, kind) of a periodic nx x ny lattice, each unordered pair once: kind 0 nearest
    neighbours (x+1, y) and (x, y+1), kind 1 next-nearest (x+1, y+1) and (x+1, y-1).  Site
    index x + y*nx, as the lattice hopping generator (model.square_lattice_hopping)."""
    if nx < 3 or ny < 3:
        raise ValueError("bond amplitudes need nx, ny >= 3 (no doubly counted periodic bonds)")
    idx = lambda x, y: (x % nx) + (y % ny) * nx  # noqa: E731
    out = []
    for y in range(ny):
        for x in range(nx):
            out.append((idx(x, y), idx(x + 1, y), 0))
            out.append((idx(x, y), idx(x, y + 1), 0))
    if nnn:
        for y in range(ny):
            for x in range(nx):
                out.append((idx(x, y), idx(x + 1, y + 1), 1))
                out.append((idx(x, y), idx(x + 1, y - 1), 1))
    return out


def expm_derivative_adjoint(k, dt, bbar):
    """dE/dK from dE/dB for B = exp(-dt K), K real symmetric: -dt V ((V^T Bbar V) o F) V^T with
    F_ab the divided differences of exp at mu = -dt lambda."""
    lam, V = np.linalg.eigh(k)
    mu = -dt * lam
    e = np.exp(mu)
    d = mu[:, None] - mu[None, :]
    with np.errstate(divide="ignore", invalid="ignore"):
        F = np.where(np.abs(d) > 1e-12, (e[:, None] - e[None, :]) / np.where(np.abs(d) > 1e-12, d, 1.0),
                     0.5 * (e[:, None] + e[None, :]))
    return -dt * (V @ ((V.T @ bbar @ V) * F) @ V.T)


class FixedFieldHubbard:
    """Fixed-field CP segment on nx x ny periodic Hubbard: nup + ndn electrons, nw walkers, nsteps
    steps of size dt, CholeskyQR every mgs steps, checkpoints every ckpt steps in the adjoint."""

    def __init__(self, ctx, nx, ny, nup, ndn, t1=1.0, t2=0.0, U=4.0, dt=0.01, nw=64, nsteps=100, mgs=10, ckpt=50,
                 field_seed=12345):
        self.ctx = ctx
        self.nx, self.ny, self.nup, self.ndn = nx, ny, nup, ndn
        self.L = nx * ny
        self.dt, self.nw, self.nsteps, self.mgs, self.ckpt, self.seed = dt, nw, nsteps, mgs, ckpt, field_seed
        self.bonds = lattice_bonds(nx, ny, nnn=True)
        self.t0 = np.array([t1 if kind == 0 else t2 for _, _, kind in self.bonds])
        self.U0 = float(U)
        base = HubbardModel.square(nx, ny, nup + ndn, U, t1, t2)
        self.phi = spin_block_trial(base, nup, ndn)  # held fixed: not a function of the parameters

    def hopping(self, t):
        k = np.zeros((self.L, self.L))
        for (i, j, _), tb in zip(self.bonds, t):
            k[i, j] += -tb
            k[j, i] += -tb
        return k

    def _setup(self, t, U):
        L = self.L
        k = self.hopping(t)
        K = np.zeros((2 * L, 2 * L), dtype=np.complex128)
        K[:L, :L] = k
        K[L:, L:] = k
        model = HubbardModel(L=L, N=self.nup + self.ndn, K=K, U=np.full(L, float(U)))
        m = CPMethod(dt=self.dt, walkerSizePerThread=self.nw, decompType="densitySpin", ET=0.0)
        prop = CPPropagator(self.ctx, model, m, self.phi, nup=self.nup, spin_blocks=True, precision=0,
                            max_walkers=self.nw)
        ws = prop.walkers(self.nw)
        ws.set_from_trial(self.phi)
        return k, prop, ws

    def energy(self, t, U):
        k, prop, ws = self._setup(t, U)
        e = np.zeros(2)
        check(lib().afqmc_cp_fixed_field_energy(prop.h, ws.h, self.nsteps, self.mgs, C.c_uint64(self.seed), dptr(e)))
        ws.close()
        prop.close()
        return complex(e[0], e[1])

    def energy_and_grad(self, t, U):
        """E, dE/dt (per bond, real), dE/dU (uniform U, real)."""
        L = self.L
        k, prop, ws = self._setup(t, U)
        e = np.zeros(2)
        bbar = np.zeros(2 * 2 * L * L)
        mbar = np.zeros(2 * 2 * L * L)
        ubar = np.zeros(2 * L)
        check(lib().afqmc_cp_fixed_field_adjoint(prop.h, ws.h, self.nsteps, self.mgs, self.ckpt, C.c_uint64(self.seed),
                                                 dptr(e), dptr(bbar), dptr(mbar), dptr(ubar)))
        ws.close()
        prop.close()
        bb = (bbar[0::2] + 1j * bbar[1::2]).reshape(2, L, L).transpose(0, 2, 1)  # column-major blocks
        mb = (mbar[0::2] + 1j * mbar[1::2]).reshape(2, L, L).transpose(0, 2, 1)
        dK = np.zeros((L, L), dtype=np.complex128)
        for b in range(2):
            dK += expm_derivative_adjoint(k, self.dt, bb[b]) + mb[b]
        gt = np.array([-(dK[i, j] + dK[j, i]).real for i, j, _ in self.bonds])
        gU = float(np.sum(ubar[0::2]))
        return complex(e[0], e[1]), gt, gU


try:
    import torch

    class FixedFieldEnergy(torch.autograd.Function):
        """E(t, U) of a FixedFieldHubbard engine; backward runs the device adjoint."""

        @staticmethod
        def forward(ctx, t, U, engine):
            ctx.engine = engine
            ctx.save_for_backward(t, U)
            E = engine.energy(t.detach().cpu().numpy(), float(U.detach().cpu()))
            return torch.tensor(E.real, dtype=torch.float64)

        @staticmethod
        def backward(ctx, grad_out):
            t, U = ctx.saved_tensors
            _, gt, gU = ctx.engine.energy_and_grad(t.detach().cpu().numpy(), float(U.detach().cpu()))
            return (grad_out * torch.from_numpy(gt).to(t.dtype), grad_out * torch.tensor(gU, dtype=U.dtype), None)
except ImportError:  # torch is the harness only; the engine works without it
    FixedFieldEnergy = None
