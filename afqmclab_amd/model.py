"""Host-side model set-up (done once, before the projection loop).

The reference also builds these on the host at start-up: the lattice
hopping matrix K (scripts/supercubic/setHoping.py semantics), the one-body
propagators exp(-alpha K) by eigendecomposition (HubbardSOC::returnExpMinusAlphaK,
HubbardSOC.cpp:97-118), and the trial wave function (fillWalkerFromModel,
HubbardSOCSDOperation.cpp:19-30, or a UHF solution read from file).
"""
from dataclasses import dataclass, field

import numpy as np


def square_lattice_hopping(nx, ny, t1=1.0, t2=0.0):
    """Ls x Ls hopping of a periodic nx x ny lattice, site = x + y*nx.

    -t1 to the four nearest neighbours and -t2 to the four next-nearest,
    accumulated with += (a 2-wide dimension double counts, as the
    reference's generator does)."""
    Ls = nx * ny
    k = np.zeros((Ls, Ls), dtype=np.complex128)
    for y in range(ny):
        for x in range(nx):
            i = x + y * nx
            for d in (-1, 1):
                k[i, (x + d) % nx + y * nx] += -t1
                k[i, x + ((y + d) % ny) * nx] += -t1
            if t2 != 0.0:
                for dx in (-1, 1):
                    for dy in (-1, 1):
                        k[i, (x + dx) % nx + ((y + dy) % ny) * nx] += -t2
    return k


@dataclass
class HubbardModel:
    """HubbardSOC: L sites, N particles, 2L x 2L spin-orbit K, on-site fields."""
    L: int
    N: int
    K: np.ndarray
    U: np.ndarray
    mu: np.ndarray = None
    hx: np.ndarray = None
    hy: np.ndarray = None
    hz: np.ndarray = None
    _eig: tuple = field(default=None, repr=False)

    def __post_init__(self):
        z = np.zeros(self.L)
        self.K = np.asfortranarray(np.asarray(self.K, dtype=np.complex128))
        self.U = np.asarray(self.U, dtype=np.float64) * np.ones(self.L)
        for name in ("mu", "hx", "hy", "hz"):
            v = getattr(self, name)
            setattr(self, name, z.copy() if v is None else np.asarray(v, dtype=np.float64) * np.ones(self.L))

    @classmethod
    def square(cls, nx, ny, N, U, t1=1.0, t2=0.0):
        k = square_lattice_hopping(nx, ny, t1, t2)
        Ls = nx * ny
        K = np.zeros((2 * Ls, 2 * Ls), dtype=np.complex128, order="F")
        K[:Ls, :Ls] = k
        K[Ls:, Ls:] = k
        return cls(Ls, N, K, U)

    def eig(self):
        if self._eig is None:
            if not np.allclose(self.K, self.K.conj().T, atol=1e-12):
                raise ValueError("K is not Hermitian")
            self._eig = np.linalg.eigh(self.K)
        return self._eig

    def exp_minus_alpha_k(self, alpha):
        w, v = self.eig()
        return np.asfortranarray(v @ (np.exp(-alpha * w)[:, None] * v.conj().T))

    def spin_block_hopping(self):
        L = self.L
        return self.K[:L, :L], self.K[L:, L:]

    def is_spin_block(self):
        L = self.L
        return (np.all(self.K[:L, L:] == 0) and np.all(self.K[L:, :L] == 0)
                and np.all(self.hx == 0) and np.all(self.hy == 0))


def free_electron_trial(model: HubbardModel):
    """Lowest N eigenvectors of the 2L x 2L K (fillWalkerFromModel)."""
    w, v = model.eig()
    return np.asfortranarray(v[:, :model.N])


def spin_block_trial(model: HubbardModel, nup, ndn):
    """Closed-shell free-electron trial built per spin: lowest nup / ndn
    eigenvectors of the up and down hopping blocks, laid out as a 2L x N SD
    with the up block in rows 0..L-1 x columns 0..nup-1."""
    L = model.L
    ku, kd = model.spin_block_hopping()
    phi = np.zeros((2 * L, nup + ndn), dtype=np.complex128, order="F")
    wu, vu = np.linalg.eigh(ku)
    wd, vd = np.linalg.eigh(kd)
    phi[:L, :nup] = vu[:, :nup]
    phi[L:, nup:] = vd[:, :ndn]
    return phi


def uhf_trial(model: HubbardModel, nup, ndn, iters=300, mix=0.5, seed_amp=0.3, tol=1e-10):
    """Unrestricted Hartree-Fock trial of a spin-block Hubbard model.

    Self-consistent mean field H_s = K_s + U n_{-s} (diagonal), seeded with
    a checkerboard (zafm) density n_up,i = n/2 + a(-1)^(x+y), n_dn = n/2 - ...
    as the reference's initialOrderParameter.py:23-29 seeds its GHF/UHF.
    Returns the 2L x N spin-block SD of the converged occupied orbitals."""
    L = model.L
    nx = int(round(np.sqrt(L)))
    ku, kd = model.spin_block_hopping()
    n = (nup + ndn) / L
    sign = np.array([(-1) ** ((i % nx) + (i // nx)) for i in range(L)], dtype=np.float64)
    nu = n / 2 + seed_amp * sign
    nd = n / 2 - seed_amp * sign
    for _ in range(iters):
        hu = ku + np.diag(model.U * nd)
        hd = kd + np.diag(model.U * nu)
        wu, vu = np.linalg.eigh(hu)
        wd, vd = np.linalg.eigh(hd)
        nu_new = np.sum(np.abs(vu[:, :nup]) ** 2, axis=1)
        nd_new = np.sum(np.abs(vd[:, :ndn]) ** 2, axis=1)
        diff = max(np.max(np.abs(nu_new - nu)), np.max(np.abs(nd_new - nd)))
        nu = mix * nu_new + (1 - mix) * nu
        nd = mix * nd_new + (1 - mix) * nd
        if diff < tol:
            break
    phi = np.zeros((2 * L, nup + ndn), dtype=np.complex128, order="F")
    phi[:L, :nup] = vu[:, :nup]
    phi[L:, nup:] = vd[:, :ndn]
    return phi
