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


class ScfNotConverged(RuntimeError):
    """A self-consistent mean field that did not reach its tolerance."""


def _anderson(xs, rs, beta):
    """Anderson mixing of the iterates xs with residuals rs = F(x) - x: the step from the last
    iterate by beta times its residual, corrected along the iterate differences by the
    least-squares coefficients that minimise the extrapolated residual."""
    x, r = xs[-1], rs[-1]
    if len(xs) < 2:
        return x + beta * r
    dX = np.array([xs[i + 1] - xs[i] for i in range(len(xs) - 1)]).T
    dR = np.array([rs[i + 1] - rs[i] for i in range(len(rs) - 1)]).T
    gamma = np.linalg.lstsq(dR, r, rcond=None)[0]
    return x + beta * r - (dX + beta * dR) @ gamma


def _mix_step(x, r, xs, rs, beta, history, accelerate):
    """One mixing step: x + beta r, or (accelerate and history > 1) Anderson mixing over the
    iterates since the last restart, restarted every `history` iterates."""
    if not accelerate or history <= 1:
        return x + beta * r, [], []
    xs, rs = xs + [x], rs + [r]
    out = _anderson(xs, rs, beta)
    if len(xs) >= history:
        xs, rs = [], []
    return out, xs, rs


def zafm_densities(model, nup, ndn, seed_amp=0.3):
    """Checkerboard (z antiferromagnetic) seed densities n_up,i = n_up/L + a(-1)^(x+y),
    n_dn,i = n_dn/L - a(-1)^(x+y) on a square nx x nx lattice, site = x + y nx."""
    L = model.L
    nx = int(round(np.sqrt(L)))
    sign = np.array([(-1) ** ((i % nx) + (i // nx)) for i in range(L)], dtype=np.float64)
    return nup / L + seed_amp * sign, ndn / L - seed_amp * sign


def uhf_trial(model: HubbardModel, nup, ndn, iters=600, mix=0.5, seed_amp=0.3, tol=1e-8, return_info=False,
              history=5, accelerate_below=3e-3):
    """Unrestricted Hartree-Fock trial of a spin-block Hubbard model.

    Self-consistent mean field H_s = K_s + U n_{-s} (diagonal), seeded with a checkerboard
    (zafm) density as the reference's initialOrderParameter.py:23-29 seeds its GHF, iterated to
    max |n_out - n_in| < tol: linear mixing by `mix` while the change is above accelerate_below,
    then Anderson mixing restarted every `history` iterates (linear mixing only when history <= 1).
    Raises ScfNotConverged if `iters` iterations do not reach tol.  Returns the 2L x N spin-block SD of the occupied orbitals (and, with
    return_info, a dict with the iterations, final change and the mean-field energy
    sum(eigenvalues) - sum_i U_i n_up,i n_dn,i)."""
    L = model.L
    ku, kd = model.spin_block_hopping()
    nu, nd = zafm_densities(model, nup, ndn, seed_amp)
    xs, rs = [], []
    diff = np.inf
    for it in range(1, iters + 1):
        wu, vu = np.linalg.eigh(ku + np.diag(model.U * nd))
        wd, vd = np.linalg.eigh(kd + np.diag(model.U * nu))
        nu_new = np.sum(np.abs(vu[:, :nup]) ** 2, axis=1)
        nd_new = np.sum(np.abs(vd[:, :ndn]) ** 2, axis=1)
        x = np.concatenate([nu, nd])
        r = np.concatenate([nu_new, nd_new]) - x
        diff = np.max(np.abs(r))
        if diff < tol:
            break
        x, xs, rs = _mix_step(x, r, xs, rs, mix, history, diff < accelerate_below)
        nu, nd = x[:L], x[L:]
    if diff >= tol:
        raise ScfNotConverged(f"UHF did not converge in {iters} iterations (max density change {diff:.3e})")
    phi = np.zeros((2 * L, nup + ndn), dtype=np.complex128, order="F")
    phi[:L, :nup] = vu[:, :nup]
    phi[L:, nup:] = vd[:, :ndn]
    if return_info:
        energy = float(np.sum(wu[:nup]) + np.sum(wd[:ndn]) - np.sum(model.U * nu_new * nd_new))
        return phi, {"iterations": it, "max_density_change": float(diff), "energy": energy,
                     "n_up": nu_new, "n_dn": nd_new}
    return phi


def order_parameter_seed(model: HubbardModel, kind, rng=None):
    """Initial GHF order parameters (density, spinOrbit), 2L complex each, as the reference's
    tutorials/HartreeFockHubbardRealSpaceSOC/script/1BandHubbardModel/initialOrderParameter.py:19-50
    writes them for a square nx x nx lattice (site = x + y nx, coordinate sum x + y):
    zafm puts the mean filling N/L on spin up of even sites and spin down of odd ones; xafm and
    yafm put +-N/(2L) (real, or imaginary with opposite signs on the two halves) into spinOrbit;
    random draws uniform densities and spinOrbit(i + L) = conj(spinOrbit(i))."""
    L, N = model.L, model.N
    nx = int(round(np.sqrt(L)))
    n = N / L
    density = np.zeros(2 * L, dtype=np.complex128)
    spin_orbit = np.zeros(2 * L, dtype=np.complex128)
    for i in range(L):
        even = ((i % nx) + (i // nx)) % 2 == 0
        if kind == "zafm":
            density[i if even else i + L] = n
        elif kind == "xafm":
            spin_orbit[i] = spin_orbit[i + L] = (n / 2.0) if even else (-n / 2.0)
        elif kind == "yafm":
            spin_orbit[i] = (1j if even else -1j) * n / 2.0
            spin_orbit[i + L] = (-1j if even else 1j) * n / 2.0
        elif kind == "random":
            rng = rng or np.random.default_rng()
            density[i] = rng.random()
            density[i + L] = rng.random()
            spin_orbit[i] = rng.random() + 1j * rng.random()
            spin_orbit[i + L] = np.conj(spin_orbit[i])
        else:
            raise ValueError("unknown order parameter type " + str(kind))
    return density, spin_orbit


def ghf_trial(model: HubbardModel, density, spin_orbit, converge_type="energy", tol=1e-10, max_iter=300,
              relax=1.0, history=5, accelerate_below=3e-3):
    """Generalised Hartree-Fock of HubbardSOC (tutorials/HartreeFockHubbardRealSpaceSOC/source/ghfDetail.cpp).

    H0 = K plus the on-site -mu +- hz/2 and the (hx -+ i hy)/2 spin mixing (setH0, :9-25); the
    mean field adds U_i rho_{i+L} and U_i rho_i on the diagonal and -U_i s_{i+L}, -U_i s_i on the
    (i, i+L) and (i+L, i) entries, from density rho = <n_a> and spinOrbit s_i = <c+_i c_{i+L}>,
    s_{i+L} = <c+_{i+L} c_i> (setMeanFieldVectorsValuesFromOrderParameter, :65-80); the state is
    its lowest N eigenvectors (:82-90); the variational energy is the sum of the occupied mean-field
    eigenvalues plus the U terms of :103-121.  Each iteration relaxes the order parameters by
    `relax` (relaxOrderParameter, :123-131); once the largest change is below accelerate_below,
    Anderson mixing restarted every `history` iterates takes over (the fixed point is the same).  Convergence: converge_type "energy" (relative
    change of the variational energy, ghf.cpp:79-84) or "orderParameter" (:85-97).  Raises
    ScfNotConverged after max_iter iterations.  Returns (state 2L x N, info dict)."""
    L, N = model.L, model.N
    U = model.U
    H0 = np.array(model.K, dtype=np.complex128)
    idx = np.arange(L)
    H0[idx, idx] += -model.mu + 0.5 * model.hz
    H0[idx + L, idx + L] += -model.mu - 0.5 * model.hz
    H0[idx, idx + L] += 0.5 * (model.hx - 1j * model.hy)
    H0[idx + L, idx] += 0.5 * (model.hx + 1j * model.hy)
    rho = np.array(density, dtype=np.complex128)
    so = np.array(spin_orbit, dtype=np.complex128)

    def mean_field(rho, so):
        M = H0.copy()
        M[idx, idx] += U * rho[idx + L]
        M[idx + L, idx + L] += U * rho[idx]
        M[idx, idx + L] -= U * so[idx + L]
        M[idx + L, idx] -= U * so[idx]
        w, v = np.linalg.eigh(M)
        return w, v[:, :N]

    def order(state):
        G = state.conj() @ state.T  # G(a, b) = <c+_a c_b> of the orthonormal determinant
        return np.diag(G).copy(), np.concatenate([np.diag(G[:L, L:]), np.diag(G[L:, :L])])

    energy = None
    xs, rs = [], []
    conv = np.inf
    for it in range(1, max_iter + 1):
        w, state = mean_field(rho, so)
        rho_new, so_new = order(state)
        e = float(np.sum(w[:N]).real + np.sum(U * (-rho[idx] * rho_new[idx + L] - rho[idx + L] * rho_new[idx]
                                                   + so[idx] * so_new[idx + L] + so[idx + L] * so_new[idx]
                                                   + rho_new[idx] * rho_new[idx + L]
                                                   - so_new[idx] * so_new[idx + L])).real)
        x = np.concatenate([rho, so])
        r = np.concatenate([rho_new, so_new]) - x
        if converge_type == "energy":
            conv = np.inf if energy is None else abs(e - energy) / abs(e)
        elif converge_type == "orderParameter":
            conv = np.sum(np.abs(r)) / np.sum(np.abs(np.concatenate([rho_new, so_new])))
        else:
            raise ValueError("unknown convergeType " + str(converge_type))
        energy = e
        x, xs, rs = _mix_step(x, r, xs, rs, relax, history, np.max(np.abs(r)) < accelerate_below)
        rho, so = x[:2 * L], x[2 * L:]
        if conv < tol and np.max(np.abs(r)) < max(np.sqrt(tol), 1e-8):
            break
    else:
        raise ScfNotConverged(f"GHF did not converge in {max_iter} iterations ({converge_type} change {conv:.3e})")
    return np.asfortranarray(state), {"iterations": it, "energy": energy, "density": rho_new,
                                      "spin_orbit": so_new, "converge_value": conv}


def heterogeneous_hubbard_instances(n_each=6, seed=20241, sizes=(6, 8, 10, 12), fills=(0.75, 0.875, 1.0)):
    """A batch of independent Hubbard lattices: n_each periodic n x n lattices for each n in sizes,
    filling fills[j % 3] with the electron count rounded to the nearest even number (Nup = Ndn),
    U uniform in [2, 8] (seeded) and t' = -0.2 on every odd instance.  Returns a list of
    (name, model, closed-shell spin-block trial, Nup)."""
    rng = np.random.default_rng(seed)
    out = []
    for n in sizes:
        for j in range(n_each):
            idx = len(out)
            Ls = n * n
            ntot = int(round(fills[j % len(fills)] * Ls / 2.0)) * 2
            nup = ntot // 2
            U = float(rng.uniform(2.0, 8.0))
            t2 = -0.2 if idx % 2 == 1 else 0.0
            model = HubbardModel.square(n, n, ntot, U, 1.0, t2)
            out.append((f"{n}x{n} N={ntot} U={U:.2f} t'={t2}", model, spin_block_trial(model, nup, nup), nup))
    return out
