"""Exact diagonalisation of Hubbard-type models (test infrastructure only).

An own Lanczos code for H = sum_ij Kup_ij c+_i c_j + sum_ij Kdn_ij d+_i d_j
+ sum_i U_i n_i,up n_i,dn in the fixed (Nup, Ndn) sector, used as the
end-to-end energy oracle (the reference's lanczos/ directory plays the same
role).  The state is held as an (n_up_states x n_dn_states) matrix so that
H v = Hup V + V Hdn^T + (U-diagonal) * V.  Pinned by the literal values of
lanczos/model/realMaterial/test/realMaterialTest.cpp:242-291 (4-site twisted
ring, U=4, 2 up / 3 down: E = 0.877119735524673, <K> = -3.87084415356266,
<V> = 4.74796388908733) and by the U=0 sum of lowest eigenvalues
(realMaterialTest.cpp:293-318).
"""
from itertools import combinations

import numpy as np
import scipy.sparse as sp


def basis(L, N):
    """Bit-string occupation basis of N fermions on L sites, ascending."""
    states = []
    for occ in combinations(range(L), N):
        s = 0
        for i in occ:
            s |= 1 << i
        states.append(s)
    states.sort()
    return np.array(states, dtype=np.int64)


def one_body_matrix(K, states):
    """Sparse sum_ij K_ij c+_i c_j on the given basis (Jordan-Wigner by site index)."""
    index = {int(s): n for n, s in enumerate(states)}
    L = K.shape[0]
    rows, cols, vals = [], [], []
    nz = [(i, j, K[i, j]) for i in range(L) for j in range(L) if K[i, j] != 0]
    for col, s in enumerate(states):
        s = int(s)
        for i, j, kij in nz:
            if not (s >> j) & 1:
                continue
            if i == j:
                rows.append(col); cols.append(col); vals.append(kij)
                continue
            if (s >> i) & 1:
                continue
            # c+_i c_j |s>: sign from fermions strictly between, with c_j acting first
            t = s & ~(1 << j)
            sign = bin(t & ((1 << i) - 1)).count("1") + bin(s & ((1 << j) - 1)).count("1")
            t |= 1 << i
            rows.append(index[t]); cols.append(col); vals.append(kij * (-1.0 if sign % 2 else 1.0))
    n = len(states)
    return sp.csr_matrix((np.array(vals, dtype=np.complex128), (rows, cols)), shape=(n, n))


class HubbardED:
    def __init__(self, Kup, Kdn, U, Nup, Ndn):
        self.Kup = np.asarray(Kup, dtype=np.complex128)
        self.Kdn = np.asarray(Kdn, dtype=np.complex128)
        self.U = np.asarray(U, dtype=np.float64)
        L = self.Kup.shape[0]
        self.su = basis(L, Nup)
        self.sd = basis(L, Ndn)
        self.Hup = one_body_matrix(self.Kup, self.su)
        self.HdnT = one_body_matrix(self.Kdn, self.sd).T.tocsr()
        occ_u = ((self.su[:, None] >> np.arange(L)) & 1).astype(np.float64)
        occ_d = ((self.sd[:, None] >> np.arange(L)) & 1).astype(np.float64)
        self.D = (occ_u * self.U) @ occ_d.T  # sum_i U_i n_i,up n_i,dn
        self.real = np.allclose(self.Kup.imag, 0) and np.allclose(self.Kdn.imag, 0)
        if self.real:
            self.Hup = self.Hup.real.tocsr()
            self.HdnT = self.HdnT.real.tocsr()

    def apply_K(self, V):
        return self.Hup @ V + (self.HdnT.T @ V.T).T

    def apply_V(self, V):
        return self.D * V

    def apply(self, V):
        return self.apply_K(V) + self.apply_V(V)

    def ground_state(self, maxiter=400, tol=1e-13, seed=7):
        """Lanczos with full reorthogonalisation of the Krylov basis (small sectors)."""
        shape = (self.su.size, self.sd.size)
        rng = np.random.default_rng(seed)
        dtype = np.float64 if self.real else np.complex128
        v = rng.standard_normal(shape).astype(dtype)
        v /= np.linalg.norm(v)
        Q = [v]
        alphas, betas = [], []
        e_old = None
        for it in range(maxiter):
            w = self.apply(Q[-1])
            a = np.vdot(Q[-1], w).real
            alphas.append(a)
            for q in Q:
                w = w - np.vdot(q, w) * q
            for q in Q:
                w = w - np.vdot(q, w) * q
            b = np.linalg.norm(w)
            T = np.diag(alphas) + np.diag(betas, 1) + np.diag(betas, -1)
            ev, evec = np.linalg.eigh(T)
            # residual norm of the Ritz vector: |b * last component|
            if abs(b * evec[-1, 0]) < tol or b < 1e-14:
                break
            betas.append(b)
            Q.append(w / b)
        c = evec[:, 0]
        psi = sum(ci * q for ci, q in zip(c, Q[:len(c)]))
        psi /= np.linalg.norm(psi)
        E = np.vdot(psi, self.apply(psi)).real
        Kx = np.vdot(psi, self.apply_K(psi))
        Vx = np.vdot(psi, self.apply_V(psi))
        return E, Kx, Vx, psi

    def ground_energy_thick(self, iters=200, seed=7):
        """Plain three-term Lanczos energy (no stored basis) for large sectors."""
        shape = (self.su.size, self.sd.size)
        rng = np.random.default_rng(seed)
        dtype = np.float64 if self.real else np.complex128
        v = rng.standard_normal(shape).astype(dtype)
        v /= np.linalg.norm(v)
        v_prev = np.zeros_like(v)
        b_prev = 0.0
        alphas, betas = [], []
        for _ in range(iters):
            w = self.apply(v)
            a = np.vdot(v, w).real
            w = w - a * v - b_prev * v_prev
            b = np.linalg.norm(w)
            alphas.append(a)
            betas.append(b)
            v_prev, v, b_prev = v, w / b, b
        T = np.diag(alphas) + np.diag(betas[:-1], 1) + np.diag(betas[:-1], -1)
        return np.linalg.eigvalsh(T)[0]
