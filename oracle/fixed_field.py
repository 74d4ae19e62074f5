"""Fixed-field propagation and mixed-estimator energy of spin-block Hubbard walkers (test
infrastructure only: the checker of the GPU fixed-field energy and, by central finite
differences, of its adjoint gradient).

Per step s and walker w, with fields x_{s,w,i} = (u < 1/2 ? -1 : +1), u the uniform number
(s*nw + w)*L + i of the stream keyed by seed_mix(seed):
  W_up <- B diag(exp(+gamma_i x) c_i) W_up,  W_dn <- B diag(exp(-gamma_i x) c_i) W_dn,
B = exp(-dt K), gamma_i = solveCoshxEqExpy(dt U_i / 2), c_i = exp(-dt U_i / 2) (densitySpin
with mu = h = 0: NiupNidn.cpp:201-250); every mgs steps W = QR (numpy), logw += log prod R_ii.
E = sum_w O_w E_L,w / sum_w O_w with O_w = exp(logw) det(phi_up^H W_up) det(phi_dn^H W_dn) and
E_L = sum_b tr(K theta_b phi_b^H) + sum_i U_i G_up,ii G_dn,ii, theta_b = W_b (phi_b^H W_b)^{-1}
(HubbardSOCMeasureFixedSDSD.cpp:178-218 restricted to spin blocks).  Gradients have no
reference counterpart: this oracle pins them only through finite differences (parity of the
gradient otherwise unpinned).
"""
import cmath

import numpy as np

from .mathfun import solveCoshxEqExpy
from .rng import seed_mix, uniforms


def fields(seed, s, nw, L):
    u = uniforms(seed_mix(seed), s * nw * L, nw * L).reshape(nw, L)
    return np.where(u < 0.5, -1.0, 1.0)


def energy(K, U, phi_up, phi_dn, nw, nsteps, mgs, dt, seed):
    K = np.asarray(K, dtype=np.float64)
    L = K.shape[0]
    U = np.asarray(U, dtype=np.float64) * np.ones(L)
    lam, V = np.linalg.eigh(K)
    B = (V * np.exp(-dt * lam)) @ V.T
    gamma = np.array([solveCoshxEqExpy(dt * u * 0.5).real for u in U])
    c = np.exp(-dt * U * 0.5)
    phis = [np.asarray(phi_up, dtype=np.complex128), np.asarray(phi_dn, dtype=np.complex128)]
    W = [[p.copy() for p in phis] for _ in range(nw)]
    logw = np.zeros(nw, dtype=np.complex128)
    for s in range(nsteps):
        x = fields(seed, s, nw, L)
        for w in range(nw):
            for b, sig in ((0, 1.0), (1, -1.0)):
                d = np.exp(sig * gamma * x[w]) * c
                W[w][b] = B @ (d[:, None] * W[w][b])
            if (s + 1) % mgs == 0:
                for b in range(2):
                    q, r = np.linalg.qr(W[w][b])
                    W[w][b] = q
                    logw[w] += cmath.log(np.prod(np.diag(r)))
    logO = np.zeros(nw, dtype=np.complex128)
    el = np.zeros(nw, dtype=np.complex128)
    for w in range(nw):
        m = []
        kin = 0j
        for b in range(2):
            o = phis[b].conj().T @ W[w][b]
            sign, ld = np.linalg.slogdet(o)
            logO[w] += ld + cmath.log(sign)
            theta = W[w][b] @ np.linalg.inv(o)
            M = theta @ phis[b].conj().T
            kin += np.trace(K @ M)
            m.append(np.diag(M))
        el[w] = kin + np.sum(U * m[0] * m[1])
        logO[w] += logw[w]
    o = np.exp(logO - logO.real.max())
    return complex(np.sum(o * el) / np.sum(o))


def expm_derivative_adjoint(K, dt, bbar):
    """dE/dK from dE/dB, B = exp(-dt K), K real symmetric (Daleckii-Krein on K's eigenbasis)."""
    lam, V = np.linalg.eigh(K)
    mu = -dt * lam
    e = np.exp(mu)
    d = mu[:, None] - mu[None, :]
    same = np.abs(d) <= 1e-12
    F = np.where(same, 0.5 * (e[:, None] + e[None, :]), (e[:, None] - e[None, :]) / np.where(same, 1.0, d))
    return -dt * (V @ ((V.T @ bbar @ V) * F) @ V.T)


def energy_and_grad(K, U, phi_up, phi_dn, nw, nsteps, mgs, dt, seed):
    """E and the reverse-mode derivatives dE/dK (L x L, holomorphic, from the propagators and the
    kinetic term) and dE/dU_i, by the adjoint of the steps above in the frame of the
    QR-normalised walkers: Wt_n = c_w [(E_L - E) conj(phi) + P - conj(phi) theta^T P] O^{-T},
    P = K^T conj(phi) + diag(U m_other) conj(phi); per step Wt <- Wt R^{-T} after a QR,
    Bbar += Wt (D Q)^T, Ybar = B^T Wt, Ubar += sum_c Ybar Q D dlogD/dU, Wt <- D Ybar."""
    K = np.asarray(K, dtype=np.float64)
    L = K.shape[0]
    Uv = np.asarray(U, dtype=np.float64) * np.ones(L)
    lam, V = np.linalg.eigh(K)
    B = (V * np.exp(-dt * lam)) @ V.T
    gam = np.array([solveCoshxEqExpy(dt * u * 0.5).real for u in Uv])
    c = np.exp(-dt * Uv * 0.5)
    e = np.exp(dt * Uv * 0.5)
    gp = np.sign(gam) * 0.5 * dt * e / np.sqrt(e * e - 1.0)  # d gamma / d U
    phis = [np.asarray(phi_up, dtype=np.complex128), np.asarray(phi_dn, dtype=np.complex128)]
    W = [[p.copy() for p in phis] for _ in range(nw)]
    logw = np.zeros(nw, dtype=np.complex128)
    Qs, Rinv = [], {}
    for s in range(nsteps):
        x = fields(seed, s, nw, L)
        Qs.append([[W[w][b].copy() for b in range(2)] for w in range(nw)])
        for w in range(nw):
            for b, sig in ((0, 1.0), (1, -1.0)):
                W[w][b] = B @ ((np.exp(sig * gam * x[w]) * c)[:, None] * W[w][b])
            if (s + 1) % mgs == 0:
                for b in range(2):
                    q, r = np.linalg.qr(W[w][b])
                    W[w][b] = q
                    logw[w] += cmath.log(np.prod(np.diag(r)))
                    Rinv[s, w, b] = np.linalg.inv(r)
    el = np.zeros(nw, dtype=np.complex128)
    logO = np.zeros(nw, dtype=np.complex128)
    th, Oi, m = {}, {}, {}
    for w in range(nw):
        kin = 0j
        for b in range(2):
            o = phis[b].conj().T @ W[w][b]
            sign, ld = np.linalg.slogdet(o)
            logO[w] += ld + cmath.log(sign)
            Oi[w, b] = np.linalg.inv(o)
            th[w, b] = W[w][b] @ Oi[w, b]
            M = th[w, b] @ phis[b].conj().T
            kin += np.trace(K @ M)
            m[w, b] = np.diag(M)
        el[w] = kin + np.sum(Uv * m[w, 0] * m[w, 1])
        logO[w] += logw[w]
    cw = np.exp(logO - logO.real.max())
    cw /= cw.sum()
    E = np.sum(cw * el)
    Wt = {}
    ubar = np.zeros(L, dtype=np.complex128)
    kbar = np.zeros((L, L), dtype=np.complex128)
    for w in range(nw):
        for b in range(2):
            cph = phis[b].conj()
            P = K.T @ cph + (Uv * m[w, 1 - b])[:, None] * cph
            P = P - cph @ (th[w, b].T @ P)
            Wt[w, b] = (cw[w] * ((el[w] - E) * cph + P)) @ Oi[w, b].T
            kbar += cw[w] * (th[w, b] @ phis[b].conj().T).T
        ubar += cw[w] * m[w, 0] * m[w, 1]
    bbar = np.zeros((L, L), dtype=np.complex128)
    for s in range(nsteps - 1, -1, -1):
        x = fields(seed, s, nw, L)
        for w in range(nw):
            for b, sig in ((0, 1.0), (1, -1.0)):
                if (s + 1) % mgs == 0:
                    Wt[w, b] = Wt[w, b] @ Rinv[s, w, b].T
                Q = Qs[s][w][b]
                d = np.exp(sig * gam * x[w]) * c
                bbar += Wt[w, b] @ (d[:, None] * Q).T
                Yb = B.T @ Wt[w, b]
                ubar += np.sum(Yb * Q, axis=1) * d * (sig * x[w] * gp - 0.5 * dt)
                Wt[w, b] = d[:, None] * Yb
    kbar += expm_derivative_adjoint(K, dt, bbar)
    return complex(E), kbar, ubar
