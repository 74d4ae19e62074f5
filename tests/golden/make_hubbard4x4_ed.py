"""Ground-state energy of the 4x4 periodic Hubbard model, t = 1, U = 4, 5 up + 5 down, by
Lanczos in the full (Nup, Ndn) sector (4368^2 = 19,079,424 states), as the end-to-end energy
check of the constraint-path run on this lattice.

Run from the repository root:  python tests/golden/make_hubbard4x4_ed.py
Writes tests/golden/hubbard4x4_u4_n5u5d_ed.json.  Three-term Lanczos without
reorthogonalisation (the lowest Ritz value converges before ghosts appear), stopped when the
lowest Ritz value changes by less than 1e-12 over 10 iterations.  The sector, basis ordering and
Hamiltonian (one-body part with Jordan-Wigner signs by site index, U n_up n_dn) are those of
oracle/lanczos.py, which is pinned by the reference's own ED test values.
"""
import json
import os
import sys
import time

import numpy as np
import scipy.linalg as sl

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle.lanczos import basis, one_body_matrix  # noqa: E402
from oracle.models import square_lattice_K  # noqa: E402


def main(nx=4, ny=4, nup=5, ndn=5, U=4.0, maxiter=600, seed=3):
    L = nx * ny
    k = square_lattice_K(nx, ny)[:L, :L].real
    su, sd = basis(L, nup), basis(L, ndn)
    hu = one_body_matrix(k.astype(np.complex128), su).real.tocsr()
    hd = one_body_matrix(k.astype(np.complex128), sd).real.tocsr()
    occ_u = ((su[:, None] >> np.arange(L)) & 1).astype(np.float64)
    occ_d = ((sd[:, None] >> np.arange(L)) & 1).astype(np.float64)
    D = U * (occ_u @ occ_d.T)

    def apply(V):
        return hu @ V + (hd @ V.T).T + D * V

    rng = np.random.default_rng(seed)
    v = rng.standard_normal(D.shape)
    v /= np.linalg.norm(v)
    v_prev = np.zeros_like(v)
    alphas, betas = [], []
    beta = 0.0
    hist = []
    t0 = time.time()
    for it in range(maxiter):
        w = apply(v)
        a = float(np.vdot(v, w))
        alphas.append(a)
        w -= a * v + beta * v_prev
        beta = float(np.linalg.norm(w))
        e0 = sl.eigh_tridiagonal(np.array(alphas), np.array(betas), select="i", select_range=(0, 0),
                                 eigvals_only=True)[0]
        hist.append(e0)
        if it >= 10 and abs(hist[-1] - hist[-11]) < 1e-12:
            break
        betas.append(beta)
        v_prev, v = v, w / beta
    out = {"lattice": f"{nx}x{ny} periodic", "t": 1.0, "U": U, "Nup": nup, "Ndn": ndn,
           "sector_dimension": int(D.size), "E0": hist[-1], "lanczos_iterations": len(alphas),
           "last_change": abs(hist[-1] - hist[-11]), "seconds": time.time() - t0}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hubbard4x4_u4_n5u5d_ed.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
