"""Dense kernels of the reference's tensor layer (test infrastructure only).

Each routine calls the LAPACK routine the reference calls, through
scipy.linalg.lapack, with the reference's conventions layered on top.
"""
import cmath
import math

import numpy as np
from scipy.linalg import lapack


def lu_construct(a: np.ndarray):
    """zgetrf of a square matrix (tensor_2d_bl_cpu.cpp:157-167).

    Returns (lu, ipiv, info) with ipiv 0-based (scipy convention)."""
    a = np.asfortranarray(np.array(a, dtype=np.complex128))
    if a.shape[0] != a.shape[1]:
        raise ValueError("Input for LU is not square matrix!")
    if a.size == 0:
        return a, np.zeros(0, dtype=np.int32), 0
    lu, ipiv, info = lapack.zgetrf(a)
    if info < 0:
        raise ValueError("illegal parameter in zgetrf")
    return lu, ipiv, info


def log_determinant(lu, ipiv, info) -> complex:
    """lognorm + log(phase) with phase = prod(+-u_ii/|u_ii|)
    (tensor_2d_common_fun.cpp:75-105).  Singular -> -1e300."""
    if info > 0:
        return complex(-1e300, 0.0)
    lognorm = 0.0
    phase = complex(1.0, 0.0)
    n = lu.shape[0]
    for i in range(n):
        u = complex(lu[i, i])
        au = abs(u)
        lognorm += math.log(au)
        if ipiv[i] != i:
            phase *= (-u / au)
        else:
            phase *= (u / au)
    return complex(lognorm, 0.0) + cmath.log(phase)


def determinant(lu, ipiv, info) -> complex:
    """prod(+-u_ii) (tensor_2d_common_fun.cpp:58-70)."""
    if info > 0:
        return 0j
    det = complex(1.0, 0.0)
    for i in range(lu.shape[0]):
        det *= (-lu[i, i]) if ipiv[i] != i else lu[i, i]
    return det


def solve_lineq(lu, ipiv, b: np.ndarray, trans: str = "N") -> np.ndarray:
    """zgetrs(TRANS) (tensor_2d_bl_cpu.cpp:218-258).  trans 'T' -> op = A^T."""
    b = np.asfortranarray(np.array(b, dtype=np.complex128))
    if b.size == 0:
        return b
    t = {"N": 0, "T": 1, "C": 2}[trans]
    x, info = lapack.zgetrs(lu, ipiv, b, trans=t)
    if info != 0:
        raise ValueError("Solve linear equation is not successful")
    return x


def qr_matrix(ph: np.ndarray):
    """zgeqrf + zungqr with the reference's sign rule (tensor_2d_bl_cpu.cpp:263-315).

    Returns (Q, det_real, det_list) where det_list holds Re(R_ii) and, if
    Re(prod R_ii) < 0, column 0 of Q and det_list[0] are negated."""
    ph = np.asfortranarray(np.array(ph, dtype=np.complex128))
    L, N = ph.shape
    qr, tau, work, info = lapack.zgeqrf(ph)
    if info != 0:
        raise ValueError("QR run is not successful")
    det = complex(1.0, 0.0)
    det_list = np.empty(N)
    for i in range(N):
        det_list[i] = qr[i, i].real
        det *= qr[i, i]
    q, work, info = lapack.zungqr(qr, tau)
    if info != 0:
        raise ValueError("Zungqr run is not successful")
    q = np.array(q[:, :N], order="F")
    if det.real < 0:
        det = -det
        det_list[0] = -det_list[0]
        q[:, 0] = -q[:, 0]
    return q, det.real, det_list


def eigen_hermitian(k: np.ndarray):
    """zheevd (ascending eigenvalues, eigenvectors in columns)."""
    w, v, info = lapack.zheevd(np.asfortranarray(np.array(k, dtype=np.complex128)), compute_v=1)
    if info != 0:
        raise ValueError("zheevd failed")
    return w, v


def eigen_symmetric(k: np.ndarray):
    """dsyevd (ascending eigenvalues, eigenvectors in columns)."""
    w, v, info = lapack.dsyevd(np.asfortranarray(np.array(k, dtype=np.float64)), compute_v=1)
    if info != 0:
        raise ValueError("dsyevd failed")
    return w, v
