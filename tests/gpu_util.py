"""Shared helpers for the GPU tests (loaded only by tests marked gpu)."""
import ctypes as C
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from afqmclab_amd._lib import check, dptr, iptr, lib  # noqa: E402
from afqmclab_amd.cp import Context  # noqa: E402

_CTX = None


def ctx():
    global _CTX
    if _CTX is None:
        _CTX = Context(0)
    return _CTX


def crand(rng, *shape):
    return rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)


def zgemm(opa, A, B, precision=0, batch=1, strideA=0, scale=None, scale_group=1, herm=False, b_upper=False,
          tile=0, split_bf16=False):
    """C_b = op(A_b) diag(scale) B_b via the device GEMM.  A, B given per batch as 3-D arrays
    (batch, rows, cols) or 2-D for shared A.  herm: the result is Hermitian (op C with A = B, only
    the upper tiles computed); b_upper: B is upper triangular."""
    A = np.asarray(A, dtype=np.complex128)
    B = np.asarray(B, dtype=np.complex128)
    if B.ndim == 2:
        B = B[None]
    batch = B.shape[0]
    K, N = B.shape[1], B.shape[2]
    if A.ndim == 2:
        A3 = A[None]
        strideA = 0
    else:
        A3 = A
        strideA = A.shape[1] * A.shape[2]
    M = A3.shape[1] if opa in (0, 3) else A3.shape[2]
    Abuf = np.ascontiguousarray(np.transpose(A3, (0, 2, 1)))  # column-major per matrix
    if opa == 3:  # real A
        Abuf = np.ascontiguousarray(Abuf.real)
    Bbuf = np.ascontiguousarray(np.transpose(B, (0, 2, 1)))
    Cbuf = np.zeros((batch, N, M), dtype=np.complex128)
    sbuf = None if scale is None else np.ascontiguousarray(scale, dtype=np.complex128)
    check(lib().afqmc_test_zgemm(ctx().h, precision, opa, M, N, K, batch, dptr(Abuf), strideA, dptr(Bbuf),
                                 dptr(sbuf), scale_group, int(herm) | (int(b_upper) << 1) | (int(split_bf16) << 2),
                                 tile, dptr(Cbuf)))
    return np.transpose(Cbuf, (0, 2, 1))


def gj_inverse(A, precision=0):
    A = np.asarray(A, dtype=np.complex128)
    batch, N, _ = A.shape
    buf = np.ascontiguousarray(np.transpose(A, (0, 2, 1)))
    out = np.zeros_like(buf)
    ld = np.zeros(batch, dtype=np.complex128)
    info = np.zeros(batch, dtype=np.int32)
    check(lib().afqmc_test_gj_inverse(ctx().h, precision, N, batch, dptr(buf), dptr(out), dptr(ld), iptr(info)))
    return np.transpose(out, (0, 2, 1)), ld, info


def gj_inverse_layout(A, layout, precision=0):
    """gj_inverse with the kernel chosen (1: lanes = rows, 2: 2-D lane layout, 3: panels of four
    columns)."""
    A = np.asarray(A, dtype=np.complex128)
    batch, N, _ = A.shape
    buf = np.ascontiguousarray(np.transpose(A, (0, 2, 1)))
    out = np.zeros_like(buf)
    ld = np.zeros(batch, dtype=np.complex128)
    info = np.zeros(batch, dtype=np.int32)
    check(lib().afqmc_test_gj_inverse_layout(ctx().h, precision, N, batch, layout, dptr(buf), dptr(out), dptr(ld),
                                             iptr(info)))
    return np.transpose(out, (0, 2, 1)), ld, info


def gj_inverse_masked(A, alive, layout=0, precision=0, fill=0.0):
    """gj_inverse over the matrices with alive != 0; the others keep A and logdet = fill."""
    A = np.asarray(A, dtype=np.complex128)
    batch, N, _ = A.shape
    buf = np.ascontiguousarray(np.transpose(A, (0, 2, 1)))
    out = np.zeros_like(buf)
    ld = np.full(batch, fill, dtype=np.complex128)
    info = np.zeros(batch, dtype=np.int32)
    al = np.ascontiguousarray(alive, dtype=np.int32)
    check(lib().afqmc_test_gj_inverse_masked(ctx().h, precision, N, batch, layout, iptr(al), dptr(buf), dptr(out),
                                             dptr(ld), iptr(info)))
    return np.transpose(out, (0, 2, 1)), ld, info


def chol_inverse(S):
    S = np.asarray(S, dtype=np.complex128)
    batch, N, _ = S.shape
    buf = np.ascontiguousarray(np.transpose(S, (0, 2, 1)))
    out = np.zeros_like(buf)
    ld = np.zeros(batch, dtype=np.complex128)
    info = np.zeros(batch, dtype=np.int32)
    check(lib().afqmc_test_chol_inverse(ctx().h, N, batch, dptr(buf), dptr(out), dptr(ld), iptr(info)))
    return np.transpose(out, (0, 2, 1)), ld, info
