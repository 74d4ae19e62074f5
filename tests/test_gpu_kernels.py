"""Device building blocks vs the oracle / numpy: MFMA GEMM operand layouts,
Gauss-Jordan inverse + log-determinant, Cholesky inverse, bit-exact comb."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import comb as ocomb  # noqa: E402
from oracle import linalg  # noqa: E402
from tests.gpu_util import chol_inverse, crand, ctx, gj_inverse, gj_inverse_layout, gj_inverse_masked, zgemm  # noqa: E402

TOL = {0: 1e-13, 1: 2e-5}


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("opa", [0, 1, 2])
@pytest.mark.parametrize("shape", [(16, 16, 4), (37, 45, 29), (64, 64, 16), (130, 70, 97)])
def test_zgemm_matches_numpy(precision, opa, shape):
    rng = np.random.default_rng(sum(shape) + 7 * opa + precision)
    M, N, K = shape
    A = crand(rng, M, K) if opa == 0 else crand(rng, K, M)
    B = crand(rng, K, N)
    C = zgemm(opa, A, B, precision)[0]
    opA = A if opa == 0 else (A.T if opa == 1 else A.conj().T)
    ref = opA @ B
    assert np.max(np.abs(C - ref)) < TOL[precision] * np.sqrt(K) * 4


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("shape", [(37, 45, 29), (130, 70, 97), (400, 8, 30)])
def test_zgemm_real_a_matches_numpy(precision, shape):
    """op N with a real A (the Cholesky-vector sample matrix): only Re(A) is read."""
    rng = np.random.default_rng(sum(shape) + precision)
    M, N, K = shape
    A = rng.uniform(-1, 1, (M, K))
    B = crand(rng, K, N)
    C = zgemm(3, A + 0j, B, precision)[0]
    assert np.max(np.abs(C - A @ B)) < TOL[precision] * np.sqrt(K) * 4


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("tile", [0, 3, 5])
@pytest.mark.parametrize("N", [40, 112, 130])
def test_zgemm_hermitian_gram(precision, tile, N):
    """W^H W with only the tiles on or above the diagonal computed: those equal the full product's
    bit for bit, the others are their exact conjugate transposes."""
    rng = np.random.default_rng(N + tile + 11 * precision)
    batch, K = 3, 96
    W = crand(rng, batch, K, N)
    full = zgemm(2, W, W, precision, tile=tile)
    herm = zgemm(2, W, W, precision, herm=True, tile=tile)
    T = 32 if tile == 3 else 64
    for b in range(batch):
        assert np.max(np.abs(herm[b] - W[b].conj().T @ W[b])) < TOL[precision] * np.sqrt(K) * 4
        for i in range(N):
            for j in range(N):
                if i // T <= j // T:
                    assert herm[b][i, j] == full[b][i, j]
                else:
                    assert herm[b][i, j] == np.conj(herm[b][j, i])


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("tile", [0, 2, 3, 6])
def test_zgemm_upper_triangular_b(precision, tile):
    """W R^-1 with R^-1 upper triangular: the K tiles below an output column tile are skipped,
    which leaves the result unchanged bit for bit."""
    rng = np.random.default_rng(5 + tile + precision)
    batch, M, N = 4, 150, 112
    A = crand(rng, batch, M, N)
    B = np.triu(crand(rng, batch, N, N))
    ref = zgemm(0, A, B, precision, tile=tile)
    tri = zgemm(0, A, B, precision, b_upper=True, tile=tile)
    assert np.array_equal(ref, tri)
    for b in range(batch):
        assert np.max(np.abs(tri[b] - A[b] @ B[b])) < TOL[precision] * np.sqrt(N) * 4


@pytest.mark.parametrize("shape", [(256, 229, 256), (130, 70, 97), (16, 8, 5)])
def test_zgemm_split_bf16(shape):
    """Split-BF16 products (each FP32 operand as two BF16 parts, hi*hi + hi*lo + lo*hi): within
    3e-5 of sum |a||b| per element (each product carries up to about 2^-16 from the split, and the
    3M form adds three real products' errors), against the
    FP64 product, with and without the fused row scaling; ragged edges in every dimension."""
    M, N, K = shape
    rng = np.random.default_rng(M + N + K)
    A = crand(rng, M, K)
    B = crand(rng, 2, K, N)
    C = zgemm(0, A, B, 1, split_bf16=True)
    for b in range(2):
        bound = np.abs(A) @ np.abs(B[b])
        assert np.max(np.abs(C[b] - A @ B[b]) / bound) < 3e-5
    s = crand(rng, 2, K)
    C = zgemm(0, A, B, 1, scale=s, scale_group=N, split_bf16=True)
    for b in range(2):
        ref = A @ (s[b][:, None] * B[b])
        bound = np.abs(A) @ np.abs(s[b][:, None] * B[b])
        assert np.max(np.abs(C[b] - ref) / bound) < 3e-5


@pytest.mark.parametrize("precision", [0, 1])
def test_zgemm_batched_strided_and_scaled(precision):
    rng = np.random.default_rng(3)
    batch, M, N, K = 5, 33, 20, 40
    A = crand(rng, batch, M, K)
    B = crand(rng, batch, K, N)
    C = zgemm(0, A, B, precision)
    for b in range(batch):
        assert np.max(np.abs(C[b] - A[b] @ B[b])) < TOL[precision] * 30
    # shared A, per column-group row scaling (the fused HS diagonal)
    A2 = crand(rng, M, K)
    groups = 4
    s = crand(rng, batch, groups, K)
    C = zgemm(0, A2, B, precision, scale=s.reshape(batch * groups, K), scale_group=N // groups)
    for b in range(batch):
        for g in range(groups):
            cols = slice(g * (N // groups), (g + 1) * (N // groups))
            ref = A2 @ (s[b, g][:, None] * B[b][:, cols])
            assert np.max(np.abs(C[b][:, cols] - ref)) < TOL[precision] * 30


@pytest.mark.parametrize("N", [1, 3, 10, 16, 17, 40, 64, 65, 112, 128, 129, 150, 160, 161, 176, 224, 256])
def test_gj_inverse_and_logdet(N):
    rng = np.random.default_rng(N)
    batch = 6
    A = crand(rng, batch, N, N) + 0.5 * np.eye(N)
    inv, ld, info = gj_inverse(A)
    for b in range(batch):
        assert info[b] == 0
        assert np.max(np.abs(inv[b] @ A[b] - np.eye(N))) < 1e-11 * N
        ref = linalg.log_determinant(*linalg.lu_construct(A[b]))
        assert abs(ld[b].real - ref.real) < 1e-11 * max(1, abs(ref.real))
        assert abs(np.exp(1j * (ld[b].imag - ref.imag)) - 1) < 1e-11
        assert -np.pi < ld[b].imag <= np.pi


@pytest.mark.parametrize("N", [112, 150, 224])
def test_gj_inverse_fp32_storage(N):
    """FP32 input and output, FP64 elimination (the mixed path): the inverse to FP32 rounding and the
    log-determinant of the FP32-rounded matrix to FP64 accuracy."""
    rng = np.random.default_rng(7 + N)
    batch = 4
    A = (crand(rng, batch, N, N) + 0.5 * np.eye(N)).astype(np.complex64).astype(np.complex128)
    inv, ld, info = gj_inverse(A, precision=1)
    for b in range(batch):
        assert info[b] == 0
        assert np.max(np.abs(inv[b] @ A[b] - np.eye(N))) < 1e-4
        ref = linalg.log_determinant(*linalg.lu_construct(A[b]))
        assert abs(ld[b].real - ref.real) < 1e-11 * max(1, abs(ref.real))
        assert abs(np.exp(1j * (ld[b].imag - ref.imag)) - 1) < 1e-11


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("N", [17, 40, 64, 80, 100, 112])
def test_gj_inverse_large_batch(N, precision):
    """More matrices than CUs take the kernels with fewer waves per matrix (4 waves up to 64
    columns; 8 waves with part of the matrix in LDS for 65..112): the same bar as with a few
    matrices, and FP32 outputs identical to those bit for bit."""
    rng = np.random.default_rng(30 + N + precision)
    batch = 600
    A = crand(rng, batch, N, N) + 0.5 * np.eye(N)
    # an exactly singular matrix (a zero last row) and a permutation matrix (rows rotated by N/2,
    # so that above 64 rows the pivots come from both row groups)
    A[7][N - 1] = 0
    A[11] = np.eye(N)[np.r_[N // 2:N, 0:N // 2]]
    if precision:
        A = A.astype(np.complex64).astype(np.complex128)
    inv, ld, info = gj_inverse(A, precision=precision)
    assert info[7] == 1 and ld[7].real == -1e300
    assert np.count_nonzero(info) == 1
    assert np.array_equal(inv[11], A[11].T)
    perm_sign = np.linalg.det(A[11]).real
    assert abs(np.exp(ld[11]) - perm_sign) < 1e-12
    tol = 1e-4 if precision else 1e-11 * N
    for b in [b for b in range(0, batch, 23) if b not in (7, 11)]:
        assert np.max(np.abs(inv[b] @ A[b] - np.eye(N))) < tol, b
        ref = linalg.log_determinant(*linalg.lu_construct(A[b]))
        assert abs(ld[b].real - ref.real) < 1e-11 * max(1, abs(ref.real)), b
        assert abs(np.exp(1j * (ld[b].imag - ref.imag)) - 1) < 1e-11, b
    few = gj_inverse(A[:5], precision=precision)
    if precision:
        assert np.array_equal(few[0], inv[:5])
    assert np.max(np.abs(few[1] - ld[:5])) < 1e-12 * N


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("N", [5, 17, 32, 33, 40, 63, 64, 72, 96, 100, 112, 128])
def test_gj_two_d_layout_matches_one_d(N, precision):
    """The 2-D lane layout (both operands of the rank-1 update from LDS) against the 1-D layout
    (lanes = rows, the pivot row by readlane) on the same matrices, with a singular and a
    permutation matrix among them: the same breakdown flags, log-determinants to 1e-13, inverses
    to 1e-12 (FP64) or to FP32 rounding (FP32 storage), and inv A = I for the 2-D layout."""
    rng = np.random.default_rng(50 + N + precision)
    batch = 300
    A = crand(rng, batch, N, N) + 0.5 * np.eye(N)
    A[7][N - 1] = 0
    A[11] = np.eye(N)[np.r_[N // 2:N, 0:N // 2]]
    if precision:
        A = A.astype(np.complex64).astype(np.complex128)
    print(f"[gj layouts] N={N} precision={precision}", flush=True)
    i1, l1, f1 = gj_inverse_layout(A, 1, precision)
    i2, l2, f2 = gj_inverse_layout(A, 2, precision)
    assert np.array_equal(f1, f2) and f2[7] == 1 and np.count_nonzero(f2) == 1
    assert l2[7].real == -1e300
    ok = [b for b in range(batch) if b != 7]
    assert np.max(np.abs(l2[ok] - l1[ok])) < 1e-13 * N
    assert np.array_equal(i2[11], A[11].T)
    scale = np.max(np.abs(i1[ok]), axis=(1, 2))
    rel = np.max(np.abs(i2[ok] - i1[ok]), axis=(1, 2)) / scale
    assert np.max(rel) < (1e-6 if precision else 1e-12)
    tol = 1e-4 if precision else 1e-11 * N
    for b in [b for b in range(0, batch, 29) if b not in (7, 11)]:
        assert np.max(np.abs(i2[b] @ A[b] - np.eye(N))) < tol, b


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("N", [1, 3, 4, 5, 13, 32, 33, 40, 63, 64, 65, 72, 96, 97, 100, 112])
def test_gj_panel_kernel_matches_one_d(N, precision):
    """The Gauss-Jordan kernel in panels of four columns (rank-4 updates, the next panel factorised
    by its own wave during the update) against the unblocked 1-D kernel on the same matrices, with
    a singular and a permutation matrix among them: the same breakdown flags, log-determinants to
    1e-12 N, inverses to 1e-11 relative (FP64) or FP32 rounding, the permutation's inverse exact,
    and inv A = I.  The two sum the updates in different orders, so they agree to rounding only."""
    rng = np.random.default_rng(70 + N + precision)
    batch = 300
    A = crand(rng, batch, N, N) + 0.5 * np.eye(N)
    A[7][N - 1] = 0
    A[11] = np.eye(N)[np.r_[N // 2:N, 0:N // 2]]
    if precision:
        A = A.astype(np.complex64).astype(np.complex128)
    print(f"[gj panels] N={N} precision={precision}", flush=True)
    i1, l1, f1 = gj_inverse_layout(A, 1, precision)
    i3, l3, f3 = gj_inverse_layout(A, 3, precision)
    assert np.array_equal(f1, f3) and f3[7] == 1 and np.count_nonzero(f3) == 1
    assert l3[7].real == -1e300
    ok = [b for b in range(batch) if b != 7]
    assert np.max(np.abs(l3[ok] - l1[ok])) < 1e-12 * N
    assert np.array_equal(i3[11], A[11].T)
    scale = np.max(np.abs(i1[ok]), axis=(1, 2))
    rel = np.max(np.abs(i3[ok] - i1[ok]), axis=(1, 2)) / scale
    assert np.max(rel) < (1e-5 if precision else 1e-11)
    tol = 1e-4 if precision else 1e-11 * N
    for b in [b for b in range(0, batch, 29) if b not in (7, 11)]:
        assert np.max(np.abs(i3[b] @ A[b] - np.eye(N))) < tol, b
    sign, logabs = np.linalg.slogdet(A[ok[:20]])
    ref = logabs + 1j * np.angle(sign)
    d = l3[ok[:20]] - ref
    d = d.real + 1j * ((d.imag + np.pi) % (2 * np.pi) - np.pi)
    assert np.max(np.abs(d)) < 1e-10 * N


@pytest.mark.parametrize("N,batch", [(112, 700), (96, 600), (72, 530)])
def test_gj_panel_kernel_staged_and_masked(N, batch):
    """The panel kernel with FP32 storage and more matrices than CUs: a workgroup walks over
    several matrices and loads each next one into LDS while it factorises the current one, except
    after a skipped (not alive) matrix, whose successor it reads from global memory.  Whichever way
    a matrix was loaded, the arithmetic is the same: alive matrices must equal those of the unmasked
    launch bit for bit, and skipped ones keep their input (the inverse is written in place) and
    the log-determinant passed in."""
    rng = np.random.default_rng(90 + N)
    A = (crand(rng, batch, N, N) + 0.5 * np.eye(N)).astype(np.complex64).astype(np.complex128)
    alive = np.ones(batch, dtype=np.int32)
    alive[rng.choice(batch, batch // 5, replace=False)] = 0
    alive[[0, 256, 257, batch - 1]] = 0
    print(f"[gj staged] N={N} batch={batch} skipped={np.count_nonzero(alive == 0)}", flush=True)
    i0, l0, f0 = gj_inverse_layout(A, 3, 1)
    fill = 123.0 + 4.0j
    i1, l1, f1 = gj_inverse_masked(A, alive, 3, 1, fill)
    ok = alive != 0
    assert np.array_equal(i1[ok], i0[ok])
    assert np.array_equal(l1[ok], l0[ok]) and np.array_equal(f1[ok], f0[ok])
    assert np.all(l1[~ok] == fill)
    assert np.array_equal(i1[~ok], A[~ok])
    sign, logabs = np.linalg.slogdet(A[ok][:10])
    d = l0[ok][:10] - (logabs + 1j * np.angle(sign))
    d = d.real + 1j * ((d.imag + np.pi) % (2 * np.pi) - np.pi)
    assert np.max(np.abs(d)) < 1e-10 * N


def test_gj_pivoting_cases():
    # permutation matrix (one swap -> det -1), exactly singular matrix, identity
    P = np.eye(8, dtype=complex)[[1, 0, 2, 3, 4, 5, 6, 7]]
    S = np.ones((8, 8), dtype=complex)
    inv, ld, info = gj_inverse(np.stack([P, S, np.eye(8, dtype=complex)]))
    assert np.allclose(inv[0], P.T)
    assert abs(np.exp(ld[0]) - (-1)) < 1e-14
    assert info[1] == 1 and ld[1].real == -1e300
    assert np.allclose(inv[2], np.eye(8)) and abs(ld[2]) == 0


def test_gj_wide_pivoting_singular_and_mask():
    """161 to 256 columns (working copies in device memory): a row permutation with an odd number
    of swaps, an exactly singular matrix, a skipped matrix and a general one against LAPACK."""
    N = 200
    rng = np.random.default_rng(5)
    perm = rng.permutation(N)
    P = np.eye(N, dtype=complex)[perm]
    S = np.ones((N, N), dtype=complex)
    A = crand(rng, 1, N, N)[0] + 0.5 * np.eye(N)
    inv, ld, info = gj_inverse_masked(np.stack([P, S, A, A]), np.array([1, 1, 0, 1], dtype=np.int32), fill=7.0)
    sign = np.linalg.det(P).real
    assert np.allclose(inv[0], P.T) and info[0] == 0
    assert abs(np.exp(ld[0]) - sign) < 1e-12
    assert info[1] == 1 and ld[1].real == -1e300
    assert ld[2] == 7.0 and np.array_equal(inv[2], A)
    ref = linalg.log_determinant(*linalg.lu_construct(A))
    assert info[3] == 0 and np.max(np.abs(inv[3] @ A - np.eye(N))) < 1e-11 * N
    assert abs(ld[3].real - ref.real) < 1e-11 * abs(ref.real)
    assert abs(np.exp(1j * (ld[3].imag - ref.imag)) - 1) < 1e-11


@pytest.mark.parametrize("N", [5, 28, 64, 112, 129, 160, 176, 224, 256])
def test_chol_inverse(N):
    rng = np.random.default_rng(100 + N)
    batch = 4
    W = crand(rng, batch, 3 * N, N)
    S = np.conj(np.transpose(W, (0, 2, 1))) @ W
    Rinv, ld, info = chol_inverse(S)
    for b in range(batch):
        R = np.linalg.inv(Rinv[b])
        assert np.allclose(np.triu(R), R, atol=1e-10)
        assert np.all(np.diag(R).real > 0)
        assert np.max(np.abs(R.conj().T @ R - S[b])) < 1e-10 * np.max(np.abs(S[b]))
        assert abs(ld[b].real - np.sum(np.log(np.diag(R).real))) < 1e-10 * N


@pytest.mark.parametrize("L", [4, 100, 2048])
def test_comb_bit_exact(L):
    from afqmclab_amd.cp import comb
    rng = np.random.default_rng(L)
    for trial in range(5):
        w = rng.uniform(0, 5, L)
        w[rng.integers(0, L, L // 10 + 1)] = 0.0
        u = rng.uniform(0, 1, L)
        if trial == 4:
            u[:] = np.nextafter(1.0, 0.0)
        t, st = comb(ctx(), w, u)
        ref = ocomb.popConfiguration(1, w, u)
        assert np.array_equal(t, ref)
        assert st == ocomb.popCheck(list(w))


@pytest.mark.parametrize("case", ["equal_u0", "equal_u_half", "one_alive", "all_dead", "ties", "max_lds", "beyond_lds"])
def test_comb_bit_exact_edges(case):
    """Teeth on cumulative sums, a single or no nonzero weight, and walker counts at and beyond the
    size the LDS comb holds (4096; larger counts take the sequential kernel)."""
    from afqmclab_amd.cp import comb
    rng = np.random.default_rng(7)
    L = {"max_lds": 4096, "beyond_lds": 5000}.get(case, 512)
    w = rng.uniform(0, 5, L)
    u = rng.uniform(0, 1, L)
    if case.startswith("equal"):
        w[:] = 1.0
        u[:] = 0.0 if case == "equal_u0" else 0.5
    elif case == "one_alive":
        w[:] = 0.0
        w[L // 3] = 2.5
    elif case == "all_dead":
        w[:] = 0.0
    elif case == "ties":
        w = rng.integers(0, 3, L).astype(float)  # many zeros and repeated values
        u[::2] = 0.0
    with np.errstate(invalid="ignore", divide="ignore"):
        ref = ocomb.popConfiguration(1, w, u)
    t, st = comb(ctx(), w, u)
    assert np.array_equal(t, ref)


def test_comb_reference_b_exact_weights():
    """The weights and tooth positions of the reference's b_exact table
    (populationControlTest.cpp:29-45: weights 1..100): the device comb (one chunk) equals the
    oracle's size-1 table, whose size-4 form reproduces b_exact slot for slot."""
    import json
    import os
    from afqmclab_amd.cp import comb
    from tests.test_oracle_models import _uniforms_for_table
    k = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                    "reference_kats.json")))["popConfiguration_b_exact"]
    w = np.array(k["weights"])
    expect = np.array(k["table_1based"]) - 1
    u = _uniforms_for_table(list(w), expect)
    assert np.array_equal(ocomb.popConfiguration(k["size"], w, u), expect)
    t, st = comb(ctx(), w, u)
    assert np.array_equal(t, ocomb.popConfiguration(1, w, u))
    assert np.array_equal(np.bincount(t, minlength=100), np.bincount(expect, minlength=100))
