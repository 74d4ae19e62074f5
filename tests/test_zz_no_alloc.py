"""afqmc_comb_noalloc (the comb on device buffers, working memory in a fixed device array, no
allocation) against afqmc_comb on the same weights and uniforms: tables and popCheck statistics
identical, on two calls in a row, the second smaller (the fixed array keeps the first call's
contents, which must not leak into the second)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from afqmclab_amd._lib import check, lib  # noqa: E402
from tests.gpu_util import ctx  # noqa: E402

# the HIP runtime the library itself is linked against (device buffers for the test only)
_hip = C.CDLL("libamdhip64.so.7")
_hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
_hip.hipFree.argtypes = [C.c_void_p]
_hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
H2D, D2H = 1, 2


class DevArray:
    def __init__(self, host):
        self.host = np.ascontiguousarray(host)
        self.p = C.c_void_p()
        assert _hip.hipMalloc(C.byref(self.p), max(self.host.nbytes, 16)) == 0
        assert _hip.hipMemcpy(self.p, self.host.ctypes.data, self.host.nbytes, H2D) == 0

    def get(self):
        out = np.empty_like(self.host)
        assert _hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        _hip.hipFree(self.p)


def comb_host(w, u):
    t = np.zeros(len(w), dtype=np.int32)
    st = np.zeros(3)
    check(lib().afqmc_comb(ctx().h, len(w), w.ctypes.data_as(C.POINTER(C.c_double)),
                           u.ctypes.data_as(C.POINTER(C.c_double)), t.ctypes.data_as(C.POINTER(C.c_int)),
                           st.ctypes.data_as(C.POINTER(C.c_double))))
    return t, st


def comb_noalloc(w, u):
    dw, du = DevArray(w), DevArray(u)
    dt, ds = DevArray(np.full(len(w), -7, dtype=np.int32)), DevArray(np.zeros(3))
    try:
        check(lib().afqmc_comb_noalloc(ctx().h, len(w), dw.p, du.p, dt.p, ds.p))
        ctx().synchronize()
        return dt.get(), ds.get()
    finally:
        for a in (dw, du, dt, ds):
            a.free()


@pytest.mark.parametrize("sizes", [(5000, 700), (4096, 100)])
def test_comb_noalloc_matches_comb(sizes):
    rng = np.random.default_rng(sum(sizes))
    for nw in sizes:
        w = rng.uniform(0, 5, nw)
        w[rng.integers(0, nw, nw // 10 + 1)] = 0.0
        u = rng.uniform(0, 1, nw)
        t_ref, st_ref = comb_host(w, u)
        t, st = comb_noalloc(w, u)
        assert np.array_equal(t, t_ref)
        assert np.array_equal(st, st_ref)


def test_comb_noalloc_rejects_sizes_beyond_its_array():
    from afqmclab_amd._lib import lib as L
    rc = L().afqmc_comb_noalloc(ctx().h, 65537, None, None, None, None)
    assert rc < 0
    rc = L().afqmc_comb_noalloc(ctx().h, 65537, 16, 16, 16, None)
    assert rc < 0 and b"outside" in L().afqmc_last_error()
