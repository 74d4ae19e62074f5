"""ctypes binding of libafqmc_mi355x.so (the C-ABI declared in include/afqmc_mi355x.h).

The library is required: importing this module without a built library, or
calling into it without a HIP device, raises -- there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libafqmc_mi355x.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p


class CPDesc(C.Structure):
    _fields_ = [
        ("L", C.c_int), ("N", C.c_int), ("Nup", C.c_int), ("spin_blocks", C.c_int), ("precision", C.c_int),
        ("decomp_type", C.c_int), ("force_type", C.c_int),
        ("dt", C.c_double), ("force_cap", C.c_double),
        ("K", _dp), ("exp_minus_dt_k", _dp), ("exp_minus_half_dt_k", _dp), ("exp_half_dt_k", _dp),
        ("hop_logw", C.c_double * 6),
        ("U", _dp), ("mu", _dp), ("hx", _dp), ("hy", _dp), ("hz", _dp),
        ("phiT", _dp), ("phiT_logw", C.c_double * 2),
        ("const_force", _dp),
        ("max_walkers", C.c_int),
    ]


class PhDesc(C.Structure):
    _fields_ = [
        ("L", C.c_int), ("Nup", C.c_int), ("Ndn", C.c_int), ("G", C.c_int), ("precision", C.c_int),
        ("force_type", C.c_int), ("dt", C.c_double), ("force_cap", C.c_double),
        ("t", _dp), ("chol", _dp), ("bg", _dp), ("phiT_up", _dp), ("phiT_dn", _dp),
        ("phiT_logw", C.c_double * 2), ("half_rot_up", _dp), ("half_rot_dn", _dp), ("const_force", _dp),
        ("max_walkers", C.c_int), ("taylor_fixed", C.c_int), ("taylor_order", C.c_int), ("taylor_base", C.c_int),
        ("taylor_accuracy", C.c_double),
    ]


# name -> (restype, argtypes)
SIGNATURES = {
    "afqmc_last_error": (C.c_char_p, []),
    "afqmc_version": (C.c_int, []),
    "afqmc_ctx_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "afqmc_ctx_destroy": (C.c_int, [_vp]),
    "afqmc_ctx_synchronize": (C.c_int, [_vp]),
    "afqmc_ctx_set_timing": (C.c_int, [_vp, C.c_int]),
    "afqmc_ctx_get_timing": (C.c_int, [_vp, _dp, C.POINTER(C.c_longlong), C.c_int]),
    "afqmc_capture_begin": (C.c_int, [_vp]),
    "afqmc_capture_begin_group": (C.c_int, [_vp, C.c_int, C.POINTER(_vp)]),
    "afqmc_capture_end": (C.c_int, [_vp, C.POINTER(_vp)]),
    "afqmc_graph_launch": (C.c_int, [_vp]),
    "afqmc_graph_node_count": (C.c_int, [_vp]),
    "afqmc_graph_destroy": (C.c_int, [_vp]),
    "afqmc_cp_set_tiles": (C.c_int, [_vp, _ip]),
    "afqmc_cp_autotune": (C.c_int, [_vp, _vp, C.c_int, _ip, _dp]),
    "afqmc_cp_group_create": (C.c_int, [C.c_int, _vp, _vp, _vp]),
    "afqmc_cp_group_destroy": (C.c_int, [_vp]),
    "afqmc_cp_group_set_tiles": (C.c_int, [_vp, _ip]),
    "afqmc_cp_group_step": (C.c_int, [_vp]),
    "afqmc_cp_group_stabilize": (C.c_int, [_vp]),
    "afqmc_cp_group_pop_control": (C.c_int, [_vp]),
    "afqmc_cp_group_table_count": (C.c_int, [_vp]),
    "afqmc_cp_create": (C.c_int, [_vp, C.POINTER(CPDesc), C.POINTER(_vp)]),
    "afqmc_cp_destroy": (C.c_int, [_vp]),
    "afqmc_cp_get_hs_constants": (C.c_int, [_vp, _dp]),
    "afqmc_cp_walkers_create": (C.c_int, [_vp, C.c_int, C.POINTER(_vp)]),
    "afqmc_walkers_destroy": (C.c_int, [_vp]),
    "afqmc_walkers_count": (C.c_int, [_vp]),
    "afqmc_walkers_upload": (C.c_int, [_vp, _dp, _dp, _ip]),
    "afqmc_walkers_download": (C.c_int, [_vp, _dp, _dp, _ip]),
    "afqmc_cp_set_seed": (C.c_int, [_vp, C.c_uint64, C.c_uint64]),
    "afqmc_cp_get_counter": (C.c_int, [_vp, C.POINTER(C.c_uint64)]),
    "afqmc_cp_get_walker_steps": (C.c_int, [_vp, C.POINTER(C.c_uint64)]),
    "afqmc_cp_set_uniforms": (C.c_int, [_vp, _dp, C.c_int64]),
    "afqmc_cp_skip_uniforms": (C.c_int, [_vp, C.c_uint64]),
    "afqmc_cp_set_ET": (C.c_int, [_vp, C.c_double]),
    "afqmc_cp_get_ET": (C.c_int, [_vp, _dp]),
    "afqmc_cp_log_overlap": (C.c_int, [_vp, _vp, _dp]),
    "afqmc_cp_green_diagonal": (C.c_int, [_vp, _vp, _dp, _dp]),
    "afqmc_cp_force": (C.c_int, [_vp, _vp, _dp]),
    "afqmc_cp_onebody": (C.c_int, [_vp, _vp, C.c_int]),
    "afqmc_cp_hs_apply": (C.c_int, [_vp, _vp, _dp, _dp]),
    "afqmc_cp_stabilize": (C.c_int, [_vp, _vp]),
    "afqmc_cp_get_qr_stats": (C.c_int, [_vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "afqmc_ph_get_qr_stats": (C.c_int, [_vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "afqmc_comb": (C.c_int, [_vp, C.c_int, _dp, _dp, _ip, _dp]),
    "afqmc_walkers_download_some": (C.c_int, [_vp, C.c_int, _ip, _dp, _dp, _ip]),
    "afqmc_walkers_set_alive": (C.c_int, [_vp, _ip]),
    "afqmc_cp_walkers_from_trial": (C.c_int, [_vp, _vp, _dp]),
    "afqmc_comb_noalloc": (C.c_int, [_vp, C.c_int, _vp, _vp, _vp, _vp]),
    "afqmc_cp_apply_table": (C.c_int, [_vp, _vp, _ip]),
    "afqmc_cp_step": (C.c_int, [_vp, _vp, _dp, _ip]),
    "afqmc_cp_pop_control": (C.c_int, [_vp, _vp, _dp, _ip, _dp]),
    "afqmc_cp_step_fields": (C.c_int, [_vp, _vp, _ip]),
    "afqmc_cp_pop_control_table": (C.c_int, [_vp, _vp, _ip, _dp]),
    "afqmc_cp_measure": (C.c_int, [_vp, _vp]),
    "afqmc_bp_create": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(_vp)]),
    "afqmc_bp_destroy": (C.c_int, [_vp]),
    "afqmc_bp_begin": (C.c_int, [_vp]),
    "afqmc_bp_step": (C.c_int, [_vp, _vp]),
    "afqmc_bp_pop_control": (C.c_int, [_vp, _vp, _dp]),
    "afqmc_bp_measure": (C.c_int, [_vp, _vp]),
    "afqmc_bp_get": (C.c_int, [_vp, _dp, _dp, _dp, _dp, _dp, _dp]),
    "afqmc_bp_reset": (C.c_int, [_vp]),
    "afqmc_bp_get_step": (C.c_int, [_vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "afqmc_cp_energy": (C.c_int, [_vp, _dp]),
    "afqmc_cp_site_measure": (C.c_int, [_vp, _dp]),
    "afqmc_cp_reset_measure": (C.c_int, [_vp]),
    "afqmc_cp_adjust_ET": (C.c_int, [_vp]),
    "afqmc_cp_fixed_field_energy": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_uint64, _dp]),
    "afqmc_cp_fixed_field_adjoint": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_uint64, _dp, _dp, _dp, _dp]),
    "afqmc_ph_create": (C.c_int, [_vp, C.POINTER(PhDesc), C.POINTER(_vp)]),
    "afqmc_ph_destroy": (C.c_int, [_vp]),
    "afqmc_ph_set_onebody": (C.c_int, [_vp, C.c_int, _dp, _dp, C.c_int]),
    "afqmc_ph_set_background": (C.c_int, [_vp, _dp]),
    "afqmc_ph_walkers_create": (C.c_int, [_vp, C.c_int, C.POINTER(_vp)]),
    "afqmc_ph_walkers_destroy": (C.c_int, [_vp]),
    "afqmc_ph_walkers_upload": (C.c_int, [_vp, _dp, _dp, _ip]),
    "afqmc_ph_walkers_download": (C.c_int, [_vp, _dp, _dp, _ip]),
    "afqmc_ph_set_seed": (C.c_int, [_vp, C.c_uint64, C.c_uint64]),
    "afqmc_ph_get_counter": (C.c_int, [_vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "afqmc_ph_set_uniforms": (C.c_int, [_vp, _dp, C.c_int64]),
    "afqmc_ph_skip_uniforms": (C.c_int, [_vp, C.c_uint64]),
    "afqmc_ph_set_ET": (C.c_int, [_vp, C.c_double]),
    "afqmc_ph_get_ET": (C.c_int, [_vp, _dp]),
    "afqmc_ph_log_overlap": (C.c_int, [_vp, _vp, _dp]),
    "afqmc_ph_background": (C.c_int, [_vp, _vp, _dp]),
    "afqmc_ph_sample_apply": (C.c_int, [_vp, _vp, _dp, _dp, _dp, _ip]),
    "afqmc_ph_sample_matrix": (C.c_int, [_vp, _vp, _dp]),
    "afqmc_ph_taylor_orders": (C.c_int, [_vp, _vp, _ip]),
    "afqmc_ph_onebody": (C.c_int, [_vp, _vp, C.c_int]),
    "afqmc_ph_stabilize": (C.c_int, [_vp, _vp]),
    "afqmc_ph_step": (C.c_int, [_vp, _vp, _dp, _dp]),
    "afqmc_ph_pop_control": (C.c_int, [_vp, _vp, _dp, _ip, _dp]),
    "afqmc_ph_measure": (C.c_int, [_vp, _vp]),
    "afqmc_ph_energy": (C.c_int, [_vp, _dp, _dp, _dp]),
    "afqmc_ph_reset_measure": (C.c_int, [_vp]),
    "afqmc_ph_adjust_ET": (C.c_int, [_vp]),
    "afqmc_test_zgemm": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _dp, C.c_longlong,
                                   _dp, _dp, C.c_int, C.c_int, C.c_int, _dp]),
    "afqmc_test_gj_inverse": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _ip]),
    "afqmc_test_gj_inverse_layout": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _ip]),
    "afqmc_test_gj_inverse_masked": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, _ip, _dp, _dp, _dp, _ip]),
    "afqmc_test_copy": (C.c_int, [_vp, C.c_longlong, C.c_int, _dp]),
    "afqmc_test_chol_inverse": (C.c_int, [_vp, C.c_int, C.c_int, _dp, _dp, _dp, _ip]),
}


class AfqmcError(RuntimeError):
    pass


_lib = None


def lib():
    """Load the in-tree library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise AfqmcError(f"{LIB_PATH} is missing: run `python -m afqmclab_amd.build` first")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        msg = lib().afqmc_last_error().decode(errors="replace")
        raise AfqmcError(f"afqmc error {rc}: {msg}")
    return rc


def dptr(a):
    if a is None:
        return None
    if a.dtype not in (np.float64, np.complex128):
        raise TypeError("expected float64 or complex128 array")
    if not (a.flags["C_CONTIGUOUS"] or a.flags["F_CONTIGUOUS"]):
        raise ValueError("array must be contiguous")
    return a.ctypes.data_as(_dp)


def iptr(a):
    if a is None:
        return None
    if a.dtype != np.int32 or not a.flags["C_CONTIGUOUS"]:
        raise TypeError("expected a contiguous int32 array")
    return a.ctypes.data_as(_ip)
