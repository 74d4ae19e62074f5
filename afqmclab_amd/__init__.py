"""MI355X-native AFQMC walker-propagation path (HIP kernels behind a C-ABI).

Python layer: ctypes bindings (``_lib``), host-side model set-up that the
reference also does on the host (``model``: lattice K, exp(-alpha K) by
eigendecomposition, trial wave functions), and the device-resident drivers
that mirror the reference's AfqmcConstraintPath loop (``cp``).  All walker
work runs in libafqmc_mi355x.so on the GPU.
"""
from ._lib import AfqmcError, lib  # noqa: F401

__all__ = ["AfqmcError", "lib"]
