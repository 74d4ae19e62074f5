"""The C-ABI library loads (no GPU needed) and exports every function that
include/*.h declares; the Python binding covers the same set."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "afqmclab_amd", "libafqmc_mi355x.so")


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*(afqmc_[A-Za-z0-9_]+)\s*\(", text, re.M):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    assert "afqmc_cp_step" in names and "afqmc_comb" in names and len(names) > 30


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built (run __graft_entry__.build())")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"declared in include/ but not exported: {missing}"


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built (run __graft_entry__.build())")
def test_binding_matches_header():
    from afqmclab_amd import _lib
    declared = set(declared_functions())
    bound = set(_lib.SIGNATURES)
    assert bound <= declared, f"bound but not declared: {sorted(bound - declared)}"
    lib = _lib.lib()  # resolves every bound symbol
    assert lib.afqmc_version() >= 1


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built (run __graft_entry__.build())")
def test_errors_are_reported_not_fatal():
    """Invalid arguments return a negative code and a message (the reference would exit(1))."""
    from afqmclab_amd import _lib
    lib = _lib.lib()
    rc = lib.afqmc_cp_create(None, None, None)
    assert rc < 0
    assert b"null" in lib.afqmc_last_error()


EXE = os.path.join(ROOT, "afqmclab_amd", "afqmc_cp_run")


@pytest.mark.skipif(not os.path.exists(EXE), reason="host driver not built (run __graft_entry__.build())")
def test_host_driver_reports_missing_input(tmp_path):
    """The C++ host driver reads afqmc_param before touching the device; a missing file is an
    error message and exit code 1, not a crash."""
    import subprocess
    r = subprocess.run([EXE, str(tmp_path / "nothing")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert "Error opening file" in r.stderr and "afqmc_param" in r.stderr


@pytest.mark.skipif(not os.path.exists(EXE), reason="host driver not built (run __graft_entry__.build())")
def test_host_driver_back_propagation_needs_a_reference_directory(tmp_path):
    import subprocess
    r = subprocess.run([EXE, "--bp", str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "--bp needs a reference run directory" in r.stderr
    r = subprocess.run([EXE, "--bp"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr
