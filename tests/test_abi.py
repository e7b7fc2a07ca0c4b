"""The C ABI library loads and exports every entry point include/oc_asm.h
declares, and its error paths behave like the reference bindings (a nonzero
status plus a retrievable message, never an abort).  No GPU work here."""
import ctypes
import os
import re

import pytest

from cm_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    names = []
    for fn in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            names += re.findall(r"^int\s+(oc_\w+)\s*\(", text, flags=re.M)
    return names


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_entry_points()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    assert set(declared_entry_points()) <= set(_lib.SIGNATURES)


def test_fortran_module_covers_header():
    text = open(os.path.join(ROOT, "include", "oc_asm.f90")).read()
    bound = set(re.findall(r'BIND\(C,\s*NAME="(oc_\w+)"\)', text))
    assert set(declared_entry_points()) <= bound, sorted(set(declared_entry_points()) - bound)


def test_error_paths_without_device():
    lib = _lib.load()
    assert lib.oc_equations_set_info(None, None, None, None, None, None) == _lib.OC_ERR_INVALID_HANDLE
    assert "NULL" in _lib.last_error()
    h = ctypes.c_void_p()
    import numpy as np
    rc = lib.oc_laplace_create(8, np.zeros(24), 1, np.zeros(8, np.int32), 0, ctypes.byref(h))
    assert rc in (_lib.OC_ERR_NOT_INITIALISED,)
    assert "oc_initialise" in _lib.last_error()
    with pytest.raises(_lib.CMISSError):
        _lib.call("oc_equations_set_destroy", None)
    buf = ctypes.create_string_buffer(4)
    assert lib.oc_last_error(buf, 4) == 0 and len(buf.value) <= 3


FORTRAN_DRIVER = os.path.join(ROOT, "tests", "fortran", "drive_laplace")


@pytest.mark.skipif(not os.path.exists(FORTRAN_DRIVER), reason="Fortran driver not built (no amdflang)")
def test_fortran_host_reaches_the_abi():
    """The ISO_C_BINDING module (include/oc_asm.f90) binds the C ABI: a Fortran
    program gets the status code and the error message of a bad call."""
    import subprocess
    out = subprocess.run([FORTRAN_DRIVER], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    assert "null handle: -1 equations set handle is NULL" in out.stdout
