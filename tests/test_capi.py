"""The C-ABI library loads and exports every entry point include/exahype_amd.h
declares; argument validation that needs no GPU."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "exahype_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(exa_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_python_binding_set(exa):
    assert declared_functions() == sorted(exa.EXPORTED)


def test_library_exports_every_declared_symbol(exa):
    out = subprocess.run(["nm", "-D", "--defined-only", exa.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (exa_[a-z0-9_]+)", out))
    missing = [f for f in declared_functions() if f not in exported]
    assert not missing, missing


def test_status_strings(exa):
    L = exa.lib()
    assert L.exa_status_string(0) == b"ok"
    assert L.exa_status_string(2).startswith(b"no compiled")


def test_supported_instantiations(exa):
    assert exa.supported(exa.euler_solver(3))
    assert exa.supported(exa.euler_solver(5))
    assert exa.supported(exa.euler_solver(3, npar=1, terms=exa.TERM_FLUX | exa.TERM_SOURCE | exa.TERM_NCP))
    assert not exa.supported(exa.euler_solver(9))
    s = exa.euler_solver(3)
    s.no_time_averaging = 1
    assert not exa.supported(s)


def test_invalid_arguments_rejected_before_any_device_work(exa):
    L = exa.lib()
    s = exa.euler_solver(3)
    h = ctypes.c_void_p()
    lo = (ctypes.c_double * 3)(0, 0, 0)
    hi = (ctypes.c_double * 3)(1, 1, 1)
    assert L.exa_grid_create(ctypes.byref(s), 0, 4, 4, lo, hi, 0, ctypes.byref(h)) == 3
    assert L.exa_grid_create(ctypes.byref(s), 4, 4, 4, lo, hi, 7, ctypes.byref(h)) == 2
    bad = exa.euler_solver(9)
    assert L.exa_grid_create(ctypes.byref(bad), 4, 4, 4, lo, hi, 0, ctypes.byref(h)) == 2
    assert L.exa_space_time_predictor_nonlinear(ctypes.byref(s), None, None, None, None, None, None, None,
                                                None, 0.1, None) == 3
    with pytest.raises(exa.ExaError):
        exa.check(3, "probe")
