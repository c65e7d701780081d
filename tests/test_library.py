"""CPU-side checks of the product library: it loads, exports every entry point include/ressim_hip.h declares,
and its host logic rejects bad input without touching a GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

import ressimstudio_amd as rs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "ressim_hip.h")).read()
    return sorted(set(re.findall(r"\b(rsh_[a-z0-9_]+)\s*\(", txt)))


def test_header_and_python_lists_agree():
    assert declared() == sorted(rs.EXPORTED)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(rs.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_version_and_error_plumbing_without_gpu():
    L = rs.lib()
    assert L.rsh_version() == 1
    o = rs.default_options()
    assert o.precond == rs.PREC_CPR and o.precision == rs.FP64 and o.restart == 30
    if rs.device_count() == 0:
        h = ctypes.c_void_p()
        assert L.rsh_context_create(0, ctypes.byref(h)) == rs.ERR_DEVICE
        assert b"no HIP device" in L.rsh_last_error()


def test_no_cpu_fallback_without_device():
    if rs.device_count() > 0:
        pytest.skip("a device is present")
    from ressimstudio_amd import gen
    with pytest.raises(rs.RshError):
        rs.System(gen.small())
