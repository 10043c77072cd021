"""The C-ABI library loads and exports every entry point include/pic_hip.h
declares; host-side name mapping mirrors the reference (CPU only: no compute
call is made here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "pic_hip.h")).read()
    return sorted(set(re.findall(r"\b(pic_hip_\w+)\s*\(", text)))


def test_header_declares_entry_points():
    names = _declared()
    assert "pic_hip_deposit_host" in names and "pic_hip_deposit" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol(hip):
    lib = ctypes.CDLL(hip.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing
    # and the binding wraps every declared entry point
    assert set(_declared()) <= set(hip.EXPORTED)


def test_library_is_built_for_gfx950():
    so = os.path.join(ROOT, "pic-deposition_amd", "libpic_hip.so")
    data = open(so, "rb").read()
    assert b"gfx950" in data


def test_names_mirror_reference(hip):
    assert hip.scheme_from_string("Esirkepov") == hip.DepositionScheme.Esirkepov
    assert hip.backend_from_string("GPU_Priv") == hip.DepositionBackend.GPUPrivatized
    assert hip.backend_name(1) == "GPU_Atomics" and hip.scheme_name(2) == "TSC"
    with pytest.raises(RuntimeError, match="Unknown deposition scheme: FOO"):
        hip.scheme_from_string("FOO")
    with pytest.raises(RuntimeError, match="Unknown backend: GPU"):
        hip.backend_from_string("GPU")


def test_cpu_backend_is_not_served(hip):
    dom = hip.make_domain(1, 1, 8, 8)
    with pytest.raises(RuntimeError, match="CPU backend"):
        hip.deposit_charge({"x": [0.1], "y": [0.1], "q": [1.0]}, dom, "CIC", "CPU")


def test_invalid_arguments_fail_without_gpu(hip):
    # argument validation happens before any device call
    dom = hip.make_domain(1, 1, 0, 8)
    with pytest.raises(hip.PicError, match="invalid domain"):
        hip.deposit_host([0.1], [0.1], [1.0], dom)
