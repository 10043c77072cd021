"""The reference's own dispatch code (deposition.cpp built with its GPU branch)
calling the HIP library through the seam adapter; mirrors main_test.cpp's GPU
tests and compares rho with the reference's CPU backend."""
import ctypes
import os

import pytest

pytestmark = pytest.mark.gpu

LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref", "libseamtest.so")


def test_reference_dispatch_through_hip_seam(hip):
    if not os.path.exists(LIB):
        pytest.skip("oracle/_ref/libseamtest.so not built (needs the reference checkout at build time)")
    print("[case] seam tests", flush=True)
    lib = ctypes.CDLL(LIB, mode=os.RTLD_LAZY)
    lib.seam_run_tests.restype = ctypes.c_int
    assert lib.seam_run_tests() == 0
