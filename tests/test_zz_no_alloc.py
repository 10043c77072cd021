"""pic_hip_deposit_noalloc (the allocation-free global-atomic deposit) through
the C ABI, on device arrays this test allocates with the HIP runtime the
library itself uses: two calls in a row (the second on a smaller grid and
particle set, on another stream), each against pic_hip_deposit with backend
GPU_Atomics on the same particles, to FP32 rounding of FP64 sums added in a
different order (rel L2 <= 1e-7, well inside the 1e-6 parity bound), and
against the FP64 oracle (rel L2 <= 1e-6); a grid beyond the scratch array is
rejected before any device work."""
import ctypes
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


class HipRuntime:
    """The few HIP runtime calls the test needs as the caller (allocation,
    copies, streams), from the runtime library libpic_hip.so is linked to."""

    def __init__(self):
        self.lib = ctypes.CDLL("libamdhip64.so.7")
        L = self.lib
        L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        L.hipFree.argtypes = [ctypes.c_void_p]
        L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        L.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        L.hipStreamDestroy.argtypes = [ctypes.c_void_p]
        L.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
        L.hipDeviceSynchronize.argtypes = []
        self.bufs = []

    def check(self, rc, what):
        assert rc == 0, f"{what} failed with HIP error {rc}"

    def upload(self, a):
        a = np.ascontiguousarray(a, np.float32)
        p = ctypes.c_void_p()
        self.check(self.lib.hipMalloc(ctypes.byref(p), max(a.nbytes, 4)), "hipMalloc")
        self.bufs.append(p.value)
        self.check(self.lib.hipMemcpy(p.value, a.ctypes.data, a.nbytes, 1), "hipMemcpy H2D")
        return p.value

    def empty(self, nfloat):
        p = ctypes.c_void_p()
        self.check(self.lib.hipMalloc(ctypes.byref(p), 4 * nfloat), "hipMalloc")
        self.bufs.append(p.value)
        return p.value

    def download(self, ptr, nfloat):
        out = np.empty(nfloat, np.float32)
        self.check(self.lib.hipMemcpy(out.ctypes.data, ptr, 4 * nfloat, 2), "hipMemcpy D2H")
        return out

    def stream(self):
        s = ctypes.c_void_p()
        self.check(self.lib.hipStreamCreate(ctypes.byref(s)), "hipStreamCreate")
        return s.value

    def close(self):
        self.lib.hipDeviceSynchronize()
        for p in self.bufs:
            self.lib.hipFree(p)
        self.bufs = []


def particles(n, L, seed, vmax):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, L[0], n).astype(np.float32)
    y = rng.uniform(0, L[1], n).astype(np.float32)
    x[x >= np.float32(L[0])] = 0
    y[y >= np.float32(L[1])] = 0
    q = (rng.uniform(0.5, 1.5, n) / 64).astype(np.float32)
    vx = rng.uniform(-vmax, vmax, n).astype(np.float32)
    vy = rng.uniform(-vmax, vmax, n).astype(np.float32)
    return x, y, q, vx, vy


def test_noalloc_deposit_matches_atomic_entry_point(hip, orc):
    rt = HipRuntime()
    try:
        cases = [  # (Nx, Ny, Lx, Ly, n, scheme, use own stream)
            (1024, 1024, 1.0, 1.0, 2_000_000, 1, False),
            (256, 128, 2.0, 1.0, 150_000, 3, True),
        ]
        for Nx, Ny, Lx, Ly, n, scheme, own in cases:
            say("noalloc", Nx, Ny, n, hip.scheme_name(scheme), "own stream" if own else "null stream")
            dom = hip.make_domain(Lx, Ly, Nx, Ny)
            dt = 0.01
            x, y, q, vx, vy = particles(n, (Lx, Ly), seed=Nx + n, vmax=0.5 * (Lx / Nx) / dt)
            dx, dy, dq, dvx, dvy = (rt.upload(a) for a in (x, y, q, vx, vy))
            drho = rt.empty(Nx * Ny)
            stream = rt.stream() if own else None
            hip.deposit_noalloc(dom, dx, dy, dq, n, drho, scheme, dvx, dvy, dt, stream)
            rt.check(rt.lib.hipStreamSynchronize(stream), "hipStreamSynchronize")
            got = rt.download(drho, Nx * Ny).reshape(Ny, Nx)
            if own:
                rt.lib.hipStreamDestroy(stream)
            with hip.Context(dom, n) as c:
                c.upload(x, y, q, vx, vy)
                c.deposit(scheme, 1, dt)
                ref = c.rho()
            od = orc.domain(Lx, Ly, Nx, Ny)
            oracle = orc.deposit(scheme, x.astype(np.float64), y.astype(np.float64), q.astype(np.float64), od,
                                 vx.astype(np.float64), vy.astype(np.float64), dt)
            e_entry, e_oracle = rel_l2(got, ref), rel_l2(got, oracle)
            print(f"noalloc {Nx}x{Ny} n={n} {hip.scheme_name(scheme)}: rel_l2 vs pic_hip_deposit {e_entry:.2e}, "
                  f"vs oracle {e_oracle:.2e}", flush=True)
            assert np.all(np.isfinite(got))
            assert e_entry <= 1e-7
            assert e_oracle <= 1e-6
        # a grid beyond the scratch array is rejected before any device work
        big = hip.make_domain(1.0, 1.0, 2048, 2048)
        with pytest.raises(hip.PicError, match="exceeds"):
            hip.deposit_noalloc(big, 0, 0, 0, 0, 0, 1)
    finally:
        rt.close()
