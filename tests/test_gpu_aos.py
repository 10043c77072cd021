"""AoS particle records (ParticleAoS, src/pic/particles.hpp:13-20: x, y, vx,
vy, q, m as six FP64 values) through the device: upload converts on the GPU to
the same FP32 streams as a host cast, download promotes exactly, and the AoS
seam (depositChargeAoS's device branch, deposition.cpp:160-179) gives the SoA
seam's rho."""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def records(n, seed, L=1.0, uniform_q=False):
    rng = np.random.default_rng(seed)
    rec = np.empty((n, 6))
    rec[:, 0] = rng.uniform(0, L, n)
    rec[:, 1] = rng.uniform(0, L, n)
    rec[:, 2] = rng.normal(0, 1, n)
    rec[:, 3] = rng.normal(0, 1, n)
    rec[:, 4] = 0.5 if uniform_q else rng.uniform(0.5, 1.5, n)
    rec[:, 5] = rng.uniform(1, 2, n)
    return rec


@pytest.mark.parametrize("n", [0, 1, 1000, (1 << 22) + 12345])
def test_upload_download_aos(hip, n):
    """Chunk boundaries included: 2^22 + 12345 records take two device copies."""
    say("aos upload", n)
    rec = records(n, seed=n)
    dom = hip.make_domain(1.0, 1.0, 64, 64)
    with hip.Context(dom, max(n, 1)) as c:
        c.upload_aos(rec)
        assert c.count == n
        s = c.download()
        for j, k in enumerate(hip.AOS_FIELDS):
            assert np.array_equal(s[k], rec[:, j].astype(np.float32)), k
        back = c.download_aos()
    assert back.shape == (n, 6)
    assert np.array_equal(back, rec.astype(np.float32).astype(np.float64))


def test_aos_context_deposit_and_sort_match_soa(hip):
    say("aos context deposit")
    n = 200_000
    rec = records(n, seed=7, uniform_q=True)
    dom = hip.make_domain(1.0, 1.0, 128, 96)
    out = {}
    for layout in ("soa", "aos"):
        with hip.Context(dom, n) as c:
            if layout == "aos":
                c.upload_aos(rec)
            else:
                f = rec.astype(np.float32)
                c.upload(f[:, 0], f[:, 1], f[:, 4], f[:, 2], f[:, 3], f[:, 5])
            c.sort()
            for s in (0, 1, 2):
                c.deposit(s, 2, 0.0)
                out[(layout, s)] = c.rho().copy()
            out[(layout, "p")] = c.download_aos()
    for s in (0, 1, 2):
        assert np.array_equal(out[("aos", s)], out[("soa", s)]), s
    assert np.array_equal(out[("aos", "p")], out[("soa", "p")])


@pytest.mark.parametrize("scheme", [0, 1, 2, 3])
def test_aos_seam_matches_soa_seam(hip, orc, scheme):
    say("aos seam", scheme)
    n = 64 * 64 * 16
    rec = records(n, seed=11 + scheme)
    rec[:, 2:4] *= 0.3  # Esirkepov moves (v dt) under one cell
    dom = hip.make_domain(1.0, 1.0, 64, 64)
    p = {k: rec[:, j].copy() for j, k in enumerate(hip.AOS_FIELDS)}
    for backend in (1, 2):
        a = hip.deposit_charge_aos(rec, dom, scheme=scheme, backend=backend, esirkepov_dt=0.01)
        b = hip.deposit_charge(p, dom, scheme=scheme, backend=backend, esirkepov_dt=0.01)
        if backend == 2:
            assert np.array_equal(a, b)  # deterministic tiles: the same bits
        else:  # FP64 atomics: order-dependent FP64 rounding only
            assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()
    # and the oracle on the FP32-rounded particles (rel L2 <= 1e-6, FP32 scatter)
    f = {k: v.astype(np.float32).astype(np.float64) for k, v in p.items()}
    od = orc.domain(1.0, 1.0, 64, 64)
    ref = orc.deposit(scheme, f["x"], f["y"], f["q"], od, vx=f["vx"], vy=f["vy"], dt=0.01)
    assert np.linalg.norm(a.ravel() - ref.ravel()) <= 1e-6 * np.linalg.norm(ref)


def test_aos_argument_errors(hip):
    say("aos errors")
    dom = hip.make_domain(1.0, 1.0, 16, 16)
    buf = np.zeros(6 * 10 + 2)
    base = buf.ctypes.data
    aligned = base if base % 16 == 0 else base + 8
    mis = aligned + 8  # 8-byte aligned, not 16-byte aligned
    with hip.Context(dom, 10) as c:
        st = hip._lib.pic_hip_upload_aos(c._h, ctypes.cast(mis, ctypes.POINTER(ctypes.c_double)), 10)
        assert st != 0 and b"aligned" in hip._lib.pic_hip_last_error()
        st = hip._lib.pic_hip_upload_aos(c._h, ctypes.cast(aligned, ctypes.POINTER(ctypes.c_double)), 100)
        assert st != 0 and b"capacity" in hip._lib.pic_hip_last_error()
    rho = np.zeros(256)
    st = hip._lib.pic_hip_deposit_host_aos(None, 5, ctypes.byref(dom), 1, 1, 0, 0.0,
                                           rho.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert st != 0 and b"records" in hip._lib.pic_hip_last_error()
