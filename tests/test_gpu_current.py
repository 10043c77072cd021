"""Charge-conserving Esirkepov current on the GPU: against the FP64 oracle,
the discrete continuity identity (to 1e-6 of |d rho/dt|) at small and full
size, and closed-form special cases."""
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
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def particles(n, N, dt, seed, frac=0.5):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, 1, n).astype(np.float32)
    y = rng.uniform(0, 1, n).astype(np.float32)
    vx = (rng.uniform(-frac, frac, n) / N / dt).astype(np.float32)
    vy = (rng.uniform(-frac, frac, n) / N / dt).astype(np.float32)
    q = rng.uniform(0.5, 1.5, n).astype(np.float32)
    return x, y, vx, vy, q


@pytest.mark.parametrize("N", [64, 100])
def test_current_matches_oracle_and_conserves_charge(hip, orc, N):
    say("current vs oracle", N)
    dt = 0.01
    n = N * N * 32
    x, y, vx, vy, q = particles(n, N, dt, seed=N)
    dom = hip.make_domain(1.0, 1.0, N, N)
    od = orc.domain(1.0, 1.0, N, N)
    with hip.Context(dom, n) as c:
        c.upload(x, y, q, vx, vy)
        c.esirkepov_current(dt)
        jx, jy, r0, r1 = c.current()
        res, drho = c.continuity_residual(dt)
    d = lambda a: a.astype(np.float64)
    o0, o1, ojx, ojy = orc.esirkepov_current(d(x), d(y), d(vx), d(vy), d(q), od, dt)
    assert rel_l2(r0, o0) <= 1e-6 and rel_l2(r1, o1) <= 1e-6
    assert rel_l2(jx, ojx) <= 1e-6 and rel_l2(jy, ojy) <= 1e-6
    assert res <= 1e-6 * drho, (res, drho)
    # rho_old is the CIC density the deposit kernels produce
    ref = orc.deposit(1, d(x), d(y), d(q), od)
    assert rel_l2(r0, ref) <= 1e-6


def test_static_particles_carry_no_current(hip):
    say("static current")
    N, n = 32, 32 * 32 * 8
    x, y, _, _, q = particles(n, N, 0.01, seed=3)
    z = np.zeros(n, np.float32)
    dom = hip.make_domain(1.0, 1.0, N, N)
    with hip.Context(dom, n) as c:
        c.upload(x, y, q, z, z)
        c.esirkepov_current(0.01)
        jx, jy, r0, r1 = c.current()
    assert np.all(jx == 0) and np.all(jy == 0)
    assert np.array_equal(r0, r1)


def test_move_of_a_full_cell_is_rejected(hip):
    say("long move")
    dom = hip.make_domain(1.0, 1.0, 16, 16)
    with hip.Context(dom, 4) as c:
        c.upload(np.float32([0.5]), np.float32([0.5]), np.float32([1.0]), np.float32([1.0 / 16 / 0.01]),
                 np.float32([0.0]))
        c.esirkepov_current(0.01)
        with pytest.raises(hip.PicError, match="one cell or more"):
            c.sync()


def test_full_size_continuity(hip):
    """1024^2, 64 ppc (67,108,864 particles), moves up to half a cell:
    the discrete continuity identity holds to 1e-6 of max |d rho/dt|."""
    say("full-size continuity")
    N = 1024
    n = N * N * 64
    dt = 0.01
    dom = hip.make_domain(1.0, 1.0, N, N)
    with hip.Context(dom, n) as c:
        c.init_particles(n, kind=1, seed=42, amp=0.01, mode=4, q=1.0 / 64, vmax=0.5 / N / dt)
        c.sort()
        c.esirkepov_current(dt)
        res, drho = c.continuity_residual(dt)
        d = c.diagnostics()
        _, _, r0, r1 = c.current()
    print(f"residual {res:.3e} max|drho/dt| {drho:.3e}", flush=True)
    assert res <= 1e-6 * drho
    # total charge unchanged by the move
    vol = 1.0 / (N * N)
    assert r0.astype(np.float64).sum() * vol == pytest.approx(d.total_charge, rel=1e-6)
    assert r1.astype(np.float64).sum() * vol == pytest.approx(d.total_charge, rel=1e-6)


@pytest.mark.parametrize("N,push", [(64, False), (100, False), (64, True), (256, True)])
def test_tiled_current_matches_oracle(hip, orc, N, push):
    """Cell-sorted particles take the tiled kernel (one lane per cell, FP64
    frame sums, LDS windows, merge); after a push without a re-sort some
    particles sit outside their sorted cell and take the global path."""
    say("tiled current", N, "after push" if push else "sorted")
    dt = 0.01
    n = N * N * 24
    x, y, vx, vy, q = particles(n, N, dt, seed=N + 7, frac=0.45)
    dom = hip.make_domain(1.0, 1.0, N, N)
    od = orc.domain(1.0, 1.0, N, N)
    with hip.Context(dom, n) as c:
        c.upload(x, y, q, vx, vy)
        c.sort()
        if push:
            c.gather_push(dt * 0.5, gather=False, push=True)
        c.esirkepov_current(dt)
        jx, jy, r0, r1 = c.current()
        res, drho = c.continuity_residual(dt)
        pt = c.download()
    d = lambda a: a.astype(np.float64)
    o0, o1, ojx, ojy = orc.esirkepov_current(d(pt["x"]), d(pt["y"]), d(pt["vx"]), d(pt["vy"]), d(pt["q"]), od, dt)
    for got, want in ((r0, o0), (r1, o1), (jx, ojx), (jy, ojy)):
        assert rel_l2(got, want) <= 1e-6
    assert res <= 1e-6 * drho, (res, drho)


def test_tiled_current_equals_global_path(hip):
    """The same particles through the tiled kernel (sorted) and the global
    path (unsorted upload of the sorted set): equal to FP32 rounding."""
    say("tiled vs global current")
    N, dt = 128, 0.01
    n = N * N * 16
    x, y, vx, vy, q = particles(n, N, dt, seed=11)
    dom = hip.make_domain(1.0, 1.0, N, N)
    with hip.Context(dom, n) as c:
        c.upload(x, y, q, vx, vy)
        c.sort()
        c.esirkepov_current(dt)
        tiled = [a.copy() for a in c.current()]
        pt = c.download()
    with hip.Context(dom, n) as c:
        c.upload(pt["x"], pt["y"], pt["q"], pt["vx"], pt["vy"])
        c.esirkepov_current(dt)
        glob = c.current()
    for a, b in zip(tiled, glob):
        assert rel_l2(a, b) <= 5e-7
