"""Gather/push, spectral Poisson, FP64 diagnostics and the on-device
initialiser against the oracle."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("dt", [0.01, 3.7])
def test_gather_push_vs_oracle(hip, orc, dt):
    say("gather/push", dt)
    nx, ny = 48, 32
    dom = hip.make_domain(1.0, 2.0, nx, ny)
    od = orc.domain(1.0, 2.0, nx, ny)
    rng = np.random.default_rng(0)
    n = 100_000
    p = {"x": rng.uniform(0, 1, n), "y": rng.uniform(0, 2, n), "vx": rng.normal(size=n), "vy": rng.normal(size=n),
         "q": rng.uniform(-1, 1, n), "m": rng.uniform(0.5, 2, n)}
    p = {k: v.astype(np.float32) for k, v in p.items()}
    ex = rng.normal(size=(ny, nx)).astype(np.float32)
    ey = rng.normal(size=(ny, nx)).astype(np.float32)
    with hip.Context(dom, n) as c:
        c.upload(p["x"], p["y"], p["q"], p["vx"], p["vy"], p["m"])
        c.set_fields(ex, ey)
        c.gather_push(dt, gather=True, push=False)
        g1 = c.download()
        c.gather_push(dt, gather=False, push=True)
        g2 = c.download()
    vx, vy = orc.gather(p["x"], p["y"], p["vx"], p["vy"], p["q"], p["m"], ex, ey, od, dt)
    # FP32 error bound per particle: a few ulps of |v0| + |(q/m) E dt| (the
    # interpolation sums four terms, the update two).
    # |(q/m)| * sum_k |E_k| w_k * dt: gathering |E| into zero velocities
    ax, ay = orc.gather(p["x"], p["y"], 0 * p["vx"], 0 * p["vy"], p["q"], p["m"], np.abs(ex), np.abs(ey), od, dt)
    eps = np.finfo(np.float32).eps
    scale_x = np.abs(p["vx"].astype(np.float64)) + np.abs(ax)
    scale_y = np.abs(p["vy"].astype(np.float64)) + np.abs(ay)
    assert np.all(np.abs(g1["vx"] - vx) <= 8 * eps * scale_x)
    assert np.all(np.abs(g1["vy"] - vy) <= 8 * eps * scale_y)
    assert rel_l2(g1["vx"], vx) < 2e-6 and rel_l2(g1["vy"], vy) < 2e-6
    x, y = orc.push(p["x"], p["y"], g1["vx"], g1["vy"], od, dt)
    # positions: compare on the torus (a wrap can land on either side of 0 / L)
    dxp = np.abs(((g2["x"] - x) + 0.5) % 1.0 - 0.5)
    dyp = np.abs(((g2["y"] - y) + 1.0) % 2.0 - 1.0)
    tol = 4 * np.finfo(np.float32).eps * (1 + np.abs(np.asarray(g1["vx"], np.float64) * dt).max())
    assert dxp.max() < max(tol, 1e-6) and dyp.max() < max(2 * tol, 2e-6)
    assert g2["x"].min() >= 0 and g2["x"].max() < 1.0 and g2["y"].min() >= 0 and g2["y"].max() < 2.0


def test_gather_constant_and_linear_fields(hip):
    say("gather closed forms")
    nx = ny = 16
    dom = hip.make_domain(1.0, 1.0, nx, ny)
    n = 4096
    rng = np.random.default_rng(1)
    x = rng.uniform(0, 15 / 16, n).astype(np.float32)  # keep off the periodic seam for the linear field
    y = rng.uniform(0, 1, n).astype(np.float32)
    q = np.full(n, -1.0, np.float32)
    with hip.Context(dom, n) as c:
        c.upload(x, y, q, np.zeros(n, np.float32), np.zeros(n, np.float32), np.ones(n, np.float32))
        c.set_fields(np.full((ny, nx), 2.0, np.float32), np.full((ny, nx), -3.0, np.float32))
        c.gather_push(0.5, gather=True, push=False)
        o = c.download()
        np.testing.assert_allclose(o["vx"], 1.0, rtol=1e-6)   # -(q/m) E dt = 2 * 0.5
        np.testing.assert_allclose(o["vy"], -1.5, rtol=1e-6)
        ii = np.tile(np.arange(nx, dtype=np.float32), (ny, 1))
        c.upload(x, y, q, np.zeros(n, np.float32), np.zeros(n, np.float32), np.ones(n, np.float32))
        c.set_fields(ii, np.zeros((ny, nx), np.float32))
        c.gather_push(1.0, gather=True, push=False)
        o = c.download()
    np.testing.assert_allclose(o["vx"], x.astype(np.float64) * nx, rtol=2e-6, atol=2e-6)


@pytest.mark.parametrize("shape", [(64, 64, 1.0, 1.0), (128, 32, 2.0, 1.0), (512, 512, 1.0, 1.0), (1024, 1024, 1.0, 1.0),
                                   (8, 4, 1.0, 0.5), (256, 1, 6.283185307179586, 1.0), (16, 2, 1.0, 1.0),
                                   (100, 60, 1.0, 1.0), (2048, 16, 1.0, 1.0)])
def test_poisson_vs_oracle(hip, orc, shape):
    """Power-of-two grids up to 1024 per axis take the LDS transforms, the
    others hipFFT; both against the FP64 oracle."""
    nx, ny, lx, ly = shape
    say("poisson", shape)
    dom = hip.make_domain(lx, ly, nx, ny)
    od = orc.domain(lx, ly, nx, ny)
    rng = np.random.default_rng(2)
    xs = np.arange(nx) / nx * lx
    ys = np.arange(ny) / ny * ly
    rho = (1.0 + 0.3 * np.cos(2 * np.pi * 3 * xs / lx)[None, :] * np.sin(2 * np.pi * ys / ly)[:, None]
           + 0.05 * rng.normal(size=(ny, nx))).astype(np.float32)
    with hip.Context(dom, 1) as c:
        c.set_rho(rho)
        c.poisson()
        phi, ex, ey = c.fields()
    rphi, rex, rey = orc.poisson(rho.astype(np.float64), od)
    assert rel_l2(phi, rphi) < 5e-6
    assert rel_l2(ex, rex) < 1e-5
    if ny == 1:
        assert np.all(ey == 0)  # phi is constant along y
    else:
        assert rel_l2(ey, rey) < 1e-5


def test_background_and_diagnostics_vs_oracle(hip, orc):
    say("diagnostics")
    dom = hip.make_domain(1.0, 1.0, 64, 64)
    od = orc.domain(1.0, 1.0, 64, 64)
    n = 200_000
    rng = np.random.default_rng(3)
    p = {k: rng.uniform(-1, 1, n).astype(np.float32) for k in ("vx", "vy", "q")}
    p["m"] = rng.uniform(0.5, 2, n).astype(np.float32)
    p["x"] = rng.uniform(0, 1, n).astype(np.float32)
    p["y"] = rng.uniform(0, 1, n).astype(np.float32)
    with hip.Context(dom, n) as c:
        c.upload(p["x"], p["y"], p["q"], p["vx"], p["vy"], p["m"])
        c.sort()
        c.deposit(1, 2, 0.0)
        c.add_background(0.25)
        c.poisson()
        d = c.diagnostics()
        rho = c.rho()
        _, ex, ey = c.fields()
        q = c.download()
    assert d.integrated_rho == pytest.approx(orc.integrated_rho(rho, od), rel=1e-12)
    assert d.field_energy == pytest.approx(orc.field_energy(ex, ey, od), rel=1e-12)
    assert d.kinetic_energy == pytest.approx(orc.kinetic_energy(q["vx"], q["vy"], q["m"]), rel=1e-12)
    assert d.total_charge == pytest.approx(orc.total_charge(q["q"]), rel=1e-10, abs=1e-9)
    assert d.total_abs_charge == pytest.approx(orc.total_abs_charge(q["q"]), rel=1e-12)
    # integrated rho = total charge + background * volume
    assert d.integrated_rho == pytest.approx(d.total_charge + 0.25, rel=1e-5, abs=1e-3)


def test_init_perturbed_density(hip):
    say("init")
    dom = hip.make_domain(1.0, 1.0, 256, 256)
    n = 256 * 256 * 64
    with hip.Context(dom, n) as c:
        c.init_particles(n, kind=1, seed=42, amp=0.05, mode=4, q=1.0 / 64)
        a = c.download()
        c.init_particles(n, kind=1, seed=42, amp=0.05, mode=4, q=1.0 / 64)
        b = c.download()
    assert np.array_equal(a["x"], b["x"]) and np.array_equal(a["y"], b["y"])
    assert a["x"].min() >= 0 and a["x"].max() < 1 and a["y"].min() >= 0 and a["y"].max() < 1
    h, _ = np.histogram(a["x"], bins=256, range=(0, 1))
    centers = (np.arange(256) + 0.5) / 256
    model = n / 256 * (1 + 0.05 * np.cos(2 * np.pi * 4 * centers))
    # amplitude of mode 4 recovered; residual is shot noise
    amp = 2 * np.mean((h / (n / 256) - 1) * np.cos(2 * np.pi * 4 * centers))
    assert amp == pytest.approx(0.05, abs=0.003)
    assert np.abs(h - model).max() < 6 * np.sqrt(n / 256)
