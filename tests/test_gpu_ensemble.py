"""Batched ensemble on the GPU: per-member parity of initialisation, deposit
(against the FP64 oracle to 1e-6 relative L2 and bit for bit against a
single-grid context with the same particles and tiling), Poisson fields and
the graph-captured step loop against single-grid runs; the autotuner and its
cache; the 64-member Langmuir ensemble at full size through per-member charge
conservation."""
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
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def mixed_members(hip):
    """Uneven grids (power-of-two and not), schemes, counts and seeds."""
    M = hip.Member
    return [
        M(Nx=32, Ny=32, num_particles=32 * 32 * 16, scheme=1, seed=11, amp=0.05, q=1 / 16.0),
        M(Nx=64, Ny=64, num_particles=64 * 64 * 24 + 5, scheme=2, seed=12, amp=0.02, q=1 / 24.0),
        M(Nx=48, Ny=40, num_particles=48 * 40 * 10, scheme=0, seed=13, amp=0.01, q=0.1),
        M(Nx=64, Ny=64, num_particles=64 * 64 * 8, scheme=1, seed=14, amp=0.03, q=0.125, Lx=2.0, Ly=0.5),
        M(Nx=128, Ny=64, num_particles=128 * 64 * 4 + 1, scheme=2, seed=15, amp=0.04, q=0.25),
        M(Nx=40, Ny=40, num_particles=40 * 40 * 16, scheme=3, seed=16, amp=0.02, vmax=0.4 / 40 / 0.01, q=1 / 16.0),
        M(Nx=32, Ny=32, num_particles=3, scheme=1, seed=17, amp=0.0, q=1.0),
        M(Nx=64, Ny=64, num_particles=64 * 64 * 24, scheme=2, seed=18, amp=0.001, q=1 / 24.0),
    ]


def test_member_init_equals_single_context(hip):
    say("ensemble init")
    members = mixed_members(hip)
    with hip.Ensemble(members) as e:
        e.init()
        for i, mb in enumerate(members):
            got = e.download(i)
            with hip.Context(mb.domain(), mb.num_particles) as c:
                c.init_particles(mb.num_particles, kind=mb.kind, seed=mb.seed, amp=mb.amp, vmax=mb.vmax,
                                 mode=mb.mode, q=mb.q, m=mb.m)
                ref = c.download()
            for k in ref:
                assert np.array_equal(got[k], ref[k]), (i, k)


def test_deposit_matches_oracle_and_single_context(hip, orc):
    say("ensemble deposit")
    members = mixed_members(hip)
    dt = 0.01
    with hip.Ensemble(members) as e:
        e.init()
        e.sort()
        e.deposit(dt)
        for i, mb in enumerate(members):
            rho = e.rho(i)
            p = e.download(i)
            t = e.tiling(i)
            d = {k: v.astype(np.float64) for k, v in p.items()}
            od = orc.domain(mb.Lx, mb.Ly, mb.Nx, mb.Ny)
            ref = orc.deposit(mb.scheme, d["x"], d["y"], d["q"], od, vx=d["vx"], vy=d["vy"], dt=dt)
            err = rel_l2(rho, ref)
            print(f"member {i} scheme {mb.scheme} {mb.Nx}x{mb.Ny} n={mb.num_particles} rel_l2={err:.2e}", flush=True)
            assert err <= 1e-6
            with hip.Context(mb.domain(), mb.num_particles) as c:
                c.upload(p["x"], p["y"], p["q"], p["vx"], p["vy"], p["m"])
                c.set_tiling(t.tile_x, t.tile_y, t.particles_per_lane, 0, scheme=mb.scheme)
                if mb.scheme in (0, 2):  # the ensemble sorts NGP and TSC members by nearest node
                    c.sort_nodes()
                else:
                    c.sort()
                c.deposit(mb.scheme, 2, dt)
                single = c.rho()
            assert np.array_equal(rho, single), i


@pytest.mark.parametrize("halo", [0, 2])
def test_drifted_particles_between_sorts(hip, orc, halo):
    """Two steps with one sort: the second deposit sees particles up to three
    cells away from their sorted tiles (overflow path), and stays exact
    against the oracle on the positions it deposited; a repeated deposit on
    the same stale order (overflow grids cleared by the merge) agrees to the
    last bits (overflow atomics land in scheduling order)."""
    say("ensemble drift", halo)
    members = mixed_members(hip)[:5]
    dt = 0.01
    for mb in members:
        mb.vmax = 3.0 * min(mb.Lx / mb.Nx, mb.Ly / mb.Ny) / dt
    with hip.Ensemble(members) as e:
        e.set_tile_halo(halo)
        e.init()
        e.run(2, dt, sort_interval=5)  # step 1 deposits without re-sorting; the last step does not push
        first = [e.rho(i) for i in range(len(members))]
        e.deposit(dt)
        for i, mb in enumerate(members):
            assert rel_l2(e.rho(i), first[i]) <= 1e-7, i
            p = {k: v.astype(np.float64) for k, v in e.download(i).items()}
            ref = orc.deposit(mb.scheme, p["x"], p["y"], p["q"], orc.domain(mb.Lx, mb.Ly, mb.Nx, mb.Ny))
            err = rel_l2(first[i], ref)
            print(f"member {i} drift rel_l2={err:.2e}", flush=True)
            assert err <= 1e-6


def test_poisson_matches_single_context(hip):
    say("ensemble poisson")
    members = mixed_members(hip)[:5]
    with hip.Ensemble(members) as e:
        e.init()
        e.sort()
        e.deposit(0.0)
        e.poisson()
        for i, mb in enumerate(members):
            ex, ey = e.fields(i)
            with hip.Context(mb.domain(), max(mb.num_particles, 1)) as c:
                c.set_rho(e.rho(i))
                c.poisson()
                _, rex, rey = c.fields()
            scale = max(np.abs(rex).max(), np.abs(rey).max(), 1e-30)
            assert np.abs(ex - rex).max() <= 1e-5 * scale, i
            assert np.abs(ey - rey).max() <= 1e-5 * scale, i


@pytest.mark.parametrize("use_graph", [True, False])
def test_run_matches_single_contexts(hip, use_graph):
    """Graph-captured ensemble steps (push, sort every 5th step, deposit,
    Poisson, gather) against one single-grid run per member (same initial
    particles, tiling and step order, no background)."""
    say("ensemble run", use_graph)
    members = hip.langmuir_ensemble(count=6, grids=(32, 64), ppc=(8, 32), seed=5)
    steps, dt = 30, 0.1
    with hip.Ensemble(members) as e:
        e.init()
        init = [e.download(i) for i in range(len(members))]
        tilings = [e.tiling(i) for i in range(len(members))]
        rec = e.run(steps, dt, sort_interval=5, use_graph=use_graph)
        final = [e.download(i) for i in range(len(members))]
    for i, mb in enumerate(members):
        p = init[i]
        with hip.Context(mb.domain(), mb.num_particles) as c:
            c.upload(p["x"], p["y"], p["q"], p["vx"], p["vy"], p["m"])
            t = tilings[i]
            c.set_tiling(t.tile_x, t.tile_y, t.particles_per_lane, 0, scheme=mb.scheme)
            ref = c.run(steps, mb.scheme, 2, dt, 0.0, hip.ORDER_PUSH_FIRST, sort_interval=5, record_from=0,
                        record_particles=True)
            ref_final = c.download()
        np.testing.assert_allclose(rec["field_energy"][i], ref["field_energy"], rtol=1e-4, err_msg=str(i))
        np.testing.assert_allclose(rec["mode1"][i], ref["mode1"], rtol=1e-4, atol=1e-12, err_msg=str(i))
        np.testing.assert_allclose(rec["kinetic"][i], ref["kinetic"], rtol=1e-4, err_msg=str(i))
        assert np.all(rec["charge_error"][i] < 1e-5)
        for k in ("x", "y", "vx", "vy"):
            np.testing.assert_allclose(np.sort(final[i][k]), np.sort(ref_final[k]), rtol=1e-4, atol=1e-6)


def test_graph_and_eager_runs_agree(hip):
    say("ensemble graph vs eager")
    members = hip.langmuir_ensemble(count=5, grids=(32, 64, 128), ppc=(4, 16), seed=9)
    out = []
    for g in (True, False, True):
        with hip.Ensemble(members) as e:
            e.init()
            out.append(e.run(25, 0.1, sort_interval=4, use_graph=g))
    for k in out[0]:
        np.testing.assert_allclose(out[1][k], out[0][k], rtol=1e-12, atol=0, err_msg=k)
        np.testing.assert_allclose(out[2][k], out[0][k], rtol=1e-12, atol=0, err_msg=k)


def test_autotune_and_cache(hip, orc, tmp_path):
    say("autotune")
    cache = str(tmp_path / "tiling.json")
    N, ppc = 128, 32
    n = N * N * ppc
    dom = hip.make_domain(1.0, 1.0, N, N)
    with hip.Context(dom, n) as c:
        c.init_particles(n, kind=1, seed=3, amp=0.01, mode=4, q=1.0 / ppc)
        c.sort()
        t1, ms1 = c.autotune(2, cache, reps=3)
        assert ms1 > 0.0
        assert c.tiling(2).tile_x == t1.tile_x and c.tiling(2).tile_y == t1.tile_y
        c.set_tiling(16, 8, 4, 0, scheme=2)
        t2, ms2 = c.autotune(2, cache, reps=3)
        assert ms2 == 0.0 and (t2.tile_x, t2.tile_y, t2.particles_per_lane) == (t1.tile_x, t1.tile_y,
                                                                                 t1.particles_per_lane)
        c.deposit(2, 2)
        rho = c.rho()
        p = {k: v.astype(np.float64) for k, v in c.download().items()}
    ref = orc.deposit(2, p["x"], p["y"], p["q"], orc.domain(1.0, 1.0, N, N))
    assert rel_l2(rho, ref) <= 1e-6
    text = open(cache).read()
    assert "/TSC/128x128/ppc32" in text

    members = mixed_members(hip)[:5]
    with hip.Ensemble(members) as e:
        e.init()
        e.sort()
        e.autotune(cache, reps=2)
        e.deposit(0.0)
        for i, mb in enumerate(members):
            p = {k: v.astype(np.float64) for k, v in e.download(i).items()}
            ref = orc.deposit(mb.scheme, p["x"], p["y"], p["q"], orc.domain(mb.Lx, mb.Ly, mb.Nx, mb.Ny))
            assert rel_l2(e.rho(i), ref) <= 1e-6
    assert text.count("ppc") < open(cache).read().count("ppc")


def test_invalid_arguments(hip):
    say("ensemble errors")
    with pytest.raises(hip.PicError, match="scheme"):
        hip.Ensemble([hip.Member(Nx=8, Ny=8, num_particles=10, scheme=7)])
    with hip.Ensemble([hip.Member(Nx=8, Ny=8, num_particles=10, scheme=3)]) as e:
        with pytest.raises(hip.PicError, match="4 particles per lane"):
            e.set_tiling(8, 4, 8)
        with pytest.raises(hip.PicError, match="out of range"):
            e.tiling(1)
        e.init()
        with pytest.raises(hip.PicError, match="cell order"):
            e.deposit(0.0)


def test_full_langmuir_ensemble_conserves_charge(hip, orc):
    """64 members, grids 64^2..512^2, 16-256 particles per cell, CIC/TSC
    mixed (about 0.64e9 particles): one sort + deposit; every member's
    integrated rho equals its total charge to 1e-6, and the smallest members
    match the oracle to 1e-6 relative L2."""
    say("full ensemble")
    members = hip.langmuir_ensemble()
    with hip.Ensemble(members) as e:
        e.init()
        e.sort()
        e.deposit(0.0)
        e.sync()
        order = sorted(range(len(members)), key=lambda i: members[i].num_particles)
        for i, mb in enumerate(members):
            rho = e.rho(i).astype(np.float64)
            total = rho.sum() * (mb.Lx / mb.Nx) * (mb.Ly / mb.Ny)
            assert total == pytest.approx(mb.q * mb.num_particles, rel=1e-6), i
        for i in order[:3]:
            mb = members[i]
            p = {k: v.astype(np.float64) for k, v in e.download(i).items()}
            ref = orc.deposit(mb.scheme, p["x"], p["y"], p["q"], orc.domain(mb.Lx, mb.Ly, mb.Nx, mb.Ny))
            assert rel_l2(e.rho(i), ref) <= 1e-6
