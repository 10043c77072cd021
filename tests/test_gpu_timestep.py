"""Device timestep (push, deposit, background, FFT Poisson, gather) against the
FP64 oracle loop and the reference-held physics scalars; hipGraph replay
against eager execution."""
import json
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def gold(name):
    return json.load(open(os.path.join(HERE, "golden", "reference_scalars.json")))[name]


def f32_round(p):
    return {k: np.asarray(v, np.float32).astype(np.float64) for k, v in p.items()}


def upload(c, p):
    c.upload(p["x"], p["y"], p["q"], p["vx"], p["vy"], p["m"])


SCHEMES = [(0, "NGP", 0), (1, "CIC", 0), (2, "TSC", 0), (3, "Esirkepov", 2)]


@pytest.mark.parametrize("backend", [1, 2])
@pytest.mark.parametrize("scheme,name,order", SCHEMES)
def test_two_stream_growth_matches_oracle_and_reference(hip, orc, scheme, name, order, backend):
    """two_stream_validation.csv configuration on the device.  Gate: growth
    rate of the GPU run equals the FP64 CPU run's (same FP32-rounded initial
    particles) to 1e-4 relative, and the reference-held value to 1e-4."""
    say("two-stream", name, backend)
    g = gold("two_stream_validation")
    n, G = g["num_particles"], g["grid"]
    od = orc.domain(1.0, 1.0, G, G)
    v = orc.resonant_beam_velocity(od, n, 1)
    p = f32_round(orc.init("two_stream", n, od, seed=g["seed"], beam_velocity=v, perturbation=g["perturbation"]))
    rs = g["steps"] // 4
    _, ref = orc.run_loop(p, od, scheme, order, g["dt"], 1.0, g["steps"], rs)
    dom = hip.make_domain(1.0, 1.0, G, G)
    with hip.Context(dom, n) as c:
        upload(c, p)
        rec = c.run(g["steps"], scheme, backend, g["dt"], 1.0, order, sort_interval=10, record_from=rs)
    wp = hip.plasma_frequency(dom, n)
    gamma_gpu = hip.growth_rate(rec["mode1"], g["dt"], rs) * wp
    gamma_cpu = orc.growth_rate(ref["mode1"], g["dt"], rs) * orc.plasma_frequency(od, n)
    print(f"gamma gpu={gamma_gpu!r} cpu={gamma_cpu!r} ref={g['growth_rate_measured'][name]!r}", flush=True)
    assert gamma_gpu == pytest.approx(gamma_cpu, rel=1e-4)
    assert gamma_gpu == pytest.approx(g["growth_rate_measured"][name], rel=1e-4)
    # NGP and the Esirkepov density have discontinuous weights: rounding to FP32
    # flips a few nearest-node / face-crossing decisions per step.
    np.testing.assert_allclose(rec["mode1"], ref["mode1"], rtol=1e-3 if scheme in (1, 2) else 2e-2)


@pytest.mark.parametrize("backend", [1, 2])
@pytest.mark.parametrize("scheme,name,order", SCHEMES)
def test_quasi1d_two_stream_matches_oracle_and_reference(hip, orc, scheme, name, order, backend):
    """two_stream_quasi1d.csv configuration (256 x 1 grid, Lx = 2 pi): the
    device loop on a one-row grid against the FP64 loop and the reference-held
    value.  Here the per-step displacement (v dt ~ 1e-5) is only ~40 FP32 ulps
    of a position near 2 pi, so every FP32 push rounds the move by ~1%
    (the unit-box configuration moves ~300 ulps per step): CIC and TSC growth
    rates agree to 5e-4 (1.1e-4 measured), and NGP and the Esirkepov density,
    whose discontinuous weights turn that rounding into flipped node and face
    decisions, to 1e-2 (2e-3 measured)."""
    say("quasi-1D two-stream", name, backend)
    g = gold("two_stream_quasi1d")
    n = g["num_particles"]
    od = orc.domain(g["Lx"], g["Ly"], g["Nx"], g["Ny"])
    v = orc.resonant_beam_velocity(od, n, 1)
    p = f32_round(orc.init("two_stream", n, od, seed=g["seed"], beam_velocity=v, perturbation=g["perturbation"],
                           quasi1d=True))
    rs = g["steps"] // 4
    _, ref = orc.run_loop(p, od, scheme, order, g["dt"], 1.0, g["steps"], rs)
    dom = hip.make_domain(g["Lx"], g["Ly"], g["Nx"], g["Ny"])
    with hip.Context(dom, n) as c:
        upload(c, p)
        rec = c.run(g["steps"], scheme, backend, g["dt"], 1.0, order, sort_interval=10, record_from=rs,
                    use_graph=True)
    wp = hip.plasma_frequency(dom, n)
    gamma_gpu = hip.growth_rate(rec["mode1"], g["dt"], rs) * wp
    gamma_cpu = orc.growth_rate(ref["mode1"], g["dt"], rs) * orc.plasma_frequency(od, n)
    print(f"gamma gpu={gamma_gpu!r} cpu={gamma_cpu!r} ref={g['growth_rate_measured'][name]!r}", flush=True)
    tol = 5e-4 if scheme in (1, 2) else 1e-2
    assert gamma_gpu == pytest.approx(gamma_cpu, rel=tol)
    assert gamma_gpu == pytest.approx(g["growth_rate_measured"][name], rel=tol)


def test_langmuir_frequency_matches_reference(hip, orc):
    """validation.csv Langmuir configuration (lattice init, 16384 steps):
    the GPU run's dominant field-energy frequency is the reference's bin."""
    say("langmuir")
    g = gold("langmuir_validation")
    n, G = g["num_particles"], g["grid"]
    od = orc.domain(1.0, 1.0, G, G)
    p = f32_round(orc.init("langmuir", n, od, amplitude=g["amplitude"], mode=g["mode"]))
    dom = hip.make_domain(1.0, 1.0, G, G)
    rs = g["steps"] // 4
    for scheme, order in ((1, 0), (3, 1)):
        with hip.Context(dom, n) as c:
            upload(c, p)
            rec = c.run(g["steps"], scheme, 1, g["dt"], 1.0, order, record_from=rs, use_graph=True)
        assert hip.dominant_frequency(rec["field_energy"], g["dt"]) == g["omega_measured"]


@pytest.mark.parametrize("scheme", [1, 2])
def test_graph_replay_matches_eager(hip, orc, scheme):
    """A captured step replayed many times (with and without re-sort) gives the
    eager run's diagnostics; eager work between graph runs does not disturb it."""
    say("graph replay", scheme)
    G, n = 64, 64 * 64 * 32
    dom = hip.make_domain(1.0, 1.0, G, G)
    od = orc.domain(1.0, 1.0, G, G)
    rng = np.random.default_rng(9)
    p = {"x": rng.uniform(0, 1, n), "y": rng.uniform(0, 1, n), "vx": rng.normal(0, 0.2, n),
         "vy": rng.normal(0, 0.2, n), "q": np.full(n, -1.0 / n), "m": np.ones(n)}
    p = f32_round(p)
    kw = dict(scheme=scheme, backend=2, dt=0.02, background=1.0, sort_interval=5, record_from=0,
              record_particles=True)
    with hip.Context(dom, n) as c:
        upload(c, p)
        eager = c.run(30, use_graph=False, **kw)
        eager_state = c.download()
        upload(c, p)
        g1 = c.run(30, use_graph=True, **kw)
        g1_state = c.download()
        c.deposit(scheme, 1, 0.0)  # unrelated eager work between graph runs
        c.poisson()
        upload(c, p)
        g2 = c.run(30, use_graph=True, **kw)
    for k in eager:
        np.testing.assert_allclose(g1[k], eager[k], rtol=1e-5, atol=1e-12, err_msg=k)
        np.testing.assert_allclose(g2[k], eager[k], rtol=1e-5, atol=1e-12, err_msg=k)
    for k in ("x", "y", "vx", "vy"):
        # particle order differs only if sorts differ; compare sorted multisets
        np.testing.assert_allclose(np.sort(g1_state[k]), np.sort(eager_state[k]), rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("steps,interval,record_from", [(23, 4, 9), (21, 10, 10), (11, 10, 0), (40, 10, 40)])
def test_interval_graphs_match_eager(hip, steps, interval, record_from):
    """Whole sort intervals replay as one graph each: with the records switched
    on inside an interval (23, 4, 9), on an interval boundary (21, 10, 10),
    with a run of one interval and a step (11, 10, 0) and with no record
    (40, 10, 40), the graph run's records and particles equal the eager run's
    bit for bit."""
    say("interval graphs", steps, interval, record_from)
    G, n = 64, 64 * 64 * 16
    dom = hip.make_domain(1.0, 1.0, G, G)
    rng = np.random.default_rng(steps)
    p = {"x": rng.uniform(0, 1, n), "y": rng.uniform(0, 1, n), "vx": rng.normal(0, 0.2, n),
         "vy": rng.normal(0, 0.2, n), "q": np.full(n, -1.0 / n), "m": np.ones(n)}
    kw = dict(scheme=2, backend=2, dt=0.02, background=1.0, sort_interval=interval, record_from=record_from,
              record_particles=True)
    runs = {}
    with hip.Context(dom, n) as c:
        for mode in (False, True):
            upload(c, p)
            rec = c.run(steps, use_graph=mode, **kw)
            runs[mode] = (rec, c.download())
    assert len(runs[True][0]["field_energy"]) == steps - record_from
    for k in runs[False][0]:
        np.testing.assert_array_equal(runs[True][0][k], runs[False][0][k], err_msg=k)
    for k in ("x", "y", "vx", "vy"):
        np.testing.assert_array_equal(runs[True][1][k], runs[False][1][k], err_msg=k)


@pytest.mark.parametrize("shape", [(64, 64), (48, 40)])  # LDS transforms, hipFFT
@pytest.mark.parametrize("use_graph", [False, True])
def test_late_records_and_fields_after_run(hip, use_graph, shape):
    """Diagnostics run only on recorded steps: records taken from step 12 on
    equal the tail of a run recorded from step 0.  The loop's solves leave the
    potential as a spectrum; reading phi after the run transforms it and gives
    the potential of a fresh solve of the final rho (the loop takes the mean of
    rho from the particle charges, a fresh solve from the grid sum)."""
    say("late records", use_graph, shape)
    nx, ny = shape
    n = nx * ny * 16
    dom = hip.make_domain(1.0, 1.0, nx, ny)
    rng = np.random.default_rng(12)
    p = {"x": rng.uniform(0, 1, n), "y": rng.uniform(0, 1, n), "vx": rng.normal(0, 0.05, n),
         "vy": rng.normal(0, 0.05, n), "q": np.full(n, -1.0 / n), "m": np.ones(n)}
    p["x"] = (p["x"] + 0.05 * np.sin(2 * np.pi * p["x"])) % 1.0
    p = f32_round(p)
    kw = dict(scheme=2, backend=2, dt=0.05, background=1.0, sort_interval=1, record_particles=True,
              use_graph=use_graph)
    with hip.Context(dom, n) as c:
        upload(c, p)
        full = c.run(30, record_from=0, **kw)
        upload(c, p)
        late = c.run(30, record_from=12, **kw)
        phi_run, ex_run, ey_run = c.fields()
        rho = c.rho()
        c.poisson()
        phi_new, ex_new, ey_new = c.fields()
        np.testing.assert_array_equal(c.rho(), rho)
    for k in full:
        assert late[k].shape == (18,)
        np.testing.assert_array_equal(late[k], full[k][12:], err_msg=k)
    for a, b in ((phi_run, phi_new), (ex_run, ex_new), (ey_run, ey_new)):
        scale = np.abs(b).max()
        assert scale > 0
        assert np.abs(a - b).max() <= 1e-5 * scale


def test_energy_and_charge_records_match_oracle_loop(hip, orc):
    """Reduced config-3 shape (64x64, 100 ppc, two beams +-0.2, thermal 0.01,
    TSC, sort every 10 steps): per-step field energy, kinetic energy and
    charge error from the device equal the FP64 loop's."""
    say("config3 reduced")
    G = 64
    n = G * G * 100
    dom = hip.make_domain(1.0, 1.0, G, G)
    od = orc.domain(1.0, 1.0, G, G)
    dt = 0.05
    with hip.Context(dom, n) as c:
        c.init_particles(n, kind=2, seed=7, amp=0.2, vth=0.01, q=-1.0 / n, m=1.0)
        p = {k: v.astype(np.float64) for k, v in c.download().items()}
        rec = c.run(40, 2, 2, dt, 1.0, 0, sort_interval=10, record_from=0, record_particles=True, use_graph=True)
    _, ref = orc.run_loop(p, od, 2, 0, dt, 1.0, 40, 0)
    np.testing.assert_allclose(rec["kinetic"], ref["kinetic"], rtol=1e-6)
    np.testing.assert_allclose(rec["field_energy"], ref["field_energy"], rtol=2e-4)
    np.testing.assert_allclose(rec["charge_error"], ref["charge_error"], rtol=1e-5)


def test_full_size_config3_timestep_properties(hip):
    """512x512, 100 ppc (26,214,400 particles), TSC, sort every 10 steps,
    200 steps, reference normalisation dt = 0.1 / omega_p: runs through the
    graph path; total charge and kinetic energy are conserved to FP32 level
    (the reference normalisation makes the dynamics ballistic)."""
    say("config3 full")
    G, ppc = 512, 100
    n = G * G * ppc
    dom = hip.make_domain(1.0, 1.0, G, G)
    wp = hip.plasma_frequency(dom, n)
    dt = 0.1 / wp
    with hip.Context(dom, n) as c:
        c.init_particles(n, kind=2, seed=42, amp=0.2, vth=0.01, q=-1.0 / n, m=1.0)
        rec = c.run(200, 2, 2, dt, 1.0, 0, sort_interval=10, record_from=0, record_particles=True, use_graph=True)
    assert np.all(np.isfinite(rec["field_energy"])) and np.all(np.isfinite(rec["kinetic"]))
    assert np.abs(rec["kinetic"] / rec["kinetic"][0] - 1).max() < 1e-4
    # rho includes the +1/V background, so integrated rho = sum q + 1 = 0 and the
    # reference's charge error is |0 - (-1)| / 1 = 1 at every step
    np.testing.assert_allclose(rec["charge_error"], 1.0, rtol=1e-5)


def test_full_size_config3_dynamics_match_oracle_loop(hip, orc):
    """512x512, 100 ppc (26,214,400 particles), TSC, in a regime with real
    dynamics: unit plasma frequency, the two beams at the mode-1 resonant
    velocity with a 1% velocity perturbation and thermal spread 0.01, dt 0.2,
    sort every step.  The mode-1 field grows and the beams' kinetic energy
    changes (mode-1 amplitude x6, field energy x19, kinetic energy +1.1e-4 in
    the oracle over these 8 steps), and the device's per-step field energy,
    mode-1 amplitude and kinetic energy equal the FP64 oracle loop's on the
    same FP32 particles."""
    say("config3 full dynamics")
    G, ppc = 512, 100
    n = G * G * ppc
    dom = hip.make_domain(1.0, 1.0, G, G)
    od = orc.domain(1.0, 1.0, G, G)
    q = hip.unit_plasma_charge(n, G, G)
    v0 = np.sqrt(3.0 / 8.0) / (2.0 * np.pi)  # k v0 = sqrt(3/8) omega_p for mode 1 at omega_p = 1
    dt, steps = 0.2, 8
    with hip.Context(dom, n) as c:
        c.init_particles(n, kind=2, seed=42, amp=v0, pert=0.01 * v0, vth=0.01, q=q, m=1.0)
        p = {k: v.astype(np.float64) for k, v in c.download().items()}
        rec = c.run(steps, 2, 2, dt, 0.0, 0, sort_interval=1, record_from=0, record_particles=True, use_graph=True)
    _, ref = orc.run_loop(p, od, 2, 0, dt, 0.0, steps, 0)
    fe = ref["field_energy"]
    print(f"field energy {fe[0]:.4e} -> {fe[-1]:.4e}, kinetic {ref['kinetic'][0]:.6e} -> {ref['kinetic'][-1]:.6e}",
          flush=True)
    # the mode-1 field grows: not a ballistic run
    assert ref["mode1"][-1] > 1.5 * ref["mode1"][0] and fe[-1] > 1.5 * fe[0]
    np.testing.assert_allclose(rec["field_energy"], fe, rtol=2e-4)
    np.testing.assert_allclose(rec["mode1"], ref["mode1"], rtol=2e-4)
    np.testing.assert_allclose(rec["kinetic"], ref["kinetic"], rtol=1e-6)
    # without a background the charge error is round-off on both sides: the
    # device's (FP32 rho, FP64 sums) stays at FP32 level
    assert np.all(rec["charge_error"] < 1e-5)


@pytest.mark.parametrize("scheme,backend,interval", [(1, 2, 1), (1, 2, 10), (2, 2, 3), (1, 1, 0), (3, 2, 2)])
def test_run_profiled_phases_and_state(hip, scheme, backend, interval):
    """pic_hip_run_profiled (Simulation::runProfiled, simulation.cpp:147-292):
    every phase time is non-negative, the total is their sum, the sort phase is
    empty without sorting, and the particles it leaves equal those of the same
    phase sequence issued through the single-call entry points."""
    say("run_profiled", scheme, backend, interval)
    N = 128
    n = 20000
    dom = hip.make_domain(1.0, 1.0, N, N)
    p = hip.reference_particles("maxwellian", n, dom, seed=42)
    f = {k: v.astype(np.float32) for k, v in p.items()}
    warm, timed, dt = 2, 6, 0.01
    with hip.Context(dom, n) as a, hip.Context(dom, n) as b:
        for c in (a, b):
            c.upload(f["x"], f["y"], f["q"], f["vx"], f["vy"], f["m"])
        prof = a.run_profiled(warm, timed, scheme, backend, dt, background=0.5, sort_interval=interval)
        print(prof, flush=True)
        for k in ("push_ms", "sort_ms", "deposit_ms", "poisson_ms", "gather_ms"):
            assert prof[k] >= 0.0, k
        total = sum(prof[k] for k in ("push_ms", "sort_ms", "deposit_ms", "poisson_ms", "gather_ms"))
        assert prof["total_ms"] == pytest.approx(total, rel=1e-12)
        assert prof["sort_interval"] == max(1, interval)
        if backend == 1 and interval == 0:
            assert prof["sort_ms"] < 0.01  # two event records, nothing between them
        else:
            assert prof["sort_ms"] > 0.0
        for step in range(warm + timed):
            if (backend == 2 or interval >= 1) and step % max(1, interval) == 0:
                if backend == 2 and scheme in (0, 2):
                    b.sort_nodes()
                else:
                    b.sort()
            if scheme != 3:
                b.gather_push(dt, gather=False, push=True)
            b.deposit(scheme, backend, dt)
            b.add_background(0.5)
            b.poisson()
            if scheme == 3:
                b.gather_push(dt, gather=False, push=True)
            b.gather_push(dt, gather=True, push=False)
        pa, pb = a.download(), b.download()
        # particles that left their tile since the sort go through FP64
        # atomics, which add in scheduling order: equal to FP32 rounding
        for k in ("x", "y", "vx", "vy"):
            np.testing.assert_allclose(pa[k], pb[k], rtol=1e-5, atol=1e-6, err_msg=k)
