"""The reference's validation and study loops on the device backend.

The reference runs these loops through depositChargeSoA with a host round trip
of every particle on every step (validation.cpp:200-443,
physics_studies.cpp:312-353,355-392,479-569).  Here each run is one
pic_hip_run: the initial particles (the reference's own mt19937 draws,
pic_hip_reference_particles) go to the device once, every step (push, deposit,
neutralising background, FFT Poisson, gather, FP64 records) runs there as
replayed hipGraphs, and only the recorded series come back.  The estimators
(dominant frequency, growth rate) and the pass rules are the reference's.

Spectral noise (spectralNoiseLevel, diagnostics.cpp:53-102) is evaluated on the
host from the one deposited rho it needs, as the reference does with FFTW:
the separable R2HC x R2HC power with |kx| >= Nx/3 or |ky| >= Ny/3 over the
total.

Each suite writes the reference's CSV (same header, 10 significant digits,
booleans as 0/1), so the files sit beside data/benchmarks/:

    validation.csv             runValidationSuite       (validation.cpp:263-307)
    two_stream_validation.csv  runTwoStreamValidationSuite (validation.cpp:390-443)
    two_stream_quasi1d.csv     runQuasi1DTwoStreamValidation (physics_studies.cpp:312-353)
    validation_multi_seed.csv  runMultiSeedValidation   (physics_studies.cpp:479-569)
    noise_vs_grid.csv          runNoiseVsGridStudy      (validation.cpp:445-472)
    langmuir_convergence.csv   runLangmuirConvergenceSweep (physics_studies.cpp:355-392)
    conservation_study.csv     runConservationStudy     (validation.cpp:309-337)

    python pic-deposition_amd/validation.py --out-dir results [--suites validation,two_stream,...]
"""
import argparse
import math
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import pichip  # noqa: E402

SCHEMES = ((0, "NGP"), (1, "CIC"), (2, "TSC"), (3, "Esirkepov"))
ESIRKEPOV = 3


def fmt(v) -> str:
    """std::setprecision(10) on a default-formatted stream (%.10g); bools as 0/1."""
    if isinstance(v, (bool, np.bool_)):
        return "1" if v else "0"
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    if isinstance(v, str):
        return v
    return format(float(v), ".10g")


def write_csv(path, header, rows):
    with open(path, "w") as f:
        f.write(",".join(header) + "\n")
        for r in rows:
            f.write(",".join(fmt(r[h]) for h in header) + "\n")


def _r2hc(a, axis):
    """FFTW_R2HC along one axis: r_0 .. r_{n/2}, then i_{(n+1)/2-1} .. i_1."""
    n = a.shape[axis]
    z = np.fft.rfft(a, axis=axis)
    out = np.empty_like(a)
    re = np.moveaxis(z.real, axis, 0)
    im = np.moveaxis(z.imag, axis, 0)
    o = np.moveaxis(out, axis, 0)
    o[: n // 2 + 1] = re[: n // 2 + 1]
    for k in range(1, (n + 1) // 2):
        o[n - k] = im[k]
    return out


def spectral_noise(rho) -> float:
    rho = np.asarray(rho, dtype=np.float64)
    ny, nx = rho.shape
    p = _r2hc(_r2hc(rho, 1), 0) ** 2
    kx = np.minimum(np.arange(nx), nx - np.arange(nx))
    ky = np.minimum(np.arange(ny), ny - np.arange(ny))
    high = (kx[None, :] >= max(1, nx // 3)) | (ky[:, None] >= max(1, ny // 3))
    total = float(p.sum())
    return float(p[high].sum()) / total if total > 1e-30 else 0.0


def charge_error(rho, q, dom) -> float:
    """chargeConservationError(integratedRho, totalCharge, totalAbsCharge), diagnostics.cpp:26-29."""
    integrated = float(np.asarray(rho, np.float64).sum()) * dom.cell_volume()
    qd = np.asarray(q, np.float64)
    return abs(integrated - float(qd.sum())) / max(float(np.abs(qd).sum()), 1e-12)


class Runner:
    """One context per domain and capacity, reused across the runs of a suite."""

    def __init__(self, backend=2, sort_interval=10, use_graph=True):
        self.backend, self.sort_interval, self.use_graph = backend, sort_interval, use_graph
        self._ctx = None
        self._key = None

    def ctx(self, dom, n):
        key = (dom.Lx, dom.Ly, dom.Nx, dom.Ny)
        if self._ctx is None or self._key != key or self._ctx_cap < n:
            self.close()
            self._ctx = pichip.Context(dom, n)
            self._key, self._ctx_cap = key, n
        return self._ctx

    def close(self):
        if self._ctx is not None:
            self._ctx.close()
            self._ctx = None

    def upload(self, dom, p):
        c = self.ctx(dom, len(p["x"]))
        f = {k: np.asarray(v, np.float32) for k, v in p.items()}
        c.upload(f["x"], f["y"], f["q"], f["vx"], f["vy"], f["m"])
        return c

    def deposit(self, dom, p, scheme, dt=0.0):
        c = self.upload(dom, p)
        if self.backend == 2:
            c.sort()
        c.deposit(scheme, self.backend, dt)
        return c.rho().astype(np.float64)

    def run(self, dom, p, scheme, order, steps, dt, background, record_from, record_particles=False):
        c = self.upload(dom, p)
        return c.run(steps, scheme, self.backend, dt, background, order, sort_interval=self.sort_interval,
                     record_from=record_from, record_particles=record_particles, use_graph=self.use_graph)


def langmuir_validation(r: Runner, scheme, grid=64, n=200, dt=0.002, steps=16384, amplitude=0.05, mode=1, seed=42):
    """runLangmuirValidation (validation.cpp:200-261)."""
    dom = pichip.make_domain(1.0, 1.0, grid, grid)
    p = pichip.reference_particles("langmuir", n, dom, seed=seed, amplitude=amplitude, mode=mode)
    order = pichip.ORDER_DEPOSIT_SOLVE_PUSH if scheme == ESIRKEPOV else pichip.ORDER_PUSH_FIRST
    rs = steps // 4
    rec = r.run(dom, p, scheme, order, steps, dt, 1.0 / (dom.Lx * dom.Ly), rs)
    fe = rec["field_energy"]
    theory = pichip.plasma_frequency(dom, n)
    measured = pichip.dominant_frequency(fe, dt)
    out = {"omega_theory": theory, "omega_measured": measured, "relative_error": 0.0, "omega_ratio": 0.0,
           "passed": False, "amplitude_decay": 0.0}
    if theory > 0:
        out["relative_error"] = abs(measured - theory) / theory
        out["omega_ratio"] = measured / theory
    if len(fe) >= 2:
        out["amplitude_decay"] = (fe[0] - fe[-1]) / max(abs(fe[0]), 1e-12)
    return out


def two_stream_validation(r: Runner, scheme, Lx=1.0, Ly=1.0, Nx=64, Ny=64, n=5000, dt=0.002, steps=2000,
                          perturbation=0.01, seed=42, quasi1d=False):
    """runTwoStreamValidation (validation.cpp:339-388) with resonant beams."""
    dom = pichip.make_domain(Lx, Ly, Nx, Ny)
    v = pichip.resonant_beam_velocity(dom, n, 1)
    p = pichip.reference_particles("two_stream_1d" if quasi1d else "two_stream", n, dom, seed=seed,
                                   beam_velocity=v, perturbation=perturbation)
    order = pichip.ORDER_DEPOSIT_PUSH_SOLVE if scheme == ESIRKEPOV else pichip.ORDER_PUSH_FIRST
    rs = steps // 4
    rec = r.run(dom, p, scheme, order, steps, dt, 1.0 / (dom.Lx * dom.Ly), rs)
    wp = pichip.plasma_frequency(dom, n)
    theory = pichip.cold_two_stream_growth_rate(dom, n)
    measured = pichip.growth_rate(rec["mode1"], dt, rs) * wp
    return {"scheme": scheme, "growth_rate_measured": measured, "growth_rate_theory": theory,
            "growth_rate_ratio": measured / theory if theory > 0 else 0.0, "omega_p_macro": wp,
            "growth_rate_over_omega_p": measured / wp if wp > 0 else 0.0, "passed": False}


def _mark_spread(rows, key, tol, use):
    vals = [r[key] for r in rows if use(r)]
    if not vals:
        return
    ref = sum(vals) / len(vals)
    for r in rows:
        if r[key] <= 0.0:
            r["passed"] = False
        elif ref > 0:
            r["passed"] = abs(r[key] - ref) / ref <= tol


def validation_suite(r: Runner, grid=64, n=200, dt=0.002, steps=16384, amplitude=0.05, mode=1, seed=42, T=1.0):
    """runValidationSuite: per scheme, noise and charge error of a Maxwellian
    deposit (esirkepov_dt = 0, as DepositionConfig's default), then the
    Langmuir run; passed = omega within 5% of the schemes' mean."""
    dom = pichip.make_domain(1.0, 1.0, grid, grid)
    rows = []
    for s, name in SCHEMES:
        p = pichip.reference_particles("maxwellian", n, dom, seed=seed, temperature=T)
        rho = r.deposit(dom, p, s, 0.0)
        q32 = np.asarray(p["q"], np.float32)
        lang = langmuir_validation(r, s, grid, n, dt, steps, amplitude, mode, seed)
        rows.append({"scheme": name, "spectral_noise": spectral_noise(rho), "charge_error": charge_error(rho, q32, dom),
                     **{k: lang[k] for k in ("omega_theory", "omega_measured", "relative_error", "omega_ratio")},
                     "passed": False})
    omega_ref = sum(x["omega_measured"] for x in rows) / len(rows)
    if omega_ref > 0:
        for x in rows:
            x["passed"] = abs(x["omega_measured"] - omega_ref) / omega_ref <= 0.05
    return rows


def two_stream_suite(r: Runner, **kw):
    """runTwoStreamValidationSuite: passed = growth within 10% of the mean of the
    NGP/CIC/TSC rates (Esirkepov reported against it)."""
    rows = [two_stream_validation(r, s, **kw) for s, _ in SCHEMES]
    _mark_spread(rows, "growth_rate_measured", 0.10, lambda x: x["scheme"] != ESIRKEPOV and x["growth_rate_measured"] > 0)
    for x in rows:
        x["scheme"] = dict(SCHEMES)[x["scheme"]]
    return rows


def quasi1d_suite(r: Runner):
    """runQuasi1DTwoStreamValidation: Lx = 2 pi, 256 x 1; passed = within 10% of
    the mean of all positive rates."""
    rows = [two_stream_validation(r, s, Lx=2.0 * math.pi, Ly=1.0, Nx=256, Ny=1, quasi1d=True) for s, _ in SCHEMES]
    _mark_spread(rows, "growth_rate_measured", 0.10, lambda x: x["growth_rate_measured"] > 0)
    for x in rows:
        x["scheme"] = dict(SCHEMES)[x["scheme"]]
        x["geometry"] = "quasi1d"
    return rows


def multi_seed(r: Runner, seeds=(42, 43, 44, 45, 46)):
    """runMultiSeedValidation: Langmuir (16384 steps, 200 particles) and
    two-stream (2000 steps, 5000 particles) per seed and scheme."""
    rows = []
    for seed in seeds:
        lang = [{"seed": seed, "test": "langmuir", "scheme": name, "metric": "omega_measured",
                 "value": langmuir_validation(r, s, seed=seed)["omega_measured"], "passed": False}
                for s, name in SCHEMES]
        ref = sum(x["value"] for x in lang) / len(lang)
        for x in lang:
            x["passed"] = ref > 0 and abs(x["value"] - ref) / ref <= 0.05
        ts = [{"seed": seed, "test": "twostream", "scheme": name, "metric": "growth_rate_measured",
               "value": two_stream_validation(r, s, seed=seed)["growth_rate_measured"], "passed": False}
              for s, name in SCHEMES]
        pos = [x["value"] for x in ts if x["value"] > 0]
        if pos:
            g = sum(pos) / len(pos)
            for x in ts:
                if x["value"] > 0 and g > 0:
                    x["passed"] = abs(x["value"] - g) / g <= 0.10
        rows += lang + ts
    return rows


def noise_vs_grid(r: Runner, n=100000, grids=(64, 128, 256)):
    """runNoiseVsGridStudy: Maxwellian (seed 42), sorted deposit, esirkepov_dt 0."""
    rows = []
    for g in grids:
        dom = pichip.make_domain(1.0, 1.0, g, g)
        p = pichip.reference_particles("maxwellian", n, dom, seed=42, temperature=1.0)
        for s, name in SCHEMES:
            rows.append({"grid": f"{g}x{g}", "scheme": name, "spectral_noise": spectral_noise(r.deposit(dom, p, s))})
    return rows


def langmuir_convergence(r: Runner, grids=(64, 128, 256), counts=(200, 2000, 20000), steps=8192):
    """runLangmuirConvergenceSweep: 8192 steps, dt 0.002, amplitude 0.05, mode 1."""
    rows = []
    for g in grids:
        for n in counts:
            for s, name in SCHEMES:
                lang = langmuir_validation(r, s, grid=g, n=n, steps=steps)
                rows.append({"grid_n": g, "num_particles": n, "scheme": name, "omega_measured": lang["omega_measured"],
                             "omega_theory": lang["omega_theory"], "omega_ratio": lang["omega_ratio"]})
    return rows


def conservation_study(r: Runner, grid=64, n=10000, dt=0.005, steps=2000, seed=42, T=1.0):
    """runConservationStudy -> Simulation::run (simulation.cpp:69-145): Maxwellian,
    no background, sorted; max charge error and total-energy drift over the run."""
    dom = pichip.make_domain(1.0, 1.0, grid, grid)
    rows = []
    for s, name in SCHEMES:
        p = pichip.reference_particles("maxwellian", n, dom, seed=seed, temperature=T)
        order = pichip.ORDER_DEPOSIT_SOLVE_PUSH if s == ESIRKEPOV else pichip.ORDER_PUSH_FIRST
        rec = r.run(dom, p, s, order, steps, dt, 0.0, 0, record_particles=True)
        e = rec["kinetic"] + rec["field_energy"]
        rows.append({"scheme": name, "steps": steps, "max_charge_error": float(np.max(rec["charge_error"])),
                     "final_energy_drift": (e[-1] - e[0]) / max(abs(e[0]), 1e-12)})
    return rows


SUITES = {
    "validation": ("validation.csv", validation_suite,
                   ["scheme", "spectral_noise", "charge_error", "omega_theory", "omega_measured", "relative_error",
                    "omega_ratio", "passed"]),
    "two_stream": ("two_stream_validation.csv", two_stream_suite,
                   ["scheme", "growth_rate_measured", "growth_rate_theory", "growth_rate_ratio", "omega_p_macro",
                    "growth_rate_over_omega_p", "passed"]),
    "quasi1d": ("two_stream_quasi1d.csv", quasi1d_suite,
                ["geometry", "scheme", "growth_rate_measured", "growth_rate_theory", "growth_rate_ratio",
                 "omega_p_macro", "growth_rate_over_omega_p", "passed"]),
    "multi_seed": ("validation_multi_seed.csv", multi_seed, ["seed", "test", "scheme", "metric", "value", "passed"]),
    "noise_vs_grid": ("noise_vs_grid.csv", noise_vs_grid, ["grid", "scheme", "spectral_noise"]),
    "langmuir_convergence": ("langmuir_convergence.csv", langmuir_convergence,
                             ["grid_n", "num_particles", "scheme", "omega_measured", "omega_theory", "omega_ratio"]),
    "conservation": ("conservation_study.csv", conservation_study,
                     ["scheme", "steps", "max_charge_error", "final_energy_drift"]),
}


def main(argv=None):
    ap = argparse.ArgumentParser(description="the reference's validation suites on the device backend")
    ap.add_argument("--out-dir", required=True)
    ap.add_argument("--suites", default=",".join(SUITES))
    ap.add_argument("--backend", type=int, default=2, help="1 GPU_Atomics, 2 GPU_Priv")
    ap.add_argument("--sort-interval", type=int, default=10)
    args = ap.parse_args(argv)
    names = [s.strip() for s in args.suites.split(",") if s.strip()]
    unknown = [s for s in names if s not in SUITES]
    if unknown:
        ap.error(f"unknown suites {unknown}; choose from {list(SUITES)}")
    os.makedirs(args.out_dir, exist_ok=True)
    r = Runner(args.backend, args.sort_interval)
    try:
        for name in names:
            fname, fn, header = SUITES[name]
            t = time.perf_counter()
            rows = fn(r)
            write_csv(os.path.join(args.out_dir, fname), header, rows)
            print(f"{fname}: {len(rows)} rows in {time.perf_counter() - t:.2f} s", flush=True)
    finally:
        r.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
