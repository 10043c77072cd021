"""The reference's GPU benchmark matrix on this library, written in the
reference's CSV schema so the rows sit beside its archived GPU results.

BenchmarkRunner::gpuMatrix (src/benchmark/benchmark_runner.cpp:79-111):
particle counts {1e4, 1e5, 1e6} x grids {64^2, 128^2, 256^2} x schemes
{NGP, CIC, TSC} x backends {GPU_Atomics, GPU_Priv (CIC only)}, SoA layout,
unsorted input, 5 repeats.  Each row is timed like depositChargeTimedSoA
(src/deposition/deposition.cpp:257-310): 3 warm-up deposits, then the mean and
sample standard deviation of `repeats` deposits through the synchronous seam
(host FP64 particles in, FP64 rho out: conversion, copies, device sort for
GPU_Priv and the deposit are all inside the timed call, as the reference's
CUDA seam copies in and out).  Bytes and bandwidth follow estimateBytes
(deposition.cpp:136-158); charge error follows chargeConservationError
(src/pic/diagnostics.cpp:26-29).  The CSV header is BenchmarkRunner::writeCsv's
(benchmark_runner.cpp:210-224).  The CPU rows of the matrix are the
reference's own code and are not produced here.

Particles follow initializeUniformMaxwellian (src/pic/particles_aos.cpp:67-81):
uniform positions, unit-temperature Maxwellian velocities, mass 1, charges
+1 / -1 alternating; drawn with numpy (seed 42), not the reference's mt19937
stream.

With --resident, a second CSV in the same schema times the deposit alone on
particles already resident in HBM (context API, HIP-synchronised wall time,
GPU_Priv after one sort): the device-side figure behind each seam row.

gpuMatrix itself is SoA only; --layouts SoA,AoS adds the AoS rows of the CPU
matrices (quickMatrix, benchmark_runner.cpp:15-56) on the device: the seam
rows go through depositChargeAoS's device branch (pic_hip_deposit_host_aos,
48-byte records), the resident rows upload the records with
pic_hip_upload_aos (converted on the device).

    python pic-deposition_amd/gpu_matrix.py --out benchmark_gpu_mi355x.csv [--resident resident.csv] \
        [--layouts SoA,AoS]
"""
import argparse
import csv
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import pichip  # noqa: E402

HEADER = ["num_particles", "grid", "scheme", "layout", "sorted", "backend", "threads", "sort_ms", "deposit_ms",
          "deposit_std_ms", "deposition_ms", "full_step_ms", "throughput", "bandwidth_gbs", "charge_error",
          "energy_drift", "spectral_noise", "speedup"]

SCHEME_NAMES = {0: "NGP", 1: "CIC", 2: "TSC", 3: "Esirkepov"}
BACKEND_NAMES = {1: "GPU_Atomics", 2: "GPU_Priv"}
WRITES_PER_PARTICLE = {0: 1, 1: 4, 2: 9, 3: 4}  # estimateBytes, deposition.cpp:141-155


def gpu_matrix(counts=(10000, 100000, 1000000), grids=(64, 128, 256), schemes=(0, 1, 2), esirkepov=False):
    """(num_particles, grid, scheme, backend) in gpuMatrix order."""
    cases = []
    for n in counts:
        for g in grids:
            for s in list(schemes) + ([3] if esirkepov else []):
                for b in (1, 2):
                    if b == 2 and s != 1 and not esirkepov:
                        continue
                    cases.append((n, g, s, b))
    return cases


def uniform_maxwellian(n, L=1.0, temperature=1.0, seed=42):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.0, L, n)
    y = rng.uniform(0.0, L, n)
    vx = rng.normal(0.0, np.sqrt(temperature), n)
    vy = rng.normal(0.0, np.sqrt(temperature), n)
    q = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    return {"x": x, "y": y, "vx": vx, "vy": vy, "q": q, "m": np.ones(n)}


def estimate_bytes(n, scheme, g, layout="SoA"):
    """bytes read + written of estimateBytes (40-byte SoA, 48-byte AoS record)."""
    return n * (48 if layout == "AoS" else 40) + n * WRITES_PER_PARTICLE[scheme] * 8 + g * g * 8


def charge_error(rho, p, dom):
    cv = (dom.Lx / dom.Nx) * (dom.Ly / dom.Ny)
    return abs(float(rho.sum()) * cv - float(p["q"].sum())) / max(float(np.abs(p["q"]).sum()), 1e-12)


def row(n, g, s, b, times_ms, err, layout="SoA"):
    mean = float(np.mean(times_ms))
    std = float(np.std(times_ms, ddof=1)) if len(times_ms) > 1 else 0.0
    return {"num_particles": n, "grid": f"{g}x{g}", "scheme": SCHEME_NAMES[s], "layout": layout, "sorted": "off",
            "backend": BACKEND_NAMES[b], "threads": 1, "sort_ms": 0, "deposit_ms": mean, "deposit_std_ms": std,
            "deposition_ms": mean, "full_step_ms": 0, "throughput": n / (mean / 1000.0),
            "bandwidth_gbs": estimate_bytes(n, s, g, layout) / 1e9 / (mean / 1000.0), "charge_error": err,
            "energy_drift": 0, "spectral_noise": 0, "speedup": 1}


def seam_case(n, g, s, b, repeats, layout="SoA"):
    dom = pichip.make_domain(1.0, 1.0, g, g)
    p = uniform_maxwellian(n)
    dt = 0.01  # Domain::dt, the reference's esirkepov_dt for the matrix
    if layout == "AoS":
        rec = pichip.to_aos(p)
        call = lambda: pichip.deposit_charge_aos(rec, dom, scheme=s, backend=b, esirkepov_dt=dt)  # noqa: E731
    else:
        call = lambda: pichip.deposit_charge(p, dom, scheme=s, backend=b, esirkepov_dt=dt)  # noqa: E731
    rho = None
    for _ in range(3):
        rho = call()
    times = []
    for _ in range(repeats):
        t = time.perf_counter()
        rho = call()
        times.append((time.perf_counter() - t) * 1000.0)
    return row(n, g, s, b, times, charge_error(rho, p, dom), layout)


def resident_case(n, g, s, b, repeats, layout="SoA"):
    dom = pichip.make_domain(1.0, 1.0, g, g)
    p = uniform_maxwellian(n)
    f32 = {k: v.astype(np.float32) for k, v in p.items()}
    with pichip.Context(dom, n) as c:
        if layout == "AoS":
            c.upload_aos(pichip.to_aos(p))
        else:
            c.upload(f32["x"], f32["y"], f32["q"], f32["vx"], f32["vy"], f32["m"])
        if b == 2:
            c.sort()
        for _ in range(3):
            c.deposit(s, b, 0.01)
        c.sync()
        times = []
        for _ in range(repeats):
            t = time.perf_counter()
            c.deposit(s, b, 0.01)
            c.sync()
            times.append((time.perf_counter() - t) * 1000.0)
        rho = c.rho().astype(np.float64)
    return row(n, g, s, b, times, charge_error(rho, p, dom), layout)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", required=True, help="CSV path for the seam rows")
    ap.add_argument("--resident", default="", help="CSV path for device-resident deposit rows")
    ap.add_argument("--repeats", type=int, default=5)
    ap.add_argument("--esirkepov", action="store_true", help="also Esirkepov, and GPU_Priv for every scheme")
    ap.add_argument("--counts", default="10000,100000,1000000")
    ap.add_argument("--grids", default="64,128,256")
    ap.add_argument("--layouts", default="SoA", help="comma list of SoA, AoS")
    args = ap.parse_args(argv)
    cases = gpu_matrix(tuple(int(v) for v in args.counts.split(",")), tuple(int(v) for v in args.grids.split(",")),
                       esirkepov=args.esirkepov)
    layouts = [v.strip() for v in args.layouts.split(",") if v.strip()]
    if any(v not in ("SoA", "AoS") for v in layouts):
        ap.error("--layouts takes SoA and/or AoS")
    outs = [(args.out, seam_case)] + ([(args.resident, resident_case)] if args.resident else [])
    for path, fn in outs:
        with open(path, "w", newline="") as f:
            w = csv.DictWriter(f, fieldnames=HEADER)
            w.writeheader()
            for n, g, s, b in cases:
                for layout in layouts:
                    r = fn(n, g, s, b, args.repeats, layout)
                    w.writerow(r)
                    f.flush()
                    print(f"{os.path.basename(path)}: {n} {g}x{g} {SCHEME_NAMES[s]} {layout} {BACKEND_NAMES[b]} "
                          f"{r['deposit_ms']:.4f} ms", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
