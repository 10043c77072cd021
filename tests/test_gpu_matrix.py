"""The reference's GPU benchmark matrix (BenchmarkRunner::gpuMatrix) through
pic-deposition_amd/gpu_matrix.py: rows in the reference's CSV schema, charge
conserved, seam and device-resident figures both present."""
import csv
import os
import sys

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "pic-deposition_amd"))

# BenchmarkRunner::writeCsv, src/benchmark/benchmark_runner.cpp:210-211
REFERENCE_HEADER = ("num_particles,grid,scheme,layout,sorted,backend,threads,sort_ms,deposit_ms,deposit_std_ms,"
                    "deposition_ms,full_step_ms,throughput,bandwidth_gbs,charge_error,energy_drift,spectral_noise,"
                    "speedup")


@pytest.mark.parametrize("esirkepov", [False, True])
def test_matrix_rows(tmp_path, esirkepov):
    import gpu_matrix
    print("[case] gpu matrix", esirkepov, flush=True)
    seam, res = tmp_path / "seam.csv", tmp_path / "resident.csv"
    args = ["--out", str(seam), "--resident", str(res), "--counts", "10000,40000", "--grids", "64",
            "--repeats", "3"] + (["--esirkepov"] if esirkepov else [])
    assert gpu_matrix.main(args) == 0
    expect = [(n, s, b) for n in (10000, 40000) for s in (["NGP", "CIC", "TSC"] + (["Esirkepov"] if esirkepov else []))
              for b in ("GPU_Atomics", "GPU_Priv") if esirkepov or b == "GPU_Atomics" or s == "CIC"]
    for path in (seam, res):
        text = path.read_text().splitlines()
        assert text[0] == REFERENCE_HEADER
        rows = list(csv.DictReader(text))
        assert [(int(r["num_particles"]), r["scheme"], r["backend"]) for r in rows] == expect
        for r in rows:
            assert r["grid"] == "64x64" and r["layout"] == "SoA" and r["sorted"] == "off"
            assert float(r["deposit_ms"]) > 0 and float(r["throughput"]) > 0
            assert float(r["charge_error"]) < 1e-6
            n, s = int(r["num_particles"]), r["scheme"]
            writes = {"NGP": 1, "CIC": 4, "TSC": 9, "Esirkepov": 4}[s]
            nbytes = n * 40 + n * writes * 8 + 64 * 64 * 8
            assert float(r["bandwidth_gbs"]) == pytest.approx(nbytes / 1e9 / (float(r["deposit_ms"]) / 1000), rel=1e-6)


def test_matrix_aos_rows(tmp_path):
    """--layouts SoA,AoS: every case twice, the AoS row with the 48-byte
    record of estimateBytes (deposition.cpp:138)."""
    import gpu_matrix
    print("[case] gpu matrix AoS", flush=True)
    seam, res = tmp_path / "seam.csv", tmp_path / "resident.csv"
    args = ["--out", str(seam), "--resident", str(res), "--counts", "20000", "--grids", "64", "--repeats", "2",
            "--layouts", "SoA,AoS"]
    assert gpu_matrix.main(args) == 0
    for path in (seam, res):
        rows = list(csv.DictReader(path.read_text().splitlines()))
        assert [(r["scheme"], r["layout"], r["backend"]) for r in rows] == [
            (s, lay, b) for s in ("NGP", "CIC", "TSC") for b in ("GPU_Atomics", "GPU_Priv") for lay in ("SoA", "AoS")
            if b == "GPU_Atomics" or s == "CIC"]
        for r in rows:
            assert float(r["charge_error"]) < 1e-6
            rec = 48 if r["layout"] == "AoS" else 40
            writes = {"NGP": 1, "CIC": 4, "TSC": 9}[r["scheme"]]
            nbytes = 20000 * rec + 20000 * writes * 8 + 64 * 64 * 8
            assert float(r["bandwidth_gbs"]) == pytest.approx(nbytes / 1e9 / (float(r["deposit_ms"]) / 1000), rel=1e-6)
