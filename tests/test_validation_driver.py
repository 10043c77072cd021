"""Host side of the device validation driver (pic-deposition_amd/validation.py)
on the CPU: the R2HC spectral-noise evaluation against the oracle's
restatement of spectralNoiseLevel, and the reference's CSV formatting
(std::setprecision(10), booleans 0/1) and headers.  No GPU."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def val(hip):
    import validation

    return validation


@pytest.mark.parametrize("shape", [(64, 64), (48, 40), (1, 256), (33, 17), (2, 3)])
def test_spectral_noise_matches_oracle(val, orc, shape):
    rng = np.random.default_rng(sum(shape))
    rho = rng.normal(1.0, 0.3, shape)
    assert val.spectral_noise(rho) == pytest.approx(orc.spectral_noise_level(rho), rel=1e-12, abs=1e-15)


def test_spectral_noise_closed_forms(val):
    ny, nx = 24, 30
    x = np.arange(nx)[None, :] * np.ones((ny, 1))
    assert val.spectral_noise(np.cos(2 * np.pi * 2 * x / nx)) == pytest.approx(0.0, abs=1e-20)
    assert val.spectral_noise(np.cos(2 * np.pi * 12 * x / nx)) == pytest.approx(1.0, rel=1e-12)
    assert val.spectral_noise(np.zeros((4, 4))) == 0.0


def test_csv_number_format_matches_reference_text(val):
    # strings as written by the reference (data/benchmarks/*.csv)
    for text in ("0.6099871407", "6.661338148e-18", "0", "1.726334915", "11.50889943", "0.0008040122131",
                 "0.005068541579", "3.689404995e-05", "-1.802267236e-06"):
        assert val.fmt(float(text)) == text
    assert val.fmt(True) == "1" and val.fmt(False) == "0" and val.fmt(2000) == "2000" and val.fmt("64x64") == "64x64"


def test_csv_headers_are_the_reference_headers(val):
    expect = {  # first lines of the reference's data/benchmarks/*.csv
        "validation.csv": "scheme,spectral_noise,charge_error,omega_theory,omega_measured,relative_error,omega_ratio,passed",
        "two_stream_validation.csv": "scheme,growth_rate_measured,growth_rate_theory,growth_rate_ratio,omega_p_macro,"
                                     "growth_rate_over_omega_p,passed",
        "two_stream_quasi1d.csv": "geometry,scheme,growth_rate_measured,growth_rate_theory,growth_rate_ratio,"
                                  "omega_p_macro,growth_rate_over_omega_p,passed",
        "validation_multi_seed.csv": "seed,test,scheme,metric,value,passed",
        "noise_vs_grid.csv": "grid,scheme,spectral_noise",
        "langmuir_convergence.csv": "grid_n,num_particles,scheme,omega_measured,omega_theory,omega_ratio",
        "conservation_study.csv": "scheme,steps,max_charge_error,final_energy_drift",
    }
    got = {fname: ",".join(header) for fname, _, header in val.SUITES.values()}
    assert got == expect


def test_unknown_suite_is_rejected(val, tmp_path):
    with pytest.raises(SystemExit):
        val.main(["--out-dir", str(tmp_path), "--suites", "landau"])
