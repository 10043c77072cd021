"""The library's host-side copies of the reference's initialisers
(pic_hip_reference_particles; ParticlesSoA::initialize*,
src/pic/particles_aos.cpp:67-180) on the CPU: bit-exact against the oracle's
restatement and, where the reference build oracle/_ref is present, against the
reference's own functions; closed-form properties of each kind.  No GPU."""
import numpy as np
import pytest

CASES = [
    ("maxwellian", 0, dict(temperature=2.0), dict(p1=2.0)),
    ("langmuir", 1, dict(amplitude=0.05, mode=2), dict(p1=0.05, mode=2)),
    ("two_stream", 2, dict(beam_velocity=0.2, perturbation=0.01), dict(p1=0.2, p2=0.01)),
    ("two_stream_1d", 3, dict(beam_velocity=0.2, perturbation=0.01), dict(p1=0.2, p2=0.01)),
    ("warm_langmuir", 4, dict(amplitude=0.05, mode=1, temperature=0.01), dict(p1=0.05, p2=0.01, mode=1)),
]


@pytest.mark.parametrize("kind,ref_kind,kw,ref_kw", CASES)
@pytest.mark.parametrize("seed", [42, 7])
def test_library_initialisers_match_reference_build(hip, orc, kind, ref_kind, kw, ref_kw, seed):
    if not orc.ref_available():
        pytest.skip("reference build oracle/_ref not present (needs the reference checkout at build time)")
    dom = hip.make_domain(1.3, 0.7, 32, 16)
    od = orc.domain(1.3, 0.7, 32, 16)
    n = 1001
    mine = hip.reference_particles(kind, n, dom, seed=seed, **kw)
    ref = orc.ref_init(ref_kind, n, od, seed=seed, **ref_kw)
    for k in mine:
        assert np.array_equal(mine[k], ref[k]), k


@pytest.mark.parametrize("quasi1d", [False, True])
def test_library_initialisers_match_oracle(hip, orc, quasi1d):
    dom = hip.make_domain(1.0, 1.0, 64, 64)
    od = orc.domain(1.0, 1.0, 64, 64)
    n = 5000
    v = hip.resonant_beam_velocity(dom, n, 1)
    assert v == orc.resonant_beam_velocity(od, n, 1)
    mine = hip.reference_particles("two_stream_1d" if quasi1d else "two_stream", n, dom, seed=43, beam_velocity=v,
                                   perturbation=0.01)
    ref = orc.init("two_stream", n, od, seed=43, beam_velocity=v, perturbation=0.01, quasi1d=quasi1d)
    for k in mine:
        assert np.array_equal(mine[k], ref[k]), k
    lm = hip.reference_particles("langmuir", 200, dom, amplitude=0.05, mode=1)
    lr = orc.init("langmuir", 200, od, amplitude=0.05, mode=1)
    for k in lm:
        assert np.array_equal(lm[k], lr[k]), k
    mx = hip.reference_particles("maxwellian", 300, dom, seed=42, temperature=1.0)
    mr = orc.init("maxwellian", 300, od, seed=42, T=1.0)
    for k in mx:
        assert np.array_equal(mx[k], mr[k]), k


def test_initialiser_closed_forms(hip):
    dom = hip.make_domain(2.0, 1.0, 16, 8)
    n = 1000
    m = hip.reference_particles("maxwellian", n, dom, seed=1, temperature=4.0)
    assert np.all(m["q"][0::2] == 1.0) and np.all(m["q"][1::2] == -1.0) and np.all(m["m"] == 1.0)
    assert np.all((m["x"] >= 0) & (m["x"] < 2.0)) and np.all((m["y"] >= 0) & (m["y"] < 1.0))
    assert 1.6 < np.std(m["vx"]) < 2.4  # sqrt(T) = 2
    lang = hip.reference_particles("langmuir", n, dom, amplitude=0.1, mode=3)
    i = np.arange(n)
    assert np.array_equal(lang["x"], 2.0 * i / n) and np.all(lang["y"] == 0.5)
    assert np.array_equal(lang["vx"], 0.1 * np.sin(2 * np.pi * 3 / 2.0 * lang["x"]))
    assert np.all(lang["q"] == -1.0 / n) and np.all(lang["vy"] == 0)
    ts = hip.reference_particles("two_stream_1d", n, dom, seed=5, beam_velocity=0.3, perturbation=0.0)
    assert np.all(ts["vx"][: n // 2] == 0.3) and np.all(ts["vx"][n // 2:] == -0.3) and np.all(ts["y"] == 0.5)
    d64 = hip.make_domain(1.0, 1.0, 64, 64)
    wp = hip.plasma_frequency(d64, 5000)
    assert hip.cold_two_stream_growth_rate(d64, 5000) == pytest.approx(wp / (2 * np.sqrt(2)), rel=1e-15)
    assert hip.resonant_beam_velocity(d64, 5000, 2) == pytest.approx(wp / (4 * np.pi), rel=1e-15)


def test_initialiser_errors(hip):
    dom = hip.make_domain(1.0, 1.0, 8, 8)
    with pytest.raises(KeyError):
        hip.reference_particles("plasma", 10, dom)
    with pytest.raises(hip.PicError):
        hip.reference_particles("maxwellian", 10, hip.make_domain(1.0, 1.0, 0, 8))
