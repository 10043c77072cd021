"""Learned density denoiser on the GPU: BF16-MFMA forward and backward
against the FP64 oracle network (with the device's BF16 roundings emulated),
Adam training, and the physics check: trained on pairs deposited by this
library (512 particles per cell vs a 1/64 subsample), the learned 8-ppc
density is closer to the exact expected density than the plain 8-ppc
deposit; renormalised charge exact to FP64 round-off; the Gauss-law gate
falls back to the plain deposit when the network is corrupted or the
tolerance is zero."""
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


BLOCKS = {"W1": (0, 288), "b1": (288, 320), "W2": (320, 9536), "b2": (9536, 9568), "W3": (9568, 18784),
          "b3": (18784, 18816), "W4": (18816, 19104), "b4": (19104, 19105)}


@pytest.mark.parametrize("dil", [(1, 2, 4, 8), (1, 1, 1, 1)])
def test_forward_matches_oracle(hip, orc, dil):
    say("cnn forward", dil)
    rng = np.random.default_rng(5)
    H, W, B = 32, 64, 2
    x = rng.normal(0, 0.3, (B, H, W)).astype(np.float32)
    with hip.Denoiser(W, H, max_batch=B, max_pairs=0, dilation=dil, seed=7) as net:
        p = net.params().astype(np.float64)
        out = net.forward(x)
    ref = orc.cnn_forward(p, x, dil, emulate_bf16=True)
    plain = orc.cnn_forward(p, x, dil, emulate_bf16=False)
    err, err_plain = rel_l2(out, ref), rel_l2(out, plain)
    print(f"forward rel_l2 vs bf16-emulating oracle {err:.2e}, vs pure FP64 {err_plain:.2e}", flush=True)
    assert err <= 2e-3


def test_gradient_matches_oracle(hip, orc):
    say("cnn gradient")
    rng = np.random.default_rng(6)
    H, W, B = 32, 64, 2
    dil = (1, 2, 4, 8)
    x = rng.normal(0, 0.3, (B, H, W)).astype(np.float32)
    t = rng.normal(0, 0.1, (B, H, W)).astype(np.float32)
    with hip.Denoiser(W, H, max_batch=B, max_pairs=0, dilation=dil, seed=8) as net:
        # a larger last layer so that every block gets a sizeable gradient
        p = net.params()
        p[18816:19105] *= 10.0
        net.set_params(p)
        loss, g = net.gradient(x, t)
    p = p.astype(np.float64)
    rloss, rg = orc.cnn_backward(p, x, t, dil, emulate_bf16=True)
    assert loss == pytest.approx(rloss, rel=2e-3)
    for name, (a, b) in BLOCKS.items():
        err = rel_l2(g[a:b], rg[a:b])
        print(f"grad {name} rel_l2 {err:.2e}", flush=True)
        assert err <= 5e-3, name


def test_training_reduces_loss(hip):
    """Denoising a smooth random field under white noise: Adam on the device
    lowers the loss well below the identity's."""
    say("cnn training")
    rng = np.random.default_rng(9)
    H, W = 32, 64
    ky, kx = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")

    def smooth():
        f = np.zeros((H, W))
        for _ in range(3):
            a, m, n, ph = rng.uniform(0.05, 0.2), rng.integers(0, 3), rng.integers(1, 4), rng.uniform(0, 2 * np.pi)
            f += a * np.cos(2 * np.pi * (m * ky / H + n * kx / W) + ph)
        return f

    with hip.Denoiser(W, H, max_batch=4, max_pairs=16, dilation=(1, 2, 4, 8), seed=1) as net, \
            hip.Context(hip.make_domain(1.0, 1.0, W, H), 64) as c:
        sig = []
        for k in range(16):
            s = smooth()
            noisy = s + rng.normal(0, 0.2, (H, W))
            # samples enter as densities: rho = 1 + field (delta = rho / mean - 1)
            c.set_rho((1.0 + s) / np.mean(1.0 + s))
            net.add_sample(c, True)
            c.set_rho((1.0 + noisy) / np.mean(1.0 + noisy))
            net.add_sample(c, False)
            sig.append(s)
        assert net.pairs == 16
        hist = net.train(300, batch=4, lr=2e-3, seed=3)
    start, end = hist[:10].mean(), hist[-20:].mean()
    print(f"loss {start:.4e} -> {end:.4e}", flush=True)
    assert np.all(np.isfinite(hist))
    assert end < 0.25 * start


def deposit_pairs(hip, net, N, states, ppc_hi=512, stride=64, salts=(1, 2)):
    n = N * N * ppc_hi
    with hip.Context(hip.make_domain(1.0, 1.0, N, N), n) as c:
        for (seed, amp, mode) in states:
            c.init_particles(n, kind=1, seed=seed, amp=amp, mode=mode, q=-1.0 / n)
            c.sort()
            for salt in salts:
                c.deposit(2, 2)
                net.add_sample(c, True)
                c.deposit_subsample(2, stride, salt + 1000 * seed)
                net.add_sample(c, False)


def rms(a):
    return float(np.sqrt(np.mean(np.square(a))))


@pytest.fixture(scope="module")
def trained(hip):
    """Network trained on 48 Langmuir states at 64^2 (512-ppc TSC targets,
    1/64 subsample inputs, two salts each)."""
    N = 64
    rng = np.random.default_rng(11)
    train_states = [(int(s), float(rng.uniform(0.001, 0.05)), int(rng.integers(1, 5))) for s in range(100, 148)]
    net = hip.Denoiser(N, N, max_batch=8, max_pairs=96, seed=2)
    deposit_pairs(hip, net, N, train_states)
    hist = net.train(1000, batch=8, lr=3e-3, seed=5)
    print(f"train loss {hist[:10].mean():.3e} -> {hist[-20:].mean():.3e}", flush=True)
    yield net
    net.close()


def test_learned_density_beats_plain(hip, trained):
    say("learned density")
    N = 64
    net = trained
    errs = {"plain8": [], "learned8": [], "plain512": []}
    for seed, amp, mode in [(900, 0.04, 2), (901, 0.01, 1), (902, 0.03, 3), (903, 0.002, 4)]:
        truth = hip.langmuir_delta_truth(N, N, amp, mode, 2)
        for ppc, key in ((512, "plain512"), (8, "plain8")):
            n = N * N * ppc
            with hip.Context(hip.make_domain(1.0, 1.0, N, N), n) as c:
                c.init_particles(n, kind=1, seed=seed, amp=amp, mode=mode, q=-1.0 / n)
                c.sort()
                c.deposit(2, 2)
                rho = c.rho().astype(np.float64)
                errs[key].append(rms(rho / rho.mean() - 1 - truth))
                if ppc == 8:
                    fb, ratio = net.denoise(c, 2, tolerance=4.0)
                    assert not fb, ratio
                    out = net.output_f64()
                    q = c.download()["q"].astype(np.float64).sum()
                    assert abs(out.sum() / (N * N) - q) <= 1e-13 * abs(q)
                    lr = c.rho().astype(np.float64)
                    np.testing.assert_allclose(lr, out, rtol=1e-6)
                    errs["learned8"].append(rms(out / out.mean() - 1 - truth))
    m = {k: float(np.mean(v)) for k, v in errs.items()}
    print("rms delta error vs exact:", m, flush=True)
    assert m["learned8"] < 0.2 * m["plain8"]


def test_gate_falls_back(hip, trained):
    say("denoiser gate")
    N = 64
    net = trained
    n = N * N * 8
    keep = net.params()
    try:
        with hip.Context(hip.make_domain(1.0, 1.0, N, N), n) as c:
            c.init_particles(n, kind=1, seed=77, amp=0.02, mode=1, q=-1.0 / n)
            c.sort()
            c.deposit(2, 2)
            plain = c.rho()
            # zero tolerance keeps the plain deposit bit for bit
            fb, ratio = net.denoise(c, 2, tolerance=0.0)
            assert fb and ratio > 0.0
            assert np.array_equal(c.rho(), plain)
            np.testing.assert_array_equal(net.output_f64(), plain.astype(np.float64))
            # a corrupted network (last layer x 100) is rejected at the default tolerance
            p = keep.copy()
            p[18816:19104] *= 100.0
            net.set_params(p)
            fb, ratio = net.denoise(c, 2, tolerance=4.0)
            print(f"corrupted network residual ratio {ratio:.1f}", flush=True)
            assert fb and ratio > 4.0
            assert np.array_equal(c.rho(), plain)
    finally:
        net.set_params(keep)


def test_learned_step_loop(hip, trained):
    """Step loop at 8 particles per cell with the learned density after every
    deposit (TSC, sort every 10 steps, unit plasma frequency, dt 0.1, 20
    steps from a 1% Langmuir perturbation, which the solver's force sign
    amplifies about e^2-fold): graph replay and eager execution give identical
    records, no step falls back, the renormalised density carries the
    particles' charge; the energy series are printed beside the plain 8- and
    512-ppc runs."""
    say("learned step loop")
    N = 64
    net = trained
    runs = {}
    for ppc, mode in ((512, "plain"), (8, "plain"), (8, "graph"), (8, "eager")):
        n = N * N * ppc
        with hip.Context(hip.make_domain(1.0, 1.0, N, N), n) as c:
            c.init_particles(n, kind=1, seed=31, amp=0.01, mode=1, q=hip.unit_plasma_charge(n, N, N), m=1.0)
            if mode == "plain":
                rec = c.run(20, 2, 2, 0.1, 0.0, hip.ORDER_PUSH_FIRST, sort_interval=10, record_particles=True,
                            use_graph=True)
            else:
                rec, fallbacks = net.run(c, 20, 2, 2, dt=0.1, sort_interval=10, tolerance=4.0,
                                         use_graph=mode == "graph")
                print(f"{mode}: {fallbacks} gate fallbacks", flush=True)
                assert fallbacks == 0
        runs[(ppc, mode)] = rec
    for k in runs[(8, "graph")]:
        np.testing.assert_array_equal(runs[(8, "graph")][k], runs[(8, "eager")][k], err_msg=k)
    learned = runs[(8, "graph")]
    assert all(np.all(np.isfinite(v)) for v in learned.values())
    assert np.all(learned["charge_error"] < 1e-6)
    for key in ("field_energy", "kinetic"):
        ref = runs[(512, "plain")][key] / 512
        print(f"{key} per particle rel_l2 vs 512 ppc: plain8 {rel_l2(runs[(8, 'plain')][key] / 8, ref):.3f} "
              f"learned8 {rel_l2(learned[key] / 8, ref):.3f}", flush=True)


@pytest.mark.parametrize("background", [0.0, 1.0])
def test_learned_run_matches_denoise(hip, trained, background):
    """One learned step against pic_hip_denoise on the same plain deposit,
    with and without a background (added to rho after the learned density is
    written): the FP64 learned density equals the single call's to 1e-12
    relative (the run takes the density mean from the particles' charge sum,
    the single call from the grid sum), FP32 rho is its rounding plus the
    background, and the total charge is exact."""
    say("learned run vs denoise, background", background)
    N = 64
    n = N * N * 8
    dom = hip.make_domain(1.0, 1.0, N, N)
    with hip.Context(dom, n) as a, hip.Context(dom, n) as b:
        for c in (a, b):
            c.init_particles(n, kind=1, seed=41, amp=0.02, mode=2, q=hip.unit_plasma_charge(n, N, N), m=1.0)
        _, fallbacks = trained.run(a, 1, 2, 2, dt=0.1, background=background, sort_interval=10, tolerance=4.0)
        assert fallbacks == 0
        learned = trained.output_f64().copy()
        rho_a = a.rho()
        # the same particles deposited the same way (the background only enters after the deposit)
        b.run(1, 2, 2, 0.1, 0.0, hip.ORDER_PUSH_FIRST, sort_interval=10)
        fb, ratio = trained.denoise(b, 2, tolerance=4.0)
        assert not fb, ratio
        np.testing.assert_allclose(learned, trained.output_f64(), rtol=1e-12, atol=0.0)
        expect_rho = learned.astype(np.float32) + np.float32(background)
        assert np.array_equal(rho_a, expect_rho)
        q = a.download()["q"].astype(np.float64).sum()
        assert abs(learned.sum() / (N * N) - q) <= 1e-13 * abs(q)


@pytest.mark.parametrize("halo", [0, 1])
def test_learned_run_fallback_equals_plain_run(hip, trained, halo):
    """A learned run whose gate always falls back (tolerance 0) keeps the plain
    deposit every step, so it must reproduce the plain step loop bit for bit:
    the first layer's fused window merge (with particles drifting past the
    tiles' halo into the FP64 overflow grid, folded there and cleared by the
    second layer) writes the same rho as the deposit's own merge."""
    say("learned fallback vs plain run, halo", halo)
    N = 64
    n = N * N * 8
    dom = hip.make_domain(1.0, 1.0, N, N)
    steps = 12
    with hip.Context(dom, n) as a, hip.Context(dom, n) as b:
        for c in (a, b):
            c.set_tile_halo(halo)
            # thermal speed ~1.5 cells per step: many particles leave their tile's reach between sorts
            c.init_particles(n, kind=1, seed=43, amp=0.02, mode=2, vth=1.5 / N / 0.1,
                             q=hip.unit_plasma_charge(n, N, N), m=1.0)
        rec_a, fallbacks = trained.run(a, steps, 2, 2, dt=0.1, sort_interval=5, tolerance=0.0)
        assert fallbacks == steps
        rec_b = b.run(steps, 2, 2, 0.1, 0.0, hip.ORDER_PUSH_FIRST, sort_interval=5, use_graph=True)
        assert np.array_equal(a.rho(), b.rho())
        pa, pb = a.download(), b.download()
        for k in ("x", "y", "vx", "vy"):
            assert np.array_equal(pa[k], pb[k]), k
        # same fields; the learned run also records particle sums, which
        # changes the diagnostics' reduction layout
        np.testing.assert_allclose(rec_a["field_energy"], rec_b["field_energy"], rtol=1e-12, atol=0.0)


def test_invalid_denoiser_arguments(hip):
    say("denoiser errors")
    with pytest.raises(hip.PicError, match="multiple of 32"):
        hip.Denoiser(48, 32)
    with hip.Denoiser(64, 64, max_batch=2, max_pairs=1) as net:
        with pytest.raises(hip.PicError, match="no training pairs"):
            net.train(1, batch=1)
        with hip.Context(hip.make_domain(1.0, 1.0, 32, 64), 16) as c:
            with pytest.raises(hip.PicError, match="grid differs"):
                net.add_sample(c, True)
