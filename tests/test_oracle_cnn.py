"""Pins the FP64 denoiser-network oracle (or_cnn_forward / or_cnn_backward):
an independent numpy restatement with periodic rolls, delta-kernel closed
forms, and central finite differences of the loss."""
import numpy as np
import pytest


def np_forward(p, x, dil):
    """Independent restatement: periodic 3x3 correlation by np.roll."""
    W1 = p[0:288].reshape(32, 9)
    b1 = p[288:320]
    W2 = p[320:9536].reshape(9, 32, 32)
    b2 = p[9536:9568]
    W3 = p[9568:18784].reshape(9, 32, 32)
    b3 = p[18784:18816]
    W4 = p[18816:19104].reshape(9, 32)
    b4 = p[19104]

    def shift(a, t, d):
        ty, tx = divmod(t, 3)
        return np.roll(a, (-(ty - 1) * d, -(tx - 1) * d), axis=(1, 2))

    a0 = x[..., None]
    z1 = sum(shift(a0, t, dil[0]) * W1[:, t] for t in range(9)) + b1
    a1 = np.maximum(z1, 0)
    z2 = sum(shift(a1, t, dil[1]) @ W2[t].T for t in range(9)) + b2
    a2 = np.maximum(z2, 0)
    z3 = sum(shift(a2, t, dil[2]) @ W3[t].T for t in range(9)) + b3
    a3 = np.maximum(z3, 0)
    return sum(shift(a3, t, dil[3]) @ W4[t] for t in range(9)) + b4


def random_params(rng, scale=1.0):
    p = np.zeros(19105)
    p[0:288] = rng.normal(0, np.sqrt(2 / 9), 288)
    p[288:320] = rng.normal(0, 0.1, 32)
    p[320:9536] = rng.normal(0, np.sqrt(2 / 288), 9216)
    p[9536:9568] = rng.normal(0, 0.1, 32)
    p[9568:18784] = rng.normal(0, np.sqrt(2 / 288), 9216)
    p[18784:18816] = rng.normal(0, 0.1, 32)
    p[18816:19104] = rng.normal(0, np.sqrt(2 / 288), 288)
    p[19104] = 0.05
    return p * scale


@pytest.mark.parametrize("dil", [(1, 1, 1, 1), (1, 2, 4, 8), (2, 1, 3, 1)])
def test_forward_matches_numpy_restatement(orc, dil):
    rng = np.random.default_rng(1)
    p = random_params(rng)
    x = rng.normal(0, 0.3, (2, 24, 20))
    got = orc.cnn_forward(p, x, dil)
    ref = np_forward(p, x, dil)
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("tap,shift", [(4, (0, 0)), (5, (0, -1)), (1, (1, 0))])
def test_delta_kernels_are_identity_and_rolls(orc, tap, shift):
    """Only channel 0 carries a unit weight at one tap in every layer: a
    positive image comes out rolled by four times that tap's offset (times the
    dilation of each layer)."""
    p = np.zeros(19105)
    p[0 + 0 * 9 + tap] = 1.0
    p[320 + (tap * 32 + 0) * 32 + 0] = 1.0
    p[9568 + (tap * 32 + 0) * 32 + 0] = 1.0
    p[18816 + tap * 32 + 0] = 1.0
    rng = np.random.default_rng(2)
    x = rng.uniform(0.1, 1.0, (1, 16, 32))
    dil = (1, 2, 1, 3)
    out = orc.cnn_forward(p, x, dil)
    total = sum(dil)
    ref = np.roll(x, (shift[0] * total, shift[1] * total), axis=(1, 2))
    np.testing.assert_allclose(out, ref, rtol=0, atol=1e-15)


def test_backward_matches_finite_differences(orc):
    rng = np.random.default_rng(3)
    p = random_params(rng)
    x = rng.normal(0, 0.5, (1, 6, 8))
    t = rng.normal(0, 0.5, (1, 6, 8))
    dil = (1, 2, 1, 2)
    loss, grad = orc.cnn_backward(p, x, t, dil)
    assert loss == pytest.approx(np.mean((np_forward(p, x, dil) - t) ** 2), rel=1e-12)
    # sampled parameters from every weight and bias block
    blocks = [(0, 288), (288, 320), (320, 9536), (9536, 9568), (9568, 18784), (18784, 18816), (18816, 19104),
              (19104, 19105)]
    idx = np.concatenate([rng.choice(np.arange(a, b), size=min(8, b - a), replace=False) for a, b in blocks])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (orc.cnn_backward(pp, x, t, dil)[0] - orc.cnn_backward(pm, x, t, dil)[0]) / (2 * h)
        assert grad[i] == pytest.approx(fd, rel=1e-5, abs=1e-9), i


def test_bf16_emulation_rounds_as_documented(orc):
    """Weights and inputs that are exact in BF16, with activations that stay
    exact, give identical emulated and plain outputs; a value between two
    BF16 numbers moves the emulated output."""
    p = np.zeros(19105)
    p[0 + 4] = 0.5
    p[320 + (4 * 32) * 32] = 2.0
    p[9568 + (4 * 32) * 32] = 0.25
    p[18816 + 4 * 32] = 4.0
    x = np.full((1, 4, 32), 0.75)
    a = orc.cnn_forward(p, x, (1, 1, 1, 1), emulate_bf16=False)
    b = orc.cnn_forward(p, x, (1, 1, 1, 1), emulate_bf16=True)
    np.testing.assert_array_equal(a, b)
    np.testing.assert_allclose(a, 0.75, rtol=0, atol=0)
    # 1 + 2^-9 lies below the midpoint of the BF16 neighbours 1 and 1 + 2^-7
    x2 = np.full((1, 4, 32), 1.0 + 2.0 ** -9)
    c = orc.cnn_forward(p, x2, (1, 1, 1, 1), emulate_bf16=True)
    np.testing.assert_allclose(c, 1.0, rtol=0, atol=0)
