"""Pins of the radiance-cache oracle (oracle/nrc_oracle.hpp). The reference has no such network, so
the oracle is checked against what the mathematics fixes: the encoding at points where sin / cos
are known exactly, a network that reduces to a product of matrices (all activations positive),
a network that routes one chosen feature to one chosen output (pins the [out][in] indexing and
every layer's orientation), the analytic gradient against central finite differences, and Adam
against its closed form under a constant gradient."""
import numpy as np
import pytest

from oracle import pyoracle as po

N_PARAMS = 17923
DIMS = [(81, 64), (64, 64), (64, 64), (64, 64), (64, 3)]


def offsets():
    off, out = 0, []
    for ni, no in DIMS:
        out.append((off, off + ni * no))  # W_l at off, b_l at off + ni*no
        off += ni * no + no
    return out, off


def unpack(p):
    offs, _ = offsets()
    Ws, bs = [], []
    for (wo, bo), (ni, no) in zip(offs, DIMS):
        Ws.append(p[wo:wo + ni * no].reshape(no, ni).astype(np.float64))
        bs.append(p[bo:bo + no].astype(np.float64))
    return Ws, bs


def test_param_count_and_layout():
    offs, total = offsets()
    assert po.nrc_nparams() == N_PARAMS == total


def test_encoding_known_points():
    lo, hi = np.array([-1.0, 0.0, 2.0], np.float32), np.array([1.0, 4.0, 6.0], np.float32)
    # axis 0 at its lower bound (p = 0), axis 1 at the middle (p = 0.5), axis 2 at a quarter
    raw = np.array([[-1.0, 2.0, 3.0, 0.1, 0.2, 0.3, 0.4, 0.5, 0.6, 0.7, 0.8, 0.9]], np.float32)
    x = po.nrc_encode(raw, lo, hi)[0]
    f = x[:72].reshape(3, 12, 2)
    np.testing.assert_allclose(f[0, :, 0], 0.0, atol=1e-12)   # sin(2^k pi 0)
    np.testing.assert_allclose(f[0, :, 1], 1.0, atol=1e-12)   # cos(2^k pi 0)
    # p = 1/2: sin(2^k pi / 2) = 1, 0, 0, ...; cos = 0, -1, 1, 1, ...
    np.testing.assert_allclose(f[1, :, 0], [1.0] + [0.0] * 11, atol=1e-6)
    np.testing.assert_allclose(f[1, :, 1], [0.0, -1.0] + [1.0] * 10, atol=1e-6)
    # p = 1/4: sin(pi/4), sin(pi/2), sin(pi), 0...; cos(pi/4), 0, -1, 1, ...
    s2 = np.sqrt(0.5)
    np.testing.assert_allclose(f[2, :, 0], [s2, 1.0] + [0.0] * 10, atol=1e-6)
    np.testing.assert_allclose(f[2, :, 1], [s2, 0.0, -1.0] + [1.0] * 9, atol=1e-6)
    np.testing.assert_array_equal(x[72:], raw[0, 3:])


def test_positive_network_is_a_matrix_product():
    """With non-negative inputs, weights and biases every ReLU is the identity, so the network is
    W4 (W3 (W2 (W1 (W0 x + b0) + b1) + b2) + b3) + b4."""
    rng = np.random.default_rng(3)
    p = rng.uniform(0.0, 0.05, N_PARAMS).astype(np.float32)
    x = rng.uniform(0.0, 1.0, (7, 81)).astype(np.float32)
    y = po.nrc_forward(p, x, bf16=False)
    Ws, bs = unpack(p)
    h = x.astype(np.float64).T
    for W, b in zip(Ws, bs):
        h = W @ h + b[:, None]
    np.testing.assert_allclose(y, h.T, rtol=1e-12)


@pytest.mark.parametrize("feature,channel", [(0, 0), (40, 2), (80, 1)])
def test_routing_network(feature, channel):
    """One path of weight 1 from input `feature` through hidden unit 5 to output `channel`, an output
    bias 0.25: y[channel] = relu(x[feature]) + 0.25, the other outputs 0 (bias 0)."""
    offs, _ = offsets()
    p = np.zeros(N_PARAMS, np.float32)
    (w0, _), (w1, _), (w2, _), (w3, _), (w4, b4) = offs
    p[w0 + 5 * 81 + feature] = 1.0
    for w in (w1, w2, w3):
        p[w + 5 * 64 + 5] = 1.0
    p[w4 + channel * 64 + 5] = 1.0
    p[b4 + channel] = 0.25
    rng = np.random.default_rng(feature)
    x = rng.uniform(-1.0, 1.0, (16, 81)).astype(np.float32)
    for bf16 in (False, True):
        y = po.nrc_forward(p, x, bf16=bf16)
        xf = x[:, feature].astype(np.float64)
        if bf16:
            xf = (x[:, feature].view(np.uint32).astype(np.uint64) + 0x7FFF
                  + ((x[:, feature].view(np.uint32) >> 16) & 1)) >> 16 << 16
            xf = xf.astype(np.uint32).view(np.float32).astype(np.float64)
        expect = np.zeros((16, 3))
        expect[:, channel] = np.maximum(xf, 0.0) + 0.25
        np.testing.assert_allclose(y, expect, rtol=0, atol=1e-12 if not bf16 else 4e-3)
        if bf16:  # the last hidden activation is BF16-rounded again; y - 0.25 is a BF16 value
            np.testing.assert_array_equal(y[:, channel] - 0.25, np.maximum(xf, 0.0))


def loss_fixed_denominator(p, x, t, y0):
    y = po.nrc_forward(p, x, bf16=False)
    return float(np.mean((y - t) ** 2 / (y0 ** 2 + 0.01)))


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(11)
    p = (rng.standard_normal(N_PARAMS) * 0.15).astype(np.float32)
    x = rng.uniform(-1.0, 1.0, (24, 81)).astype(np.float32)
    t = rng.uniform(0.0, 2.0, (24, 3)).astype(np.float32)
    loss, g = po.nrc_loss_grad(p, x, t, bf16=False)
    y0 = po.nrc_forward(p, x, bf16=False)
    assert loss == pytest.approx(float(np.mean((y0 - t) ** 2 / (y0 ** 2 + 0.01))), rel=1e-12)
    # the denominator is held constant (stop-gradient), so differentiate the loss with it frozen
    offs, _ = offsets()
    picks = []
    for (wo, bo), (ni, no) in zip(offs, DIMS):
        picks += list(rng.integers(wo, bo, 6)) + list(rng.integers(bo, bo + no, 2))
    h = 2.0 ** -12
    for i in picks:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        dp = float(pp[i]) - float(p[i])
        dm = float(p[i]) - float(pm[i])
        fd = (loss_fixed_denominator(pp, x, t, y0) - loss_fixed_denominator(pm, x, t, y0)) / (dp + dm)
        assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(g[i])) + 2e-4 * abs(g[i]), (i, fd, g[i])


def test_bf16_gradient_close_to_exact():
    rng = np.random.default_rng(5)
    p = (rng.standard_normal(N_PARAMS) * 0.15).astype(np.float32)
    x = rng.uniform(-1.0, 1.0, (64, 81)).astype(np.float32)
    t = rng.uniform(0.0, 2.0, (64, 3)).astype(np.float32)
    l64, g64 = po.nrc_loss_grad(p, x, t, bf16=False)
    l16, g16 = po.nrc_loss_grad(p, x, t, bf16=True)
    assert abs(l16 - l64) / l64 < 0.02
    cos = float(g16 @ g64 / np.linalg.norm(g16) / np.linalg.norm(g64))
    assert cos > 0.995, cos


def test_adam_constant_gradient_closed_form():
    """Under a constant gradient g, m_t = (1 - b1^t) g and v_t = (1 - b2^t) g^2, so the bias-corrected
    step is lr g / (|g| + eps) every time: after T steps p = p0 - T lr g / (|g| + eps)."""
    rng = np.random.default_rng(2)
    n = 1000
    p0 = rng.standard_normal(n).astype(np.float32)
    g = (rng.standard_normal(n) * 10.0 ** rng.uniform(-3, 1, n)).astype(np.float32)
    p, m, v = p0.copy(), np.zeros(n, np.float32), np.zeros(n, np.float32)
    lr, eps, T = 1e-3, 1e-8, 20
    for step in range(1, T + 1):
        po.nrc_adam(p, m, v, g, step, lr=lr, eps=eps)
    expect = p0.astype(np.float64) - T * lr * g / (np.abs(g) + eps)
    np.testing.assert_allclose(p, expect, rtol=0, atol=2e-6)
    np.testing.assert_allclose(m, (1 - 0.9 ** T) * g.astype(np.float64), rtol=2e-5)
    np.testing.assert_allclose(v, (1 - 0.99 ** T) * g.astype(np.float64) ** 2, rtol=2e-5)
