"""Parity of the device radiance cache with the oracle (oracle/nrc_oracle.hpp).

Bars: the matrix-core product exact against FP64 (BF16 products are exact in FP32; 32-term sums of
values chosen on a coarse grid are exact too); the encoding within one BF16 step of the oracle's,
with at least 99 % of features equal; the forward pass within 2e-2 relative RMSE (and 1e-2 of the
output scale at most per value) of the oracle with the device's BF16 rounding points — both round
the same activations, the device accumulates in FP32 where the oracle uses FP64, so an
activation near a BF16 rounding boundary can round the other way; gradients within 2e-2 relative
L2 of the oracle's; the Adam update bit-exact against the oracle's update applied to the device's
own gradient; training that lowers the loss."""
import numpy as np
import pytest

import pbrt_amd
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

NP = pbrt_amd.NRC_PARAMS


def pad_features(x81):
    x = np.zeros((len(x81), pbrt_amd.NRC_FEATURES), np.float32)
    x[:, :81] = x81
    return pbrt_amd.bf16_bits(x)


def random_batch(rng, n, scale=1.0):
    x = rng.uniform(-1.0, 1.0, (n, 81)).astype(np.float32) * scale
    t = rng.uniform(0.0, 1.5, (n, 3)).astype(np.float32)
    return x, t


def test_mfma_operand_layout(gpu_ctx):
    rng = np.random.default_rng(0)
    a = rng.integers(-8, 9, (16, 32)).astype(np.float32) / 4.0
    b = rng.integers(-8, 9, (32, 16)).astype(np.float32) / 8.0
    d = pbrt_amd.mfma_check(gpu_ctx, pbrt_amd.bf16_bits(a), pbrt_amd.bf16_bits(b))
    np.testing.assert_array_equal(d.astype(np.float64), a.astype(np.float64) @ b.astype(np.float64))


def test_encoding_matches_oracle(gpu_ctx):
    rng = np.random.default_rng(1)
    lo, hi = np.array([-2.0, -1.0, 0.5], np.float32), np.array([2.0, 3.0, 1.5], np.float32)
    n = 4096
    raw = np.empty((n, 12), np.float32)
    raw[:, :3] = rng.uniform(lo, hi, (n, 3))
    raw[:, 3:] = rng.uniform(-1.0, 1.0, (n, 9))
    raw[0, :3] = lo
    raw[1, :3] = hi
    nrc = pbrt_amd.RadianceCache(gpu_ctx, seed=1)
    nrc.set_bounds(lo, hi)
    dev = nrc.encode_host(raw)
    ref = pad_features(po.nrc_encode(raw, lo, hi))
    assert (dev[:, 81:] == 0).all()
    diff = np.abs(pbrt_amd.bf16_float(dev).astype(np.float64) - pbrt_amd.bf16_float(ref))
    # one BF16 step of the oracle's value, with a floor of 2^-24 for values at or near an exact zero
    step = np.maximum(np.abs(pbrt_amd.bf16_float(ref).astype(np.float64)) * 2.0 ** -7, 2.0 ** -24)
    assert (diff <= step).all(), diff.max()
    assert (dev == ref).mean() >= 0.99
    np.testing.assert_array_equal(dev[:, 72:81], pbrt_amd.bf16_bits(raw[:, 3:]))


@pytest.mark.parametrize("n", [1, 17, 64, 1000, 5000])
def test_forward_matches_oracle(gpu_ctx, n):
    rng = np.random.default_rng(n)
    nrc = pbrt_amd.RadianceCache(gpu_ctx, seed=3)
    p = nrc.params
    x, _ = random_batch(rng, n)
    xb = pad_features(x)
    y = nrc.infer_host(xb).astype(np.float64)
    ref = po.nrc_forward(p, pbrt_amd.bf16_float(xb)[:, :81], bf16=True)
    scale = max(float(np.sqrt(np.mean(ref ** 2))), 1e-3)
    assert np.sqrt(np.mean((y - ref) ** 2)) / scale < 2e-2
    assert np.abs(y - ref).max() / scale < 1e-1


def test_forward_routing_network_exact(gpu_ctx):
    """One path of weight 1 from input 40 to output 2 through hidden unit 5, output bias 0.25: every
    intermediate value is a BF16 number, so the device result is exact."""
    nrc = pbrt_amd.RadianceCache(gpu_ctx, seed=3)
    p = np.zeros(NP, np.float32)
    off = 0
    offs = []
    for ni, no in [(81, 64), (64, 64), (64, 64), (64, 64), (64, 3)]:
        offs.append((off, off + ni * no))
        off += ni * no + no
    p[offs[0][0] + 5 * 81 + 40] = 1.0
    for l in (1, 2, 3):
        p[offs[l][0] + 5 * 64 + 5] = 1.0
    p[offs[4][0] + 2 * 64 + 5] = 1.0
    p[offs[4][1] + 2] = 0.25
    nrc.params = p
    rng = np.random.default_rng(4)
    x, _ = random_batch(rng, 300)
    xb = pad_features(x)
    y = nrc.infer_host(xb)
    expect = np.zeros((300, 3), np.float32)
    expect[:, 2] = np.maximum(pbrt_amd.bf16_float(xb[:, 40]), 0.0) + np.float32(0.25)
    np.testing.assert_array_equal(y, expect)


@pytest.mark.parametrize("n", [64, 200, 4096])
def test_gradients_match_oracle(gpu_ctx, n):
    rng = np.random.default_rng(100 + n)
    nrc = pbrt_amd.RadianceCache(gpu_ctx, seed=5)
    p = nrc.params
    x, t = random_batch(rng, n)
    xb = pad_features(x)
    loss, g = nrc.gradients_host(xb, t)
    ref_loss, ref_g = po.nrc_loss_grad(p, pbrt_amd.bf16_float(xb)[:, :81], t, bf16=True)
    assert abs(loss - ref_loss) / ref_loss < 1e-2, (loss, ref_loss)
    err = np.linalg.norm(g - ref_g) / np.linalg.norm(ref_g)
    assert err < 2e-2, err
    # per layer, so a wrong block of the padded layout cannot hide behind the others
    off = 0
    for ni, no in [(81, 64), (64, 64), (64, 64), (64, 64), (64, 3)]:
        for a, b in ((off, off + ni * no), (off + ni * no, off + ni * no + no)):
            e = np.linalg.norm(g[a:b] - ref_g[a:b]) / max(np.linalg.norm(ref_g[a:b]), 1e-12)
            assert e < 5e-2, (a, b, e)
        off += ni * no + no


def test_gradients_deterministic(gpu_ctx):
    rng = np.random.default_rng(7)
    nrc = pbrt_amd.RadianceCache(gpu_ctx, seed=5)
    x, t = random_batch(rng, 3000)
    xb = pad_features(x)
    l1, g1 = nrc.gradients_host(xb, t)
    l2, g2 = nrc.gradients_host(xb, t)
    assert l1 == l2 and g1.tobytes() == g2.tobytes()


def test_adam_step_matches_oracle(gpu_ctx):
    rng = np.random.default_rng(9)
    nrc = pbrt_amd.RadianceCache(gpu_ctx, seed=7)
    nrc.set_optim(lr=3e-3, beta1=0.9, beta2=0.99, eps=1e-8)
    p = nrc.params
    m = np.zeros(NP, np.float32)
    v = np.zeros(NP, np.float32)
    for step in range(1, 4):
        x, t = random_batch(rng, 512)
        xb = pad_features(x)
        _, g = nrc.gradients_host(xb, t)  # the gradient train_step computes on this batch
        nrc.train_step_host(xb, t)
        po.nrc_adam(p, m, v, g, step, lr=3e-3, b1=0.9, b2=0.99, eps=1e-8)
        np.testing.assert_array_equal(nrc.params, p)
    assert nrc.steps == 3


def test_training_lowers_loss(gpu_ctx):
    """Fit a smooth radiance-like function of the encoded position: the relative L2 loss falls
    by an order of magnitude within 300 steps of 4096 samples."""
    rng = np.random.default_rng(21)
    nrc = pbrt_amd.RadianceCache(gpu_ctx, seed=11)
    nrc.set_optim(lr=5e-3)
    lo, hi = np.zeros(3, np.float32), np.ones(3, np.float32)
    nrc.set_bounds(lo, hi)

    def batch(n):
        raw = np.zeros((n, 12), np.float32)
        raw[:, :3] = rng.uniform(0.0, 1.0, (n, 3))
        raw[:, 3:6] = [0.0, 0.0, 1.0]
        raw[:, 6:9] = [0.0, 1.0, 0.0]
        raw[:, 9:] = 0.5
        q = raw[:, :3].astype(np.float64)
        t = np.stack([1.0 + np.sin(3 * q[:, 0]) * q[:, 1], 0.5 + 0.5 * q[:, 2] ** 2, 0.2 + q[:, 0] * q[:, 1]], 1)
        return raw, t.astype(np.float32)

    raw_eval, t_eval = batch(4096)
    x_eval = nrc.encode_host(raw_eval)
    first = nrc.gradients_host(x_eval, t_eval)[0]
    for _ in range(300):
        raw, t = batch(4096)
        nrc.train_step_host(nrc.encode_host(raw), t)
    last = nrc.gradients_host(x_eval, t_eval)[0]
    assert last < first / 10, (first, last)


@pytest.mark.parametrize("n", [1, 100, 20000])
def test_query_equals_encode_then_infer(gpu_ctx, n):
    rng = np.random.default_rng(50 + n)
    lo, hi = np.array([-2.0, -1.0, 0.5], np.float32), np.array([2.0, 3.0, 1.5], np.float32)
    raw = np.empty((n, 12), np.float32)
    raw[:, :3] = rng.uniform(lo, hi, (n, 3))
    raw[:, 3:] = rng.uniform(-1.0, 1.0, (n, 9))
    nrc = pbrt_amd.RadianceCache(gpu_ctx, seed=8)
    p = nrc.params
    p[-(64 * 3 + 3):] = rng.uniform(-0.2, 0.2, 64 * 3 + 3)  # a non-zero output layer
    nrc.params = p
    nrc.set_bounds(lo, hi)
    np.testing.assert_array_equal(nrc.query_host(raw), nrc.infer_host(nrc.encode_host(raw)))


def test_inference_reads_the_parameter_average(gpu_ctx):
    """With ema = 0.5 the inference weights after one step are the mean of the parameters before
    and after it (0.5 a + 0.5 b, one rounding, as the device computes it)."""
    rng = np.random.default_rng(77)
    nrc = pbrt_amd.RadianceCache(gpu_ctx, seed=9)
    p = nrc.params
    p[-(64 * 3 + 3):] = rng.uniform(-0.2, 0.2, 64 * 3 + 3)
    nrc.params = p
    nrc.set_optim(lr=1e-2, ema=0.5)
    before = nrc.params
    x, t = random_batch(rng, 1024)
    xb = pad_features(x)
    nrc.train_step_host(xb, t)
    after = nrc.params
    avg = (np.float32(0.5) * before + np.float32(0.5) * after).astype(np.float32)
    y = nrc.infer_host(xb).astype(np.float64)
    ref = po.nrc_forward(avg, pbrt_amd.bf16_float(xb)[:, :81], bf16=True)
    other = po.nrc_forward(after, pbrt_amd.bf16_float(xb)[:, :81], bf16=True)
    scale = float(np.sqrt(np.mean(ref ** 2)))
    assert np.sqrt(np.mean((y - ref) ** 2)) / scale < 2e-2
    assert np.sqrt(np.mean((y - other) ** 2)) > 5 * np.sqrt(np.mean((y - ref) ** 2))
