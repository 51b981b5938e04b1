"""The Sobol sampler (pbrt_render_config.sampler = PBRT_SAMPLER_SOBOL) in the oracle, pinned by what
its construction guarantees and what a plausible mistake breaks:

  * the first 2^m samples of a pixel form a (0,m,2)-net in every stratified pair, for every m up to
    12: each elementary interval of area 2^-m (every split 2^a x 2^(m-a)) holds exactly one point;
  * different pairs of one pixel, and one pair of different pixels, are decorrelated (the shuffles
    are independent): correlation below 0.05 over 4096 samples;
  * renders are unbiased: the room's mean radiance with the Sobol sampler agrees with the
    independent sampler's, and at equal spp the error against a 4096-spp render is lower.
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from pbrt_amd import scenes


def _is_net(x, y, m):
    for a in range(m + 1):
        cx = (x >> (32 - a)) if a else np.zeros_like(x)
        cy = (y >> (32 - (m - a))) if m - a else np.zeros_like(y)
        cell = cx.astype(np.int64) * (1 << (m - a)) + cy
        if len(np.unique(cell)) != len(x):
            return False
    return True


@pytest.mark.parametrize("pair", [0, 1, 2, 3, 4])
def test_prefixes_are_nets(pair):
    for pixel in (0, 7, 123456):
        x = po.sobol_bits(11, pixel, pair, 0, 4096)
        y = po.sobol_bits(11, pixel, pair, 1, 4096)
        for m in range(0, 13):
            n = 1 << m
            assert _is_net(x[:n], y[:n], m), (pair, pixel, m)
        # a net of its own at every aligned block too (the shuffle maps blocks to blocks)
        assert _is_net(x[256:512], y[256:512], 8)


def test_independent_counter_draws_are_not_nets():
    # the property is not automatic: 256 independent points miss it
    rng = np.random.default_rng(1)
    x = rng.integers(0, 1 << 32, 256, dtype=np.uint64).astype(np.uint32)
    y = rng.integers(0, 1 << 32, 256, dtype=np.uint64).astype(np.uint32)
    assert not _is_net(x, y, 8)


def test_pairs_and_pixels_are_decorrelated():
    n = 4096
    u = lambda b: b.astype(np.float64) * 2.0 ** -32  # noqa: E731
    base = u(po.sobol_bits(5, 42, 0, 0, n))
    for other in (u(po.sobol_bits(5, 42, 1, 0, n)), u(po.sobol_bits(5, 42, 2, 1, n)), u(po.sobol_bits(5, 43, 0, 0, n)),
                  u(po.sobol_bits(6, 42, 0, 0, n))):
        assert abs(np.corrcoef(base, other)[0, 1]) < 0.05
    # and the two coordinates of one pair
    assert abs(np.corrcoef(base, u(po.sobol_bits(5, 42, 0, 1, n)))[0, 1]) < 0.05


def test_sobol_renders_are_unbiased_and_converge_faster():
    d, cam = scenes.room(subdiv=4)
    o = po.Scene(d, precision=64)
    args = (32, 32, cam["fov"], cam["eye"], cam["look"])
    ref, _ = o.render(*args, 4096, 8, 99)
    ind, _ = o.render(*args, 64, 8, 3)
    sob, _ = o.render(*args, 64, 8, 3, sampler="sobol")
    e_ind = np.sqrt(((ind - ref) ** 2).mean())
    e_sob = np.sqrt(((sob - ref) ** 2).mean())
    assert e_sob < 0.9 * e_ind, (e_sob, e_ind)
    assert abs(sob.mean() - ref.mean()) / ref.mean() < 0.01
    # FP32 oracle, same sampler: close to f64
    sob32, _ = po.Scene(d, precision=32, bins=12, t_trav=1.0, max_leaf=4).render(*args, 64, 8, 3, sampler="sobol")
    assert np.sqrt(((sob32 - sob) ** 2).mean()) / sob.mean() < 0.05
