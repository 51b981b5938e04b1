"""Russian roulette in the oracle's path integrator (integrators/path.rs:93-100, commented out in
the reference; an option here, off by default): after each scattering the path continues with
probability q = min(1, max component of the throughput) and the throughput is divided by q.

Pins:
- a mirror (Metal, fuzz 0, albedo (0.9, 0.8, 0.7)) under the background light: every path is the
  camera ray, the light sample and the reflected ray, and roulette ends the reflected ray with
  probability 1 - q = 0.1 exactly, so the ray count per path is 2.9 (q from the smallest or the
  mean component would give 2.7 or 2.8), and the image mean is unchanged (a missing division by q
  would lower it by 10 %);
- the room lit by its ceiling panel (eight bounces, GGX): roulette halves the rays and keeps the
  image mean within 1 % (seed-to-seed noise here is 0.3 %; a missing division scales every
  indirect term down by its survival probability, about 0.7 per bounce here).
"""
import numpy as np

from oracle import pyoracle as po
from pbrt_amd import scenes
from pbrt_amd.scene import SceneDesc

EYE, LOOK, FOV = (0.3, 1.5, 2.0), (0.0, 0.0, 0.0), 0.6


def mirror_scene():
    d = SceneDesc()
    m = d.add_material("metal", (0.9, 0.8, 0.7), fuzz=0.0)
    d.add_object("rectangle", m, (50.0, 50.0))
    d.add_background_light((0.2, 0.2, 0.2), (0.4, 0.9, 1.0))
    return d


def test_roulette_on_a_mirror_ends_one_path_in_ten():
    sc = po.Scene(mirror_scene(), precision=64)
    w = h = 32
    spp = 256
    plain, rays0 = sc.render(w, h, FOV, EYE, LOOK, spp, 3, 4)
    rr, rays1 = sc.render(w, h, FOV, EYE, LOOK, spp, 3, 4, roulette=True)
    paths = w * h * spp
    assert rays0 == 3 * paths
    # ends ~ Binomial(paths, 0.1): 4 standard deviations
    ends = 3 * paths - rays1
    assert abs(ends - 0.1 * paths) < 4 * np.sqrt(paths * 0.09), ends / paths
    np.testing.assert_allclose(rr.reshape(-1, 3).mean(0), plain.reshape(-1, 3).mean(0), rtol=3e-3)
    # each sample is 0 or albedo / 0.9 times the background: the per-pixel spread grows
    assert rr.std() > plain.std()


def test_roulette_keeps_the_room_unbiased_with_half_the_rays():
    desc, cam = scenes.room(subdiv=3, ggx=True)
    sc = po.Scene(desc, precision=64)
    args = (32, 32, cam["fov"], cam["eye"], cam["look"], 256, 8, 5)
    plain, rays0 = sc.render(*args)
    rr, rays1 = sc.render(*args, roulette=True)
    rel = (rr.mean(axis=(0, 1)) - plain.mean(axis=(0, 1))) / plain.mean(axis=(0, 1))
    assert np.all(np.abs(rel) < 0.01), rel
    assert rays1 < 0.6 * rays0, (rays1, rays0)
    assert np.isfinite(rr).all()
