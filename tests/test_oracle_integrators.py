"""The oracle's other two integrators of the reference, pinned by closed forms on the CPU:

  * NormalIntegrator (integrators/normal.rs:17-30): a pixel on a rectangle of normal +y is exactly
    (0.5, 1, 0.5); a pixel on a sphere is (n + 1) / 2 for a unit n, so 2c - 1 has unit length, and
    the pixel facing the camera is (0.5, 0.5, 1); a miss is black (integrators.rs:9-11);
  * NaiveIntegrator (integrators/naive.rs:30-54): over a Lambertian floor of albedo rho under the
    integrator's own sky (naive.rs:21-26) every floor pixel converges to rho * E[sky] with E over
    cosine-distributed directions: E[t] = (1 + E[cos]) / 2 = 5/6, so rho * (7/12, 3/4, 1); at depth 1
    the floor is black and a miss is the sky itself;
  * naive and path integrators converge to the same image in a closed room lit only by an emissive
    panel (the check ideas/area_light.md:131-137 plans): every light path of up to max_depth
    vertices is counted once by either estimator.
"""
import math

import numpy as np

from oracle import pyoracle as po
from pbrt_amd.scene import SceneDesc, Transform


def _plane_and_sphere():
    d = SceneDesc()
    m = d.add_material("lambertian", (0.5, 0.5, 0.5))
    d.add_object("rectangle", m, (8.0, 8.0), transform=Transform.translation(0.0, -1.0, 0.0))
    d.add_object("sphere", m, param=(0.7,), transform=Transform.translation(0.0, 0.0, 0.0))
    return d


def test_normal_integrator_closed_forms():
    d = _plane_and_sphere()
    eye, look, fov = (0.0, 0.0, 4.0), (0.0, 0.0, 0.0), 0.9
    w = h = 33
    for precision in (64, 32):
        sc = po.Scene(d, precision=precision)
        film, rays = sc.render(w, h, fov, eye, look, 1, 8, 5, integrator="normal")
        assert rays == w * h  # one camera ray per sample, nothing else
        c = film.reshape(-1, 3)
        black = (c == 0).all(1)
        plane = np.isclose(c, (0.5, 1.0, 0.5), atol=1e-6).all(1)
        sphere = ~black & ~plane
        assert black.any() and plane.any() and sphere.sum() > 100
        n = 2 * c[sphere] - 1
        np.testing.assert_allclose(np.linalg.norm(n, axis=1), 1.0, atol=1e-5 if precision == 32 else 1e-12)
        # the centre pixel's sample is jittered within the pixel: up to ~0.06 off the axis normal
        mid = film[h // 2, w // 2]
        np.testing.assert_allclose(mid, (0.5, 0.5, 1.0), atol=0.04)
        assert mid[2] > 0.99


def _floor_under_sky(rho):
    d = SceneDesc()
    m = d.add_material("lambertian", (rho, rho, rho))
    d.add_object("rectangle", m, (1e4, 1e4))
    return d


def test_naive_integrator_sky_over_lambertian_floor():
    rho = 0.6
    d = _floor_under_sky(rho)
    eye, look, fov = (0.0, 3.0, 0.0), (0.0, 0.0, 0.001), 0.5  # straight down at the floor
    sc = po.Scene(d, precision=64)
    film, _ = sc.render(8, 8, fov, eye, look, 4096, 8, 3, integrator="naive")
    expect = rho * np.array([7.0 / 12.0, 0.75, 1.0])
    mean = film.reshape(-1, 3).mean(0)
    # per sample the value is rho * sky(d) with sky within [0.5, 1]: std <= rho / 4 per channel
    np.testing.assert_allclose(mean, expect, atol=4 * rho * 0.25 / math.sqrt(64 * 4096))
    one, _ = sc.render(8, 8, fov, eye, look, 4, 1, 3, integrator="naive")
    assert (one == 0).all()  # depth 1: the floor's (zero) emission only
    up, _ = sc.render(4, 4, 0.05, (0.0, 3.0, 0.0), (0.0, 10.0, 0.001), 1, 8, 3, integrator="naive")
    # looking straight up (within 0.04 rad): the sky at y = 1 is (0.5, 0.7, 1.0)
    np.testing.assert_allclose(up.reshape(-1, 3), np.tile([0.5, 0.7, 1.0], (16, 1)), atol=1e-3)


def closed_room():
    """A closed box (six inward walls, albedo 0.7, one red) lit only by an emissive ceiling panel
    covering a quarter of the ceiling; the camera stands inside."""
    d = SceneDesc()
    H = 1.0
    white = d.add_material("lambertian", (0.7, 0.7, 0.7))
    red = d.add_material("lambertian", (0.7, 0.1, 0.1))
    wall = (2 * H, 2 * H)
    d.add_object("rectangle", white, wall, transform=Transform.translation(0.0, -H, 0.0))
    d.add_object("rectangle", white, wall, transform=Transform.translation(0.0, H, 0.0) * Transform.rotation("x", math.pi))
    d.add_object("rectangle", white, wall,
                 transform=Transform.translation(0.0, 0.0, -H) * Transform.rotation("x", math.pi / 2))
    d.add_object("rectangle", white, wall,
                 transform=Transform.translation(0.0, 0.0, H) * Transform.rotation("x", -math.pi / 2))
    d.add_object("rectangle", red, wall, transform=Transform.translation(-H, 0.0, 0.0) * Transform.rotation("z", -math.pi / 2))
    d.add_object("rectangle", white, wall, transform=Transform.translation(H, 0.0, 0.0) * Transform.rotation("z", math.pi / 2))
    light = d.add_material("diffuse_light", (4.0, 4.0, 4.0))
    panel = d.add_object("rectangle", light, (H, H),
                         transform=Transform.translation(0.0, H - 1e-3, 0.0) * Transform.rotation("x", math.pi))
    d.add_area_light(panel, (4.0, 4.0, 4.0))
    cam = dict(eye=(0.0, 0.0, 0.9), look=(0.0, 0.0, -1.0), fov=1.2)
    return d, cam


def test_naive_and_path_converge_to_the_same_image():
    d, cam = closed_room()
    sc = po.Scene(d, precision=64)
    w = h = 8
    spp, depth = 2048, 4
    a, _ = sc.render(w, h, cam["fov"], cam["eye"], cam["look"], spp, depth, 7, integrator="path")
    b, _ = sc.render(w, h, cam["fov"], cam["eye"], cam["look"], spp, depth, 8, integrator="naive")
    # image means within 2 %, and pixels within 10 % where the naive estimate is not light-noisy
    assert abs(a.mean() - b.mean()) / a.mean() < 0.02, (a.mean(), b.mean())
    rel = np.abs(a - b).mean(-1) / a.mean(-1)
    assert np.median(rel) < 0.1
    # depth 1: both see only the panel itself
    a1, _ = sc.render(w, h, cam["fov"], cam["eye"], cam["look"], 4, 1, 7, integrator="path")
    b1, _ = sc.render(w, h, cam["fov"], cam["eye"], cam["look"], 4, 1, 7, integrator="naive")
    assert np.array_equal(a1, b1)
