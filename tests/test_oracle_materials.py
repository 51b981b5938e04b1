"""The oracle's Metal, CheckerBoard and BackgroundInfiniteLight (materials/metal.rs,
textures/checker_board.rs, lights/background_infinite_light.rs), pinned on closed forms:

- a checkerboard emitter seen directly: every pixel whose footprint lies inside one texel has that
  texel's colour, the texel given by an independent restatement of CheckerBoard::shade at the
  pixel centre's hit point (u = x, v = z on a rectangle);
- the background seen directly: le(d) = down (1 - f) + up f, f = (d.y + 1) / 2;
- a Lambertian floor under a uniform background of radiance 1, one bounce: next-event estimation
  over the sphere gives radiance = albedo (the irradiance of the upper hemisphere is pi), with or
  without MIS;
- a mirror (Metal, fuzz 0) under the background: the pixel is albedo x le(the mirrored camera ray).
"""
import math

import numpy as np

import pbrt_amd
from oracle import pyoracle as po
from pbrt_amd.scene import SceneDesc

EYE, LOOK, FOV = (0.3, 1.5, 2.0), (0.0, 0.0, 0.0), 0.6


def checker_texel_is_first(u, v, scale):
    su = math.fmod(abs(u) * scale, 1.0)
    sv = math.fmod(abs(v) * scale, 1.0)
    use_color = (su < 0.5 and sv < 0.5) or (su >= 0.5 and sv >= 0.5)
    positive = math.copysign(1.0, u) * math.copysign(1.0, v) >= 0.0
    return use_color != positive


def centre_rays(w, h):
    r = pbrt_amd.pinhole_rays(w, h, FOV, EYE, LOOK)
    o = np.stack([r["ox"], r["oy"], r["oz"]], 1).astype(np.float64)
    d = np.stack([r["dx"], r["dy"], r["dz"]], 1).astype(np.float64)
    return o.reshape(h, w, 3), d.reshape(h, w, 3)


def test_checkerboard_emitter_seen_directly():
    c1, c2, scale = (0.9, 0.2, 0.1), (0.1, 0.3, 0.8), 0.8
    d = SceneDesc()
    m = d.add_material("diffuse_light", c1, checker=(c2, scale))
    d.add_object("rectangle", m, (8.0, 8.0))
    w = h = 48
    film, _ = po.Scene(d, precision=64).render(w, h, FOV, EYE, LOOK, 4, 2, 1)
    o, dr = centre_rays(w, h)
    t = -o[..., 1] / dr[..., 1]
    p = o + dr * t[..., None]
    first = np.vectorize(lambda u, v: checker_texel_is_first(u, v, scale))(p[..., 0], p[..., 2])
    expect = np.where(first[..., None], np.array(c1), np.array(c2))
    # pixels whose 8 neighbours share the texel: their whole footprint is inside it
    same = np.ones_like(first)
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            same &= np.roll(np.roll(first, dy, 0), dx, 1) == first
    same[[0, -1], :] = same[:, [0, -1]] = False
    assert same.mean() > 0.3 and first[same].any() and (~first[same]).any()
    np.testing.assert_array_equal(film[same], expect[same])


def test_background_seen_directly():
    d = SceneDesc()
    d.add_material("lambertian", (0.5, 0.5, 0.5))
    d.add_background_light((1.0, 1.0, 1.0), (0.5, 0.7, 1.0))
    w = h = 32
    film, _ = po.Scene(d, precision=64).render(w, h, FOV, EYE, LOOK, 8, 3, 2)
    _, dr = centre_rays(w, h)
    f = 0.5 * (dr[..., 1] / np.linalg.norm(dr, axis=-1) + 1.0)
    expect = np.array([1.0, 1.0, 1.0]) * (1 - f[..., None]) + np.array([0.5, 0.7, 1.0]) * f[..., None]
    np.testing.assert_allclose(film, expect, atol=2e-3)  # jitter within the pixel


def test_uniform_background_lights_a_lambertian_floor():
    albedo = np.array([0.2, 0.5, 0.8])
    d = SceneDesc()
    m = d.add_material("lambertian", albedo)
    d.add_object("rectangle", m, (50.0, 50.0))
    d.add_background_light((1.0, 1.0, 1.0), (1.0, 1.0, 1.0))
    for mis in (False, True):
        film, _ = po.Scene(d, precision=64).render(16, 16, FOV, EYE, LOOK, 1024, 2, 3, mis=mis)
        mean = film.reshape(-1, 3).mean(0)
        np.testing.assert_allclose(mean, albedo, rtol=0.02)


def test_mirror_reflects_the_background():
    albedo = np.array([0.9, 0.8, 0.7])
    d = SceneDesc()
    m = d.add_material("metal", albedo, fuzz=0.0)
    d.add_object("rectangle", m, (50.0, 50.0))
    d.add_background_light((0.2, 0.2, 0.2), (0.4, 0.9, 1.0))
    w = h = 32
    film, _ = po.Scene(d, precision=64).render(w, h, FOV, EYE, LOOK, 8, 3, 4)
    _, dr = centre_rays(w, h)
    dr = dr / np.linalg.norm(dr, axis=-1, keepdims=True)
    refl_y = -dr[..., 1]  # the mirror about the floor flips y
    f = 0.5 * (refl_y + 1.0)
    le = np.array([0.2, 0.2, 0.2]) * (1 - f[..., None]) + np.array([0.4, 0.9, 1.0]) * f[..., None]
    hit = dr[..., 1] < 0
    np.testing.assert_allclose(film[hit], (albedo * le)[hit], atol=3e-3)
