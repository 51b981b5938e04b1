"""The renderer's integrator seam (pbrt_render_config.integrator, the reference's
integrators::Type, src/integrators.rs:23-28 chosen in src/main.rs:16-20) on the device:

  * NORMAL (integrators/normal.rs:17-30) against the FP32 oracle: the same arithmetic on the same
    hits, so films agree to 1e-6 (bit-identical pixels counted), one ray per sample;
  * NAIVE (integrators/naive.rs:30-54) against the FP32 oracle at equal spp under the same draws,
    with the bars of the path integrator's film tests, and the closed form of a Lambertian floor
    under the integrator's own sky;
  * the options of the path integrator (mis, roulette, the radiance cache) and unknown kinds are
    refused with the other integrators; a job batch must agree on the integrator and equals its
    single renders.
"""
import math

import numpy as np
import pytest

import pbrt_amd
from oracle import pyoracle as po
from pbrt_amd import scenes
from test_oracle_integrators import _floor_under_sky, _plane_and_sphere, closed_room

pytestmark = pytest.mark.gpu


def rmse(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)))


@pytest.mark.parametrize("which", ["plane_sphere", "room"])
def test_normal_integrator_matches_oracle(gpu_ctx, which):
    if which == "room":
        desc, cam = scenes.room(subdiv=4)
        eye, look, fov = cam["eye"], cam["look"], cam["fov"]
    else:
        desc, eye, look, fov = _plane_and_sphere(), (0.0, 0.0, 4.0), (0.0, 0.0, 0.0), 0.9
    w, h, spp = 64, 48, 2
    scene = pbrt_amd.Scene(gpu_ctx, desc)
    film, st = scene.render(w, h, spp, 8, eye, look, fov=fov, seed=5, integrator="normal")
    ref, rays = po.Scene(desc, precision=32, **pbrt_amd.DEVICE_BVH).render(w, h, fov, eye, look, spp, 8, 5,
                                                                           integrator="normal")
    same = float((film == ref).all(-1).mean())
    print(f"[normal {which}] rmse {rmse(film, ref):.2e}, bit-identical pixels {same:.3f}", flush=True)
    assert st["rays"] == w * h * spp == rays
    assert np.abs(film - ref).max() < 1e-6
    assert (film > 0).any() and film.max() <= 1.0


def test_naive_integrator_matches_oracle_and_closed_form(gpu_ctx):
    desc, cam = scenes.room(subdiv=4)
    w = h = 64
    spp = 8
    scene = pbrt_amd.Scene(gpu_ctx, desc)
    film, st = scene.render(w, h, spp, cam["max_depth"], cam["eye"], cam["look"], fov=cam["fov"], seed=3,
                            integrator="naive")
    f32, rays = po.Scene(desc, precision=32, **pbrt_amd.DEVICE_BVH).render(w, h, cam["fov"], cam["eye"], cam["look"],
                                                                           spp, cam["max_depth"], 3,
                                                                           integrator="naive")
    scale = float(f32.mean())
    print(f"[naive room] rmse/mean {rmse(film, f32) / scale:.4f}, rays {st['rays']} vs {rays}", flush=True)
    assert scale > 0 and np.isfinite(film).all()
    assert rmse(film, f32) / scale < 0.02
    assert abs(film.mean() - f32.mean()) / scale < 0.01
    assert abs(st["rays"] - rays) < 0.005 * rays
    # no shadow rays: fewer rays than the path integrator's render of the same view
    _, st_path = scene.render(w, h, spp, cam["max_depth"], cam["eye"], cam["look"], fov=cam["fov"], seed=3, film=False)
    assert st["rays"] < st_path["rays"]
    # a Lambertian floor (albedo rho) under the naive sky: rho * (7/12, 3/4, 1)
    rho = 0.6
    floor = pbrt_amd.Scene(gpu_ctx, _floor_under_sky(rho))
    f, _ = floor.render(8, 8, 4096, 8, (0.0, 3.0, 0.0), (0.0, 0.0, 0.001), fov=0.5, seed=3, integrator="naive")
    np.testing.assert_allclose(f.reshape(-1, 3).mean(0), rho * np.array([7.0 / 12.0, 0.75, 1.0]),
                               atol=4 * rho * 0.25 / math.sqrt(64 * 4096))


def test_naive_and_path_agree_in_a_closed_lit_room(gpu_ctx):
    d, cam = closed_room()
    scene = pbrt_amd.Scene(gpu_ctx, d)
    args = dict(fov=cam["fov"])
    a, _ = scene.render(16, 16, 2048, 4, cam["eye"], cam["look"], seed=7, integrator="path", **args)
    b, _ = scene.render(16, 16, 2048, 4, cam["eye"], cam["look"], seed=8, integrator="naive", **args)
    assert abs(a.mean() - b.mean()) / a.mean() < 0.02, (a.mean(), b.mean())


def test_integrator_options_are_refused(gpu_ctx):
    desc, cam = scenes.small()
    scene = pbrt_amd.Scene(gpu_ctx, desc)
    args = (8, 8, 1, 4, cam["eye"], cam["look"])
    with pytest.raises(pbrt_amd.PbrtError, match="integrator must be"):
        scene.render(*args, integrator=3)
    for kind in ("naive", "normal"):
        with pytest.raises(pbrt_amd.PbrtError, match="options of the path integrator"):
            scene.render(*args, integrator=kind, mis=True)
        with pytest.raises(pbrt_amd.PbrtError, match="options of the path integrator"):
            scene.render(*args, integrator=kind, roulette=True)
        cache = pbrt_amd.RadianceCache(gpu_ctx, seed=1)
        with pytest.raises(pbrt_amd.PbrtError, match="options of the path integrator"):
            scene.render_ex(*args, integrator=kind, cache=cache)
    job = dict(width=16, height=8, spp=2, max_depth=4, eye=cam["eye"], look=cam["look"], fov=cam["fov"], seed=2)
    with pytest.raises(pbrt_amd.PbrtError, match="same for every job"):
        scene.render_jobs([job, dict(job, integrator="naive")])
    for kind in ("naive", "normal"):
        jobs = [dict(job, integrator=kind), dict(job, width=24, seed=5, integrator=kind)]
        films, _ = scene.render_jobs(jobs)
        for j, f in zip(jobs, films):
            alone, _ = scene.render(j["width"], j["height"], j["spp"], j["max_depth"], j["eye"], j["look"],
                                    fov=j["fov"], seed=j["seed"], integrator=kind)
            assert f.tobytes() == alone.tobytes()
