"""The Sobol sampler on the device: the same stratified draws as the oracle's (rng.hpp restated in
oracle/scene_oracle.hpp), so device films match the FP32 oracle's at equal spp under the Sobol
sampler with the bars of the other film tests; a job batch equals its single renders; the option is
refused when out of range or mixed within a batch; at equal spp the Sobol render is closer to a
high-spp reference than the independent one."""
import numpy as np
import pytest

import pbrt_amd
from oracle import pyoracle as po
from pbrt_amd import scenes

pytestmark = pytest.mark.gpu


def rmse(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)))


@pytest.mark.parametrize("which", ["room_ggx", "small"])
def test_sobol_render_matches_oracle(gpu_ctx, which):
    desc, cam = scenes.room(subdiv=4, ggx=True) if which == "room_ggx" else scenes.small()
    w = h = 64
    spp = 8
    scene = pbrt_amd.Scene(gpu_ctx, desc)
    film, _ = scene.render(w, h, spp, cam["max_depth"], cam["eye"], cam["look"], fov=cam["fov"], seed=3,
                           sampler="sobol")
    ref, _ = po.Scene(desc, precision=32, **pbrt_amd.DEVICE_BVH).render(w, h, cam["fov"], cam["eye"], cam["look"], spp,
                                                                        cam["max_depth"], 3, sampler="sobol")
    scale = float(ref.mean())
    same = float((film == ref).all(-1).mean())
    print(f"[sobol {which}] rmse/mean {rmse(film, ref) / scale:.4f}, bit-identical pixels {same:.3f}", flush=True)
    assert scale > 0 and np.isfinite(film).all()
    assert rmse(film, ref) / scale < 0.02
    assert abs(film.mean() - ref.mean()) / scale < 0.01
    indep, _ = scene.render(w, h, spp, cam["max_depth"], cam["eye"], cam["look"], fov=cam["fov"], seed=3)
    assert not np.array_equal(indep, film)


def test_sobol_converges_faster_and_batches_agree(gpu_ctx):
    desc, cam = scenes.room(subdiv=4, ggx=True)
    scene = pbrt_amd.Scene(gpu_ctx, desc)
    w = h = 96
    args = (cam["max_depth"], cam["eye"], cam["look"])
    ref, _ = scene.render(w, h, 4096, *args, fov=cam["fov"], seed=99)
    for spp in (16, 64, 256):
        ind, _ = scene.render(w, h, spp, *args, fov=cam["fov"], seed=3)
        sob, _ = scene.render(w, h, spp, *args, fov=cam["fov"], seed=3, sampler="sobol")
        print(f"[sobol vs independent at {spp} spp] rmse {rmse(sob, ref):.4f} vs {rmse(ind, ref):.4f}", flush=True)
        assert rmse(sob, ref) < rmse(ind, ref)
    job = dict(width=32, height=24, spp=6, max_depth=6, eye=cam["eye"], look=cam["look"], fov=cam["fov"], seed=2,
               sampler="sobol")
    jobs = [job, dict(job, width=48, spp=3, seed=5)]
    films, _ = scene.render_jobs(jobs)
    for j, f in zip(jobs, films):
        alone, _ = scene.render(j["width"], j["height"], j["spp"], j["max_depth"], j["eye"], j["look"], fov=j["fov"],
                                seed=j["seed"], sampler="sobol")
        assert f.tobytes() == alone.tobytes()
    with pytest.raises(pbrt_amd.PbrtError, match="same for every job"):
        scene.render_jobs([job, dict(job, sampler="independent")])
    with pytest.raises(pbrt_amd.PbrtError, match="sampler must be"):
        scene.render(8, 8, 1, 4, cam["eye"], cam["look"], sampler=2)
