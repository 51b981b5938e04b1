"""The batched-jobs and radiance-cache workloads at their own sizes, through the C ABI.

Batched jobs: 256 instanced meshes (9,066,296 triangles: displaced icospheres and tori, 1k-200k
each, under rigid transforms) and 32 CSG objects (scenes.batch_scene), eight jobs from 256^2 to
1024^2 (scenes.batch_jobs):
  * random rays and job 0's camera rays bit-exact against Oracle B (object, triangle, t), any hit
    equal;
  * the eight jobs at 1 spp rendered as one batch (per-bounce graphs, and eagerly) equal their
    single renders bit for bit;
  * a 128^2 render of job 0's view against Oracle B's film: RMSE / mean < 0.005, 98 % of the
    pixels within 1e-3 relative and 70 % bit-identical (measured 0.0004, 99.8 %, 87.7 %).

Radiance cache: the room with the 1,310,720-triangle mesh and GGX materials at 1280x720
(scenes.nrc_room):
  * 16 frames of the cache render: a finite film, one training step per frame, queries on
    30-70 % of the paths, and more radiance than the same render with a zero cache;
  * plain path tracing at 4 spp on a band of rows 344-375 against Oracle B's band: RMSE / mean
    < 0.01, 98 % of the pixels within 1e-3 relative and 40 % bit-identical (measured 0.0019,
    100 %, 61 %: glossy paths amplify the ulp differences of the device and host libm).
The fraction of pixels within 1e-3 relative of the oracle's is printed for each film comparison.
"""
import os

import numpy as np
import pytest

import pbrt_amd
from oracle import pyoracle as po
from pbrt_amd import scenes

pytestmark = pytest.mark.gpu

THREADS = min(16, len(os.sched_getaffinity(0)))


def _rows(r):
    return np.ascontiguousarray(r).view(np.float32).reshape(-1, 8)


def _compare_scene_hits(h, ref):
    np.testing.assert_array_equal(h["object"].astype(np.int64), ref["obj"])
    hit = ref["obj"] >= 0
    np.testing.assert_array_equal(h["prim"][hit].astype(np.int64), ref["prim"][hit])
    np.testing.assert_array_equal(h["t"][hit].astype(np.float64), ref["t"][hit])
    return int(hit.sum())


def _film_agreement(film, ref, what):
    scale = float(ref.mean())
    rel = float(np.sqrt(np.mean((film - ref) ** 2))) / scale
    close = np.abs(film - ref) <= 1e-3 * np.maximum(np.abs(ref), 1e-3 * scale)
    frac = float(close.all(-1).mean())
    same = float((film == ref).all(-1).mean())
    print(f"[{what}] rmse/mean {rel:.4f}; mean {film.mean():.5f} vs {ref.mean():.5f}; pixels within 1e-3: "
          f"{frac:.3f}; bit-identical: {same:.3f}", flush=True)
    return scale, rel, frac, same


@pytest.fixture(scope="module")
def c4(gpu_ctx):
    desc = scenes.batch_scene()
    assert len(desc.meshes) == 256 and sum(len(f) for _, f in desc.meshes) == 9066296
    scene = pbrt_amd.Scene(gpu_ctx, desc)
    orc = po.Scene(desc, precision=32, **pbrt_amd.DEVICE_BVH)
    return desc, scene, orc


def test_c4_rays_bit_exact(c4):
    _, scene, orc = c4
    rays = _rows(pbrt_amd.random_rays((0.0, 0.0, 0.0), 30.0, 300000, seed=4))
    assert _compare_scene_hits(scene.trace_host(rays), orc.trace(rays, threads=THREADS)) > 50000
    np.testing.assert_array_equal(scene.trace_host(rays, any_hit=True).astype(bool),
                                  orc.trace(rays, any_hit=True, threads=THREADS)["obj"] >= 0)
    j = scenes.batch_jobs()[0]
    cam = po.pinhole_rays(j["width"], j["height"], j["fov"], j["eye"], j["look"]).astype(np.float32)
    assert _compare_scene_hits(scene.trace_host(cam), orc.trace(cam, threads=THREADS)) > 0.3 * len(cam)
    assert scene.status() == 0


@pytest.mark.parametrize("use_graphs", [True, False])
def test_c4_eight_jobs_batch_equals_single_renders(c4, use_graphs):
    _, scene, _ = c4
    jobs = [dict(j, spp=1) for j in scenes.batch_jobs()]
    assert len(jobs) == 8 and len({(j["width"], j["height"]) for j in jobs}) >= 2
    films, stats = scene.render_jobs(jobs, use_graphs=use_graphs)
    assert stats["paths"] == sum(j["width"] * j["height"] for j in jobs)
    for j, f in zip(jobs, films):
        assert np.isfinite(f).all() and f.mean() > 0
        alone, _ = scene.render(j["width"], j["height"], 1, j["max_depth"], j["eye"], j["look"], fov=j["fov"],
                                seed=j["seed"])
        assert f.tobytes() == alone.tobytes()


def test_c4_sixteen_small_jobs_batch_equals_single_renders(c4):
    """The API's largest batch (16 jobs), each too small to fill the GPU alone: every film equals the
    job rendered alone, bit for bit."""
    _, scene, _ = c4
    jobs = scenes.small_jobs()
    assert len(jobs) == 16
    films, stats = scene.render_jobs(jobs, use_graphs=True)
    assert stats["paths"] == sum(j["width"] * j["height"] * j["spp"] for j in jobs)
    for j, f in zip(jobs, films):
        alone, _ = scene.render(j["width"], j["height"], j["spp"], j["max_depth"], j["eye"], j["look"], fov=j["fov"],
                                seed=j["seed"])
        assert f.tobytes() == alone.tobytes()


def test_c4_render_matches_oracle(c4):
    _, scene, orc = c4
    j = scenes.batch_jobs()[0]
    w = h = 128
    spp = 4
    film, _ = scene.render(w, h, spp, j["max_depth"], j["eye"], j["look"], fov=j["fov"], seed=j["seed"])
    ref, _ = orc.render(w, h, j["fov"], j["eye"], j["look"], spp, j["max_depth"], j["seed"], threads=THREADS)
    scale, rel, frac, same = _film_agreement(film, ref, "c4 job 0, 128^2 x 4 spp")
    assert scale > 0 and rel < 0.005 and frac > 0.98 and same > 0.7
    assert abs(film.mean() - ref.mean()) / scale < 0.002


@pytest.fixture(scope="module")
def c5(gpu_ctx):
    desc, cam = scenes.nrc_room()
    assert sum(len(f) for _, f in desc.meshes) == 1310720
    return desc, cam, pbrt_amd.Scene(gpu_ctx, desc)


def test_c5_cache_render_at_full_size(gpu_ctx, c5):
    _, cam, scene = c5
    W, H, D = cam["width"], cam["height"], cam["max_depth"]
    assert (W, H) == (1280, 720)
    frames = 16
    view = dict(fov=cam["fov"], seed=8)
    cache = pbrt_amd.RadianceCache(gpu_ctx, seed=4)
    film, st, prog = scene.render_ex(W, H, frames, D, cam["eye"], cam["look"], cache=cache, samples_per_pass=8,
                                     query_vertex=3, train_stride=33, train_batch=65536, train_steps=1, **view)
    zero = pbrt_amd.RadianceCache(gpu_ctx, seed=4)
    zero.params = np.zeros(pbrt_amd.NRC_PARAMS, np.float32)
    cut, _, _ = scene.render_ex(W, H, frames, D, cam["eye"], cam["look"], cache=zero, samples_per_pass=8,
                                train_stride=0, **view)
    q = st["cache_queries"] / (W * H * frames)
    print(f"[c5] {frames} frames: {st['rays']} rays, queries per path {q:.3f}, cache steps {cache.steps}, "
          f"mean {film.mean():.4f} (zero cache {cut.mean():.4f}), {prog['ms'][-1]:.1f} ms", flush=True)
    assert film.shape == (H, W, 3) and np.isfinite(film).all()
    assert cache.steps == frames
    assert 0.3 < q < 0.7
    assert film.mean() > cut.mean() * 1.05
    assert scene.status() == 0


def test_c5_plain_band_matches_oracle(c5):
    desc, cam, scene = c5
    W, H, D = cam["width"], cam["height"], cam["max_depth"]
    spp, seed = 4, 9
    film, _ = scene.render(W, H, spp, D, cam["eye"], cam["look"], fov=cam["fov"], seed=seed)
    orc = po.Scene(desc, precision=32, **pbrt_amd.DEVICE_BVH)
    y0, y1 = 344, 376
    ref, _ = orc.render(W, H, cam["fov"], cam["eye"], cam["look"], spp, D, seed, rows=(y0, y1), threads=THREADS)
    scale, rel, frac, same = _film_agreement(film[y0:y1], ref, "c5 plain band, 1280 x 32 x 4 spp")
    assert scale > 0 and rel < 0.01 and frac > 0.98 and same > 0.4
    assert abs(film[y0:y1].mean() - ref.mean()) / scale < 0.003
