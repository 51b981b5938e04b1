"""The radiance cache inside the wavefront renderer.

Pins of the mechanism, bit-exact against the plain renderer under the same counter-based sampler:
- every path a training path (stride 1) and queries at vertex 1: no path ends into the cache, so
  the film is the plain film, and the target of vertex 1 — the radiance gathered over the window
  of max_depth - 1 vertices — is each path's own radiance (the small scene has no emitter the
  camera sees, so nothing is added before the window starts);
- no training paths and a cache that answers 0: every path ends at its third vertex with nothing
  added, which is the plain path tracer cut at max_depth 3 (the small scene is all Lambertian, so
  no path passes a specular vertex there).
And what the trained cache buys: on the room, a cache-terminated render after the cache has
trained is closer to the converged plain render than the same render whose paths end at the
third vertex with nothing (the method without the learned part), at equal samples per pixel.
"""
import numpy as np
import pytest

import pbrt_amd
from pbrt_amd import scenes

pytestmark = pytest.mark.gpu


def small_scene(ctx):
    desc, cam = scenes.small()
    return pbrt_amd.Scene(ctx, desc), cam


def test_all_training_paths_equal_plain_and_targets_are_path_radiance(gpu_ctx):
    scene, cam = small_scene(gpu_ctx)
    w = h = 32
    plain, _ = scene.render(w, h, 1, cam["max_depth"], cam["eye"], cam["look"], fov=cam["fov"], seed=5)
    cache = pbrt_amd.RadianceCache(gpu_ctx, seed=2)
    film, stats, prog = scene.render_ex(w, h, 1, cam["max_depth"], cam["eye"], cam["look"], fov=cam["fov"], seed=5,
                                        cache=cache, query_vertex=1, train_stride=1, record=False,
                                        debug_slots=w * h)
    assert film.tobytes() == plain.tobytes()
    raw, tgt = prog["debug"]
    valid = tgt[:, 3] == 1.0
    assert valid.mean() > 0.2  # primary hits on Lambertian surfaces
    # stride 1: the pass offset is 0 and slot r is path r, i.e. pixel r
    flat = plain.reshape(-1, 3).astype(np.float32)
    np.testing.assert_array_equal(tgt[valid, :3], flat[valid])
    # the cache inputs of vertex 1: outgoing direction unit length, albedo in the scene's range
    wo = raw[valid, 3:6]
    np.testing.assert_allclose((wo * wo).sum(1), 1.0, rtol=1e-5)
    assert (raw[valid, 9:] >= 0.2).all() and (raw[valid, 9:] <= 0.8).all()
    assert cache.steps == 1


def test_zero_cache_without_training_is_plain_cut_at_the_query_vertex(gpu_ctx):
    scene, cam = small_scene(gpu_ctx)
    w = h = 48
    cut, _ = scene.render(w, h, 4, 3, cam["eye"], cam["look"], fov=cam["fov"], seed=9)
    cache = pbrt_amd.RadianceCache(gpu_ctx, seed=2)
    cache.params = np.zeros(pbrt_amd.NRC_PARAMS, np.float32)
    film, stats, _ = scene.render_ex(w, h, 4, cam["max_depth"], cam["eye"], cam["look"], fov=cam["fov"], seed=9,
                                     cache=cache, query_vertex=3, train_stride=0, record=False)
    assert film.tobytes() == cut.tobytes()
    assert cache.steps == 0


def glass_box_scene():
    """A closed Lambertian box (six rectangles facing inwards) around a glass sphere; no lights."""
    import math
    from pbrt_amd.scene import SceneDesc, Transform
    d = SceneDesc()
    wall = d.add_material("lambertian", (0.5, 0.5, 0.5))
    glass = d.add_material("dielectric", (1.0, 1.0, 1.0), ior=1.5)
    H = 2.0
    for xf in [Transform.translation(0.0, -H, 0.0),
               Transform.translation(0.0, H, 0.0) * Transform.rotation("x", math.pi),
               Transform.translation(0.0, 0.0, -H) * Transform.rotation("x", math.pi / 2),
               Transform.translation(0.0, 0.0, H) * Transform.rotation("x", -math.pi / 2),
               Transform.translation(-H, 0.0, 0.0) * Transform.rotation("z", -math.pi / 2),
               Transform.translation(H, 0.0, 0.0) * Transform.rotation("z", math.pi / 2)]:
        d.add_object("rectangle", wall, (2 * H, 2 * H), transform=xf)
    d.add_object("sphere", glass, param=(0.9,), transform=Transform.translation(0.0, 0.0, -0.5))
    return d


def test_queries_past_a_specular_vertex_are_answered(gpu_ctx):
    """A cache answering the constant c, queried from vertex 1 on, no training: a path ends at its
    first Lambertian vertex with throughput 1 (Lambertian vertices before it: none; glass vertices:
    factors (1 / cos) cos), so every pixel is c, including those seen through the glass sphere,
    whose paths query at a later bounce than the first."""
    scene = pbrt_amd.Scene(gpu_ctx, glass_box_scene())
    w = h = 48
    c = np.array([0.25, 0.5, 0.75], np.float32)
    cache = pbrt_amd.RadianceCache(gpu_ctx, seed=2)
    params = np.zeros(pbrt_amd.NRC_PARAMS, np.float32)
    params[-3:] = c  # output bias; every weight 0
    cache.params = params
    film, _, _ = scene.render_ex(w, h, 4, 8, (0.0, 0.0, 1.8), (0.0, 0.0, -0.5), fov=1.2, seed=3, cache=cache,
                                 query_vertex=1, train_stride=0, record=False)
    # the sphere (radius 0.9, 2.3 from the eye) covers about a third of the view; a path caught
    # inside it past the depth limit gives 0, which is rare; (1 / cos) cos differs from 1 by
    # rounding (the cosines come from local and world coordinates)
    ratio = film / c
    assert np.mean(np.all(np.abs(ratio - 1.0) < 1e-4, axis=2)) > 0.97
    assert np.all(ratio <= 1.0 + 1e-4)


def test_cache_render_is_deterministic(gpu_ctx):
    """One frame per pass, so the three path sets rotate: consecutive passes' fronts run on two
    streams and each pass's queries are answered on a third while later fronts queue their own: the
    film is the same, bit for bit, on every run (each set owns its query buffers)."""
    scene = pbrt_amd.Scene(gpu_ctx, glass_box_scene())
    cache = pbrt_amd.RadianceCache(gpu_ctx, seed=2)
    params = np.zeros(pbrt_amd.NRC_PARAMS, np.float32)
    params[-3:] = [0.25, 0.5, 0.75]
    cache.params = params
    films = [scene.render_ex(48, 48, 6, 8, (0.0, 0.0, 1.8), (0.0, 0.0, -0.5), fov=1.2, seed=3, cache=cache,
                             query_vertex=1, train_stride=0, record=False, samples_per_pass=1)[0] for _ in range(3)]
    assert films[0].tobytes() == films[1].tobytes() == films[2].tobytes()


def test_fixed_cache_film_does_not_depend_on_the_pass_size(gpu_ctx):
    """Without training the cache is a fixed function, so the film cannot depend on how the frames
    are grouped into passes: one pass of six frames (one path set), six passes of one (three sets
    rotating over three streams), three of two and two of three give the same film, bit for bit."""
    scene = pbrt_amd.Scene(gpu_ctx, glass_box_scene())
    cache = pbrt_amd.RadianceCache(gpu_ctx, seed=2)
    params = np.zeros(pbrt_amd.NRC_PARAMS, np.float32)
    params[-3:] = [0.25, 0.5, 0.75]
    cache.params = params
    films = [scene.render_ex(48, 48, 6, 8, (0.0, 0.0, 1.8), (0.0, 0.0, -0.5), fov=1.2, seed=3, cache=cache,
                             query_vertex=2, train_stride=0, record=False, samples_per_pass=k)[0] for k in (6, 1, 2, 3)]
    assert films[0].mean() > 0
    for f in films[1:]:
        assert f.tobytes() == films[0].tobytes()


def test_plain_film_does_not_depend_on_the_pass_size(gpu_ctx):
    """Plain renders of several passes run consecutive passes on two streams with two path sets; the
    film takes the passes in order, so it equals the one-pass film bit for bit."""
    scene, cam = small_scene(gpu_ctx)
    args = (40, 24, 12, cam["max_depth"], cam["eye"], cam["look"])
    films = [scene.render(*args, fov=cam["fov"], seed=5, samples_per_pass=k)[0] for k in (12, 1, 5, 4)]
    assert films[0].mean() > 0
    for f in films[1:]:
        assert f.tobytes() == films[0].tobytes()


def test_progress_record(gpu_ctx):
    scene, cam = small_scene(gpu_ctx)
    w = h = 32
    ref, _ = scene.render(w, h, 64, cam["max_depth"], cam["eye"], cam["look"], fov=cam["fov"], seed=1)
    film, _, prog = scene.render_ex(w, h, 16, cam["max_depth"], cam["eye"], cam["look"], fov=cam["fov"], seed=3,
                                    samples_per_pass=2, reference=ref)
    assert len(prog["ms"]) == 8 and np.all(np.diff(prog["ms"]) > 0)
    final = np.sqrt(np.mean((film - ref) ** 2))
    assert prog["rmse"][-1] == pytest.approx(final, rel=1e-9)
    assert prog["rmse"][-1] < prog["rmse"][0]


@pytest.mark.parametrize("w,h,spp", [(32, 32, 16), (40, 24, 1), (48, 48, 7)])
def test_default_pass_size_record(gpu_ctx, w, h, spp):
    """A plain render left to choose its pass size (half its samples per pass, at most about 8M
    paths): the convergence record has one entry per pass the library ran, the last at the full
    sample count, and its RMSE is that of the returned film."""
    scene, cam = small_scene(gpu_ctx)
    ref, _ = scene.render(w, h, 32, cam["max_depth"], cam["eye"], cam["look"], fov=cam["fov"], seed=1)
    film, st, prog = scene.render_ex(w, h, spp, cam["max_depth"], cam["eye"], cam["look"], fov=cam["fov"], seed=3,
                                     reference=ref)
    assert len(prog["ms"]) == st["passes"] == (1 if spp == 1 else 2)
    assert prog["spp"][-1] == spp
    assert prog["rmse"][-1] == pytest.approx(float(np.sqrt(np.mean((film - ref) ** 2))), rel=1e-9)


def test_trained_cache_beats_termination_without_it(gpu_ctx):
    desc, cam = scenes.room(subdiv=4)
    scene = pbrt_amd.Scene(gpu_ctx, desc)
    w = h = 64
    depth = cam["max_depth"]
    args = dict(fov=cam["fov"], seed=21)
    ref, _ = scene.render(w, h, 2048, depth, cam["eye"], cam["look"], **args)
    spp = 96
    # the method without the learned part: paths end at the third vertex with nothing added
    zero = pbrt_amd.RadianceCache(gpu_ctx, seed=4)
    zero.params = np.zeros(pbrt_amd.NRC_PARAMS, np.float32)
    cut, _, _ = scene.render_ex(w, h, spp, depth, cam["eye"], cam["look"], cache=zero, train_stride=0,
                                record=False, **args)
    # the radiance cache: one path in 8 trains (a 64x64 image has few paths), 4 steps after every pass
    cache = pbrt_amd.RadianceCache(gpu_ctx, seed=4)
    cache.set_optim(lr=1e-2)
    nrc, stats, prog = scene.render_ex(w, h, spp, depth, cam["eye"], cam["look"], cache=cache, train_stride=8,
                                       train_batch=4096, train_steps=4, reference=ref, **args)
    rm_cut = np.sqrt(np.mean((cut - ref) ** 2))
    rm_nrc = np.sqrt(np.mean((nrc - ref) ** 2))
    bias_cut = abs(cut.mean() - ref.mean()) / ref.mean()
    bias_nrc = abs(nrc.mean() - ref.mean()) / ref.mean()
    print(f"rmse cut {rm_cut:.4f} nrc {rm_nrc:.4f}; mean bias cut {bias_cut:.3f} nrc {bias_nrc:.3f}; "
          f"cache steps {cache.steps}")
    assert cache.steps == spp * 4
    assert rm_nrc < rm_cut
    assert bias_nrc < bias_cut / 2
