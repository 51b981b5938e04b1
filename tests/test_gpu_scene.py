"""Parity of the device scene traversal and the wavefront renderer with the CPU oracle.

Scene hits: object, triangle and distance bit-exact against the FP32 oracle (same statements, same
tree shapes, order-independent ties); any-hit equal. Published scene-mode hit counts
(docs/mesures_bvh.md:729-738) on the device. Images: the device film against the FP32 and f64
oracle films at equal spp under the same counter-based sampler, within the per-pixel RMSE stated
in each test (only transcendental functions differ between the device and the host libm, so most
paths are identical and the residual comes from the few that diverge).
"""
import json
import math
import os

import numpy as np
import pytest

import pbrt_amd
from conftest import ROOT, load_mesh
from oracle import pyoracle as po
from pbrt_amd import scenes
from pbrt_amd.scene import SceneDesc, Transform, from_stage_json

pytestmark = pytest.mark.gpu

STAGES = {"cornell_box_canonical": 5027, "bunny_mesh": 16336, "many_spheres": 16296, "many_meshes": 6831}


def stage(name):
    with open(os.path.join(ROOT, "tests", "golden", "stages", name + ".json")) as f:
        sc = json.load(f)
    return sc, from_stage_json(sc, {n + ".ply": load_mesh(n) for n in ("bunny", "bun_zipper_res4")})


def compare_hits(h, ref):
    np.testing.assert_array_equal(h["object"].astype(np.int64), ref["obj"])
    hit = ref["obj"] >= 0
    np.testing.assert_array_equal(h["prim"][hit].astype(np.int64), ref["prim"][hit])
    np.testing.assert_array_equal(h["t"][hit].astype(np.float64), ref["t"][hit])


@pytest.mark.parametrize("name", list(STAGES))
def test_stage_scenes_bit_exact(gpu_ctx, name):
    sc, desc = stage(name)
    c = sc["camera"]
    rays = po.pinhole_rays(200, 150, 90.0 * math.pi / 180.0, c["pos"], c["look"], c["up"]).astype(np.float32)
    scene = pbrt_amd.Scene(gpu_ctx, desc, **pbrt_amd.DEVICE_BVH)
    orc = po.Scene(desc, precision=32, **pbrt_amd.DEVICE_BVH)
    h = scene.trace_host(rays)
    ref = orc.trace(rays)
    assert scene.status() == 0
    compare_hits(h, ref)
    assert abs(int((h["object"] >= 0).sum()) - STAGES[name]) <= 4
    # shadow rays towards the loader's light at (0, 2, 1) from the oracle's FP32 hit points
    hit = np.nonzero(ref["obj"] >= 0)[0]
    p = ref["p"][hit].astype(np.float32).astype(np.float64)
    to = np.array([0.0, 2.0, 1.0]) - p
    dist = np.sqrt((to * to).sum(1))
    d = to / dist[:, None]
    d = d / np.sqrt((d * d).sum(1))[:, None]
    sh = np.concatenate([p, np.full((len(p), 1), 1e-5), d, dist[:, None]], 1).astype(np.float32)
    occ = scene.trace_host(sh, any_hit=True)
    ref_occ = orc.trace(sh, any_hit=True)["obj"] >= 0
    np.testing.assert_array_equal(occ.astype(bool), ref_occ)


@pytest.mark.parametrize("which", ["small", "room"])
def test_synthetic_scenes_random_rays(gpu_ctx, which):
    desc, cam = scenes.small() if which == "small" else scenes.room(subdiv=5)
    center = (0.0, 0.0, 0.0)
    radius = 12.0 if which == "small" else 1.9
    rays = np.ascontiguousarray(pbrt_amd.random_rays(center, radius, 60000, seed=5)).view(np.float32).reshape(-1, 8)
    scene = pbrt_amd.Scene(gpu_ctx, desc)
    orc = po.Scene(desc, precision=32, **pbrt_amd.DEVICE_BVH)
    compare_hits(scene.trace_host(rays), orc.trace(rays))
    np.testing.assert_array_equal(scene.trace_host(rays, any_hit=True).astype(bool), orc.trace(rays, any_hit=True)["obj"] >= 0)


def rmse(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)))


@pytest.mark.parametrize("which", ["small", "room"])
def test_render_matches_oracle(gpu_ctx, which):
    """The small scene (5,120-triangle icosphere, floor, 64 spheres, point light) and the room with
    its 81,920-triangle dielectric icosphere, at 64^2 x 8 spp, against the FP32 oracle's film under
    the same sampler: RMSE below 0.2 % of the mean radiance, 99.5 % of the pixels within 1e-3
    relative, and most pixels bit-identical on the small scene (measured 0.033 %, 99.98 %, 94 %; the
    room 0.001 %, 100 %, 27 %: its dielectric and area-light paths pass through more transcendental
    functions, where the device and host libm differ in the last bit); against the f64 oracle within
    5 % (measured 1.9 % and 0.09 %)."""
    desc, cam = scenes.small() if which == "small" else scenes.room()
    w = h = 64
    spp = 8
    scene = pbrt_amd.Scene(gpu_ctx, desc)
    film, stats = scene.render(w, h, spp, cam["max_depth"], cam["eye"], cam["look"], fov=cam["fov"], seed=3)
    f32, _ = po.Scene(desc, precision=32, **pbrt_amd.DEVICE_BVH).render(w, h, cam["fov"], cam["eye"], cam["look"], spp,
                                                                    cam["max_depth"], 3)
    f64, _ = po.Scene(desc, precision=64).render(w, h, cam["fov"], cam["eye"], cam["look"], spp, cam["max_depth"], 3)
    scale = float(f64.mean())
    assert scale > 0
    # stated bars, relative to the mean radiance of the image
    close = np.abs(film - f32) <= 1e-3 * np.maximum(np.abs(f32), 1e-3 * scale)
    frac, same = float(close.all(-1).mean()), float((film == f32).all(-1).mean())
    print(f"[{which}] vs FP32 oracle: rmse/mean {rmse(film, f32) / scale:.5f}, within 1e-3 {frac:.4f}, "
          f"bit-identical {same:.4f}; vs f64 rmse/mean {rmse(film, f64) / scale:.5f}", flush=True)
    assert rmse(film, f32) / scale < 0.002, rmse(film, f32) / scale
    assert frac > 0.995 and same > (0.8 if which == "small" else 0.2), (frac, same)
    assert rmse(film, f64) / scale < 0.05, rmse(film, f64) / scale
    assert abs(film.mean() - f64.mean()) / scale < 0.01
    assert stats["paths"] == w * h * spp and stats["rays"] > stats["paths"]
    # determinism: the same seed gives the same film, bit for bit
    film2, _ = scene.render(w, h, spp, cam["max_depth"], cam["eye"], cam["look"], fov=cam["fov"], seed=3,
                            samples_per_pass=3)
    assert film.tobytes() == film2.tobytes()


def test_point_lit_plane_closed_form(gpu_ctx):
    """A Lambertian plane (albedo rho) under a point light of intensity I, seen straight down with
    max_depth 2 (one vertex, one shadow ray): every pixel is rho / pi * I * cos(theta) / d^2 at the
    point the pixel's ray hits (path.rs:69-79 with point_light.rs:23-35)."""
    desc = SceneDesc()
    rho = 0.6
    m = desc.add_material("lambertian", (rho, rho, rho))
    desc.add_object("rectangle", m, (4.0, 4.0))
    I, lp = 5.0, np.array([0.3, 2.0, -0.2])
    desc.add_point_light(tuple(lp), (I, I, I))
    scene = pbrt_amd.Scene(gpu_ctx, desc)
    w = 32
    eye, look = (0.0, 3.0, 0.0), (0.0, 0.0, 0.0)
    film, _ = scene.render(w, w, 1, 2, eye, look, up=(0.0, 0.0, -1.0), fov=0.6, seed=9)
    rays = po.pinhole_rays(w, w, 0.6, eye, look, (0.0, 0.0, -1.0))
    # pixel-centre closed form; jitter moves the point by under a pixel, hence the tolerance
    o, d = rays[:, :3], rays[:, 4:7]
    t = -o[:, 1] / d[:, 1]
    p = o + d * t[:, None]
    wi = lp - p
    dist2 = (wi * wi).sum(1)
    cos = wi[:, 1] / np.sqrt(dist2)
    expect = rho / math.pi * I * cos / dist2
    got = film[..., 0].reshape(-1)
    assert np.abs(got - expect).max() / expect.max() < 0.02


def csg_scene(seed=3):
    """Union, intersection and subtraction objects of spheres and boxes, some under a rotation and
    translation of the whole object, over a floor; a point light and Lambertian materials."""
    rng = np.random.default_rng(seed)
    d = SceneDesc()
    floor = d.add_material("lambertian", (0.7, 0.7, 0.7))
    d.add_object("rectangle", floor, (20.0, 20.0), transform=Transform.translation(0.0, -1.5, 0.0))
    for k in range(12):
        op = ("union", "intersection", "subtraction")[k % 3]
        els = []
        for j in range(2 + k % 3):
            c = rng.uniform(-0.5, 0.5, 3) * (0.4 if op == "intersection" else 1.0)
            xf = Transform.translation(*c) * Transform.rotation("y", float(rng.uniform(-1, 1)))
            if rng.uniform() < 0.5:
                els.append(("sphere", (float(rng.uniform(0.4, 0.8)),), xf))
            else:
                els.append(("aabox", tuple(rng.uniform(0.5, 1.2, 3)), xf))
        pos = Transform.translation(*(rng.uniform(-6, 6, 3) * np.array([1.0, 0.2, 1.0])))
        obj_xf = pos * Transform.rotation("x", float(rng.uniform(-1, 1))) if k % 2 else pos
        d.add_csg(op, els, d.add_material("lambertian", rng.uniform(0.2, 0.8, 3)), transform=obj_xf)
    d.add_point_light((0.0, 8.0, 4.0), (200.0, 200.0, 200.0))
    return d


def test_csg_stage_scene_bit_exact(gpu_ctx):
    with open(os.path.join(ROOT, "tests", "golden", "stages", "csg_union_spheres.json")) as f:
        sc = json.load(f)
    desc = from_stage_json(sc, {})
    c = sc["camera"]
    rays = po.pinhole_rays(200, 150, 90.0 * math.pi / 180.0, c["pos"], c["look"], c["up"]).astype(np.float32)
    scene = pbrt_amd.Scene(gpu_ctx, desc)
    orc = po.Scene(desc, precision=32, **pbrt_amd.DEVICE_BVH)
    h = scene.trace_host(rays)
    ref = orc.trace(rays)
    compare_hits(h, ref)
    csg_hit = ref["obj"] == 1
    assert csg_hit.sum() > 2000 and len(np.unique(ref["prim"][csg_hit])) == 9  # every sphere shows
    # f64 (the reference's arithmetic) agrees on the object of every hit but a handful at silhouettes
    f64 = po.Scene(desc, precision=64).trace(rays.astype(np.float64))
    assert (f64["obj"] != h["object"]).sum() <= 8


def test_csg_random_rays_bit_exact(gpu_ctx):
    desc = csg_scene()
    rays = np.ascontiguousarray(pbrt_amd.random_rays((0.0, 0.0, 0.0), 7.0, 80000, seed=11)).view(np.float32)
    rays = rays.reshape(-1, 8)
    scene = pbrt_amd.Scene(gpu_ctx, desc)
    orc = po.Scene(desc, precision=32, **pbrt_amd.DEVICE_BVH)
    h = scene.trace_host(rays)
    ref = orc.trace(rays)
    compare_hits(h, ref)
    hit_csg = ref["obj"] >= 1
    assert hit_csg.sum() > 1000
    np.testing.assert_array_equal(scene.trace_host(rays, any_hit=True).astype(bool), orc.trace(rays, any_hit=True)["obj"] >= 0)


def test_csg_render_matches_oracle(gpu_ctx):
    desc = csg_scene()
    w = h = 64
    eye, look = (0.0, 6.0, 14.0), (0.0, -0.5, 0.0)
    scene = pbrt_amd.Scene(gpu_ctx, desc)
    film, _ = scene.render(w, h, 8, 6, eye, look, fov=0.9, seed=4)
    f32, _ = po.Scene(desc, precision=32, **pbrt_amd.DEVICE_BVH).render(w, h, 0.9, eye, look, 8, 6, 4)
    scale = float(f32.mean())
    assert scale > 0
    assert rmse(film, f32) / scale < 0.02, rmse(film, f32) / scale


@pytest.mark.parametrize("use_graphs", [False, True])
def test_job_batch_equals_single_renders(gpu_ctx, use_graphs):
    """Jobs of different sizes, sample counts, depths, seeds and cameras rendered as one batch give,
    bit for bit, the films of rendering each alone (with per-bounce graphs replayed for every
    bounce of every pass, or launched eagerly)."""
    desc, cam = scenes.small()
    scene = pbrt_amd.Scene(gpu_ctx, desc)
    jobs = [dict(width=64, height=48, spp=5, max_depth=8, eye=cam["eye"], look=cam["look"], fov=cam["fov"], seed=3),
            dict(width=32, height=32, spp=3, max_depth=4, eye=(5.0, 4.0, 20.0), look=(0.0, 0.0, 0.0), fov=0.7,
                 seed=9),
            dict(width=80, height=20, spp=1, max_depth=6, eye=(-8.0, 2.0, 15.0), look=(0.0, -1.0, 0.0), fov=0.9,
                 seed=1)]
    films, stats = scene.render_jobs(jobs, use_graphs=use_graphs)
    for j, f in zip(jobs, films):
        alone, _ = scene.render(j["width"], j["height"], j["spp"], j["max_depth"], j["eye"], j["look"], fov=j["fov"],
                                seed=j["seed"])
        assert f.tobytes() == alone.tobytes()
    assert stats["paths"] == sum(j["width"] * j["height"] * j["spp"] for j in jobs)


def test_batch_scene_random_rays_bit_exact(gpu_ctx):
    """A reduced batched-jobs scene (instanced displaced icospheres and tori under rigid transforms,
    CSG objects, an emissive panel): random rays bit-exact against the FP32 oracle."""
    desc = scenes.batch_scene(n_meshes=24, n_csg=8, seed=7, tri_range=(500, 20000))
    rays = np.ascontiguousarray(pbrt_amd.random_rays((0.0, 0.0, 0.0), 12.0, 60000, seed=4)).view(np.float32)
    rays = rays.reshape(-1, 8)
    scene = pbrt_amd.Scene(gpu_ctx, desc)
    orc = po.Scene(desc, precision=32, **pbrt_amd.DEVICE_BVH)
    compare_hits(scene.trace_host(rays), orc.trace(rays))
    np.testing.assert_array_equal(scene.trace_host(rays, any_hit=True).astype(bool), orc.trace(rays, any_hit=True)["obj"] >= 0)


@pytest.mark.parametrize("which", ["room", "glossy_floor"])
def test_mis_render_matches_oracle(gpu_ctx, which):
    """Multiple importance sampling on the device against the FP32 oracle's MIS at equal spp under
    the same sampler (the bars of test_render_matches_oracle), and the default estimator unchanged
    by the option's code path."""
    from test_oracle_mis import glossy_floor
    if which == "room":
        desc, cam = scenes.room(subdiv=4, ggx=True)
        depth = cam["max_depth"]
    else:
        desc, cam = glossy_floor()
        depth = 4
    w = h = 64
    spp = 8
    scene = pbrt_amd.Scene(gpu_ctx, desc)
    film, _ = scene.render(w, h, spp, depth, cam["eye"], cam["look"], fov=cam["fov"], seed=3, mis=True)
    orc = po.Scene(desc, precision=32, **pbrt_amd.DEVICE_BVH)
    f32, _ = orc.render(w, h, cam["fov"], cam["eye"], cam["look"], spp, depth, 3, mis=True)
    scale = float(f32.mean())
    assert scale > 0 and np.isfinite(film).all()
    assert rmse(film, f32) / scale < 0.02, rmse(film, f32) / scale
    assert abs(film.mean() - f32.mean()) / scale < 0.01
    plain, _ = scene.render(w, h, spp, depth, cam["eye"], cam["look"], fov=cam["fov"], seed=3)
    p32, _ = orc.render(w, h, cam["fov"], cam["eye"], cam["look"], spp, depth, 3, mis=False)
    assert rmse(plain, p32) / scale < 0.02
    assert not np.array_equal(plain, film)


@pytest.mark.parametrize("which", ["room", "mirror"])
def test_roulette_render_matches_oracle(gpu_ctx, which):
    """Russian roulette on the device against the FP32 oracle's roulette at equal spp under the
    same sampler (the draw that ends a path is the same counter-based sample on both sides), with
    the bars of test_render_matches_oracle; the ray counts agree within 0.5 % and roulette traces
    fewer rays than the plain render. Jobs that disagree on the option are refused."""
    from test_oracle_roulette import EYE, FOV, LOOK, mirror_scene
    if which == "room":
        desc, cam = scenes.room(subdiv=4, ggx=True)
        eye, look, fov, depth = cam["eye"], cam["look"], cam["fov"], cam["max_depth"]
    else:
        desc, eye, look, fov, depth = mirror_scene(), EYE, LOOK, FOV, 3
    w = h = 64
    spp = 8
    scene = pbrt_amd.Scene(gpu_ctx, desc)
    film, st = scene.render(w, h, spp, depth, eye, look, fov=fov, seed=3, roulette=True)
    orc = po.Scene(desc, precision=32, **pbrt_amd.DEVICE_BVH)
    f32, orays = orc.render(w, h, fov, eye, look, spp, depth, 3, roulette=True)
    scale = float(f32.mean())
    assert scale > 0 and np.isfinite(film).all()
    assert rmse(film, f32) / scale < 0.02, rmse(film, f32) / scale
    assert abs(film.mean() - f32.mean()) / scale < 0.01
    assert abs(st["rays"] - orays) < 0.005 * orays, (st["rays"], orays)
    _, st0 = scene.render(w, h, spp, depth, eye, look, fov=fov, seed=3, film=False)
    assert st["rays"] < st0["rays"]
    job = dict(width=8, height=8, spp=1, max_depth=depth, eye=eye, look=look, fov=fov)
    with pytest.raises(pbrt_amd.PbrtError, match="roulette"):
        scene.render_jobs([job, dict(job, roulette=True)])


def spheres_on_checkerboard():
    """default.stage's content built by hand: a fuzzy metal, a glass and a Lambertian sphere on a
    checkerboard floor, an emissive panel above, the loader's point light and background light."""
    d = SceneDesc()
    floor = d.add_material("lambertian", (0.65, 0.0, 0.0), checker=((0.65, 0.65, 0.65), 2.0))
    metal = d.add_material("metal", (0.0, 0.6470588235294118, 1.0), fuzz=0.5)
    glass = d.add_material("dielectric", (0.0, 0.6470588235294118, 1.0), ior=1.5)
    lamb = d.add_material("lambertian", (0.0, 0.6470588235294118, 1.0))
    light = d.add_material("diffuse_light", (5.0, 5.0, 5.0))
    d.add_object("sphere", metal, param=(0.4,), transform=Transform.translation(-1.0, 0.0, 0.0))
    d.add_object("sphere", glass, param=(0.4,), transform=Transform.translation(0.0, 0.0, 0.0))
    d.add_object("sphere", lamb, param=(0.4,), transform=Transform.translation(1.0, 0.0, 0.0))
    d.add_object("rectangle", floor, (3.0, 3.0), transform=Transform.translation(0.0, -0.6, 0.0))
    d.add_object("rectangle", light, (3.0, 3.0),
                 transform=Transform.translation(0.0, 3.0, 0.0) * Transform.rotation("y", 3.1416))
    d.add_point_light((0.0, 2.0, 1.0), (15.0, 15.0, 15.0))
    d.add_background_light((1.0, 1.0, 1.0), (0.5, 0.7, 1.0))
    return d, dict(eye=(-1.5, 0.0, 1.5), look=(0.0, 0.0, 0.0), fov=math.pi / 2)


@pytest.mark.parametrize("mis", [False, True])
def test_metal_checkerboard_background_match_oracle(gpu_ctx, mis):
    desc, cam = spheres_on_checkerboard()
    w = h = 64
    spp, depth = 8, 8
    scene = pbrt_amd.Scene(gpu_ctx, desc)
    film, _ = scene.render(w, h, spp, depth, cam["eye"], cam["look"], fov=cam["fov"], seed=3, mis=mis)
    f32, _ = po.Scene(desc, precision=32, **pbrt_amd.DEVICE_BVH).render(w, h, cam["fov"], cam["eye"], cam["look"],
                                                                      spp, depth, 3, mis=mis)
    scale = float(f32.mean())
    assert scale > 0 and np.isfinite(film).all()
    assert rmse(film, f32) / scale < 0.02, rmse(film, f32) / scale
    assert abs(film.mean() - f32.mean()) / scale < 0.01


def test_device_checkerboard_and_background_closed_forms(gpu_ctx):
    """The oracle's closed forms (tests/test_oracle_materials.py) on the device."""
    import test_oracle_materials as tm
    # checkerboard emitter: interior pixels exactly the texel colour
    c1, c2, sc = (0.9, 0.2, 0.1), (0.1, 0.3, 0.8), 0.8
    d = SceneDesc()
    d.add_object("rectangle", d.add_material("diffuse_light", c1, checker=(c2, sc)), (8.0, 8.0))
    film, _ = pbrt_amd.Scene(gpu_ctx, d).render(48, 48, 4, 2, tm.EYE, tm.LOOK, fov=tm.FOV, seed=1)
    ref, _ = po.Scene(d, precision=32).render(48, 48, tm.FOV, tm.EYE, tm.LOOK, 4, 2, 1)
    assert np.mean(np.all(film == ref, axis=2)) > 0.97  # texel edges aside, the same texel
    # a Lambertian floor under a uniform background: radiance = albedo
    albedo = np.array([0.2, 0.5, 0.8])
    d = SceneDesc()
    d.add_object("rectangle", d.add_material("lambertian", albedo), (50.0, 50.0))
    d.add_background_light((1.0, 1.0, 1.0), (1.0, 1.0, 1.0))
    scene = pbrt_amd.Scene(gpu_ctx, d)
    for mis in (False, True):
        film, _ = scene.render(16, 16, 1024, 2, tm.EYE, tm.LOOK, fov=tm.FOV, seed=3, mis=mis)
        np.testing.assert_allclose(film.reshape(-1, 3).mean(0), albedo, rtol=0.02)


@pytest.mark.parametrize("name", ["bunny_mesh", "many_meshes", "room"])
@pytest.mark.parametrize("device_build", [False, True])
def test_wide_mesh_walk_same_hits_and_films(gpu_ctx, name, device_build):
    """The scene walking 4-wide trees inside its meshes (pbrt_scene_set_mesh_width 4) gives the
    binary walk's hits, occlusion and films bit for bit."""
    if name == "room":
        desc, cam = scenes.room()
        eye, look, fov = cam["eye"], cam["look"], cam["fov"]
    else:
        sc, desc = stage(name)
        eye, look, fov = sc["camera"]["pos"], sc["camera"]["look"], math.pi / 2
    a = pbrt_amd.Scene(gpu_ctx, desc, mesh_width=2)
    b = pbrt_amd.Scene(gpu_ctx, desc, mesh_width=4, device_build=device_build)
    rays = po.pinhole_rays(160, 120, fov, eye, look).astype(np.float32)
    rng = np.random.default_rng(4)
    extra = rays.copy()
    extra[:, 4:7] = rng.normal(size=(len(rays), 3)).astype(np.float32)  # same origins, random directions
    for rr in (rays, extra):
        ha, hb = a.trace_host(rr), b.trace_host(rr)
        np.testing.assert_array_equal(ha["object"], hb["object"])
        hit = ha["object"] >= 0
        for field in ("prim", "t", "u", "v"):
            np.testing.assert_array_equal(ha[field][hit], hb[field][hit], err_msg=field)
        np.testing.assert_array_equal(a.trace_host(rr, any_hit=True), b.trace_host(rr, any_hit=True))
    assert b.status() == 0
    fa, _ = a.render(48, 36, 4, 6, eye, look, fov=fov, seed=9)
    fb, _ = b.render(48, 36, 4, 6, eye, look, fov=fov, seed=9)
    np.testing.assert_array_equal(fa, fb)


@pytest.mark.parametrize("device_build", [False, True])
def test_wide_mesh_walk_leaf_roots_and_mixed_meshes(gpu_ctx, device_build):
    """Meshes whose tree is a single leaf (one triangle, a two-triangle quad) beside deep ones
    (bunny, a displaced icosphere) under transforms, in one scene: the 4-wide trees, collapsed on
    the device level by level over all meshes, give the binary walk's hits, occlusion and films,
    and Oracle B's hits."""
    d = SceneDesc()
    white = d.add_material("lambertian", (0.7, 0.7, 0.7))
    tri = d.add_mesh(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), np.array([[0, 1, 2]], np.uint32))
    quad = d.add_mesh(np.array([[-1, 0, -1], [1, 0, -1], [1, 0, 1], [-1, 0, 1]], np.float32),
                      np.array([[0, 1, 2], [0, 2, 3]], np.uint32))
    bv, bf = load_mesh("bunny")
    bunny = d.add_mesh(bv * 10.0, bf)
    iv, iff = pbrt_amd.icosphere(4, noise="fbm", amp=0.1, octaves=3, seed=3)
    ico = d.add_mesh(iv, iff)
    d.add_object("mesh", white, mesh=quad, transform=Transform.scale(6.0, 1.0, 6.0) * Transform.translation(0, -1, 0))
    d.add_object("mesh", white, mesh=tri, transform=Transform.translation(-1.5, 0.0, 0.5))
    d.add_object("mesh", white, mesh=bunny, transform=Transform.translation(0.5, -1.3, 0.0))
    d.add_object("mesh", white, mesh=ico, transform=Transform.translation(2.0, 0.2, -0.5))
    d.add_object("mesh", white, mesh=tri, transform=Transform.rotation("y", 0.7) * Transform.translation(1.0, 1.0, 1.0))
    d.add_point_light((0.0, 4.0, 3.0), (20.0, 20.0, 20.0))
    eye, look, fov = (0.0, 1.5, 6.0), (0.0, 0.0, 0.0), math.pi / 3
    a = pbrt_amd.Scene(gpu_ctx, d, mesh_width=2)
    b = pbrt_amd.Scene(gpu_ctx, d, mesh_width=4, device_build=device_build)
    rays = _mixed_rays(eye, look, fov)
    ha, hb = a.trace_host(rays), b.trace_host(rays)
    np.testing.assert_array_equal(ha["object"], hb["object"])
    hit = ha["object"] >= 0
    assert hit.mean() > 0.2
    assert len(np.unique(ha["object"][hit])) == 5  # every object is hit somewhere
    for field in ("prim", "t", "u", "v"):
        np.testing.assert_array_equal(ha[field][hit], hb[field][hit], err_msg=field)
    np.testing.assert_array_equal(a.trace_host(rays, any_hit=True), b.trace_host(rays, any_hit=True))
    compare_hits(b.trace_host(rays), po.Scene(d, precision=32, **pbrt_amd.DEVICE_BVH).trace(rays))
    assert b.status() == 0
    fa, _ = a.render(48, 36, 4, 6, eye, look, fov=fov, seed=9)
    fb, _ = b.render(48, 36, 4, 6, eye, look, fov=fov, seed=9)
    np.testing.assert_array_equal(fa, fb)


def _mixed_rays(eye, look, fov, w=160, h=120, seed=4):
    """Camera rays and the same origins with seeded random directions."""
    rays = po.pinhole_rays(w, h, fov, eye, look).astype(np.float32)
    extra = rays.copy()
    d = np.random.default_rng(seed).normal(size=(len(rays), 3))
    extra[:, 4:7] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    return np.concatenate([rays, extra])


@pytest.mark.parametrize("launch", [dict(lds_stack=8, blocks_per_cu=2, width=2, min_waves=0),
                                    dict(lds_stack=24, blocks_per_cu=0, width=4, min_waves=5),
                                    dict(lds_stack=8, blocks_per_cu=8, width=4, min_waves=0),
                                    dict(lds_stack=16, blocks_per_cu=32, width=4, min_waves=0)],
                         ids=lambda l: f"s{l['lds_stack']}b{l['blocks_per_cu']}w{l['width']}m{l['min_waves']}")
def test_scene_launch_settings_same_hits_and_films(gpu_ctx, launch):
    """Every scene launch setting (LDS stack depth, workgroups per CU up to a grid four times what
    fits at once, mesh tree width, set after commit) gives the default's hits, occlusion and films
    bit for bit."""
    desc = scenes.batch_scene(n_meshes=24, n_csg=6)
    eye, look, fov = (0.0, 8.0, 30.0), (0.0, 0.0, 0.0), math.pi / 3
    a = pbrt_amd.Scene(gpu_ctx, desc)
    b = pbrt_amd.Scene(gpu_ctx, desc)
    b.set_launch(**launch)
    assert b.get_launch() == launch
    rays = _mixed_rays(eye, look, fov)
    ha, hb = a.trace_host(rays), b.trace_host(rays)
    np.testing.assert_array_equal(ha["object"], hb["object"])
    hit = ha["object"] >= 0
    assert hit.mean() > 0.05
    for field in ("prim", "t", "u", "v"):
        np.testing.assert_array_equal(ha[field][hit], hb[field][hit], err_msg=field)
    np.testing.assert_array_equal(a.trace_host(rays, any_hit=True), b.trace_host(rays, any_hit=True))
    fa, _ = a.render(48, 36, 4, 6, eye, look, fov=fov, seed=9)
    fb, _ = b.render(48, 36, 4, 6, eye, look, fov=fov, seed=9)
    np.testing.assert_array_equal(fa, fb)
    assert b.status() == 0


def test_scene_autotune_keeps_results(gpu_ctx):
    desc, cam = scenes.room()
    a = pbrt_amd.Scene(gpu_ctx, desc)
    b = pbrt_amd.Scene(gpu_ctx, desc)
    rays = _mixed_rays(cam["eye"], cam["look"], cam["fov"], 256, 256)
    dr = gpu_ctx.upload(rays)
    dh = gpu_ctx.alloc(32 * len(rays))
    tuned = b.autotune(dr, len(rays), dh, reps=1)
    assert tuned["lds_stack"] in (8, 16, 24) and tuned["width"] in (2, 4) and tuned["ms"] > 0
    assert b.get_launch() == {k: tuned[k] for k in ("lds_stack", "blocks_per_cu", "width", "min_waves")}
    fa, _ = a.render(64, 64, 4, 8, cam["eye"], cam["look"], fov=cam["fov"], seed=2)
    fb, _ = b.render(64, 64, 4, 8, cam["eye"], cam["look"], fov=cam["fov"], seed=2)
    np.testing.assert_array_equal(fa, fb)


def test_scene_autotune_on_renders_keeps_results(gpu_ctx):
    desc, cam = scenes.small()
    a = pbrt_amd.Scene(gpu_ctx, desc)
    b = pbrt_amd.Scene(gpu_ctx, desc)
    tuned = b.autotune_render(64, 64, 2, 8, cam["eye"], cam["look"], fov=cam["fov"], reps=1)
    assert b.get_launch() == {k: tuned[k] for k in ("lds_stack", "blocks_per_cu", "width", "min_waves")}
    fa, _ = a.render(64, 64, 4, 8, cam["eye"], cam["look"], fov=cam["fov"], seed=3)
    fb, _ = b.render(64, 64, 4, 8, cam["eye"], cam["look"], fov=cam["fov"], seed=3)
    np.testing.assert_array_equal(fa, fb)
