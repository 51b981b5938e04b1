"""The mesh BVH built on the device (pbrt_mesh_create_device) against the host builder
(pbrt_mesh_create), whose trees are pinned on the oracle and on the reference's published tree
shapes (tests/test_host_product.py): the flattened node and triangle arrays, the root and the
build statistics must be equal byte for byte, for the reference's meshes, synthetic meshes that
cross the device builder's internal thresholds (256 triangles per subtree, 4096 per chunk), the
1.3M-triangle displaced icosphere, and every builder parameter set the host tests use.
Errors must agree too: a leaf above the device bound of 32 triangles, coincident centroids.
Scenes whose meshes are built in one batched device build (pbrt_scene_set_mesh_build) must trace
and render exactly as the same scenes built on the host.
"""
import math

import numpy as np
import pytest

import pbrt_amd
from conftest import grid_mesh, load_mesh
from oracle import pyoracle as po
from pbrt_amd import scenes
from pbrt_amd.scene import SceneDesc, Transform

pytestmark = pytest.mark.gpu

PARAMS = [dict(bins=12, max_leaf=4, t_trav=1.0), dict(bins=12, max_leaf=2, t_trav=3.0),
          dict(bins=16, max_leaf=8, t_trav=0.5), dict(bins=256, max_leaf=4, t_trav=1.0),
          dict(bins=2, max_leaf=32, t_trav=0.0), dict(bins=8, max_leaf=0, t_trav=0.0)]


def soup(n, seed):
    """n random small triangles in the unit cube (a mesh with no shared vertices)."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(-1.0, 1.0, (n, 1, 3))
    v = (c + rng.normal(0.0, 0.02, (n, 3, 3))).astype(np.float32).reshape(-1, 3)
    return v, np.arange(3 * n, dtype=np.uint32).reshape(-1, 3)


def assert_same_build(ctx, v, f, params):
    print(f"mesh {len(f)} triangles, {params}", flush=True)
    h = pbrt_amd.TriangleMesh(ctx, v, f, **params)
    d = pbrt_amd.TriangleMesh(ctx, v, f, device_build=True, **params)
    assert d.build_stats() == h.build_stats()
    hn, ht, hb, hr = h.layout()
    dn, dt, db, dr = d.layout()
    assert hr == dr
    np.testing.assert_array_equal(db, hb)
    assert hn.shape == dn.shape and ht.shape == dt.shape
    np.testing.assert_array_equal(dn, hn)
    np.testing.assert_array_equal(dt, ht)
    return h, d


@pytest.mark.parametrize("name", ["cube", "bun_zipper_res4", "dragon_vrip_res4", "bunny"])
@pytest.mark.parametrize("params", PARAMS, ids=lambda p: f"b{p['bins']}l{p['max_leaf']}t{p['t_trav']}")
def test_reference_meshes(gpu_ctx, name, params):
    v, f = load_mesh(name)
    assert_same_build(gpu_ctx, v, f, params)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 64, 65, 255, 256, 257, 4096, 4097, 9000, 70001])
def test_sizes_across_the_thresholds(gpu_ctx, n):
    v, f = soup(n, n)
    assert_same_build(gpu_ctx, v, f, PARAMS[0])


def test_grid_mesh_and_shared_vertices(gpu_ctx):
    v, f = grid_mesh(12)
    assert_same_build(gpu_ctx, v, f, PARAMS[1])


def test_zero_coordinates_and_axis_aligned_triangles(gpu_ctx):
    # triangles in the coordinate planes: many +-0 box bounds and degenerate extents
    rng = np.random.default_rng(3)
    v = rng.integers(-3, 4, (3000, 3)).astype(np.float32)
    v[rng.random(v.shape) < 0.3] = -0.0
    f = rng.integers(0, len(v), (5000, 3)).astype(np.uint32)
    assert_same_build(gpu_ctx, v, f, PARAMS[0])


def test_large_displaced_icosphere(gpu_ctx):
    v, f = pbrt_amd.icosphere(8, radius=1.0, noise="fbm", amp=0.10, octaves=5, freq=1.5, seed=7)
    assert len(f) == 1310720
    h, d = assert_same_build(gpu_ctx, v, f, PARAMS[0])
    rays = pbrt_amd.random_rays((0.0, 0.0, 0.0), 1.2, 20000, seed=5)
    dr = gpu_ctx.upload(rays)
    out_h, out_d = gpu_ctx.alloc(20000 * 16), gpu_ctx.alloc(20000 * 16)
    h.query(dr, 20000, out_h)
    d.query(dr, 20000, out_d)
    np.testing.assert_array_equal(out_d.read(np.uint8, 20000 * 16), out_h.read(np.uint8, 20000 * 16))


def test_coincident_centroids(gpu_ctx):
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    f = np.tile(np.array([[0, 1, 2]], np.uint32), (200, 1))  # split by index in halves (bottom levels)
    assert_same_build(gpu_ctx, v, f, PARAMS[0])
    f = np.tile(np.array([[0, 1, 2]], np.uint32), (3000, 1))  # too many for the device builder
    pbrt_amd.TriangleMesh(gpu_ctx, v, f, **PARAMS[0])
    with pytest.raises(pbrt_amd.PbrtError, match="share one centroid"):
        pbrt_amd.TriangleMesh(gpu_ctx, v, f, device_build=True, **PARAMS[0])


@pytest.mark.parametrize("n", [100, 2000])  # one leaf of n triangles: in a subtree, at the top
def test_errors_agree(gpu_ctx, n):
    v, f = soup(n, 11)
    costly = dict(bins=12, max_leaf=0, t_trav=1e9)  # no split pays for itself: a single leaf
    with pytest.raises(pbrt_amd.PbrtError) as eh:
        pbrt_amd.TriangleMesh(gpu_ctx, v, f, **costly)
    with pytest.raises(pbrt_amd.PbrtError) as ed:
        pbrt_amd.TriangleMesh(gpu_ctx, v, f, device_build=True, **costly)
    assert eh.value.code == ed.value.code == -5
    v, f = load_mesh("bun_zipper_res4")
    with pytest.raises(pbrt_amd.PbrtError):
        pbrt_amd.TriangleMesh(gpu_ctx, v, np.array([[0, 1, len(v)]], np.uint32), device_build=True)
    with pytest.raises(pbrt_amd.PbrtError) as e:
        pbrt_amd.TriangleMesh(gpu_ctx, v, np.zeros((0, 3), np.uint32), device_build=True)
    assert e.value.code == -2


def many_meshes():
    """Meshes of 1 to 1.3M triangles (the reference's, soups, an icosphere), transformed."""
    d = SceneDesc()
    mat = d.add_material("lambertian", (0.6, 0.6, 0.6))
    d.add_object("rectangle", mat, (40.0, 40.0), transform=Transform.translation(0.0, -1.5, 0.0))
    meshes = [load_mesh(n) for n in ("cube", "bun_zipper_res4", "dragon_vrip_res4", "bunny")]
    meshes += [soup(n, n) for n in (1, 2, 255, 257, 5000)]
    meshes.append(pbrt_amd.icosphere(6, radius=1.0, noise="fbm", amp=0.1, octaves=3, seed=3))
    for k, (v, f) in enumerate(meshes):
        v = np.asarray(v, np.float32)
        scale = 1.0 / float(np.abs(v).max())
        xf = (Transform.translation(-6.0 + 3.0 * (k % 5), 0.0, -3.0 + 3.0 * (k // 5))
              * Transform.rotation("y", 0.3 * k) * Transform.scale(scale, scale, scale))
        d.add_object("mesh", mat, mesh=d.add_mesh(v, f), transform=xf)
    d.add_point_light((0.0, 8.0, 4.0), (200.0, 200.0, 200.0))
    return d


@pytest.mark.parametrize("which", ["many_meshes", "batch_scene"])
def test_scene_meshes_built_on_the_device(gpu_ctx, which):
    desc = many_meshes() if which == "many_meshes" else scenes.batch_scene(n_meshes=24, n_csg=4)
    host = pbrt_amd.Scene(gpu_ctx, desc)
    dev = pbrt_amd.Scene(gpu_ctx, desc, device_build=True)
    assert host.info() == dev.info()
    eye, look = (0.0, 6.0, 14.0), (0.0, 0.0, 0.0)
    rays = po.pinhole_rays(160, 120, math.pi / 3, eye, look).astype(np.float32)
    a, b = host.trace_host(rays), dev.trace_host(rays)
    np.testing.assert_array_equal(a["object"], b["object"])
    hit = a["object"] >= 0
    assert hit.mean() > 0.2
    for field in ("prim", "t", "u", "v"):
        np.testing.assert_array_equal(a[field][hit], b[field][hit], err_msg=field)
    np.testing.assert_array_equal(host.trace_host(rays, any_hit=True), dev.trace_host(rays, any_hit=True))
    fa, _ = host.render(64, 48, 4, 5, eye, look, fov=math.pi / 3, seed=2)
    fb, _ = dev.render(64, 48, 4, 5, eye, look, fov=math.pi / 3, seed=2)
    np.testing.assert_array_equal(fa, fb)
