"""Parity of the device traversal kernels with the FP32 oracle (Oracle B), through the C ABI.

Bar: triangle indices bit-exact, t/u/v bit-exact (same FP32 operation sequence), per-ray work
counters equal to the oracle's on the same tree, any-hit equal; hit t within 1e-5 relative of an
f64 re-evaluation of the hit triangle. Published hit counts (docs/mesures_bvh.md:706-713) are
reproduced on the device. The 4-wide tree (launch width 4) must give the same hits and occlusion
as the binary one (the counters are taken on the binary tree only).
"""
import numpy as np
import pytest

import pbrt_amd
from conftest import PUBLISHED, grid_mesh, grid_ray_set, load_mesh
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def as_rows(rays):
    return np.ascontiguousarray(rays).view(np.float32).reshape(-1, 8)


def check_closest(ctx, v, f, rays_rows, params, stats=True, launch=None):
    rays_rows = np.ascontiguousarray(rays_rows, np.float32)
    mesh = pbrt_amd.TriangleMesh(ctx, v, f, **params)
    if launch:
        mesh.set_launch(**launch)
    ref = po.Mesh(v, f, precision=32, **params).trace(rays_rows, mode=po.MODE_CANONICAL, threads=8,
                                                      per_ray_stats=stats)
    if stats:
        hits, st = mesh.trace_host(rays_rows, stats=True)
    else:
        hits = mesh.trace_host(rays_rows)
    assert mesh.status() == 0
    np.testing.assert_array_equal(hits["prim"].astype(np.int64), ref["prim"])
    hit = ref["prim"] >= 0
    np.testing.assert_array_equal(hits["t"][hit].astype(np.float64), ref["t"][hit])
    np.testing.assert_array_equal(hits["u"][hit].astype(np.float64), ref["u"][hit])
    np.testing.assert_array_equal(hits["v"][hit].astype(np.float64), ref["v"][hit])
    if stats:
        np.testing.assert_array_equal(st, ref["stats"])
    return mesh, hits, ref


@pytest.mark.parametrize("name", list(PUBLISHED))
@pytest.mark.parametrize("which", ["reference", "device"])
def test_published_meshes_bit_exact(gpu_ctx, name, which):
    v, f = load_mesh(name)
    params = pbrt_amd.REFERENCE_BVH if which == "reference" else pbrt_amd.DEVICE_BVH
    rays = po.orbit_rays(v).astype(np.float32)
    _, hits, _ = check_closest(gpu_ctx, v, f, rays, params)
    assert int((hits["prim"] >= 0).sum()) == PUBLISHED[name][0]


def test_hit_t_against_f64_reevaluation(gpu_ctx):
    """FP32 hit t versus the same triangle re-intersected in f64 by the f64 oracle's brute force on
    a one-triangle mesh, for a sample of hits on the bunny."""
    v, f = load_mesh("bunny")
    rays = po.orbit_rays(v).astype(np.float32)
    mesh = pbrt_amd.TriangleMesh(gpu_ctx, v, f)
    hits = mesh.trace_host(rays)
    idx = np.nonzero(hits["prim"] >= 0)[0][::97]
    worst = 0.0
    for i in idx:
        tri = f[hits["prim"][i]]
        m = po.Mesh(v[tri], np.array([[0, 1, 2]], np.uint32), precision=64)
        r = m.trace(rays[i:i + 1].astype(np.float64), mode=po.MODE_BRUTE)
        if r["prim"][0] == 0:
            worst = max(worst, abs(float(hits["t"][i]) - r["t"][0]) / r["t"][0])
    assert worst < 1e-5, worst


@pytest.mark.parametrize("precision_params", [pbrt_amd.REFERENCE_BVH, pbrt_amd.DEVICE_BVH])
def test_reference_grid_fixture_and_brute_force(gpu_ctx, precision_params):
    """The reference's own fixture (triangle_mesh/bvh.rs:746-854): grid mesh, centroid rays and
    axis-aligned pencils (exactly-zero direction components take the literal slab path)."""
    v, f = grid_mesh(3)
    rays = grid_ray_set(v, f).astype(np.float32)
    _, hits, _ = check_closest(gpu_ctx, v, f, rays, precision_params)
    brute = po.Mesh(v, f, precision=32).trace(rays, mode=po.MODE_BRUTE)
    np.testing.assert_array_equal(hits["prim"].astype(np.int64), brute["prim"])
    assert (hits["prim"] >= 0).sum() > 20


@pytest.mark.parametrize("width", [2, 4])
@pytest.mark.parametrize("refill", [0, 1, 64])
def test_any_hit_and_far_sweep(gpu_ctx, refill, width):
    """triangle_mesh/bvh.rs:937-975 on the device: any-hit equals closest-hit existence, also with
    `far` cut at factors of the nearest distance; and equals the oracle's any-hit. For each refill
    policy (64: lanes wait until the whole wave is done)."""
    v, f = load_mesh("dragon_vrip_res4")
    rays = po.orbit_rays(v).astype(np.float32)[::3]
    mesh = pbrt_amd.TriangleMesh(gpu_ctx, v, f)
    mesh.set_launch(refill=refill, width=width)
    orc = po.Mesh(v, f, precision=32, **pbrt_amd.DEVICE_BVH)
    hits = mesh.trace_host(rays)
    occ = mesh.trace_host(rays, mode="any")
    np.testing.assert_array_equal(occ.astype(bool), hits["prim"] >= 0)
    hit = np.nonzero(hits["prim"] >= 0)[0]
    for factor in [0.5, 0.99, 0.999999, 1.0, 1.000001]:
        rr = rays[hit].copy()
        rr[:, 7] = (hits["t"][hit].astype(np.float64) * factor).astype(np.float32)
        occ = mesh.trace_host(rr, mode="any")
        ref = orc.trace(rr, mode=po.MODE_ANY)
        np.testing.assert_array_equal(occ.astype(np.int64), (ref["prim"] >= 0).astype(np.int64))
        closest = mesh.trace_host(rr)
        np.testing.assert_array_equal(occ.astype(bool), closest["prim"] >= 0)


@pytest.mark.parametrize("block", [64, 128, 256])
@pytest.mark.parametrize("lds_stack", [8, 12, 16, 32])
@pytest.mark.parametrize("refill", [0, 4])
def test_launch_shapes_agree(gpu_ctx, block, lds_stack, refill):
    """Every launch shape and refill policy gives the oracle's answer, including an 8-entry LDS
    stack that spills to HBM on the deep reference-mode tree."""
    v, f = load_mesh("dragon_vrip_res3")
    rays = np.concatenate([as_rows(pbrt_amd.orbit_rays(v, 1, 96)),
                           as_rows(pbrt_amd.random_rays(v.mean(0), 0.1, 20000, seed=9))])
    check_closest(gpu_ctx, v, f, rays, pbrt_amd.REFERENCE_BVH, stats=True,
                  launch=dict(block=block, lds_stack=lds_stack, waves_per_cu=8, refill=refill))


@pytest.mark.parametrize("name", list(PUBLISHED))
@pytest.mark.parametrize("which", ["reference", "device"])
def test_wide_tree_published_meshes(gpu_ctx, name, which):
    v, f = load_mesh(name)
    params = pbrt_amd.REFERENCE_BVH if which == "reference" else pbrt_amd.DEVICE_BVH
    rays = po.orbit_rays(v).astype(np.float32)
    _, hits, _ = check_closest(gpu_ctx, v, f, rays, params, stats=False, launch=dict(width=4))
    assert int((hits["prim"] >= 0).sum()) == PUBLISHED[name][0]


@pytest.mark.parametrize("block", [64, 256])
@pytest.mark.parametrize("lds_stack", [8, 12, 32])
@pytest.mark.parametrize("refill", [0, 4])
def test_wide_tree_launch_shapes(gpu_ctx, block, lds_stack, refill):
    """The 4-wide tree over launch shapes, with the 8-entry LDS stack spilling on the deep
    reference-mode tree, and on incoherent rays."""
    v, f = load_mesh("dragon_vrip_res3")
    rays = np.concatenate([as_rows(pbrt_amd.orbit_rays(v, 1, 96)),
                           as_rows(pbrt_amd.random_rays(v.mean(0), 0.1, 20000, seed=9))])
    launch = dict(block=block, lds_stack=lds_stack, waves_per_cu=8, refill=refill, width=4)
    check_closest(gpu_ctx, v, f, rays, pbrt_amd.REFERENCE_BVH, stats=False, launch=launch)
    v, f = pbrt_amd.icosphere(6, noise="fbm", amp=0.08, octaves=3, seed=2)
    rays = as_rows(pbrt_amd.random_rays((0, 0, 0), 1.1, 50000, seed=4))
    check_closest(gpu_ctx, v, f, rays, pbrt_amd.DEVICE_BVH, stats=False, launch=launch)


def test_wide_tree_edge_cases_and_device_build(gpu_ctx):
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    f = np.array([[0, 1, 2]], np.uint32)
    mesh = pbrt_amd.TriangleMesh(gpu_ctx, v, f)  # the root is a leaf
    mesh.set_launch(width=4)
    rays = np.array([[0.25, 0.25, 1.0, 1e-4, 0.0, 0.0, -1.0, 1000.0],
                     [0.25, 0.25, 1.0, 1e-4, 0.0, 0.0, 1.0, 1000.0]], np.float32)
    assert mesh.trace_host(rays)["prim"].tolist() == [0, -1]
    assert mesh.trace_host(rays, mode="any").tolist() == [1, 0]
    v, f = load_mesh("bunny")  # a device-built tree collapses the same way
    rays = po.orbit_rays(v).astype(np.float32)
    h = pbrt_amd.TriangleMesh(gpu_ctx, v, f)
    d = pbrt_amd.TriangleMesh(gpu_ctx, v, f, device_build=True)
    d.set_launch(width=4)
    assert d.trace_host(rays).tobytes() == h.trace_host(rays).tobytes()


@pytest.mark.parametrize("refill", [0, 1])
def test_wide_tree_overflowing_slab_products(gpu_ctx, refill):
    """Rays that run almost parallel to a coordinate plane across a mesh wider than 3.4 units: a
    direction component of 1e-38 makes the reciprocal 1e38, so the slab distance through a plane
    more than 3.4 units away overflows to infinity. Such rays take the 4-wide step's per-plane form
    (the widen-once form is exact only without infinite plane distances), exactly-zero components
    the literal form, the rest the widen-once form; all must equal Oracle B (and so the binary
    walk) bit for bit, in a mix inside one wave."""
    v, f = pbrt_amd.icosphere(5, radius=10.0, noise="fbm", amp=0.1, octaves=3, seed=5)
    rng = np.random.default_rng(11)
    n = 40000
    o = rng.uniform(-12.0, 12.0, (n, 3)).astype(np.float32)
    phi = rng.uniform(0.0, 2.0 * np.pi, n)
    tiny = rng.choice(np.array([1e-38, -1e-38, 3e-37, -2e-36, 1e-30, -1e-33, 0.0, 5e-39, 0.3], np.float32), n)
    axis = rng.integers(0, 3, n)
    d = np.empty((n, 3), np.float32)
    for k in range(3):
        a, b = (k + 1) % 3, (k + 2) % 3
        m = axis == k
        d[m, k] = tiny[m]
        d[m, a] = np.cos(phi[m])
        d[m, b] = np.sin(phi[m])
    rays = np.concatenate([o, np.full((n, 1), 1e-4, np.float32), d, np.full((n, 1), 1000.0, np.float32)], 1)
    _, hits, ref = check_closest(gpu_ctx, v, f, rays, pbrt_amd.DEVICE_BVH, stats=False,
                                 launch=dict(width=4, refill=refill))
    # the mix really holds every class: hits among the overflowing rays, the parallel ones, the rest
    with np.errstate(divide="ignore", over="ignore"):
        inv = np.float32(1.0) / np.abs(tiny).astype(np.float32)
    overflow = (tiny != 0) & (np.abs(tiny) < 1e-37) & np.isfinite(inv)
    hit = ref["prim"] >= 0
    assert hit[overflow].sum() > 100 and hit[tiny == 0].sum() > 100 and hit[np.abs(tiny) > 1e-35].sum() > 100
    occ = pbrt_amd.TriangleMesh(gpu_ctx, v, f, **pbrt_amd.DEVICE_BVH)
    occ.set_launch(width=4, refill=refill)
    np.testing.assert_array_equal(occ.trace_host(rays, mode="any").astype(bool), hit)


@pytest.mark.parametrize("name", list(PUBLISHED))
def test_packet_traversal_published_meshes(gpu_ctx, name):
    """Wave packets (64 consecutive rays on one wave-wide stack) give the per-ray hits, on the
    orbit rays (coherent) and on random rays (incoherent), with a batch that is not a multiple of
    64 rays."""
    v, f = load_mesh(name)
    rays = po.orbit_rays(v).astype(np.float32)[:-5]
    _, hits, _ = check_closest(gpu_ctx, v, f, rays, pbrt_amd.DEVICE_BVH, stats=False,
                               launch=dict(packet=1, waves_per_cu=8))
    assert int((hits["prim"] >= 0).sum()) >= PUBLISHED[name][0] - 5
    rays = as_rows(pbrt_amd.random_rays(v.mean(0), 0.1, 30001, seed=3))
    check_closest(gpu_ctx, v, f, rays, pbrt_amd.DEVICE_BVH, stats=False, launch=dict(packet=1))


def test_packet_traversal_edge_cases(gpu_ctx):
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    f = np.array([[0, 1, 2]], np.uint32)
    mesh = pbrt_amd.TriangleMesh(gpu_ctx, v, f)  # the root is a leaf
    mesh.set_launch(packet=1)
    rays = np.array([[0.25, 0.25, 1.0, 1e-4, 0.0, 0.0, -1.0, 1000.0],
                     [0.25, 0.25, 1.0, 1e-4, 0.0, 0.0, 1.0, 1000.0],
                     [0.25, 0.25, 1.0, 2.0, 0.0, 0.0, -1.0, 1000.0]], np.float32)
    assert mesh.trace_host(rays)["prim"].tolist() == [0, -1, -1]
    with pytest.raises(pbrt_amd.PbrtError):
        mesh.set_launch(packet=1, width=4)


def test_incoherent_rays_on_displaced_icosphere(gpu_ctx):
    v, f = pbrt_amd.icosphere(6, noise="fbm", amp=0.08, octaves=3, seed=2)
    rays = as_rows(pbrt_amd.random_rays((0, 0, 0), 1.1, 100000, seed=4))
    _, hits, _ = check_closest(gpu_ctx, v, f, rays, pbrt_amd.DEVICE_BVH)
    # origins fill a ball of radius 1.1 around a closed surface of radius about 1: at least the
    # (1/1.1)^3 = 75% that start inside it must hit
    assert (hits["prim"] >= 0).mean() > 0.75


def test_edge_cases(gpu_ctx):
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    f = np.array([[0, 1, 2]], np.uint32)
    mesh = pbrt_amd.TriangleMesh(gpu_ctx, v, f)  # the root is a leaf
    nan = np.float32("nan")
    rays = np.array([
        [0.25, 0.25, 1.0, 1e-4, 0.0, 0.0, -1.0, 1000.0],   # straight down: t = 1
        [0.25, 0.25, 1.0, 1e-4, 0.0, 0.0, 1.0, 1000.0],    # away
        [0.25, 0.25, 1.0, 1e-4, 0.0, 0.0, -1.0, 0.5],      # far before the hit
        [0.25, 0.25, 1.0, 2.0, 0.0, 0.0, -1.0, 1000.0],    # near past the hit
        [0.25, 0.25, 1.0, 1e-4, nan, nan, nan, 1000.0],    # NaN direction: no hit, terminates
        [0.25, 0.25, 1.0, 1e-4, 0.0, 0.0, -1.0, np.inf],   # infinite far
    ], np.float32)
    hits = mesh.trace_host(rays)
    assert hits["prim"].tolist() == [0, -1, -1, -1, -1, 0]
    assert hits["t"][0] == 1.0
    # an empty batch is a no-op
    d = gpu_ctx.alloc(64)
    mesh.query(d, 0, d)
    assert mesh.status() == 0


def test_repeated_runs_are_bitwise_identical(gpu_ctx):
    v, f = pbrt_amd.icosphere(5, noise="fbm", amp=0.1, octaves=5, seed=1)
    rays = as_rows(pbrt_amd.random_rays((0, 0, 0), 1.0, 50000, seed=8))
    mesh = pbrt_amd.TriangleMesh(gpu_ctx, v, f)
    a = mesh.trace_host(rays)
    b = mesh.trace_host(rays)
    assert a.tobytes() == b.tobytes()


@pytest.mark.parametrize("w,h", [(64, 32), (48, 16)])
def test_tiled_ray_order_gives_batch_order_results(gpu_ctx, w, h):
    """pbrt_mesh_set_ray_tiles: an image batch taken in 8 x 16 tiles gives every ray the hit,
    counters and occlusion of batch order, for each kernel family (binary with per-lane and wave
    refill, 4-wide, instrumented, any hit); every result slot is written (outputs start filled
    with 0xff). Widths not a multiple of 16, and batches not a multiple of sixteen rows, fail."""
    v, f = pbrt_amd.icosphere(5, noise="fbm", amp=0.08, octaves=3, seed=3)
    mesh = pbrt_amd.TriangleMesh(gpu_ctx, v, f, **pbrt_amd.DEVICE_BVH)
    rays = pbrt_amd.pinhole_rays(w, h, 0.9, (0.3, 0.4, 2.6), (0.0, 0.0, 0.0))
    n = len(rays)
    dr = gpu_ctx.upload(rays)
    hits, occ, st = gpu_ctx.alloc(16 * n), gpu_ctx.alloc(n), gpu_ctx.alloc(12 * n)

    def run(tiles, launch):
        mesh.set_ray_tiles(tiles)
        mesh.set_launch(**launch)
        for d in (hits, occ, st):
            d.fill(0xFF)
        mesh.query(dr, n, hits)
        a = hits.read(pbrt_amd.HIT_DTYPE, n)
        mesh.intersect_p(dr, n, occ)
        b = occ.read(np.uint8, n)
        mesh.query_instrumented(dr, n, hits, st)
        c = st.read(np.uint32, 3 * n)
        return a, b, c

    for launch in (dict(block=256, lds_stack=8, refill=0), dict(block=128, lds_stack=16, refill=4),
                   dict(block=128, lds_stack=8, refill=1, width=4)):
        base = run(0, launch)
        tiled = run(w, launch)
        assert (base[0]["prim"] >= 0).any() and (base[0]["prim"] < 0).any()
        for x, y in zip(base, tiled):
            np.testing.assert_array_equal(x.view(np.uint8), y.view(np.uint8), err_msg=str(launch))
    mesh.set_ray_tiles(0)
    with pytest.raises(pbrt_amd.PbrtError, match="multiple of 16"):
        mesh.set_ray_tiles(20)
    mesh.set_ray_tiles(16)
    with pytest.raises(pbrt_amd.PbrtError, match="sixteen image rows"):
        mesh.query(dr, n - 16, hits)
    mesh.set_ray_tiles(0)
    assert mesh.status() == 0


@pytest.mark.parametrize("width", [2, 4])
def test_traces_on_several_streams_at_once(gpu_ctx, width):
    """Traces of one mesh on six streams, queued back to back without a host wait, at an 8-entry LDS
    stack that spills to HBM on the deep reference-mode tree: each stream traces on a spill area of
    its own, and past the mesh's four areas one changes hands (its new stream waits on the device
    for the old one's launch). Every batch gets the hits it gets alone, on every repetition."""
    v, f = load_mesh("dragon_vrip_res3")
    mesh = pbrt_amd.TriangleMesh(gpu_ctx, v, f, **pbrt_amd.REFERENCE_BVH)
    mesh.set_launch(block=128, lds_stack=8, waves_per_cu=8, width=width)
    batches = [as_rows(pbrt_amd.random_rays(v.mean(0), 0.1 + 0.01 * i, 100000, seed=21 + i)) for i in range(6)]
    refs = [mesh.trace_host(b) for b in batches]
    streams = [gpu_ctx.new_stream() for _ in batches]
    dev = [(gpu_ctx.upload(b), gpu_ctx.alloc(16 * len(b))) for b in batches]
    for _ in range(3):
        for _, h in dev:
            h.fill(0xFF)
        gpu_ctx.synchronize()  # the fills (on the context's stream) before the other streams' traces
        for s, (d, h), b in zip(streams, dev, batches):
            mesh.query(d, len(b), h, stream=s.h)
        for s in streams:
            s.synchronize()
        for (_, h), b, ref in zip(dev, batches, refs):
            assert h.read(pbrt_amd.HIT_DTYPE, len(b)).tobytes() == ref.tobytes()
    assert mesh.status() == 0


def test_ray_tile_width_bound_at_trace(gpu_ctx):
    """A batch must hold a whole number of sixteen-row bands of the declared image width."""
    v, f = pbrt_amd.icosphere(2)
    mesh = pbrt_amd.TriangleMesh(gpu_ctx, v, f)
    d = gpu_ctx.alloc(64)
    mesh.set_ray_tiles((1 << 28) - 16)
    with pytest.raises(pbrt_amd.PbrtError, match="sixteen image rows"):
        mesh.query(d, 16, d)
    mesh.set_ray_tiles(0)
