"""The headline traversal workload at full size, checked through the C ABI with the benchmark's own
launches: the 1,310,720-triangle displaced icosphere (bench.make_workload), 4,194,304 coherent
rays of a 2048^2 pinhole taken in 8 x 16 tiles, and 4,194,304 incoherent rays.

Bar, for each batch:
  * closest hit of the timed kernels (4-wide tree; coherent: block 128, a grid of 128 waves per CU
    (four times what fits at once: workgroups queue), LDS stack 12, per-lane refill, tiles;
    incoherent: block 128, a grid of 32 waves per CU, LDS stack 8 spilling to HBM, wave refill)
    bit-exact against Oracle B: triangle index, t, u, v;
  * the two batches traced at once on two streams, as the benchmark's step does, give the hits
    each gives alone;
  * the reference's instrumented walk on the binary tree gives byte-identical hits;
  * any hit of the timed launches equals Oracle B's any hit and closest-hit existence;
  * the fetch-counting walk (pbrt_mesh_trace_work) gives the same hits, and on the binary tree its
    node and triangle counts are the instrumented counters' (interior visits, triangle tests);
  * against Oracle A (f64, the reference's BVH and tie order) every differing triangle is a crack,
    an epsilon flip or a rounding-order near tie (oracle.parity.classify_vs_reference), at a rate
    below 1e-4;
  * hit t against the f64 re-evaluation of the hit triangle: within 1e-5 relative wherever the ray
    meets the triangle at |cos| >= 0.01; at more grazing incidence the FP32 statement sequence the
    reference's arithmetic fixes is ill-conditioned, and the bound there is 1e-4.
"""
import os

import numpy as np
import pytest

import bench
import pbrt_amd
from oracle import parity
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

THREADS = min(16, len(os.sched_getaffinity(0)))


def _launch(spec):
    vals = [int(x) for x in spec.split(",")]
    return dict(bench.LAUNCH_DEFAULTS, **dict(zip(bench.LAUNCH_KEYS, vals)))


@pytest.fixture(scope="module")
def c3(gpu_ctx):
    v, f, coherent, incoherent = bench.make_workload(pbrt_amd)
    mesh = pbrt_amd.TriangleMesh(gpu_ctx, v, f, **pbrt_amd.DEVICE_BVH)
    ob = parity.oracle_b(v, f, pbrt_amd.DEVICE_BVH)
    oa = parity.oracle_a(v, f)
    launches = [_launch(x) for x in bench.DEFAULT_BATCH_LAUNCHES.split(";")]
    batches = {}
    for (name, rays), launch, tiles in zip((("coherent", coherent), ("incoherent", incoherent)), launches, (2048, 0)):
        rows = bench.rows(rays)
        batches[name] = dict(rays=rays, rows=rows, launch=launch, tiles=tiles,
                             b=ob.trace(rows, mode=po.MODE_CANONICAL, threads=THREADS),
                             b_any=ob.trace(rows, mode=po.MODE_ANY, threads=THREADS),
                             a=oa.trace(rows, mode=po.MODE_REFERENCE, threads=THREADS))
    return dict(v=v, f=f, mesh=mesh, batches=batches)


def _device(ctx, c3, name):
    """Hits, occlusion, fetch counts and instrumented counters of one batch on the device."""
    mesh, b = c3["mesh"], c3["batches"][name]
    n = len(b["rays"])
    dr = ctx.upload(b["rays"])
    dh, do, dw = ctx.alloc(16 * n), ctx.alloc(n), ctx.alloc(8 * n)
    mesh.set_launch(**b["launch"])
    mesh.set_ray_tiles(b["tiles"])
    for d in (dh, do):
        d.fill(0xFF)
    mesh.query(dr, n, dh)
    hits = dh.read(pbrt_amd.HIT_DTYPE, n)
    mesh.intersect_p(dr, n, do)
    occ = do.read(np.uint8, n)
    dh.fill(0xFF)
    mesh.query_work(dr, n, dh, dw)
    work_hits = dh.read(pbrt_amd.HIT_DTYPE, n)
    work = dw.read(np.uint32, 2 * n).reshape(n, 2)
    mesh.query_work(dr, n, do, dw, any_hit=True)
    work_occ = do.read(np.uint8, n)
    mesh.set_ray_tiles(0)
    assert mesh.status() == 0
    return dict(hits=hits, occ=occ, work_hits=work_hits, work=work, work_occ=work_occ, dr=dr, n=n)


@pytest.mark.parametrize("name", ["coherent", "incoherent"])
def test_headline_batches_bit_exact(gpu_ctx, c3, name):
    b = c3["batches"][name]
    d = _device(gpu_ctx, c3, name)
    cmp = parity.compare_hits(d["hits"], b["b"])
    print(f"[c3 {name}] vs Oracle B: {cmp}", flush=True)
    assert cmp["total"] == 0, cmp
    assert cmp["hits"] > 0.2 * cmp["rays"]
    assert d["work_hits"].tobytes() == d["hits"].tobytes()
    hit_b = b["b"]["prim"] >= 0
    np.testing.assert_array_equal(d["occ"].astype(bool), b["b_any"]["prim"] >= 0)
    np.testing.assert_array_equal(d["occ"].astype(bool), hit_b)
    np.testing.assert_array_equal(d["work_occ"], d["occ"])
    # the 4-wide walk fetches records; a hit needs at least one triangle fetched
    assert (d["work"][hit_b, 1] >= 1).all()
    assert d["work"][:, 0].mean() > 1.0


@pytest.mark.parametrize("name", ["coherent", "incoherent"])
def test_headline_binary_walk_and_counters(gpu_ctx, c3, name):
    """The reference's instrumented walk on the binary tree (untiled) gives the timed kernels' hits
    byte for byte; the binary fetch counts equal its counters (interior visits = (box tests - 1) / 2,
    triangle fetches = triangle tests)."""
    mesh, b = c3["mesh"], c3["batches"][name]
    d = _device(gpu_ctx, c3, name)
    n = d["n"]
    dh, ds, dw = gpu_ctx.alloc(16 * n), gpu_ctx.alloc(12 * n), gpu_ctx.alloc(8 * n)
    mesh.set_launch(block=128, waves_per_cu=16, lds_stack=16, width=2)
    mesh.query_instrumented(d["dr"], n, dh, ds)
    st = ds.read(np.uint32, 3 * n).reshape(n, 3).astype(np.int64)
    assert dh.read(pbrt_amd.HIT_DTYPE, n).tobytes() == d["hits"].tobytes()
    mesh.query_work(d["dr"], n, dh, dw)
    w2 = dw.read(np.uint32, 2 * n).reshape(n, 2).astype(np.int64)
    np.testing.assert_array_equal(w2[:, 0], (st[:, 1] - 1) // 2)
    np.testing.assert_array_equal(w2[:, 1], st[:, 2])
    assert mesh.status() == 0


@pytest.mark.parametrize("name", ["coherent", "incoherent"])
def test_headline_against_reference_arithmetic(gpu_ctx, c3, name):
    b = c3["batches"][name]
    d = _device(gpu_ctx, c3, name)
    h = d["hits"]
    cls = parity.classify_vs_reference(b["a"], h["prim"], h["t"], h["u"], h["v"], c3["v"], c3["f"], b["rows"])
    err = parity.t_error_by_incidence(c3["v"], c3["f"], b["rows"], h["prim"].astype(np.int64), h["t"])
    print(f"[c3 {name}] vs Oracle A: {cls}; t vs f64: {err}", flush=True)
    assert cls["unexplained"] == 0, cls
    assert cls["rate"] < 1e-4, cls
    assert err["max_rel_non_grazing"] <= 1e-5, err
    assert err["max_rel"] <= 1e-4, err


def test_headline_batches_at_once_on_two_streams(gpu_ctx, c3):
    """The benchmark's step: the coherent batch on the context's stream and the incoherent one on a
    second stream, forked and joined with events, three steps back to back; every step's hits equal
    each batch's hits alone, byte for byte, and no traversal stack overflowed."""
    mesh = c3["mesh"]
    alone = {name: _device(gpu_ctx, c3, name)["hits"] for name in ("coherent", "incoherent")}
    side = gpu_ctx.new_stream()
    dev = {}
    for name in ("coherent", "incoherent"):
        b = c3["batches"][name]
        dev[name] = (gpu_ctx.upload(b["rays"]), gpu_ctx.alloc(16 * len(b["rays"])), len(b["rays"]))
    fork, joined = gpu_ctx.event(), gpu_ctx.event()
    on = {"coherent": gpu_ctx.stream, "incoherent": side.h}
    for _ in range(3):
        for _, dh, _ in dev.values():
            dh.fill(0xFF)
        fork.record(gpu_ctx.stream)
        fork.wait(gpu_ctx, side.h)
        for name, (dr, dh, n) in dev.items():
            b = c3["batches"][name]
            mesh.set_launch(**b["launch"])
            mesh.set_ray_tiles(b["tiles"])
            mesh.query(dr, n, dh, stream=on[name])
        joined.record(side.h)
        joined.wait(gpu_ctx, gpu_ctx.stream)
        gpu_ctx.synchronize()
        for name, (_, dh, n) in dev.items():
            assert dh.read(pbrt_amd.HIT_DTYPE, n).tobytes() == alone[name].tobytes(), name
    assert mesh.status() == 0
