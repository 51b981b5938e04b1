"""The reference's own unit tests for the mesh hot path, run against the CPU oracle.

Ported from src/shapes/triangle_mesh/bvh.rs:867-995 and src/geom/aabound.rs:476-682, plus closed
forms the mathematics fixes (Möller–Trumbore against an exact f64 re-evaluation).
"""
import numpy as np
import pytest

from conftest import grid_mesh, grid_ray_set
from oracle import pyoracle as po


@pytest.mark.parametrize("bins", [8, 12])
def test_binned_cost_matches_exhaustive_scan(bins):
    """triangle_mesh/bvh.rs:867-894: the prefix/suffix scan equals a full recount, exactly."""
    v, f = grid_mesh(4)
    ranges = [(0, 64), (0, 33), (7, 40), (10, 20), (0, 9), (56, 8), (30, 2)]
    checked = 0
    for lf, cnt in ranges:
        c = po.split_costs(v, f, bins, lf, cnt)
        if c is not None:
            assert c[0] == c[1]
            checked += 1
    assert checked == len(ranges)


@pytest.mark.parametrize("precision", [64, 32])
def test_query_matches_brute_force(precision):
    """triangle_mesh/bvh.rs:904-926: the tree returns exactly what testing every triangle returns."""
    v, f = grid_mesh(3)
    m = po.Mesh(v, f, precision=precision)
    rays = grid_ray_set(v, f)
    if precision == 32:
        rays = rays.astype(np.float32)
    mode = po.MODE_REFERENCE if precision == 64 else po.MODE_CANONICAL
    tree = m.trace(rays, mode=mode)
    brute = m.trace(rays, mode=po.MODE_BRUTE)
    assert np.array_equal(tree["prim"], brute["prim"])
    hit = tree["prim"] >= 0
    assert np.array_equal(tree["t"][hit], brute["t"][hit])
    assert hit.sum() > 20


@pytest.mark.parametrize("precision", [64, 32])
def test_intersect_p_agrees_with_query(precision):
    """triangle_mesh/bvh.rs:937-975, with the far sweep across the nearest hit."""
    v, f = grid_mesh(3)
    m = po.Mesh(v, f, precision=precision)
    rays = grid_ray_set(v, f)
    dt = np.float64 if precision == 64 else np.float32
    r = rays.astype(dt)
    q = m.trace(r, mode=po.MODE_CANONICAL)
    anyh = m.trace(r, mode=po.MODE_ANY)
    assert np.array_equal(anyh["prim"] >= 0, q["prim"] >= 0)
    hit = np.nonzero(q["prim"] >= 0)[0]
    assert len(hit) > 20
    for factor in [0.5, 0.99, 0.999999, 1.0, 1.000001]:
        rr = r[hit].copy()
        rr[:, 7] = (q["t"][hit] * factor).astype(dt)
        qa = m.trace(rr, mode=po.MODE_ANY)
        qc = m.trace(rr, mode=po.MODE_CANONICAL)
        assert np.array_equal(qa["prim"] >= 0, qc["prim"] >= 0), factor


def test_separated_triangles_land_in_distinct_leaves():
    """triangle_mesh/bvh.rs:980-995."""
    v, f = grid_mesh(4)
    s = po.Mesh(v, f).build_stats()
    assert s["total_leaf_tri_count"] == 64
    assert s["max_leaf_tri_count"] <= 4


def _n(d):
    d = np.array(d, np.float64)
    return d / np.sqrt((d * d).sum())


UNIT = [-1, -1, -1, 1, 1, 1]
COLLIDE = [  # aabound.rs:476-505
    ((0, 0, 0), (0, 0, 1), True), ((0, 0, 0), (0, 1, 0), True), ((0, 0, 0), (1, 0, 0), True),
    ((0, 0, 0), (0, 0, -1), True), ((0, 0, 0), (0, -1, 0), True), ((0, 0, 0), (-1, 0, 0), True),
    ((1.01, 0, 0), (0, 0, 1), False), ((0, 0, -1.1), (0, 1, 0), False), ((0, 0, 1.01), (1, 0, 0), False),
    ((-1.01, 0, 0), (0, 0, -1), False), ((1.01, 0, -1.01), (0, -1, 0), False), ((1.10, 1.01, 0), (-1, 0, 0), False),
    ((0.5, -0.5, 0), (1, -1, 1), True), ((0.5, -0.5, 0), (-0.63, -1, 2.13), True),
    ((0.5, -0.5, 0), (12.2, 0.004, -0.0003), True), ((-2, 0, 0), (3, 1, 0), True), ((-2, 0, 0), (3, 1, -2), True),
]
FLAT = [-1, 0, -1, 1, 0, 1]
FLAT_CASES = [  # aabound.rs:516-547
    ((0, 1, 0), (0, -1, 0), True), ((0, -1, 0), (0, 1, 0), True), ((-0.5, 1, -0.5), (0.3, -1, 0.2), True),
    ((-0.5, 0, 0), (1, 0, 0), True), ((-0.5, 0.5, 0), (1, 0, 0), False), ((2, 1, 0), (0, -1, 0), False),
    ((0, 1, 0), (0, 1, 0), False),
]


@pytest.mark.parametrize("precision", [64, 32])
def test_slab_cases(precision):
    for box, cases in [(UNIT, COLLIDE), (FLAT, FLAT_CASES)]:
        for o, d, expect in cases:
            got = po.box_hit(precision, box, [*o, *_n(d)], 0.0, 1000.0)
            assert (got is not None) == expect, (box, o, d)
    # aabound.rs:551-558: grazing a face is a hit
    assert po.box_hit(precision, UNIT, [-2, 1, 0, 1, 0, 0], 0.0, 1000.0) is not None


def test_rounding_bound_magnitude():
    """aabound.rs:677-682, and the FP32 analogue: gamma(3) between 3u and 4u."""
    for precision, u in [(64, 2.0 ** -53), (32, 2.0 ** -24)]:
        g3 = po.slab_widening(precision) / 2
        assert 3 * u < g3 < 4 * u
    assert po.slab_widening(64) < 1e-15
    assert abs(po.slab_widening(32) - 3.5762793e-07) < 1e-13


def test_moller_trumbore_closed_forms():
    """A ray down the z axis onto a triangle in z = 0 hits at t = height with the barycentrics of
    the hit point; FP32 t stays within 1e-6 of the f64 value over random well-posed triangles."""
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    f = np.array([[0, 1, 2]], np.uint32)
    m64 = po.Mesh(v, f, precision=64)
    m32 = po.Mesh(v, f, precision=32)
    ray = np.array([[0.25, 0.5, 2.0, 1e-4, 0, 0, -1, 1000.0]])
    for m in (m64, m32):
        r = m.trace(ray.astype(np.float64 if m is m64 else np.float32), mode=po.MODE_BRUTE)
        assert r["prim"][0] == 0 and r["t"][0] == 2.0
        assert r["u"][0] == 0.25 and r["v"][0] == 0.5
    rng = np.random.default_rng(7)
    tri = rng.uniform(-1, 1, (200, 3, 3)).astype(np.float32)
    ok = 0
    for t in tri:
        m64 = po.Mesh(t, np.array([[0, 1, 2]]), precision=64)
        m32 = po.Mesh(t, np.array([[0, 1, 2]]), precision=32)
        c = t.astype(np.float64).mean(0)
        o = c + rng.normal(size=3) * 3
        d = _n(c - o).astype(np.float32).astype(np.float64)
        row = np.array([[*o.astype(np.float32), 1e-4, *d, 1000.0]])
        a = m64.trace(row, mode=po.MODE_BRUTE)
        b = m32.trace(row.astype(np.float32), mode=po.MODE_BRUTE)
        if a["prim"][0] == 0 and b["prim"][0] == 0:
            ok += 1
            assert abs(a["t"][0] - b["t"][0]) <= 1e-5 * a["t"][0]
    assert ok > 150


def test_empty_mesh_is_rejected():
    """triangle_mesh.rs:276-280."""
    with pytest.raises(ValueError):
        po.Mesh(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32))
