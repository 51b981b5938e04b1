"""Pins of the oracle's CSG (shapes/csg/*.rs) against interval arithmetic.

For a ray starting outside every element, each convex element (sphere, box) covers one interval
[entry, exit] of the ray, computed here independently in closed form. The first point of a union
is the smallest entry; of an intersection the largest entry when it precedes the smallest exit;
of a subtraction the first point of the base's interval left uncovered by the subtracted ones (a
sweep). The oracle instead filters every element's recorded intersections by point membership, as
the reference does, so the two formulations check each other; the element the hit belongs to is
pinned as well (the subtracted element's exit when the base's entry is covered).
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as po
from pbrt_amd.scene import SceneDesc, Transform


def sphere_interval(o, d, c, r):
    oc = o - c
    b = float(np.dot(oc, d))
    disc = b * b - (float(np.dot(oc, oc)) - r * r)
    if disc <= 0.0:
        return None
    s = math.sqrt(disc)
    t0, t1 = -b - s, -b + s
    return (t0, t1) if t1 > 0 else None


def box_interval(o, d, center, ext, angle):
    # box rotated by `angle` about y, then translated to `center`
    ca, sa = math.cos(-angle), math.sin(-angle)
    rot = np.array([[ca, 0.0, sa], [0.0, 1.0, 0.0], [-sa, 0.0, ca]])
    lo_, ld = rot @ (o - center), rot @ d
    t0, t1 = -math.inf, math.inf
    for a in range(3):
        h = ext[a] / 2.0
        if abs(ld[a]) < 1e-300:
            if abs(lo_[a]) > h:
                return None
            continue
        u, v = (-h - lo_[a]) / ld[a], (h - lo_[a]) / ld[a]
        t0, t1 = max(t0, min(u, v)), min(t1, max(u, v))
    return (t0, t1) if t0 < t1 and t1 > 0 else None


def element_list(rng, n, spread=0.6):
    els = []
    for _ in range(n):
        c = rng.uniform(-spread, spread, 3)
        if rng.uniform() < 0.5:
            els.append(("sphere", (float(rng.uniform(0.3, 0.7)),), c, None))
        else:
            ext = rng.uniform(0.4, 1.2, 3)
            ang = float(rng.uniform(-1.0, 1.0))
            els.append(("aabox", tuple(ext), c, ang))
    return els


def intervals(els, o, d):
    out = []
    for shape, param, c, ang in els:
        out.append(sphere_interval(o, d, c, param[0]) if shape == "sphere" else box_interval(o, d, c, param, ang))
    return out


def expected(op, iv):
    if op == "union":
        hits = [(a, k) for k, x in enumerate(iv) if x is not None for a in [x[0]]]
        return min(hits) if hits else None
    if op == "intersection":
        if any(x is None for x in iv):
            return None
        a = max(x[0] for x in iv)
        k = max(range(len(iv)), key=lambda i: iv[i][0])
        return (a, k) if a < min(x[1] for x in iv) else None
    base = iv[0]
    if base is None:
        return None
    # sweep: while the point is covered by a subtracted interval (other than the one it is the
    # exit of), move to the farthest exit among the covering ones
    t, k = base[0], 0
    while True:
        cover = [(x[1], j) for j, x in enumerate(iv[1:], start=1) if j != k and x is not None and x[0] <= t <= x[1]]
        if not cover:
            break
        t, k = max(cover)
    return (t, k) if t <= base[1] else None


def make_scene(op, els):
    desc = SceneDesc()
    m = desc.add_material("lambertian", (0.5, 0.5, 0.5))
    elements = []
    for shape, param, c, ang in els:
        xf = Transform.translation(*c)
        if ang is not None:
            xf = xf * Transform.rotation("y", ang)
        elements.append((shape, param, xf))
    desc.add_csg(op, elements, m)
    return desc


@pytest.mark.parametrize("op", ["union", "intersection", "subtraction"])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_csg_first_hit_matches_intervals(op, seed):
    rng = np.random.default_rng(seed)
    els = element_list(rng, 3, spread=0.6 if op != "intersection" else 0.15)
    scene = po.Scene(make_scene(op, els), precision=64)
    n = 3000
    # rays from a sphere of radius 6 towards random points near the origin
    u = rng.normal(size=(n, 3))
    o = 6.0 * u / np.linalg.norm(u, axis=1)[:, None]
    target = rng.uniform(-0.8, 0.8, (n, 3))
    d = target - o
    d /= np.linalg.norm(d, axis=1)[:, None]
    rays = np.concatenate([o, np.full((n, 1), 1e-4), d, np.full((n, 1), 1000.0)], 1)
    got = scene.trace(rays)
    hits = 0
    for i in range(n):
        exp = expected(op, intervals(els, o[i], d[i]))
        if exp is None:
            assert got["obj"][i] == -1, (i, got["t"][i])
            continue
        hits += 1
        assert got["obj"][i] == 0, i
        assert got["t"][i] == pytest.approx(exp[0], rel=1e-9, abs=1e-12), i
        assert got["prim"][i] == exp[1], i
    assert hits > n // 10


def test_subtraction_exit_of_subtracted_element():
    """A unit box minus a sphere around its near face: a ray along -z first meets the sphere's exit,
    which belongs to the subtracted element (element 1)."""
    els = [("aabox", (2.0, 2.0, 2.0), np.zeros(3), None), ("sphere", (0.5,), np.array([0.0, 0.0, 1.0]), None)]
    scene = po.Scene(make_scene("subtraction", els), precision=64)
    rays = np.array([[0.0, 0.0, 5.0, 1e-4, 0.0, 0.0, -1.0, 1000.0], [0.9, 0.0, 5.0, 1e-4, 0.0, 0.0, -1.0, 1000.0]])
    got = scene.trace(rays)
    assert got["prim"][0] == 1 and got["t"][0] == pytest.approx(5.0 - 0.5)
    assert got["prim"][1] == 0 and got["t"][1] == pytest.approx(4.0)
