"""Host-side checks of the product library that need no GPU: the C ABI loads and exports every
entry point the header declares, the host BVH builder reproduces the oracle's tree, the flattened
layout respects its bounds, and the synthetic generators have the shapes they promise."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from conftest import ROOT, grid_mesh, load_mesh
from oracle import pyoracle as po

import pbrt_amd
from pbrt_amd import _native


def header_symbols():
    text = open(os.path.join(ROOT, "include", "pbrt_gpu.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(pbrt_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", _native.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (pbrt_\w+)", out))
    declared = header_symbols()
    assert len(declared) >= 30
    missing = [s for s in declared if s not in exported]
    assert not missing, missing


def test_python_binding_covers_header():
    assert set(header_symbols()) <= set(_native.SIGNATURES)


def _canon_oracle(m):
    boxes, lf, tc, ti = m.export()

    def rec(i):
        if tc[i] > 0:
            return ("L", tuple(boxes[i]), tuple(sorted(ti[lf[i]:lf[i] + tc[i]].tolist())))
        return ("N", tuple(boxes[i]), rec(lf[i]), rec(lf[i] + 1))

    return rec(0)


def _canon_product(t):
    def rec(i):
        if t["leaf"][i]:
            s = t["first"][i]
            return ("L", tuple(t["boxes"][i]), tuple(sorted(t["tris"][s:s + t["count"][i]].tolist())))
        return ("N", tuple(t["boxes"][i]), rec(t["left"][i]), rec(t["right"][i]))

    return rec(0)


PARAMS = [dict(bins=8, max_leaf=0, t_trav=0.0), dict(bins=12, max_leaf=4, t_trav=1.0),
          dict(bins=12, max_leaf=2, t_trav=3.0), dict(bins=16, max_leaf=8, t_trav=0.5)]


@pytest.mark.parametrize("name", ["cube", "bun_zipper_res4", "dragon_vrip_res4"])
@pytest.mark.parametrize("params", PARAMS, ids=lambda p: f"b{p['bins']}l{p['max_leaf']}t{p['t_trav']}")
def test_builder_reproduces_oracle_tree(name, params):
    """Same nodes, same boxes (exact), same triangle set per leaf, same build statistics."""
    sys.setrecursionlimit(100000)
    v, f = load_mesh(name)
    o = po.Mesh(v, f, precision=32, **params)
    p = pbrt_amd.TriangleMesh(None, v, f, **params)
    assert o.build_stats() == p.build_stats()
    assert _canon_oracle(o) == _canon_product(p.export_tree())


def test_builder_reference_mode_matches_published_tree_shape():
    v, f = load_mesh("bunny")
    s = pbrt_amd.TriangleMesh(None, v, f, **pbrt_amd.REFERENCE_BVH).build_stats()
    assert (s["node_count"], s["leaf_count"]) == (138881, 69441)


def test_coincident_centroids_are_split_by_index_when_over_leaf_bound():
    # 40 copies of one triangle: no plane separates them; max_leaf forces halving.
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    f = np.tile(np.array([[0, 1, 2]], np.uint32), (40, 1))
    o = po.Mesh(v, f, precision=32, bins=12, max_leaf=4, t_trav=1.0)
    p = pbrt_amd.TriangleMesh(None, v, f, bins=12, max_leaf=4, t_trav=1.0)
    assert p.build_stats()["max_leaf_tri_count"] <= 4
    assert _canon_oracle(o) == _canon_product(p.export_tree())


def test_empty_mesh_and_bad_params_are_errors():
    with pytest.raises(pbrt_amd.PbrtError) as e:
        pbrt_amd.TriangleMesh(None, np.zeros((0, 3)), np.zeros((0, 3)))
    assert e.value.code == -2
    v, f = grid_mesh(2)
    with pytest.raises(pbrt_amd.PbrtError):
        pbrt_amd.TriangleMesh(None, v, f, bins=1)
    with pytest.raises(pbrt_amd.PbrtError):
        pbrt_amd.TriangleMesh(None, v, np.array([[0, 1, 99]], np.uint32))
    with pytest.raises(pbrt_amd.PbrtError):  # leaves above the device bound of 32
        pbrt_amd.TriangleMesh(None, v, f, max_leaf=33)


def test_reference_mode_leaf_above_device_bound_is_a_layout_error():
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    f = np.tile(np.array([[0, 1, 2]], np.uint32), (40, 1))
    with pytest.raises(pbrt_amd.PbrtError) as e:
        pbrt_amd.TriangleMesh(None, v, f, **pbrt_amd.REFERENCE_BVH)
    assert e.value.code == -5


@pytest.mark.parametrize("subdiv", [0, 1, 4])
def test_icosphere_shape_winding_and_displacement(subdiv):
    v, f = pbrt_amd.icosphere(subdiv, radius=2.0, noise="uniform", amp=0.05, seed=11)
    assert len(v) == 10 * 4 ** subdiv + 2 and len(f) == 20 * 4 ** subdiv
    r = np.sqrt((v.astype(np.float64) ** 2).sum(1))
    assert r.min() >= 2.0 * 0.95 - 1e-6 and r.max() <= 2.0 * 1.05 + 1e-6
    # the reference's mesh normal -normalize(cross(e1, e2)) points outward (triangle_mesh.rs:119)
    e1 = v[f[:, 1]] - v[f[:, 0]]
    e2 = v[f[:, 2]] - v[f[:, 0]]
    n = -np.cross(e1, e2)
    c = v[f].mean(1)
    assert ((n * c).sum(1) > 0).all()
    # every edge is shared by exactly two faces: a closed surface
    edges = np.sort(np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]]), axis=1)
    _, counts = np.unique(edges, axis=0, return_counts=True)
    assert (counts == 2).all()


def test_icosphere_is_seeded():
    a = pbrt_amd.icosphere(3, noise="fbm", amp=0.1, octaves=5, seed=5)[0]
    b = pbrt_amd.icosphere(3, noise="fbm", amp=0.1, octaves=5, seed=5)[0]
    c = pbrt_amd.icosphere(3, noise="fbm", amp=0.1, octaves=5, seed=6)[0]
    assert np.array_equal(a, b) and not np.array_equal(a, c)


def test_orbit_rays_match_oracle_protocol():
    """The product's bvh_stats framing equals the oracle's (f64, then rounded to FP32)."""
    v, _ = load_mesh("bun_zipper_res4")
    ref = po.orbit_rays(v).astype(np.float32)
    for view in range(6):
        got = pbrt_amd.orbit_rays(v, view, 200).view(np.float32).reshape(-1, 8)
        assert np.array_equal(got, ref[view * 40000:(view + 1) * 40000])


def test_random_rays_are_unit_and_inside_the_ball():
    r = pbrt_amd.random_rays((1.0, 2.0, 3.0), 0.5, 10000, seed=3).view(np.float32).reshape(-1, 8)
    d = r[:, 4:7].astype(np.float64)
    assert np.abs(np.sqrt((d * d).sum(1)) - 1).max() < 1e-6
    o = r[:, :3].astype(np.float64) - [1, 2, 3]
    assert np.sqrt((o * o).sum(1)).max() <= 0.5 + 1e-6
    assert abs(d.mean(0)).max() < 0.05


def test_ray_tiles_setting_is_validated_on_the_host():
    """pbrt_mesh_set_ray_tiles accepts 0 and multiples of 16 and refuses other widths (no device
    call: the setting is checked when it is made)."""
    v, f = grid_mesh(2)
    m = pbrt_amd.TriangleMesh(None, v, f)
    m.set_ray_tiles(0)
    m.set_ray_tiles(2048)
    for bad in (8, 20, 1000):
        with pytest.raises(pbrt_amd.PbrtError, match="multiple of 16"):
            m.set_ray_tiles(bad)


def test_ray_tile_width_bounds():
    """Sixteen image rows must fit a 32-bit ray count: widths from 2^28 up are refused when set
    (16 x 2^28 wraps to 0 in 32 bits), as are widths not a multiple of 16."""
    v, f = pbrt_amd.icosphere(2)
    mesh = pbrt_amd.TriangleMesh(None, v, f)
    mesh.set_ray_tiles((1 << 28) - 16)
    for w in (1 << 28, (1 << 32) - 16):
        with pytest.raises(pbrt_amd.PbrtError, match="too large"):
            mesh.set_ray_tiles(w)
    with pytest.raises(pbrt_amd.PbrtError, match="multiple of 16"):
        mesh.set_ray_tiles(24)
    mesh.set_ray_tiles(0)
