"""The checks of oracle/parity.py, pinned on the CPU before the GPU tests and the benchmark rely on
them:

  * mt_eval in FP32 reproduces Oracle B's t, u, v bit for bit on its own hits (so the FP32 half of
    the re-evaluation is the device's arithmetic), and in f64 Oracle A's;
  * every triangle Oracle A and Oracle B disagree on, over the reference's 240,000-ray protocol on
    dragon_vrip_res4, is explained, and a planted wrong triangle on a hit ray is not;
  * compare_hits counts a planted change of each field;
  * the t error against f64 is zero for hits that f64 computes itself, and flags a planted error.
"""
import numpy as np

from conftest import load_mesh
from oracle import parity
from oracle import pyoracle as po


def _setup():
    v, f = load_mesh("dragon_vrip_res4")
    rays = po.orbit_rays(v).astype(np.float32)
    rb = po.Mesh(v, f, precision=32, bins=12, t_trav=1.0, max_leaf=4).trace(rays, mode=po.MODE_CANONICAL, threads=8)
    ra = po.Mesh(v, f, precision=64).trace(rays, mode=po.MODE_REFERENCE, threads=8)
    return v, f, rays, ra, rb


def test_mt_eval_reproduces_both_oracles():
    v, f, rays, ra, rb = _setup()
    for ref, dtype in ((rb, np.float32), (ra, np.float64)):
        sel = np.nonzero(ref["prim"] >= 0)[0]
        d, uo, vo, t, u, w = parity.mt_eval(v, f, rays[sel], ref["prim"][sel], dtype)
        assert d.all() and uo.all() and vo.all()
        for x, k in ((t, "t"), (u, "u"), (w, "v")):
            np.testing.assert_array_equal(np.asarray(x, np.float64), ref[k][sel])


def test_classification_explains_fp32_and_flags_planted_triangles():
    v, f, rays, ra, rb = _setup()
    cls = parity.classify_vs_reference(ra, rb["prim"], rb["t"], rb["u"], rb["v"], v, f, rays)
    assert cls["mismatches"] > 0 and cls["unexplained"] == 0, cls
    # plant: 20 hit rays name a triangle far from their own
    prim = rb["prim"].copy()
    rng = np.random.default_rng(3)
    hit = np.nonzero((ra["prim"] >= 0) & (ra["prim"] == rb["prim"]))[0]
    k = rng.choice(hit, 20, replace=False)
    prim[k] = (prim[k] + len(f) // 2) % len(f)
    cls2 = parity.classify_vs_reference(ra, prim, rb["t"], rb["u"], rb["v"], v, f, rays)
    assert cls2["unexplained"] == 20, cls2
    assert set(cls2["unexplained_rays"]) <= set(k.tolist())
    # plant: a hit ray reported as a miss (the f64 hit is not at an edge)
    prim = rb["prim"].copy()
    u, w = ra["u"], ra["v"]
    inner = hit[np.minimum(np.minimum(u[hit], w[hit]), 1 - u[hit] - w[hit])[:] > 0.1][:5]
    prim[inner] = -1
    cls3 = parity.classify_vs_reference(ra, prim, rb["t"], rb["u"], rb["v"], v, f, rays)
    assert cls3["unexplained"] == 5, cls3


def test_compare_hits_and_t_error_flag_planted_changes():
    v, f, rays, _, rb = _setup()
    import pbrt_amd
    h = np.zeros(len(rays), pbrt_amd.HIT_DTYPE)
    for k in ("prim", "t", "u", "v"):
        h[k] = rb[k]
    assert parity.compare_hits(h, rb)["total"] == 0
    i = int(np.nonzero(rb["prim"] >= 0)[0][7])
    for k in ("t", "u", "v"):
        g = h.copy()
        g[k][i] = np.nextafter(g[k][i], np.float32(np.inf))
        assert parity.compare_hits(g, rb)[k] == 1
    g = h.copy()
    g["prim"][i] = -1
    assert parity.compare_hits(g, rb)["prim"] == 1
    ra64 = po.Mesh(v, f, precision=64).trace(rays.astype(np.float64), mode=po.MODE_BRUTE, threads=8)
    sel = ra64["prim"] >= 0
    # f64 distances of the same triangles, widened from the FP32 rays: no error against themselves
    prim = np.where(sel, ra64["prim"], -1)
    _, _, _, t64, _, _ = parity.mt_eval(v, f, rays, np.where(sel, prim, 0), np.float64)
    e = parity.t_error_by_incidence(v, f, rays, prim, np.where(sel, t64, 0.0))
    assert e["max_rel"] == 0.0 and e["hits"] == int(sel.sum())
    e32 = parity.t_error_by_incidence(v, f, rays, rb["prim"], rb["t"])
    assert 0.0 < e32["max_rel"] < 1e-4
    bad = np.where(sel, t64, 0.0)
    bad[np.nonzero(sel)[0][3]] *= 1.0 + 3e-5
    assert parity.t_error_by_incidence(v, f, rays, prim, bad)["above_1e-5"] == 1
