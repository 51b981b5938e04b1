"""ORACLE — test infrastructure, not product code.

Checks of device closest hits against the oracle, shared by the tests and by the CPU-baseline leg
of bench.py (the only other place allowed to use the oracle, and there only as the checker):

  * `compare_hits`: device hits against Oracle B (FP32, canonical ties) field by field;
  * `classify_vs_reference`: the rays on which Oracle A (f64, the reference's arithmetic and tie
    order) names another triangle than the FP32 path, each explained by re-running both
    candidate triangles through Möller–Trumbore in FP32 and in f64: a near tie of t, a crack (the
    barycentric bounds, which are not watertight, flip between precisions at an edge) or an
    epsilon flip (|det| at the reference's 1e-8 on a grazing ray);
  * `fp64_t_error`, `t_error_by_incidence`: the relative error of the FP32 hit distance against
    the same triangle re-intersected in f64 (triangle_mesh.rs:183-217 evaluated in numpy f64).
"""
import numpy as np

from oracle import pyoracle as po


def compare_hits(hits, ref):
    """hits: device pbrt_hit records; ref: Oracle B's trace dict. Returns mismatch counts."""
    prim = hits["prim"].astype(np.int64)
    hit = ref["prim"] >= 0
    out = {"rays": int(len(prim)), "hits": int(hit.sum()),
           "prim": int((prim != ref["prim"]).sum())}
    same = hit & (prim == ref["prim"])
    for k in ("t", "u", "v"):
        out[k] = int((hits[k][same].astype(np.float64) != ref[k][same]).sum())
    out["total"] = out["prim"] + out["t"] + out["u"] + out["v"]
    return out


def fp64_t_error(vertices, faces, rays_rows, prim, t):
    """Max over hit rays of |t_fp32 - t_f64| / t_f64, with t_f64 from the f64 Möller–Trumbore of the
    hit triangle (operands widened from FP32) on the same ray. Returns (max, rays checked)."""
    e = t_error_by_incidence(vertices, faces, rays_rows, prim, t)
    return e["max_rel"], e["hits"]


def mt_eval(vertices, faces, rays_rows, prim, dtype):
    """Möller–Trumbore (triangle_mesh.rs:183-217) of triangle prim[i] on ray i in `dtype`
    (np.float32: the FP32 statement sequence of the device and Oracle B, one rounding per
    operation; np.float64: the reference's arithmetic on the FP32 inputs widened). Returns
    (det_ok, u_ok, v_ok, t, u, v); t, u, v are meaningful where the earlier tests pass."""
    R = dtype
    v = np.asarray(vertices, np.float32).astype(R)
    tri = np.asarray(faces, np.int64)[np.asarray(prim, np.int64)]
    r = np.asarray(rays_rows, np.float32).astype(R)
    p0, p1, p2 = v[tri[:, 0]], v[tri[:, 1]], v[tri[:, 2]]
    e1, e2 = p1 - p0, p2 - p0
    ox, oy, oz, dx, dy, dz = r[:, 0], r[:, 1], r[:, 2], r[:, 4], r[:, 5], r[:, 6]
    pvx = dy * e2[:, 2] - dz * e2[:, 1]
    pvy = dz * e2[:, 0] - dx * e2[:, 2]
    pvz = dx * e2[:, 1] - dy * e2[:, 0]
    det = e1[:, 0] * pvx + e1[:, 1] * pvy + e1[:, 2] * pvz
    det_ok = np.abs(det) >= R(1e-8)
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        inv = R(1.0) / det
        sx, sy, sz = ox - p0[:, 0], oy - p0[:, 1], oz - p0[:, 2]
        u = inv * (sx * pvx + sy * pvy + sz * pvz)
        u_ok = ~((u < 0) | (u > 1))
        qx = sy * e1[:, 2] - sz * e1[:, 1]
        qy = sz * e1[:, 0] - sx * e1[:, 2]
        qz = sx * e1[:, 1] - sy * e1[:, 0]
        vv = inv * (dx * qx + dy * qy + dz * qz)
        v_ok = ~((vv < 0) | (u + vv > 1))
        t = inv * (e2[:, 0] * qx + e2[:, 1] * qy + e2[:, 2] * qz)
    return det_ok, u_ok, v_ok, t, u, vv


def classify_vs_reference(ref_a, prim, t, u, v, vertices, faces, rays_rows, tie_rel=1e-5):
    """ref_a: Oracle A's trace dict; prim/t/u/v: the FP32 path's results on the same (FP32) rays.
    Every ray whose triangle differs is re-examined: both candidate triangles (Oracle A's and the
    FP32 path's) go through Möller–Trumbore in FP32 and in f64. A mismatch is
      * a near tie when both precisions see both hits and their f64 distances are closer than the
        FP32 rounding error of either distance (or within tie_rel), so the order flips;
      * a crack when one precision's barycentric bounds reject a triangle the other accepts (the
        hit lies at an edge: the test is not watertight);
      * an epsilon flip when one precision's |det| < 1e-8 rejects a triangle the other accepts (a
        ray grazing the triangle at the reference's determinant epsilon);
    and unexplained otherwise. Returns the counts and the first unexplained rays."""
    pa = np.asarray(ref_a["prim"], np.int64)
    pb = np.asarray(prim, np.int64)
    bad = np.nonzero(pa != pb)[0]
    out = {"rays": int(len(pa)), "mismatches": int(len(bad)), "rate": float(len(bad)) / max(1, len(pa)),
           "near_tie": 0, "crack": 0, "epsilon": 0, "unexplained": 0,
           "hit_only_reference": int(((pa[bad] >= 0) & (pb[bad] < 0)).sum()),
           "hit_only_fp32": int(((pa[bad] < 0) & (pb[bad] >= 0)).sum()), "unexplained_rays": []}
    if len(bad) == 0:
        return out
    rr = np.asarray(rays_rows)[bad]
    kind = np.full(len(bad), "", dtype=object)
    flips_det = np.zeros(len(bad), bool)   # a full hit in one precision, |det| < 1e-8 in the other
    flips_bary = np.zeros(len(bad), bool)  # a full hit in one precision, outside the bounds in the other
    seen_all = np.ones(len(bad), bool)     # both candidates hit in both precisions
    t32s, t64s = [], []
    for cand in (pa[bad], pb[bad]):
        has = cand >= 0
        c = np.where(has, cand, 0)
        d32, u32, v32, t32, _, _ = mt_eval(vertices, faces, rr, c, np.float32)
        d64, u64, v64, t64, _, _ = mt_eval(vertices, faces, rr, c, np.float64)
        hit32, hit64 = d32 & u32 & v32, d64 & u64 & v64
        flips_det |= has & ((hit32 & ~d64) | (hit64 & ~d32))
        flips_bary |= has & ((hit32 & d64 & ~hit64) | (hit64 & d32 & ~hit32))
        seen_all &= has & hit32 & hit64
        t32s.append(np.asarray(t32, np.float64))
        t64s.append(t64)
    # both triangles seen in both precisions: the order of their distances flips only through the
    # FP32 rounding of t when their f64 distances are closer than that rounding (or than tie_rel)
    with np.errstate(invalid="ignore"):
        gap = np.abs(t64s[0] - t64s[1])
        err = np.maximum(np.abs(t32s[0] - t64s[0]), np.abs(t32s[1] - t64s[1]))
        tie = seen_all & ((gap <= err) | (gap < tie_rel * np.abs(t64s[0])))
    kind[tie] = "near_tie"
    kind[(kind == "") & flips_bary] = "crack"
    kind[(kind == "") & flips_det] = "epsilon"
    kind[kind == ""] = "unexplained"
    for k in ("near_tie", "crack", "epsilon", "unexplained"):
        out[k] = int((kind == k).sum())
    out["unexplained_rays"] = bad[kind == "unexplained"][:16].tolist()
    return out


def t_error_by_incidence(vertices, faces, rays_rows, prim, t, grazing=0.01):
    """fp64_t_error split by incidence: the largest relative t error over hits whose |cos| between
    the ray and the triangle's normal is at least `grazing`, and over all hits, with the count of
    hits above 1e-5 relative."""
    sel = np.nonzero(np.asarray(prim) >= 0)[0]
    out = {"hits": int(len(sel)), "max_rel": 0.0, "max_rel_non_grazing": 0.0, "above_1e-5": 0, "grazing_cos": grazing}
    v64 = np.asarray(vertices, np.float64)
    f = np.asarray(faces, np.int64)
    for c0 in range(0, len(sel), 1 << 20):
        s = sel[c0:c0 + (1 << 20)]
        _, _, _, t64, _, _ = mt_eval(vertices, faces, np.asarray(rays_rows)[s], np.asarray(prim)[s], np.float64)
        rel = np.abs(np.asarray(t, np.float64)[s] - t64) / np.abs(t64)
        tri = f[np.asarray(prim)[s]]
        nrm = np.cross(v64[tri[:, 1]] - v64[tri[:, 0]], v64[tri[:, 2]] - v64[tri[:, 0]])
        d = np.asarray(rays_rows, np.float64)[s][:, 4:7]
        cos = np.abs((nrm * d).sum(1)) / (np.linalg.norm(nrm, axis=1) * np.linalg.norm(d, axis=1))
        out["max_rel"] = max(out["max_rel"], float(rel.max()))
        ng = cos >= grazing
        if ng.any():
            out["max_rel_non_grazing"] = max(out["max_rel_non_grazing"], float(rel[ng].max()))
        out["above_1e-5"] += int((rel > 1e-5).sum())
    return out


def oracle_b(vertices, faces, params):
    return po.Mesh(vertices, faces, precision=32, **params)


def oracle_a(vertices, faces):
    return po.Mesh(vertices, faces, precision=64, bins=8, t_trav=0.0)
