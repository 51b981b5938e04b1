"""ORACLE — test infrastructure, not product code.

ctypes binding of liboracle.so (the CPU restatement of the reference's ray hot path). Only the
tests, the smoke check and the CPU-baseline leg of bench.py import this module, as a checker.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

P = ctypes.c_void_p
U64 = ctypes.c_uint64


def lib():
    global _LIB
    if _LIB is None:
        variant = os.environ.get("ORACLE_LIB_VARIANT", "")  # "asan": the host-sanitised build
        path = os.path.join(_HERE, f"liboracle_{variant}.so" if variant else "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: {path} (run `make -C oracle`)")
        L = ctypes.CDLL(path)
        L.orc_mesh_create.restype = P
        L.orc_mesh_create.argtypes = [P, U64, P, U64, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int]
        L.orc_mesh_destroy.argtypes = [P]
        L.orc_mesh_build_stats.argtypes = [P, P]
        L.orc_mesh_node_count.restype = U64
        L.orc_mesh_node_count.argtypes = [P]
        L.orc_mesh_export.argtypes = [P, P, P, P, P]
        for fn in (L.orc_mesh_trace_f32, L.orc_mesh_trace_f64):
            fn.restype = ctypes.c_int
            fn.argtypes = [P, P, U64, ctypes.c_int, P, P, P, P, P, P, ctypes.c_int]
        L.orc_split_costs.restype = ctypes.c_int
        L.orc_split_costs.argtypes = [P, U64, P, U64, ctypes.c_int, U64, U64, P, P]
        L.orc_orbit_rays.argtypes = [P, U64, ctypes.c_uint, ctypes.c_double, ctypes.c_int, P]
        L.orc_box_hit.restype = ctypes.c_int
        L.orc_box_hit.argtypes = [ctypes.c_int, P, P, ctypes.c_double, ctypes.c_double, P]
        L.orc_slab_widening.restype = ctypes.c_double
        L.orc_slab_widening.argtypes = [ctypes.c_int]
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(P) if a is not None else None


MODE_REFERENCE, MODE_CANONICAL, MODE_BRUTE, MODE_ANY = 0, 1, 2, 3


class Mesh:
    """A mesh BVH built by the oracle. precision=64: Oracle A; precision=32: Oracle B."""

    def __init__(self, vertices, faces, precision=64, bins=8, t_trav=0.0, max_leaf=0):
        self.v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
        self.f = np.ascontiguousarray(faces, np.uint32).reshape(-1, 3)
        self.h = lib().orc_mesh_create(_p(self.v), len(self.v), _p(self.f), len(self.f), precision, bins, t_trav,
                                       max_leaf)
        if not self.h:
            raise ValueError("a mesh needs at least one triangle")
        self.precision = precision

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_mesh_destroy(self.h)
            self.h = None

    def build_stats(self):
        out = np.zeros(5, np.uint64)
        lib().orc_mesh_build_stats(self.h, _p(out))
        keys = ("node_count", "leaf_count", "max_depth", "max_leaf_tri_count", "total_leaf_tri_count")
        return {k: int(x) for k, x in zip(keys, out)}

    def export(self):
        n = int(lib().orc_mesh_node_count(self.h))
        boxes = np.zeros((n, 6), np.float64)
        lf = np.zeros(n, np.uint64)
        tc = np.zeros(n, np.uint64)
        ti = np.zeros(len(self.f), np.uint64)
        lib().orc_mesh_export(self.h, _p(boxes), _p(lf), _p(tc), _p(ti))
        return boxes, lf, tc, ti

    def trace(self, rays, mode=MODE_CANONICAL, threads=1, per_ray_stats=False):
        """rays: (n, 8) float32 or float64 rows (ox, oy, oz, tmin, dx, dy, dz, tmax).

        Returns dict(prim (int64, -1 = miss), t, u, v (float64), stats_sum (3,), stats (n,3)?)."""
        rays = np.ascontiguousarray(rays)
        n = len(rays)
        prim = np.empty(n, np.int64)
        t = np.empty(n, np.float64)
        u = np.empty(n, np.float64)
        v = np.empty(n, np.float64)
        ssum = np.zeros(3, np.uint64)
        spr = np.zeros((n, 3), np.uint32) if per_ray_stats else None
        fn = lib().orc_mesh_trace_f32 if rays.dtype == np.float32 else lib().orc_mesh_trace_f64
        assert rays.dtype in (np.float32, np.float64) and rays.shape[1] == 8
        fn(self.h, _p(rays), n, mode, _p(prim), _p(t), _p(u), _p(v), _p(spr), _p(ssum), threads)
        out = dict(prim=prim, t=t, u=u, v=v, stats_sum=ssum)
        if per_ray_stats:
            out["stats"] = spr
        return out


def orbit_rays(vertices, side=200, fov=40.0 * np.pi / 180.0, views=6):
    """The bvh_stats protocol's rays (bvh_stats.rs:345-401) as (n, 8) float64 rows."""
    v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
    out = np.empty((views * side * side, 8), np.float64)
    lib().orc_orbit_rays(_p(v), len(v), side, fov, views, _p(out))
    return out


def split_costs(vertices, faces, bins, left_first, count):
    v = np.ascontiguousarray(vertices, np.float32)
    f = np.ascontiguousarray(faces, np.uint32)
    a = ctypes.c_double()
    b = ctypes.c_double()
    ok = lib().orc_split_costs(_p(v), len(v), _p(f), len(f), bins, left_first, count, ctypes.byref(a),
                               ctypes.byref(b))
    return (a.value, b.value) if ok else None


def box_hit(precision, box, ray, tmin, tmax):
    b = np.ascontiguousarray(box, np.float64)
    r = np.ascontiguousarray(ray, np.float64)
    e = ctypes.c_double()
    ok = lib().orc_box_hit(precision, _p(b), _p(r), tmin, tmax, ctypes.byref(e))
    return e.value if ok else None


def slab_widening(precision):
    return lib().orc_slab_widening(precision)


# ---- scene level ----------------------------------------------------------------------------------
def _scene_sigs(L):
    if getattr(L, "_scene_sigs", False):
        return
    L.orc_scene_create.restype = P
    L.orc_scene_create.argtypes = [ctypes.c_int, P, P, P, P, ctypes.c_uint32, P, ctypes.c_uint32, P, ctypes.c_uint32, P,
                                   ctypes.c_uint32, P, ctypes.c_uint32, ctypes.c_int, ctypes.c_double, ctypes.c_int]
    L.orc_desc_sizes.restype = ctypes.c_uint32
    L.orc_scene_destroy.argtypes = [P]
    for fn in (L.orc_scene_trace_f32, L.orc_scene_trace_f64):
        fn.restype = ctypes.c_int
        fn.argtypes = [P, P, U64, ctypes.c_int, P, P, P, P, P, P, P, ctypes.c_int]
    L.orc_scene_set_mis.restype = ctypes.c_int
    L.orc_scene_set_mis.argtypes = [P, ctypes.c_int]
    L.orc_scene_set_roulette.restype = ctypes.c_int
    L.orc_scene_set_roulette.argtypes = [P, ctypes.c_int]
    L.orc_scene_set_integrator.restype = ctypes.c_int
    L.orc_scene_set_integrator.argtypes = [P, ctypes.c_int]
    L.orc_scene_set_sampler.restype = ctypes.c_int
    L.orc_scene_set_sampler.argtypes = [P, ctypes.c_int]
    L.orc_scene_render.restype = ctypes.c_int
    L.orc_scene_render.argtypes = [P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_double, P, P, ctypes.c_uint32,
                                   ctypes.c_uint32, U64, ctypes.c_double, ctypes.c_double, ctypes.c_uint32,
                                   ctypes.c_uint32, P, P, ctypes.c_int]
    L.orc_pinhole_rays.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_double, P, P, P, ctypes.c_double,
                                   ctypes.c_double, P]
    L.orc_scene_bvh_stats.restype = ctypes.c_int
    L.orc_scene_bvh_stats.argtypes = [P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_double, P, P, P, ctypes.c_double,
                                      ctypes.c_double, P, ctypes.c_double, P, P]
    L.orc_desc_sizes.argtypes = [P, P, P]
    L._scene_sigs = True


def pinhole_rays(w, h, fov, eye, look, up=(0.0, 1.0, 0.0), near=1e-4, far=1000.0):
    L = lib()
    _scene_sigs(L)
    out = np.empty((w * h, 8), np.float64)
    a = [np.ascontiguousarray(x, np.float64) for x in (eye, look, up)]
    L.orc_pinhole_rays(w, h, fov, _p(a[0]), _p(a[1]), _p(a[2]), near, far, _p(out))
    return out


class Scene:
    """The oracle's scene (objects, median-split scene BVH, materials, lights) from a SceneDesc-like
    object exposing object_array / material_array / light_array / mesh_arrays.
    precision=64: the reference's arithmetic and ties; 32: FP32, order-independent ties."""

    def __init__(self, desc, precision=64, bins=8, t_trav=0.0, max_leaf=0):
        L = lib()
        _scene_sigs(L)
        so, sm, sl = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        sc = L.orc_desc_sizes(ctypes.byref(so), ctypes.byref(sm), ctypes.byref(sl))
        self.objs = desc.object_array()
        self.mats = desc.material_array()
        self.lights = desc.light_array()
        self.csg = desc.csg_array()
        assert (self.objs.dtype.itemsize, self.mats.dtype.itemsize, self.lights.dtype.itemsize,
                self.csg.dtype.itemsize) == (so.value, sm.value, sl.value, sc), "descriptor layout mismatch"
        self.v, self.voff, self.f, self.foff = desc.mesh_arrays()
        self.h = L.orc_scene_create(precision, _p(self.v), _p(self.voff), _p(self.f), _p(self.foff), len(self.voff) - 1,
                                    _p(self.objs), len(self.objs), _p(self.mats), len(self.mats), _p(self.lights),
                                    len(self.lights), _p(self.csg), len(desc.csg), bins, t_trav, max_leaf)
        self.precision = precision

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_scene_destroy(self.h)
            self.h = None

    def trace(self, rays, any_hit=False, threads=8):
        rays = np.ascontiguousarray(rays)
        n = len(rays)
        obj = np.empty(n, np.int64)
        prim = np.empty(n, np.int64)
        t, u, v = (np.empty(n, np.float64) for _ in range(3))
        p = np.empty((n, 3), np.float64)
        st = np.zeros(3, np.uint64)
        fn = lib().orc_scene_trace_f32 if rays.dtype == np.float32 else lib().orc_scene_trace_f64
        fn(self.h, _p(rays), n, int(any_hit), _p(obj), _p(prim), _p(t), _p(u), _p(v), _p(p), _p(st), threads)
        return dict(obj=obj, prim=prim, t=t, u=u, v=v, p=p, stats_sum=st)

    def bvh_stats(self, w, h, fov, eye, look, up, near=1e-4, far=1000.0, target=(0.0, 2.0, 1.0), shadow_near=1e-5):
        a = [np.ascontiguousarray(x, np.float64) for x in (eye, look, up, target)]
        out = np.zeros(3, np.uint64)
        st = np.zeros(6, np.uint64)
        lib().orc_scene_bvh_stats(self.h, w, h, fov, _p(a[0]), _p(a[1]), _p(a[2]), near, far, _p(a[3]), shadow_near,
                                  _p(out), _p(st))
        return dict(rays=int(out[0]), primary_hits=int(out[1]), shadow_hits=int(out[2]),
                    primary=st[:3] / out[0], shadow=st[3:] / max(int(out[1]), 1))

    INTEGRATORS = {"path": 0, "naive": 1, "normal": 2}

    SAMPLERS = {"independent": 0, "sobol": 1}

    def render(self, w, h, fov, eye, look, spp, max_depth, seed, near=1e-4, far=1000.0, rows=None, threads=8,
               mis=False, roulette=False, integrator="path", sampler="independent"):
        lib().orc_scene_set_mis(self.h, int(bool(mis)))
        lib().orc_scene_set_roulette(self.h, int(bool(roulette)))
        lib().orc_scene_set_integrator(self.h, self.INTEGRATORS[integrator])
        lib().orc_scene_set_sampler(self.h, self.SAMPLERS[sampler])
        y0, y1 = rows if rows is not None else (0, h)
        film = np.zeros(((y1 - y0) * w, 3), np.float64)
        rays = ctypes.c_uint64()
        e = np.ascontiguousarray(eye, np.float64)
        l = np.ascontiguousarray(look, np.float64)
        lib().orc_scene_render(self.h, w, h, fov, _p(e), _p(l), spp, max_depth, seed, near, far, y0, y1, _p(film),
                               ctypes.byref(rays), threads)
        return film.reshape(y1 - y0, w, 3), rays.value


# ---- neural radiance cache (CPU restatement of the network, FP64 accumulation) ------------------
def _nrc_sigs(L):
    if getattr(L, "_nrc_sigs", False):
        return
    L.orc_nrc_nparams.restype = ctypes.c_int
    L.orc_nrc_encode.argtypes = [P, U64, P, P, P]
    L.orc_nrc_forward.argtypes = [P, P, U64, ctypes.c_int, P, ctypes.c_int]
    L.orc_nrc_loss_grad.restype = ctypes.c_double
    L.orc_nrc_loss_grad.argtypes = [P, P, P, U64, ctypes.c_int, P]
    L.orc_nrc_adam.argtypes = [P, P, P, P, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                               ctypes.c_float, ctypes.c_float]
    L._nrc_sigs = True


NRC_IN = 81


def nrc_nparams():
    L = lib()
    _nrc_sigs(L)
    return L.orc_nrc_nparams()


def nrc_encode(raw12, lo, hi):
    """n x 12 floats (p, wo, n, albedo) -> n x 81 FP32 features (before BF16 rounding)."""
    L = lib()
    _nrc_sigs(L)
    raw = np.ascontiguousarray(raw12, np.float32).reshape(-1, 12)
    lo = np.ascontiguousarray(lo, np.float32)
    hi = np.ascontiguousarray(hi, np.float32)
    out = np.empty((len(raw), NRC_IN), np.float32)
    L.orc_nrc_encode(_p(raw), len(raw), _p(lo), _p(hi), _p(out))
    return out


def nrc_forward(params, x81, bf16=True, threads=8):
    L = lib()
    _nrc_sigs(L)
    params = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(x81, np.float32).reshape(-1, NRC_IN)
    out = np.empty((len(x), 3), np.float64)
    L.orc_nrc_forward(_p(params), _p(x), len(x), int(bf16), _p(out), threads)
    return out


def nrc_loss_grad(params, x81, targets, bf16=True):
    L = lib()
    _nrc_sigs(L)
    params = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(x81, np.float32).reshape(-1, NRC_IN)
    t = np.ascontiguousarray(targets, np.float32).reshape(-1, 3)
    g = np.empty(len(params), np.float64)
    loss = L.orc_nrc_loss_grad(_p(params), _p(x), _p(t), len(x), int(bf16), _p(g))
    return loss, g


def nrc_adam(params, m, v, grad, step, lr=1e-3, b1=0.9, b2=0.99, eps=1e-8):
    """One Adam step in place on FP32 arrays."""
    L = lib()
    _nrc_sigs(L)
    grad = np.ascontiguousarray(grad, np.float32)
    for a in (params, m, v):
        assert a.dtype == np.float32 and a.flags.c_contiguous
    L.orc_nrc_adam(_p(params), _p(m), _p(v), _p(grad), len(params), step, lr, b1, b2, eps)


# ---- shading primitives (closed-form checks) -----------------------------------------------------
def _shade_sigs(L):
    if getattr(L, "_shade_sigs", False):
        return
    D = ctypes.c_double
    L.orc_fresnel.restype = D
    L.orc_fresnel.argtypes = [ctypes.c_int, P, P, D, D]
    L.orc_refract.restype = ctypes.c_int
    L.orc_refract.argtypes = [ctypes.c_int, P, P, D, P]
    L.orc_ggx_d.restype = D
    L.orc_ggx_d.argtypes = [ctypes.c_int, D, D]
    L.orc_ggx_g1.restype = D
    L.orc_ggx_g1.argtypes = [ctypes.c_int, D, D]
    L.orc_material_eval.restype = None
    L.orc_material_eval.argtypes = [ctypes.c_int, P, P, P, P, P]
    L.orc_material_scatter.restype = ctypes.c_int
    L.orc_material_scatter.argtypes = [ctypes.c_int, P, P, U64, U64, ctypes.c_uint32, P, P, P, P]
    L.orc_scene_sample_light.restype = ctypes.c_int
    L.orc_scene_sample_light.argtypes = [P, ctypes.c_uint32, P, P, U64, U64, ctypes.c_uint32, P, P, P, P]
    L._shade_sigs = True


def _v3(x):
    return np.ascontiguousarray(x, np.float64).reshape(3)


def fresnel(wo, n, n1, n2, precision=64):
    """dielectric.rs:168-182 in the oracle (local frame: only the z components enter)."""
    L = lib()
    _shade_sigs(L)
    return L.orc_fresnel(precision, _p(_v3(wo)), _p(_v3(n)), n1, n2)


def refract(wo, n, eta, precision=64):
    """dielectric.rs:129-157: the refracted direction, or None (total internal reflection)."""
    L = lib()
    _shade_sigs(L)
    t = np.zeros(3)
    ok = L.orc_refract(precision, _p(_v3(wo)), _p(_v3(n)), eta, _p(t))
    return t if ok else None


def ggx_d(cos_h, a, precision=64):
    L = lib()
    _shade_sigs(L)
    return L.orc_ggx_d(precision, cos_h, a)


def ggx_g1(cos_v, a, precision=64):
    L = lib()
    _shade_sigs(L)
    return L.orc_ggx_g1(precision, cos_v, a)


def material(kind, rgb, ior=1.5, roughness=0.0):
    """A one-record material descriptor (the layout of pbrt_amd.scene.MATERIAL_DTYPE)."""
    from pbrt_amd.scene import MATERIAL, MATERIAL_DTYPE
    m = np.zeros(1, MATERIAL_DTYPE)
    m["kind"], m["rgb"], m["ior"], m["roughness"] = MATERIAL[kind], rgb, ior, roughness
    return m


def material_eval(m, wo, wi, precision=64):
    """Material::f(wo, wi) and the density of the material's own sampling of wi, in the canonical
    frame (normal +z, tangent +x)."""
    L = lib()
    _shade_sigs(L)
    f = np.zeros(3)
    pdf = np.zeros(1)
    L.orc_material_eval(precision, _p(m), _p(_v3(wo)), _p(_v3(wi)), _p(f), _p(pdf))
    return f, float(pdf[0])


def material_scatter(m, wo, seed, path, bounce=0, precision=64):
    """Material::scatter in the canonical frame under the counter sampler's draws: dict(att, dir,
    pdf, specular), or None when the material does not scatter."""
    L = lib()
    _shade_sigs(L)
    att, d, pdf, spec = np.zeros(3), np.zeros(3), np.zeros(1), ctypes.c_int()
    ok = L.orc_material_scatter(precision, _p(m), _p(_v3(wo)), seed, path, bounce, _p(att), _p(d), _p(pdf),
                                ctypes.byref(spec))
    return dict(att=att, dir=d, pdf=float(pdf[0]), specular=bool(spec.value)) if ok else None


def scene_sample_light(scene, light, p, n, seed, path, bounce=0):
    """Light::sample_li of the oracle scene's light `light` from p (normal n): dict(li, wi, pdf,
    dist), or None when the sample is unusable."""
    L = lib()
    _shade_sigs(L)
    li, wi, pdf, dist = np.zeros(3), np.zeros(3), np.zeros(1), np.zeros(1)
    ok = L.orc_scene_sample_light(scene.h, light, _p(_v3(p)), _p(_v3(n)), seed, path, bounce, _p(li), _p(wi), _p(pdf),
                                  _p(dist))
    if ok < 0:
        raise IndexError("no such light")
    return dict(li=li, wi=wi, pdf=float(pdf[0]), dist=float(dist[0])) if ok else None


def sobol_bits(seed, pixel, pair, axis, n, first=0):
    """The Sobol sampler's 32-bit coordinates of samples first .. first + n - 1 of a pixel."""
    L = lib()
    if not getattr(L, "_sobol_sigs", False):
        L.orc_sobol_bits.restype = None
        L.orc_sobol_bits.argtypes = [U64, U64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, P]
        L._sobol_sigs = True
    out = np.zeros(n, np.uint32)
    L.orc_sobol_bits(seed, pixel, first, n, pair, axis, _p(out))
    return out
