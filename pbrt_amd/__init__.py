"""MI355X-native ray hot path of nmichel/pbrt: host-side mirror of the reference's mesh interface.

Names follow the reference (src/shapes/triangle_mesh/{triangle_mesh.rs, bvh.rs}): a `TriangleMesh`
is built from vertices and indices, reports `build_stats()` and `triangle_count()`, and answers
`query` (closest hit), `query_instrumented` (closest hit with work counters) and `intersect_p`
(any hit) — here over batches of rays resident in device memory, through the C ABI of
include/pbrt_gpu.h. There is no CPU fallback: without the native library every call raises.
"""
import ctypes
import os

import numpy as np

from . import _native
from ._native import BuildStats, BvhParams, Launch, PbrtError, call, lib

__all__ = ["Context", "DeviceArray", "TriangleMesh", "PbrtError", "icosphere", "pinhole_rays", "orbit_rays",
           "random_rays", "RAY_DTYPE", "HIT_DTYPE", "REFERENCE_BVH", "DEVICE_BVH", "RadianceCache", "bf16_bits",
           "bf16_float", "mfma_check", "NRC_PARAMS", "NRC_FEATURES"]

# pbrt_ray / pbrt_hit as numpy record layouts (32 and 16 bytes)
RAY_DTYPE = np.dtype([("ox", "<f4"), ("oy", "<f4"), ("oz", "<f4"), ("tmin", "<f4"), ("dx", "<f4"), ("dy", "<f4"),
                      ("dz", "<f4"), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("prim", "<i4"), ("t", "<f4"), ("u", "<f4"), ("v", "<f4")])
# integrators::Type (src/integrators.rs:23-28) -> pbrt_render_config.integrator
INTEGRATOR = {"path": 0, "naive": 1, "normal": 2}
# pbrt_render_config.sampler
SAMPLER = {"independent": 0, "sobol": 1}

# The reference's mesh BVH parameters (triangle_mesh/bvh.rs:129, 733-735) and the device default.
REFERENCE_BVH = dict(bins=8, max_leaf=0, t_trav=0.0)
DEVICE_BVH = dict(bins=12, max_leaf=4, t_trav=1.0)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Context:
    """One device, one stream (pbrt_ctx)."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        call("pbrt_ctx_create", device, ctypes.byref(h))
        self.h = h
        s = ctypes.c_void_p()
        call("pbrt_ctx_stream", self.h, ctypes.byref(s))
        self.stream = s
        cu = ctypes.c_int()
        name = ctypes.create_string_buffer(64)
        call("pbrt_ctx_device_info", self.h, ctypes.byref(cu), name, 64)
        self.cu_count = cu.value
        self.arch = name.value.decode()

    def close(self):
        if self.h:
            lib().pbrt_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        call("pbrt_stream_sync", self.h, self.stream)

    def alloc(self, nbytes):
        return DeviceArray(self, nbytes)

    def upload(self, array):
        a = np.ascontiguousarray(array)
        d = DeviceArray(self, a.nbytes)
        d.write(a)
        return d

    def event(self):
        return Event(self)

    def new_stream(self):
        """A further non-blocking stream on the device (pbrt_stream_create)."""
        return Stream(self)


class Stream:
    def __init__(self, ctx):
        self.ctx = ctx
        h = ctypes.c_void_p()
        call("pbrt_stream_create", ctx.h, ctypes.byref(h))
        self.h = h

    def synchronize(self):
        call("pbrt_stream_sync", self.ctx.h, self.h)

    def __del__(self):
        try:
            if self.h:
                lib().pbrt_stream_destroy(self.ctx.h, self.h)
                self.h = None
        except Exception:
            pass


class Event:
    def __init__(self, ctx):
        h = ctypes.c_void_p()
        call("pbrt_event_create", ctx.h, ctypes.byref(h))
        self.h = h

    def record(self, stream):
        call("pbrt_event_record", self.h, stream)

    def wait(self, ctx, stream):
        """Work queued on `stream` from now on waits on the device for this event's last record."""
        call("pbrt_stream_wait_event", ctx.h, stream, self.h)

    def elapsed_ms(self, end):
        ms = ctypes.c_float()
        call("pbrt_event_elapsed_ms", self.h, end.h, ctypes.byref(ms))
        return ms.value

    def __del__(self):
        try:
            lib().pbrt_event_destroy(self.h)
        except Exception:
            pass


class DeviceArray:
    """A device allocation (HBM) owned by a context."""

    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        call("pbrt_dev_alloc", ctx.h, self.nbytes, ctypes.byref(p))
        self.ptr = p

    def write(self, array):
        a = np.ascontiguousarray(array)
        if a.nbytes > self.nbytes:
            raise ValueError("array larger than the device allocation")
        call("pbrt_memcpy_h2d", self.ctx.h, self.ptr, _ptr(a), a.nbytes)

    def read(self, dtype, count=None):
        dtype = np.dtype(dtype)
        count = self.nbytes // dtype.itemsize if count is None else count
        out = np.empty(count, dtype)
        call("pbrt_memcpy_d2h", self.ctx.h, _ptr(out), self.ptr, out.nbytes)
        return out

    def view(self, offset, nbytes):
        """A non-owning window [offset, offset + nbytes) of this allocation."""
        if offset < 0 or offset + nbytes > self.nbytes:
            raise ValueError("view outside the allocation")
        v = DeviceArray.__new__(DeviceArray)
        v.ctx, v.nbytes, v.owner = self.ctx, int(nbytes), self
        v.ptr = ctypes.c_void_p(self.ptr.value + offset)
        v._view = True
        return v

    def fill(self, byte):
        call("pbrt_memset_dev", self.ctx.h, self.ptr, byte, self.nbytes)

    def free(self):
        if getattr(self, "_view", False):
            self.ptr = None
            return
        if self.ptr:
            lib().pbrt_dev_free(self.ctx.h, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class TriangleMesh:
    """TriangleMesh::new (src/shapes/triangle_mesh/triangle_mesh.rs:27-41) with its BVH on the device.

    vertices: (nv, 3) float32; indices: (nt, 3) uint32. An empty mesh raises PbrtError
    (PBRT_ERR_EMPTY_MESH), as the reference's constructor panics (triangle_mesh.rs:28)."""

    def __init__(self, ctx, vertices, indices, bins=12, max_leaf=4, t_trav=1.0, device_build=False):
        """device_build: build the BVH on the GPU (pbrt_mesh_create_device), the same tree."""
        self.ctx = ctx
        v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
        f = np.ascontiguousarray(indices, np.uint32).reshape(-1, 3)
        params = BvhParams(bins, max_leaf, t_trav)
        h = ctypes.c_void_p()
        fn = "pbrt_mesh_create_device" if device_build else "pbrt_mesh_create"
        call(fn, ctx.h if ctx is not None else None, _ptr(v), len(v), _ptr(f), len(f), ctypes.byref(params), ctypes.byref(h))
        self.h = h
        self.params = dict(bins=bins, max_leaf=max_leaf, t_trav=t_trav)

    def __del__(self):
        try:
            if self.h:
                lib().pbrt_mesh_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def build_stats(self):
        s = BuildStats()
        call("pbrt_mesh_build_stats", self.h, ctypes.byref(s))
        return {k: int(getattr(s, k)) for k, _ in BuildStats._fields_}

    def triangle_count(self):
        n = ctypes.c_uint32()
        call("pbrt_mesh_triangle_count", self.h, ctypes.byref(n))
        return n.value

    def device_bytes(self):
        a, b = ctypes.c_uint64(), ctypes.c_uint64()
        call("pbrt_mesh_device_bytes", self.h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def layout(self):
        """Host copies of the device arrays: (nodes (n, 64) uint8, triangles (nt, 48) uint8, root box
        (6,) float32, root reference)."""
        nb, tb = self.device_bytes()
        nodes = np.empty((nb // 64, 64), np.uint8)
        tris = np.empty((tb // 48, 48), np.uint8)
        box = np.empty(6, np.float32)
        ref = ctypes.c_int32()
        call("pbrt_mesh_copy_layout", self.h, _ptr(nodes), _ptr(tris), _ptr(box), ctypes.byref(ref))
        return nodes, tris, box, ref.value

    def set_ray_tiles(self, image_width=0):
        """pbrt_mesh_set_ray_tiles: later batches are images `image_width` rays wide, traversed in
        8 x 16 tiles (0: batch order); results are unchanged."""
        call("pbrt_mesh_set_ray_tiles", self.h, int(image_width))

    def set_launch(self, block=0, waves_per_cu=0, lds_stack=0, refill=0, width=0, packet=0):
        call("pbrt_mesh_set_launch", self.h,
             ctypes.byref(Launch(block, waves_per_cu, lds_stack, refill, width, packet)))

    def get_launch(self):
        l = Launch()
        call("pbrt_mesh_get_launch", self.h, ctypes.byref(l))
        return dict(block=l.block, waves_per_cu=l.waves_per_cu, lds_stack=l.lds_stack, refill=l.refill, width=l.width,
                    packet=l.packet)

    def autotune(self, rays, n, out, any_hit=False, reps=2):
        """Times every launch shape on these device rays and keeps the fastest (pbrt_mesh_autotune)."""
        l = Launch()
        ms = ctypes.c_float()
        call("pbrt_mesh_autotune", self.h, rays.ptr, n, out.ptr, int(any_hit), reps, ctypes.byref(l), ctypes.byref(ms))
        return dict(block=l.block, waves_per_cu=l.waves_per_cu, lds_stack=l.lds_stack, refill=l.refill,
                    width=l.width, packet=l.packet, ms=ms.value)

    def export_tree(self):
        s = self.build_stats()
        n, nt = s["node_count"], self.triangle_count()
        boxes = np.zeros((n, 6), np.float64)
        leaf = np.zeros(n, np.uint8)
        left, right, first, count = (np.zeros(n, np.uint32) for _ in range(4))
        tris = np.zeros(nt, np.uint32)
        call("pbrt_mesh_export_tree", self.h, _ptr(boxes), _ptr(leaf), _ptr(left), _ptr(right), _ptr(first),
             _ptr(count), _ptr(tris))
        return dict(boxes=boxes, leaf=leaf, left=left, right=right, first=first, count=count, tris=tris)

    # -- batched queries over device-resident rays ------------------------------------------------
    def query(self, rays, n, hits, stream=None):
        """BVHTree::query (triangle_mesh/bvh.rs:350-352): closest hit per ray into `hits` (pbrt_hit)."""
        call("pbrt_mesh_trace_closest", self.h, rays.ptr, n, hits.ptr, stream)

    def query_instrumented(self, rays, n, hits, stats, stream=None):
        """BVHTree::query_instrumented (bvh.rs:359-361): also n x {nodes, boxes, triangles} uint32."""
        call("pbrt_mesh_trace_closest_stats", self.h, rays.ptr, n, hits.ptr, stats.ptr, stream)

    def intersect_p(self, rays, n, occluded, stream=None):
        """BVHTree::intersect_p (bvh.rs:381-420): occluded[i] in {0, 1}."""
        call("pbrt_mesh_trace_any", self.h, rays.ptr, n, occluded.ptr, stream)

    def query_work(self, rays, n, out, work, any_hit=False, stream=None):
        """pbrt_mesh_trace_work: the launch's walk (tree width 2 or 4) with n x {node records, triangle
        records} fetched per ray into `work` (uint32); `out` as query / intersect_p."""
        call("pbrt_mesh_trace_work", self.h, rays.ptr, n, out.ptr, int(any_hit), work.ptr, stream)

    def status(self):
        s = ctypes.c_int()
        call("pbrt_mesh_status", self.h, ctypes.byref(s))
        return s.value

    # -- host convenience: upload, trace, download ------------------------------------------------
    def trace_host(self, rays, mode="closest", stats=False):
        rays = np.ascontiguousarray(rays)
        if rays.dtype != RAY_DTYPE:
            rays = np.ascontiguousarray(rays, np.float32).view(RAY_DTYPE).reshape(-1)
        n = len(rays)
        dr = self.ctx.upload(rays)
        if mode == "any":
            occ = self.ctx.alloc(max(n, 1))
            self.intersect_p(dr, n, occ)
            return occ.read(np.uint8, n)
        dh = self.ctx.alloc(max(n, 1) * 16)
        if stats:
            ds = self.ctx.alloc(max(n, 1) * 12)
            self.query_instrumented(dr, n, dh, ds)
            return dh.read(HIT_DTYPE, n), ds.read(np.uint32, 3 * n).reshape(n, 3)
        self.query(dr, n, dh)
        return dh.read(HIT_DTYPE, n)


# ---- synthetic inputs ----------------------------------------------------------------------------
def icosphere(subdiv, radius=1.0, center=(0.0, 0.0, 0.0), noise="none", amp=0.0, octaves=0, freq=1.5, seed=0):
    kinds = {"none": 0, "uniform": 1, "fbm": 2}
    nv, nt = ctypes.c_uint32(), ctypes.c_uint32()
    c = np.ascontiguousarray(center, np.float64)
    call("pbrt_gen_icosphere", subdiv, radius, _ptr(c), kinds[noise], amp, octaves, freq, seed, None, None,
         ctypes.byref(nv), ctypes.byref(nt))
    v = np.empty((nv.value, 3), np.float32)
    f = np.empty((nt.value, 3), np.uint32)
    call("pbrt_gen_icosphere", subdiv, radius, _ptr(c), kinds[noise], amp, octaves, freq, seed, _ptr(v), _ptr(f),
         ctypes.byref(nv), ctypes.byref(nt))
    return v, f


def pinhole_rays(width, height, fov, eye, look, tmin=1e-4, tmax=1000.0):
    out = np.empty(width * height, RAY_DTYPE)
    e = np.ascontiguousarray(eye, np.float64)
    l = np.ascontiguousarray(look, np.float64)
    call("pbrt_pinhole_rays", width, height, fov, _ptr(e), _ptr(l), tmin, tmax, _ptr(out))
    return out


def orbit_rays(vertices, view, side, fov=40.0 * np.pi / 180.0, tmin=1e-4, tmax=1000.0):
    v = np.ascontiguousarray(vertices, np.float32)
    out = np.empty(side * side, RAY_DTYPE)
    call("pbrt_orbit_rays", _ptr(v), len(v), view, side, fov, tmin, tmax, _ptr(out))
    return out


def random_rays(center, radius, n, seed=1, first=0, tmin=1e-4, tmax=1000.0):
    c = np.ascontiguousarray(center, np.float64)
    out = np.empty(n, RAY_DTYPE)
    call("pbrt_random_rays", _ptr(c), radius, seed, first, n, tmin, tmax, _ptr(out))
    return out


# ---- scene and renderer ----------------------------------------------------------------------------
SCENE_HIT_DTYPE = np.dtype([("t", "<f4"), ("u", "<f4"), ("v", "<f4"), ("leaf_slot", "<i4"), ("prim", "<i4"),
                            ("object", "<i4"), ("r0", "<i4"), ("r1", "<i4")])


class Scene:
    """Scene::new / add_object / add_light / commit (src/scene.rs:80-115) on the device, from a
    `pbrt_amd.scene.SceneDesc`; `intersect` / `intersect_p` are batched (scene.rs:118-132) and
    `render` is renderers::mt::render (src/renderers/mt.rs:23) with the path integrator."""

    def __init__(self, ctx, desc, bins=12, max_leaf=4, t_trav=1.0, device_build=False, mesh_width=None):
        """device_build: the meshes' BVHs built on the device at commit (pbrt_scene_set_mesh_build);
        mesh_width: the tree the scene walks inside its meshes, 4 (the library's default: 4-wide trees
        collapsed at commit) or 2 (the binary trees); pbrt_scene_set_mesh_width."""
        self.ctx = ctx
        self.desc = desc
        h = ctypes.c_void_p()
        call("pbrt_scene_create", ctx.h, ctypes.byref(BvhParams(bins, max_leaf, t_trav)), ctypes.byref(h))
        self.h = h
        if device_build:
            call("pbrt_scene_set_mesh_build", h, 1)
        if mesh_width is not None:
            call("pbrt_scene_set_mesh_width", h, mesh_width)
        out = ctypes.c_uint32()
        for v, f in desc.meshes:
            call("pbrt_scene_add_mesh", h, _ptr(v), len(v), _ptr(f), len(f), ctypes.byref(out))
        for kind, rgb, ior, rough, tex, rgb2, scale in desc.materials:
            c = np.array(rgb, np.float64)
            call("pbrt_scene_add_material", h, kind, _ptr(c), ior, rough, ctypes.byref(out))
            if tex:
                call("pbrt_scene_set_texture", h, out.value, tex, _ptr(np.array(rgb2, np.float64)), scale)
        for shape, mesh, mat, xf, param in desc.objects:
            p = np.array(param, np.float64)
            if shape == 4:  # CSG: elements [mesh, mesh + count) of desc.csg
                els = desc.csg[mesh:mesh + int(param[1])]
                shapes = np.array([e[0] for e in els], np.int32)
                prm = np.array([e[1] for e in els], np.float64)
                mats = [e[2].flat() if e[2] is not None else (np.eye(4).ravel(), np.eye(4).ravel()) for e in els]
                em = np.ascontiguousarray([m for m, _ in mats], np.float64)
                emi = np.ascontiguousarray([mi for _, mi in mats], np.float64)
                om, omi = (np.ascontiguousarray(a, np.float64) for a in xf.flat()) if xf is not None else (None, None)
                call("pbrt_scene_add_csg", h, int(param[0]), len(els), _ptr(shapes), _ptr(prm), _ptr(em), _ptr(emi), mat,
                     _ptr(om) if om is not None else None, _ptr(omi) if omi is not None else None, ctypes.byref(out))
                continue
            if xf is not None:
                m, mi = (np.ascontiguousarray(a, np.float64) for a in xf.flat())
                call("pbrt_scene_add_object", h, shape, mesh, _ptr(p), mat, _ptr(m), _ptr(mi), ctypes.byref(out))
            else:
                call("pbrt_scene_add_object", h, shape, mesh, _ptr(p), mat, None, None, ctypes.byref(out))
        for kind, obj, pos, rgb in desc.lights:
            c = np.array(rgb, np.float64)
            if kind == 0:
                call("pbrt_scene_add_point_light", h, _ptr(np.array(pos, np.float64)), _ptr(c))
            elif kind == 2:
                call("pbrt_scene_add_background_light", h, _ptr(np.array(pos, np.float64)), _ptr(c))
            else:
                call("pbrt_scene_add_area_light", h, obj, _ptr(c))
        call("pbrt_scene_commit", h)

    @classmethod
    def load_stage(cls, ctx, path, base_dir=None, bins=12, max_leaf=4, t_trav=1.0):
        """Loader::load_scene (src/loader.rs:29-46) through pbrt_scene_load_stage: the `.stage`
        file's scene on the device, mesh files read relative to `base_dir` (None: the working
        directory). Returns (scene, view), view = dict(eye, look, up, fov, near, far) for render."""
        self = cls.__new__(cls)
        self.ctx, self.desc, self.h = ctx, None, None
        h = ctypes.c_void_p()
        cfg = _native.RenderConfig()
        call("pbrt_scene_load_stage", ctx.h, os.fsencode(path), os.fsencode(base_dir) if base_dir else None,
             ctypes.byref(BvhParams(bins, max_leaf, t_trav)), ctypes.byref(h), ctypes.byref(cfg))
        self.h = h
        view = dict(eye=tuple(cfg.eye), look=tuple(cfg.look), up=tuple(cfg.up), fov=cfg.fov, near=cfg.near, far=cfg.far)
        return self, view

    def __del__(self):
        try:
            if self.h:
                lib().pbrt_scene_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def set_launch(self, lds_stack=0, blocks_per_cu=0, width=0, min_waves=0):
        """Launch settings of the scene walks (pbrt_scene_set_launch); same hits for every setting."""
        call("pbrt_scene_set_launch", self.h,
             ctypes.byref(_native.SceneLaunch(lds_stack, blocks_per_cu, width, min_waves)))

    def get_launch(self):
        l = _native.SceneLaunch()
        call("pbrt_scene_get_launch", self.h, ctypes.byref(l))
        return dict(lds_stack=l.lds_stack, blocks_per_cu=l.blocks_per_cu, width=l.width, min_waves=l.min_waves)

    def autotune(self, rays, n, hits, reps=2):
        """Times the scene walk's launch settings on these device rays and keeps the fastest."""
        l = _native.SceneLaunch()
        ms = ctypes.c_float()
        call("pbrt_scene_autotune", self.h, rays.ptr, n, hits.ptr, reps, ctypes.byref(l), ctypes.byref(ms))
        return dict(lds_stack=l.lds_stack, blocks_per_cu=l.blocks_per_cu, width=l.width, min_waves=l.min_waves,
                    ms=ms.value)

    def autotune_render(self, width, height, spp, max_depth, eye, look, up=(0.0, 1.0, 0.0), fov=np.pi / 2, seed=1,
                        near=1e-4, far=1000.0, reps=2, mis=False):
        """The scene autotuner timed on short renders of this view (pbrt_scene_autotune_render)."""
        c = self._config(width, height, spp, max_depth, eye, look, up, fov, seed, near, far, 0, mis)
        l = _native.SceneLaunch()
        ms = ctypes.c_float()
        call("pbrt_scene_autotune_render", self.h, ctypes.byref(c), reps, ctypes.byref(l), ctypes.byref(ms))
        return dict(lds_stack=l.lds_stack, blocks_per_cu=l.blocks_per_cu, width=l.width, min_waves=l.min_waves,
                    ms=ms.value)

    def build_seconds(self):
        """Wall time of the meshes' BVH builds in the commit."""
        t = ctypes.c_double()
        call("pbrt_scene_build_seconds", self.h, ctypes.byref(t))
        return t.value

    def info(self):
        a, b, c, d = (ctypes.c_uint32() for _ in range(4))
        call("pbrt_scene_info", self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c), ctypes.byref(d))
        return dict(objects=a.value, tlas_nodes=b.value, tlas_depth=c.value, mesh_depth=d.value)

    def intersect(self, rays, n, hits, stream=None):
        call("pbrt_scene_trace_closest", self.h, rays.ptr, n, hits.ptr, stream)

    def intersect_p(self, rays, n, occluded, stream=None):
        call("pbrt_scene_trace_any", self.h, rays.ptr, n, occluded.ptr, stream)

    def status(self):
        s = ctypes.c_int()
        call("pbrt_scene_status", self.h, ctypes.byref(s))
        return s.value

    def trace_host(self, rays, any_hit=False):
        rays = np.ascontiguousarray(rays)
        if rays.dtype != RAY_DTYPE:
            rays = np.ascontiguousarray(rays, np.float32).view(RAY_DTYPE).reshape(-1)
        n = len(rays)
        dr = self.ctx.upload(rays)
        if any_hit:
            occ = self.ctx.alloc(max(n, 1))
            self.intersect_p(dr, n, occ)
            return occ.read(np.uint8, n)
        dh = self.ctx.alloc(max(n, 1) * 32)
        self.intersect(dr, n, dh)
        return dh.read(SCENE_HIT_DTYPE, n)

    @staticmethod
    def _config(width, height, spp, max_depth, eye, look, up, fov, seed, near, far, samples_per_pass, mis=False,
                roulette=False, integrator="path", sampler="independent"):
        c = _native.RenderConfig()
        c.width, c.height, c.spp, c.max_depth, c.seed, c.fov = width, height, spp, max_depth, seed, fov
        c.eye[:], c.look[:], c.up[:] = list(eye), list(look), list(up)
        c.near, c.far, c.samples_per_pass = near, far, samples_per_pass
        c.mis = int(bool(mis))
        c.roulette = int(bool(roulette))
        c.integrator = INTEGRATOR[integrator] if isinstance(integrator, str) else int(integrator)
        c.sampler = SAMPLER[sampler] if isinstance(sampler, str) else int(sampler)
        return c

    def render(self, width, height, spp, max_depth, eye, look, up=(0.0, 1.0, 0.0), fov=np.pi / 2, seed=1,
               near=1e-4, far=1000.0, samples_per_pass=0, film=True, mis=False, roulette=False, integrator="path",
               sampler="independent"):
        """Per-pixel mean radiance (height, width, 3) float64 and the render statistics; `mis`:
        multiple importance sampling instead of the reference's estimator; `roulette`: Russian
        roulette after each scattering; `integrator`: "path", "naive" or "normal" (the reference's
        integrators::Type); `sampler`: "independent" (counter-based draws) or "sobol" (stratified
        pixel jitter and first-vertex light / BSDF pairs)."""
        c = self._config(width, height, spp, max_depth, eye, look, up, fov, seed, near, far, samples_per_pass, mis,
                         roulette, integrator, sampler)
        out = np.zeros((height, width, 3), np.float64) if film else None
        st = _native.RenderStats()
        call("pbrt_render", self.h, ctypes.byref(c), _ptr(out) if film else None, ctypes.byref(st))
        return out, dict(rays=st.rays, paths=st.paths, passes=st.passes, seconds=st.seconds)

    def render_jobs(self, jobs, use_graphs=True):
        """Render a batch of jobs together (pbrt_render_jobs). `jobs`: dicts with width, height,
        spp, max_depth, eye, look and optionally up, fov, seed, near, far. Returns the films (each
        (height, width, 3) float64) and the batch statistics."""
        n = len(jobs)
        cfgs = (_native.RenderConfig * n)()
        films = []
        for k, j in enumerate(jobs):
            cfgs[k] = self._config(j["width"], j["height"], j["spp"], j["max_depth"], j["eye"], j["look"],
                                   j.get("up", (0.0, 1.0, 0.0)), j.get("fov", np.pi / 2), j.get("seed", 1),
                                   j.get("near", 1e-4), j.get("far", 1000.0), 0, j.get("mis", False),
                                   j.get("roulette", False), j.get("integrator", "path"),
                                   j.get("sampler", "independent"))
            films.append(np.zeros((j["height"], j["width"], 3), np.float64))
        ptrs = (ctypes.c_void_p * n)(*[f.ctypes.data for f in films])
        st = _native.RenderStats()
        call("pbrt_render_jobs", self.h, cfgs, n, ptrs, int(bool(use_graphs)), ctypes.byref(st))
        return films, dict(rays=st.rays, paths=st.paths, passes=st.passes, seconds=st.seconds)

    def render_ex(self, width, height, spp, max_depth, eye, look, up=(0.0, 1.0, 0.0), fov=np.pi / 2, seed=1,
                  near=1e-4, far=1000.0, samples_per_pass=0, cache=None, query_vertex=3, train_stride=33,
                  train_batch=65536, train_steps=1, reference=None, record=True, debug_slots=0, mis=False,
                  roulette=False, integrator="path", sampler="independent"):
        """Render with the radiance cache `cache` (a RadianceCache, or None for plain path tracing)
        and, with `record`, the per-pass time (ms) and RMSE against `reference` (height, width, 3).
        Returns (film, stats, progress); progress has `ms`, `rmse` (or None) and `spp` per pass, and
        `debug` (raw inputs, targets) of the last pass's training slots when debug_slots > 0."""
        c = self._config(width, height, spp, max_depth, eye, look, up, fov, seed, near, far, samples_per_pass, mis,
                         roulette, integrator, sampler)
        out = np.zeros((height, width, 3), np.float64)
        st = _native.RenderStats()
        nc = None
        dbg = None
        if cache is not None:
            nc = _native.NrcRender(query_vertex, train_stride, train_batch, train_steps, None, debug_slots, 0)
            if debug_slots:
                dbg = np.zeros(debug_slots * 16, np.float32)
                nc.debug = dbg.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        prog = None
        per = samples_per_pass or (1 if cache is not None else
                                   max(1, min(max(1, (1 << 23) // (width * height)), (spp + 1) // 2)))
        per = min(per, spp)
        passes = (spp + per - 1) // per
        if record:
            ms = np.zeros(passes, np.float64)
            rmse = np.zeros(passes, np.float64)
            ref = None if reference is None else np.ascontiguousarray(reference, np.float64)
            if ref is not None and ref.shape != (height, width, 3):
                raise ValueError("reference must be (height, width, 3)")
            dp = ctypes.POINTER(ctypes.c_double)
            prog = _native.Progress(ref.ctypes.data_as(dp) if ref is not None else None, rmse.ctypes.data_as(dp),
                                    ms.ctypes.data_as(dp), passes, 0)
        call("pbrt_render_ex", self.h, ctypes.byref(c), cache.h if cache is not None else None,
             ctypes.byref(nc) if nc is not None else None, ctypes.byref(prog) if prog is not None else None,
             _ptr(out), ctypes.byref(st))
        progress = None
        if record:
            spp_done = np.minimum(np.arange(1, passes + 1) * per, spp)
            progress = dict(ms=ms[:prog.count], rmse=rmse[:prog.count] if reference is not None else None,
                            spp=spp_done[:prog.count])
        if dbg is not None:
            progress = progress or {}
            progress["debug"] = (dbg[:debug_slots * 12].reshape(debug_slots, 12),
                                 dbg[debug_slots * 12:].reshape(debug_slots, 4))
        return out, dict(rays=st.rays, paths=st.paths, passes=st.passes, seconds=st.seconds,
                        cache_queries=st.cache_queries), progress


NRC_PARAMS = 17923
NRC_FEATURES = 96


class RadianceCache:
    """Neural radiance cache (pbrt_nrc): an 81 -> 64 x 4 -> 3 ReLU MLP on the matrix cores, trained
    online with a relative L2 loss and Adam. No counterpart in the reference (its integrator is plain
    path tracing, src/integrators/path.rs). Device-pointer entry points (`encode`, `infer`,
    `train_step`) take DeviceArrays; the `*_host` helpers copy numpy arrays in and out."""

    def __init__(self, ctx, seed=1):
        self.ctx = ctx
        h = ctypes.c_void_p()
        call("pbrt_nrc_create", ctx.h, seed, ctypes.byref(h))
        self.h = h

    def __del__(self):
        try:
            if self.h:
                lib().pbrt_nrc_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def set_bounds(self, lo, hi):
        lo = np.ascontiguousarray(lo, np.float32)
        hi = np.ascontiguousarray(hi, np.float32)
        call("pbrt_nrc_set_bounds", self.h, _ptr(lo), _ptr(hi))

    def set_optim(self, lr=1e-3, beta1=0.9, beta2=0.99, eps=1e-8, ema=0.0):
        call("pbrt_nrc_set_optim", self.h, ctypes.byref(_native.NrcOptim(lr, beta1, beta2, eps, ema)))

    @property
    def params(self):
        out = np.empty(NRC_PARAMS, np.float32)
        call("pbrt_nrc_get_params", self.h, _ptr(out), NRC_PARAMS)
        return out

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        if p.shape != (NRC_PARAMS,):
            raise ValueError(f"expected {NRC_PARAMS} parameters")
        call("pbrt_nrc_set_params", self.h, _ptr(p), NRC_PARAMS)

    @property
    def steps(self):
        s = ctypes.c_uint64()
        call("pbrt_nrc_step_count", self.h, ctypes.byref(s))
        return s.value

    def encode(self, raw, n, x, stream=None):
        call("pbrt_nrc_encode", self.h, raw.ptr, n, x.ptr, stream)

    def infer(self, x, n, out, stream=None):
        call("pbrt_nrc_infer", self.h, x.ptr, n, out.ptr, stream)

    def query(self, raw, n, out, stream=None):
        """Raw inputs (n x 12 floats) to radiance (n x 3), features computed in the kernel."""
        call("pbrt_nrc_query", self.h, raw.ptr, n, out.ptr, stream)

    def query_host(self, raw12):
        raw = np.ascontiguousarray(raw12, np.float32).reshape(-1, 12)
        n = len(raw)
        dr = self.ctx.upload(raw)
        do = self.ctx.alloc(max(n, 1) * 12)
        self.query(dr, n, do)
        return do.read(np.float32, n * 3).reshape(n, 3)

    def train_step(self, x, targets, n, loss=None, stream=None):
        call("pbrt_nrc_train_step", self.h, x.ptr, targets.ptr, n, loss.ptr if loss is not None else None, stream)

    def gradients(self, x, targets, n):
        g = np.empty(NRC_PARAMS, np.float32)
        loss = ctypes.c_float()
        call("pbrt_nrc_gradients", self.h, x.ptr, targets.ptr, n, _ptr(g), ctypes.byref(loss))
        return loss.value, g

    # host-array conveniences
    def encode_host(self, raw12):
        raw = np.ascontiguousarray(raw12, np.float32).reshape(-1, 12)
        n = len(raw)
        dr = self.ctx.upload(raw)
        dx = self.ctx.alloc(max(n, 1) * NRC_FEATURES * 2)
        self.encode(dr, n, dx)
        return dx.read(np.uint16, n * NRC_FEATURES).reshape(n, NRC_FEATURES)

    def infer_host(self, x_bits):
        x = np.ascontiguousarray(x_bits, np.uint16).reshape(-1, NRC_FEATURES)
        n = len(x)
        dx = self.ctx.upload(x)
        do = self.ctx.alloc(max(n, 1) * 12)
        self.infer(dx, n, do)
        return do.read(np.float32, n * 3).reshape(n, 3)

    def gradients_host(self, x_bits, targets):
        x = np.ascontiguousarray(x_bits, np.uint16).reshape(-1, NRC_FEATURES)
        t = np.ascontiguousarray(targets, np.float32).reshape(-1, 3)
        return self.gradients(self.ctx.upload(x), self.ctx.upload(t), len(x))

    def train_step_host(self, x_bits, targets):
        x = np.ascontiguousarray(x_bits, np.uint16).reshape(-1, NRC_FEATURES)
        t = np.ascontiguousarray(targets, np.float32).reshape(-1, 3)
        dl = self.ctx.alloc(4)
        self.train_step(self.ctx.upload(x), self.ctx.upload(t), len(x), dl)
        return float(dl.read(np.float32, 1)[0])


def bf16_bits(x):
    """FP32 -> BF16 bit patterns, round to nearest even (NaN and infinities kept)."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    special = (u & 0x7F800000) == 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    return np.where(special, (u >> 16).astype(np.uint16), r)


def bf16_float(bits):
    return (np.asarray(bits, np.uint16).astype(np.uint32) << 16).view(np.float32)


def mfma_check(ctx, a_bits, b_bits):
    """D = A B on one 16x16x32 BF16 matrix-core product (A 16x32, B 32x16 bit patterns)."""
    a = np.ascontiguousarray(a_bits, np.uint16).reshape(16, 32)
    b = np.ascontiguousarray(b_bits, np.uint16).reshape(32, 16)
    d = np.empty((16, 16), np.float32)
    call("pbrt_nrc_mfma_check", ctx.h, _ptr(a), _ptr(b), _ptr(d))
    return d


def film_to_rgba8(film, tonemap=False):
    """The film (height, width, 3) as the reference's 8-bit RGBA image (pbrt_film_to_rgba8)."""
    f = np.ascontiguousarray(film, np.float64)
    h, w = f.shape[:2]
    out = np.empty((h, w, 4), np.uint8)
    call("pbrt_film_to_rgba8", _ptr(f), w, h, int(bool(tonemap)), _ptr(out))
    return out


def write_png(film, path, tonemap=False):
    """The film (height, width, 3) as an 8-bit RGBA PNG file (pbrt_film_write_png)."""
    f = np.ascontiguousarray(film, np.float64)
    h, w = f.shape[:2]
    call("pbrt_film_write_png", _ptr(f), w, h, int(bool(tonemap)), os.fsencode(path))


def read_ply(path, reverse=False):
    """load_ply_mesh (src/loader/mesh_loader.rs) through pbrt_ply_read: (vertices float32 (nv, 3),
    faces uint32 (nt, 3)) of an ASCII PLY file, each face reversed when `reverse`."""
    nv, nt = ctypes.c_uint32(), ctypes.c_uint32()
    call("pbrt_ply_read", os.fsencode(path), int(bool(reverse)), None, ctypes.byref(nv), None, ctypes.byref(nt))
    v = np.empty((nv.value, 3), np.float32)
    f = np.empty((nt.value, 3), np.uint32)
    call("pbrt_ply_read", os.fsencode(path), int(bool(reverse)), _ptr(v), ctypes.byref(nv), _ptr(f), ctypes.byref(nt))
    return v, f


def to_rgb8(film):
    """Spectrum::gamma_correct + to_rgb (src/spectrum.rs:13-30): sqrt, clamp to [0, 1], x255, truncate."""
    g = np.sqrt(np.asarray(film, np.float64))
    return (np.clip(g, 0.0, 1.0) * 255.0).astype(np.uint8)
