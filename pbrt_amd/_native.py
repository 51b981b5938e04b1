"""ctypes binding of libpbrt_gpu.so (the C ABI declared in include/pbrt_gpu.h).

The library is loaded from this package directory only. If it is missing, every call raises: there
is no CPU fallback anywhere in this package.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# PBRT_AMD_LIB_VARIANT=asan loads the host-sanitised build of the same sources (tests/sanitize_host.py)
_VARIANT = os.environ.get("PBRT_AMD_LIB_VARIANT", "")
LIB_PATH = os.path.join(HERE, f"libpbrt_gpu_{_VARIANT}.so" if _VARIANT else "libpbrt_gpu.so")

P = ctypes.c_void_p
U32 = ctypes.c_uint32
U64 = ctypes.c_uint64
I32 = ctypes.c_int
F32 = ctypes.c_float
F64 = ctypes.c_double
PP = ctypes.POINTER(ctypes.c_void_p)


class BvhParams(ctypes.Structure):
    _fields_ = [("bins", U32), ("max_leaf", U32), ("t_trav", F64)]


class BuildStats(ctypes.Structure):
    _fields_ = [("node_count", U64), ("leaf_count", U64), ("max_depth", U64), ("max_leaf_tri_count", U64),
                ("total_leaf_tri_count", U64)]


class Launch(ctypes.Structure):
    _fields_ = [("block", U32), ("waves_per_cu", U32), ("lds_stack", U32), ("refill", U32), ("width", U32), ("packet", U32)]


class SceneLaunch(ctypes.Structure):
    _fields_ = [("lds_stack", U32), ("blocks_per_cu", U32), ("width", U32), ("min_waves", U32)]


class RenderConfig(ctypes.Structure):
    _fields_ = [("width", U32), ("height", U32), ("spp", U32), ("max_depth", U32), ("seed", U64), ("fov", F64),
                ("eye", F64 * 3), ("look", F64 * 3), ("up", F64 * 3), ("near", F64), ("far", F64),
                ("samples_per_pass", U32), ("mis", U32), ("roulette", U32), ("integrator", U32),
                ("sampler", U32), ("reserved", U32)]


class RenderStats(ctypes.Structure):
    _fields_ = [("rays", U64), ("paths", U64), ("passes", U64), ("seconds", F64), ("cache_queries", U64)]


class NrcOptim(ctypes.Structure):
    _fields_ = [("lr", F32), ("beta1", F32), ("beta2", F32), ("eps", F32), ("ema", F32)]


class NrcRender(ctypes.Structure):
    _fields_ = [("query_vertex", U32), ("train_stride", U32), ("train_batch", U32), ("train_steps", U32),
                ("debug", ctypes.POINTER(F32)), ("debug_slots", U32), ("reserved", U32)]


class Progress(ctypes.Structure):
    _fields_ = [("reference", ctypes.POINTER(F64)), ("rmse", ctypes.POINTER(F64)), ("ms", ctypes.POINTER(F64)),
                ("capacity", U32), ("count", U32)]


# name -> argtypes (restype int unless listed in _RESTYPES)
SIGNATURES = {
    "pbrt_last_error": [],
    "pbrt_version": [],
    "pbrt_ctx_create": [I32, PP],
    "pbrt_ctx_destroy": [P],
    "pbrt_ctx_stream": [P, PP],
    "pbrt_ctx_device_info": [P, ctypes.POINTER(I32), ctypes.c_char_p, ctypes.c_size_t],
    "pbrt_dev_alloc": [P, ctypes.c_size_t, PP],
    "pbrt_dev_free": [P, P],
    "pbrt_memcpy_h2d": [P, P, P, ctypes.c_size_t],
    "pbrt_memcpy_d2h": [P, P, P, ctypes.c_size_t],
    "pbrt_memset_dev": [P, P, I32, ctypes.c_size_t],
    "pbrt_dev_copy": [P, P, P, ctypes.c_size_t, P],
    "pbrt_stream_create": [P, PP],
    "pbrt_stream_destroy": [P, P],
    "pbrt_stream_sync": [P, P],
    "pbrt_event_create": [P, PP],
    "pbrt_event_record": [P, P],
    "pbrt_stream_wait_event": [P, P, P],
    "pbrt_event_elapsed_ms": [P, P, ctypes.POINTER(F32)],
    "pbrt_event_destroy": [P],
    "pbrt_mesh_create": [P, P, U32, P, U32, ctypes.POINTER(BvhParams), PP],
    "pbrt_mesh_create_device": [P, P, U32, P, U32, ctypes.POINTER(BvhParams), PP],
    "pbrt_mesh_copy_layout": [P, P, P, P, ctypes.POINTER(I32)],
    "pbrt_mesh_destroy": [P],
    "pbrt_mesh_build_stats": [P, ctypes.POINTER(BuildStats)],
    "pbrt_mesh_triangle_count": [P, ctypes.POINTER(U32)],
    "pbrt_mesh_device_bytes": [P, ctypes.POINTER(U64), ctypes.POINTER(U64)],
    "pbrt_mesh_export_tree": [P, P, P, P, P, P, P, P],
    "pbrt_mesh_trace_closest": [P, P, U32, P, P],
    "pbrt_mesh_trace_closest_stats": [P, P, U32, P, P, P],
    "pbrt_mesh_trace_work": [P, P, U32, P, I32, P, P],
    "pbrt_mesh_trace_closest_noalloc": [P, P, U32, P, P],
    "pbrt_mesh_trace_any": [P, P, U32, P, P],
    "pbrt_mesh_set_launch": [P, ctypes.POINTER(Launch)],
    "pbrt_mesh_set_ray_tiles": [P, U32],
    "pbrt_mesh_get_launch": [P, ctypes.POINTER(Launch)],
    "pbrt_mesh_autotune": [P, P, U32, P, I32, U32, ctypes.POINTER(Launch), ctypes.POINTER(F32)],
    "pbrt_mesh_status": [P, ctypes.POINTER(I32)],
    "pbrt_scene_create": [P, ctypes.POINTER(BvhParams), PP],
    "pbrt_scene_destroy": [P],
    "pbrt_scene_add_mesh": [P, P, U32, P, U32, ctypes.POINTER(U32)],
    "pbrt_scene_add_material": [P, I32, P, F64, F64, ctypes.POINTER(U32)],
    "pbrt_scene_add_object": [P, I32, I32, P, U32, P, P, ctypes.POINTER(U32)],
    "pbrt_scene_add_csg": [P, I32, U32, P, P, P, P, U32, P, P, ctypes.POINTER(U32)],
    "pbrt_scene_add_point_light": [P, P, P],
    "pbrt_scene_add_area_light": [P, U32, P],
    "pbrt_scene_commit": [P],
    "pbrt_scene_set_mesh_build": [P, I32],
    "pbrt_scene_set_mesh_width": [P, U32],
    "pbrt_scene_set_launch": [P, ctypes.POINTER(SceneLaunch)],
    "pbrt_scene_get_launch": [P, ctypes.POINTER(SceneLaunch)],
    "pbrt_scene_autotune": [P, P, U32, P, U32, ctypes.POINTER(SceneLaunch), ctypes.POINTER(ctypes.c_float)],
    "pbrt_scene_autotune_render": [P, P, U32, ctypes.POINTER(SceneLaunch), ctypes.POINTER(ctypes.c_float)],
    "pbrt_scene_build_seconds": [P, ctypes.POINTER(F64)],
    "pbrt_scene_info": [P, ctypes.POINTER(U32), ctypes.POINTER(U32), ctypes.POINTER(U32), ctypes.POINTER(U32)],
    "pbrt_scene_trace_closest": [P, P, U32, P, P],
    "pbrt_scene_trace_any": [P, P, U32, P, P],
    "pbrt_scene_status": [P, ctypes.POINTER(I32)],
    "pbrt_render": [P, P, P, P],
    "pbrt_film_to_rgba8": [P, U32, U32, U32, P],
    "pbrt_scene_set_texture": [P, U32, I32, P, F64],
    "pbrt_scene_add_background_light": [P, P, P],
    "pbrt_film_write_png": [P, U32, U32, U32, ctypes.c_char_p],
    "pbrt_ply_read": [ctypes.c_char_p, I32, P, ctypes.POINTER(U32), P, ctypes.POINTER(U32)],
    "pbrt_scene_load_stage": [P, ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(BvhParams), PP,
                              ctypes.POINTER(RenderConfig)],
    "pbrt_render_ex": [P, P, P, ctypes.POINTER(NrcRender), ctypes.POINTER(Progress), P, P],
    "pbrt_render_jobs": [P, P, U32, P, I32, P],
    "pbrt_nrc_create": [P, U64, PP],
    "pbrt_nrc_destroy": [P],
    "pbrt_nrc_set_bounds": [P, P, P],
    "pbrt_nrc_set_optim": [P, ctypes.POINTER(NrcOptim)],
    "pbrt_nrc_set_params": [P, P, U32],
    "pbrt_nrc_get_params": [P, P, U32],
    "pbrt_nrc_encode": [P, P, U32, P, P],
    "pbrt_nrc_infer": [P, P, U32, P, P],
    "pbrt_nrc_query": [P, P, U32, P, P],
    "pbrt_nrc_gradients": [P, P, P, U32, P, ctypes.POINTER(F32)],
    "pbrt_nrc_train_step": [P, P, P, U32, P, P],
    "pbrt_nrc_step_count": [P, ctypes.POINTER(U64)],
    "pbrt_nrc_mfma_check": [P, P, P, P],
    "pbrt_pinhole_rays": [U32, U32, F64, P, P, F32, F32, P],
    "pbrt_orbit_rays": [P, U32, U32, U32, F64, F32, F32, P],
    "pbrt_random_rays": [P, F64, U64, U64, U32, F32, F32, P],
    "pbrt_gen_icosphere": [U32, F64, P, I32, F64, I32, F64, U64, P, P, ctypes.POINTER(U32), ctypes.POINTER(U32)],
}
_RESTYPES = {"pbrt_last_error": ctypes.c_char_p}

_lib = None


class PbrtError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"pbrt error {code}: {msg}")
        self.code = code


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"native library not built: {LIB_PATH} (run `python -m pbrt_amd.build`)")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = _RESTYPES.get(name, ctypes.c_int)
        _lib = L
    return _lib


def check(code):
    if code != 0:
        raise PbrtError(code, lib().pbrt_last_error().decode(errors="replace"))
    return code


def call(name, *args):
    return check(getattr(lib(), name)(*args))
