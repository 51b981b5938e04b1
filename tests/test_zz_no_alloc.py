"""pbrt_mesh_trace_closest_noalloc: the closest-hit entry point whose traversal spill goes to the
library's fixed device array instead of an allocation. Called through the C ABI (ctypes on the
library itself) and checked bit for bit against pbrt_mesh_trace_closest on the same rays, on two
calls in a row, the second with a smaller batch, for launches that spill to HBM (8-entry LDS
stacks on the deep reference-mode tree), whose grid the entry point must cap, and for the
benchmark's 4-wide launch."""
import numpy as np
import pytest

import pbrt_amd
from conftest import load_mesh
from pbrt_amd import _native

pytestmark = pytest.mark.gpu


def _rows(r):
    return np.ascontiguousarray(r).view(np.float32).reshape(-1, 8)


def _noalloc(mesh, dr, n, dh):
    L = _native.lib()
    rc = L.pbrt_mesh_trace_closest_noalloc(mesh.h, dr.ptr, n, dh.ptr, None)
    assert rc == 0, L.pbrt_last_error().decode()


@pytest.mark.parametrize("launch", [dict(block=128, lds_stack=8, waves_per_cu=32, width=2),
                                    dict(block=256, lds_stack=8, waves_per_cu=32, width=4, refill=4),
                                    dict(block=128, lds_stack=8, waves_per_cu=24, width=4, refill=1),
                                    dict(block=64, lds_stack=16, waves_per_cu=32, width=4),
                                    dict(packet=1)])
def test_noalloc_matches_trace_closest(gpu_ctx, launch):
    v, f = load_mesh("dragon_vrip_res3")
    mesh = pbrt_amd.TriangleMesh(gpu_ctx, v, f, **pbrt_amd.REFERENCE_BVH)
    mesh.set_launch(**launch)
    big = _rows(pbrt_amd.random_rays(v.mean(0), 0.1, 300000, seed=31))
    small = np.concatenate([_rows(pbrt_amd.orbit_rays(v, 2, 64)), _rows(pbrt_amd.random_rays(v.mean(0), 0.1, 7001, seed=32))])
    for rays in (big, small):
        n = len(rays)
        dr = gpu_ctx.upload(rays)
        ref, out = gpu_ctx.alloc(16 * n), gpu_ctx.alloc(16 * n)
        ref.fill(0xFF)
        out.fill(0xFF)
        mesh.query(dr, n, ref)
        _noalloc(mesh, dr, n, out)
        gpu_ctx.synchronize()
        a = ref.read(pbrt_amd.HIT_DTYPE, n)
        b = out.read(pbrt_amd.HIT_DTYPE, n)
        assert (a["prim"] >= 0).sum() > n // 10
        assert a.tobytes() == b.tobytes(), launch
    assert mesh.status() == 0


def test_noalloc_rejects_null_and_unset_wide_tree(gpu_ctx):
    v, f = pbrt_amd.icosphere(3)
    mesh = pbrt_amd.TriangleMesh(gpu_ctx, v, f)
    L = _native.lib()
    assert L.pbrt_mesh_trace_closest_noalloc(mesh.h, None, 16, None, None) != 0
    rays = _rows(pbrt_amd.random_rays((0, 0, 0), 1.0, 64, seed=1))
    dr, dh = gpu_ctx.upload(rays), gpu_ctx.alloc(16 * 64)
    _noalloc(mesh, dr, 0, dh)  # an empty batch is a no-op
    _noalloc(mesh, dr, 64, dh)
    assert dh.read(pbrt_amd.HIT_DTYPE, 64).tobytes() == mesh.trace_host(rays).tobytes()
