import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

MESH_DIR = os.path.join(ROOT, "tests", "golden", "meshes")

# Published known answers of the reference's bvh_stats protocol (docs/mesures_bvh.md:706-713):
# hit count, node count, leaf count, box tests / ray, triangle tests / ray.
PUBLISHED = {
    "cube": (116004, 11, 6, 5.83, 1.19),
    "bun_zipper_res4": (47693, 1811, 906, 10.04, 0.77),
    "dragon_vrip_res4": (42797, 21137, 10569, 12.01, 0.63),
    "dragon_vrip_res3": (43451, 92929, 46465, 13.75, 0.60),
    "bunny": (46904, 138881, 69441, 14.19, 0.58),
}


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU")


def load_mesh(name):
    d = np.load(os.path.join(MESH_DIR, name + ".npz"))
    return d["vertices"], d["faces"]


def grid_mesh(side):
    """triangle_mesh/bvh.rs:746-766: side^3 small, well separated, non-axis-aligned triangles."""
    v, f = [], []
    for ix in range(side):
        for iy in range(side):
            for iz in range(side):
                x, y, z = float(ix), float(iy), float(iz)
                b = len(v)
                v += [(x, y, z), (x + 0.4, y + 0.1, z), (x + 0.1, y + 0.4, z + 0.3)]
                f.append((b, b + 1, b + 2))
    return np.array(v, np.float32), np.array(f, np.uint32)


def _normalized(d):
    d = np.asarray(d, np.float64)
    n = np.sqrt((d * d).sum(-1, keepdims=True))
    return d * (1.0 / n)


def grid_ray_set(v, f, near=1e-4, far=1000.0):
    """triangle_mesh/bvh.rs:828-854: rays from three far origins at every centroid, plus 7x7x3
    axis-aligned pencils. Rows of 8 f64 (ox, oy, oz, near, dx, dy, dz, far)."""
    c = (v[f[:, 0]].astype(np.float64) + v[f[:, 1]]) + v[f[:, 2]]
    c = c * (1.0 / 3.0)
    rows = []
    origins = [(-6.0, -7.0, -8.0), (9.0, -6.0, 7.0), (-5.0, 8.0, -9.0)]
    for ci in c:
        for o in origins:
            d = _normalized(ci - np.array(o))
            rows.append([*o, near, *d, far])
    for a in range(7):
        for b in range(7):
            u = -1.0 + a * 0.7
            w = -1.0 + b * 0.7
            rows.append([-5.0, u, w, near, 1.0, 0.0, 0.0, far])
            rows.append([u, -5.0, w, near, 0.0, 1.0, 0.0, far])
            rows.append([u, w, -5.0, near, 0.0, 0.0, 1.0, far])
    return np.array(rows, np.float64)


@pytest.fixture(scope="session")
def gpu_ctx():
    import pbrt_amd
    ctx = pbrt_amd.Context(0)
    yield ctx
    ctx.close()
