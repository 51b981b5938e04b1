"""Seeded synthetic scenes of the path-tracing workloads, as `SceneDesc` + camera.

  small(): the CPU-runnable scene: an icosphere of 4 subdivisions (5,120 triangles) with +-5%
      seeded radial noise on a 2-triangle floor, 64 seeded spheres (radius U[0.2, 1], centres uniform
      in a 20^3 box), Lambertian albedo U[0.2, 0.8]^3, one point light; 256x256, 16 spp, depth 8.
  room(): a closed room of five diffuse wall rectangles (albedo 0.73, one red and one green side
      wall), two boxes turned +-17 degrees about y, an emissive ceiling rectangle (radiance 15)
      covering 1/16 of the ceiling and lighting the room as an area light, and an icosphere of
      `subdiv` subdivisions with 3-octave value-noise displacement of 8%; Lambertian and dielectric
      (IOR 1.5) materials; 512x512, 64 spp, depth 8.
"""
import math

import numpy as np

from . import icosphere
from .scene import SceneDesc, Transform


def _u(seed, *key):
    rng = np.random.default_rng([seed] + list(key))
    return rng


def small(seed=7):
    d = SceneDesc()
    rng = np.random.default_rng(seed)
    v, f = icosphere(4, radius=3.0, center=(0.0, 0.0, 0.0), noise="uniform", amp=0.05, seed=seed)
    m = d.add_material("lambertian", rng.uniform(0.2, 0.8, 3))
    d.add_object("mesh", m, mesh=d.add_mesh(v, f))
    floor_v = np.array([[-15, -10, -15], [15, -10, -15], [15, -10, 15], [-15, -10, 15]], np.float32)
    # wound so that the mesh normal -normalize(cross(e1, e2)) points up (triangle_mesh.rs:119)
    floor_f = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    d.add_object("mesh", d.add_material("lambertian", rng.uniform(0.2, 0.8, 3)), mesh=d.add_mesh(floor_v, floor_f))
    for _ in range(64):
        r = rng.uniform(0.2, 1.0)
        c = rng.uniform(-10.0, 10.0, 3)
        mat = d.add_material("lambertian", rng.uniform(0.2, 0.8, 3))
        d.add_object("sphere", mat, param=(r,), transform=Transform.translation(*c))
    d.add_point_light((0.0, 15.0, 10.0), (900.0, 900.0, 900.0))
    camera = dict(eye=(0.0, 6.0, 34.0), look=(0.0, -1.0, 0.0), fov=50.0 * math.pi / 180.0, width=256, height=256,
                  spp=16, max_depth=8)
    return d, camera


def room(subdiv=6, seed=11, ggx=False, big_mesh=None):
    """The room. `big_mesh` replaces the icosphere (vertices, faces); `ggx` gives the mesh and the
    boxes glossy GGX materials of roughness U[0.05, 0.5]."""
    d = SceneDesc()
    rng = np.random.default_rng(seed)
    H = 2.0  # half size of the room: walls at +-2
    white = d.add_material("lambertian", (0.73, 0.73, 0.73))
    red = d.add_material("lambertian", (0.65, 0.05, 0.05))
    green = d.add_material("lambertian", (0.12, 0.45, 0.15))
    # Rectangle: y = 0 plane, normal +y (rectangle.rs); each wall is turned so its normal faces inwards
    wall = (2 * H, 2 * H)
    d.add_object("rectangle", white, wall, transform=Transform.translation(0.0, -H, 0.0))  # floor
    d.add_object("rectangle", white, wall,
                 transform=Transform.translation(0.0, H, 0.0) * Transform.rotation("x", math.pi))  # ceiling
    d.add_object("rectangle", white, wall,
                 transform=Transform.translation(0.0, 0.0, -H) * Transform.rotation("x", math.pi / 2))  # back
    d.add_object("rectangle", red, wall,
                 transform=Transform.translation(-H, 0.0, 0.0) * Transform.rotation("z", -math.pi / 2))  # left
    d.add_object("rectangle", green, wall,
                 transform=Transform.translation(H, 0.0, 0.0) * Transform.rotation("z", math.pi / 2))  # right
    # emissive ceiling panel, 1/16 of the ceiling area, just below it, facing down
    light = d.add_material("diffuse_light", (15.0, 15.0, 15.0))
    panel = d.add_object("rectangle", light, (H / 2, H / 2),
                         transform=Transform.translation(0.0, H - 1e-3, 0.0) * Transform.rotation("x", math.pi))
    d.add_area_light(panel, (15.0, 15.0, 15.0))

    def glossy():
        return d.add_material("ggx", (0.9, 0.85, 0.8), roughness=float(rng.uniform(0.05, 0.5)))

    box_mat = glossy() if ggx else white
    d.add_object("aabox", box_mat, (0.9, 1.8, 0.9),
                 transform=Transform.translation(-0.8, -H + 0.9, -0.7) * Transform.rotation("y", math.radians(17)))
    d.add_object("aabox", glossy() if ggx else white, (0.8, 0.8, 0.8),
                 transform=Transform.translation(0.9, -H + 0.4, 0.6) * Transform.rotation("y", math.radians(-17)))
    if big_mesh is None:
        v, f = icosphere(subdiv, radius=0.55, center=(0.25, -0.4, 0.9), noise="fbm", amp=0.08, octaves=3, seed=seed)
    else:
        v, f = big_mesh
    mesh_mat = glossy() if ggx else d.add_material("dielectric", (1.0, 1.0, 1.0), ior=1.5)
    d.add_object("mesh", mesh_mat, mesh=d.add_mesh(v, f))
    camera = dict(eye=(0.0, 0.0, 5.8), look=(0.0, 0.0, 0.0), fov=40.0 * math.pi / 180.0, width=512, height=512,
                  spp=64, max_depth=8)
    return d, camera


def nrc_room(seed=11, mesh_seed=2024):
    """The radiance-cache workload: the room with its icosphere replaced by the 1,310,720-triangle
    displaced icosphere (8 subdivisions, 5-octave value noise of 10 % amplitude) scaled to the
    icosphere's place, and glossy GGX materials (roughness U[0.05, 0.5]) on the mesh and the boxes.
    1280x720, 1 sample per pixel per frame for 256 frames, max depth 8."""
    v, f = icosphere(8, radius=0.55, center=(0.25, -0.4, 0.9), noise="fbm", amp=0.10, octaves=5, seed=mesh_seed)
    d, cam = room(seed=seed, ggx=True, big_mesh=(v, f))
    cam = dict(cam, width=1280, height=720, spp=1, frames=256, reference_spp=4096)
    return d, cam



def torus(nu, nv, major=1.0, minor=0.35):
    """A torus around the y axis of nu x nv quads (2 nu nv triangles), wound so that the reference's
    mesh normal, -normalize((p1 - p0) x (p2 - p0)) (triangle_mesh.rs:119), points outward."""
    u = np.arange(nu) * (2.0 * math.pi / nu)
    v = np.arange(nv) * (2.0 * math.pi / nv)
    uu, vv = np.meshgrid(u, v, indexing="ij")
    r = major + minor * np.cos(vv)
    verts = np.stack([r * np.cos(uu), minor * np.sin(vv), r * np.sin(uu)], -1).reshape(-1, 3)
    i, j = np.meshgrid(np.arange(nu), np.arange(nv), indexing="ij")
    a = i * nv + j
    b = ((i + 1) % nu) * nv + j
    c = ((i + 1) % nu) * nv + (j + 1) % nv
    d = i * nv + (j + 1) % nv
    faces = np.concatenate([np.stack([a, b, d], -1).reshape(-1, 3), np.stack([b, c, d], -1).reshape(-1, 3)])
    return verts.astype(np.float32), faces.astype(np.uint32)


def batch_scene(n_meshes=256, n_csg=32, seed=5, tri_range=(1000, 200000)):
    """The batched-jobs workload: `n_meshes` mesh instances of seeded size log-uniform in
    `tri_range` triangles (displaced icospheres and tessellated tori), each under a random rigid
    transform, plus `n_csg` CSG objects (union, intersection, subtraction of unit spheres and cubes),
    over a floor, lit by a point light and an emissive panel."""
    rng = np.random.default_rng(seed)
    d = SceneDesc()
    floor = d.add_material("lambertian", (0.6, 0.6, 0.6))
    d.add_object("rectangle", floor, (120.0, 120.0), transform=Transform.translation(0.0, -2.0, 0.0))
    light = d.add_material("diffuse_light", (8.0, 8.0, 8.0))
    panel = d.add_object("rectangle", light, (20.0, 20.0),
                         transform=Transform.translation(0.0, 30.0, 0.0) * Transform.rotation("x", math.pi))
    d.add_area_light(panel, (8.0, 8.0, 8.0))
    d.add_point_light((10.0, 25.0, 20.0), (3000.0, 3000.0, 3000.0))
    side = int(math.ceil(math.sqrt(n_meshes + n_csg)))
    spots = [(x, z) for x in range(side) for z in range(side)]
    order = rng.permutation(len(spots))
    lo, hi = math.log(tri_range[0]), math.log(tri_range[1])
    for k in range(n_meshes):
        target = math.exp(rng.uniform(lo, hi))
        if rng.uniform() < 0.5:
            s = int(min(6, max(3, round(math.log(target / 20.0, 4)))))
            v, f = icosphere(s, radius=1.0, noise="fbm", amp=0.08, octaves=3, seed=int(rng.integers(1 << 30)))
        else:
            nu = max(8, int(round(math.sqrt(target))))
            nv = max(4, int(round(target / (2 * nu))))
            v, f = torus(nu, nv)
        mat = d.add_material("ggx" if rng.uniform() < 0.3 else "lambertian", rng.uniform(0.2, 0.8, 3),
                             roughness=float(rng.uniform(0.05, 0.5)))
        gx, gz = spots[order[k]]
        xf = (Transform.translation((gx - side / 2) * 3.0, rng.uniform(-1.0, 1.0), (gz - side / 2) * 3.0)
              * Transform.rotation("y", float(rng.uniform(0, 2 * math.pi)))
              * Transform.rotation("x", float(rng.uniform(-0.5, 0.5))))
        d.add_object("mesh", mat, mesh=d.add_mesh(v, f), transform=xf)
    ops = ("union", "intersection", "subtraction")
    for k in range(n_csg):
        op = ops[k % 3]
        els = []
        for _ in range(2 + int(rng.integers(0, 3))):
            c = rng.uniform(-0.5, 0.5, 3) * (0.5 if op == "intersection" else 1.0)
            xf = Transform.translation(*c) * Transform.rotation("y", float(rng.uniform(0, 2 * math.pi)))
            els.append(("sphere", (1.0,), xf) if rng.uniform() < 0.5 else ("aabox", (1.5, 1.5, 1.5), xf))
        gx, gz = spots[order[n_meshes + k]]
        d.add_csg(op, els, d.add_material("lambertian", rng.uniform(0.2, 0.8, 3)),
                  transform=Transform.translation((gx - side / 2) * 3.0, 0.0, (gz - side / 2) * 3.0))
    return d


def batch_jobs(seed=6, n=8):
    """Eight render jobs of the batched-jobs workload: resolutions drawn from 256^2, 512^2, 768x432
    and 1024^2, 16 or 64 samples per pixel, max depth 8, cameras around the scene."""
    rng = np.random.default_rng(seed)
    sizes = [(256, 256), (512, 512), (768, 432), (1024, 1024)]
    jobs = []
    for k in range(n):
        w, h = sizes[int(rng.integers(len(sizes)))]
        a = rng.uniform(0, 2 * math.pi)
        eye = (40.0 * math.cos(a), float(rng.uniform(8.0, 20.0)), 40.0 * math.sin(a))
        jobs.append(dict(width=w, height=h, spp=int(rng.choice([16, 64])), max_depth=8, eye=eye,
                         look=(0.0, 0.0, 0.0), fov=math.radians(50.0), seed=100 + k))
    return jobs


def small_jobs(seed=7, n=16):
    """Sixteen small render jobs on the batched-jobs scene (previews, tiles, thumbnails): resolutions
    drawn from 64^2, 96x64, 128^2 and 256x144, 1, 2 or 4 samples per pixel, max depth 8, cameras
    around the scene. None of them fills the GPU on its own."""
    rng = np.random.default_rng(seed)
    sizes = [(64, 64), (96, 64), (128, 128), (256, 144)]
    jobs = []
    for k in range(n):
        w, h = sizes[int(rng.integers(len(sizes)))]
        a = rng.uniform(0, 2 * math.pi)
        eye = (40.0 * math.cos(a), float(rng.uniform(8.0, 20.0)), 40.0 * math.sin(a))
        jobs.append(dict(width=w, height=h, spp=int(rng.choice([1, 2, 4])), max_depth=8, eye=eye,
                         look=(0.0, 0.0, 0.0), fov=math.radians(50.0), seed=200 + k))
    return jobs
