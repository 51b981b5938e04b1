"""`.stage` scenes loaded by the native loader (pbrt_scene_load_stage, Loader::load_scene,
src/loader.rs:29-46) against the same files parsed independently in Python
(tests/golden/make_stage_fixtures.py's parser) and built through SceneDesc in the reference
builder's order (src/loader/visitors/scene_builder.rs: objects reversed, CSG elements reversed,
transform steps applied in listed order, the point light and the background light last).

Both routes end in the same scene-construction calls, so camera-ray hits (object, triangle, t, u, v)
and rendered films must be equal bit for bit; default.stage's render is also held against the FP32
oracle. Meshes come from the reference's PLY files (tests/golden/test_files) for the loader and
from the FP32 fixtures (tests/golden/meshes) for the Python route; bunny.ply, not committed as PLY,
is written from its fixture with 9 significant digits, which strtof reads back exactly.
"""
import math
import os

import numpy as np
import pytest

import pbrt_amd
from conftest import ROOT, load_mesh
from golden.make_stage_fixtures import Parser, tokens
from oracle import pyoracle as po
from pbrt_amd.scene import SceneDesc, Transform

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(ROOT, "tests", "golden")
STAGES = ["default", "cube_mesh", "cornell_box", "cornell_box_canonical", "csg_union_spheres", "many_spheres",
          "many_meshes", "bunny_mesh"]


def steps(xf):
    return Transform.from_steps(xf) if xf else None


def described(name):
    """(SceneDesc, camera) of a `.stage` file through the Python parser."""
    with open(os.path.join(GOLDEN, "test_files", name + ".stage")) as f:
        sc = Parser(tokens(f.read())).stage()
    d = SceneDesc()
    meshes = {}
    for o in reversed(sc["objects"]):
        assert o.get("nesting", 0) <= 1
        m = o["material"]
        if "checker" in m:
            c = m["checker"]
            mat = d.add_material(m["kind"], c[0:3], ior=m.get("ior", 1.5), fuzz=m.get("fuzz"), checker=(c[3:6], c[6]))
        else:
            mat = d.add_material(m["kind"], m["color"], ior=m.get("ior", 1.5), fuzz=m.get("fuzz"))
        xf = steps(o.get("transform"))
        if o["shape"] == "csg":
            els = [(e["shape"], e["param"], steps(e["transform"]) or Transform()) for e in reversed(o["elems"])]
            d.add_csg(o["op"], els, mat, transform=xf)
        elif o["shape"] == "mesh":
            key = (o["file"], o["reverse"])
            if key not in meshes:
                v, f = load_mesh(o["file"][:-4])
                meshes[key] = d.add_mesh(v, f[:, ::-1] if o["reverse"] else f)
            d.add_object("mesh", mat, mesh=meshes[key], transform=xf)
        else:
            d.add_object(o["shape"], mat, param=o["param"], transform=xf)
    d.add_point_light((0.0, 2.0, 1.0), (15.0, 15.0, 15.0))
    d.add_background_light((1.0, 1.0, 1.0), (0.5, 0.7, 1.0))
    return d, sc["camera"]


def base_dir(name, tmp_path):
    if name != "bunny_mesh":
        return GOLDEN
    v, f = load_mesh("bunny")
    os.makedirs(tmp_path / "test_files", exist_ok=True)
    with open(tmp_path / "test_files" / "bunny.ply", "w") as out:
        out.write(f"ply\nformat ascii 1.0\nelement vertex {len(v)}\nproperty float x\nproperty float y\n"
                  f"property float z\nelement face {len(f)}\nproperty list uchar int vertex_indices\nend_header\n")
        np.savetxt(out, v.astype(np.float64), fmt="%.9g")
        np.savetxt(out, np.concatenate([np.full((len(f), 1), 3, np.int64), f.astype(np.int64)], 1), fmt="%d")
    return str(tmp_path)


@pytest.mark.parametrize("name", STAGES)
def test_loaded_scene_is_the_described_scene(gpu_ctx, tmp_path, name):
    print(f"stage {name}", flush=True)
    loaded, view = pbrt_amd.Scene.load_stage(gpu_ctx, os.path.join(GOLDEN, "test_files", name + ".stage"),
                                             base_dir(name, tmp_path))
    desc, cam = described(name)
    assert view["eye"] == tuple(cam["pos"]) and view["look"] == tuple(cam["look"]) and view["up"] == tuple(cam["up"])
    assert view["fov"] == 90.0 * math.pi / 180.0 and view["near"] == 1e-4 and view["far"] == 1000.0
    built = pbrt_amd.Scene(gpu_ctx, desc)
    assert loaded.info() == built.info()

    rays = po.pinhole_rays(200, 150, view["fov"], cam["pos"], cam["look"], cam["up"]).astype(np.float32)
    a, b = loaded.trace_host(rays), built.trace_host(rays)
    assert loaded.status() == 0 and (a["object"] >= 0).any()
    np.testing.assert_array_equal(a["object"], b["object"])
    hit = a["object"] >= 0  # a miss defines only its object (-1)
    for field in ("prim", "t", "u", "v"):
        np.testing.assert_array_equal(a[field][hit], b[field][hit], err_msg=field)

    kw = dict(fov=view["fov"], up=view["up"], seed=5, near=view["near"], far=view["far"])
    fa, _ = loaded.render(48, 36, 4, 5, view["eye"], view["look"], **kw)
    fb, _ = built.render(48, 36, 4, 5, view["eye"], view["look"], **kw)
    assert np.isfinite(fa).all() and fa.max() > 0
    np.testing.assert_array_equal(fa, fb)


def test_default_stage_render_matches_oracle(gpu_ctx):
    scene, view = pbrt_amd.Scene.load_stage(gpu_ctx, os.path.join(GOLDEN, "test_files", "default.stage"), GOLDEN)
    desc, _ = described("default")
    w = h = 64
    spp, depth = 8, 8
    film, _ = scene.render(w, h, spp, depth, view["eye"], view["look"], fov=view["fov"], seed=3)
    f32, _ = po.Scene(desc, precision=32, **pbrt_amd.DEVICE_BVH).render(w, h, view["fov"], view["eye"], view["look"],
                                                                      spp, depth, 3)
    scale = float(f32.mean())
    rmse = float(np.sqrt(np.mean((film - f32) ** 2)))
    assert scale > 0 and rmse / scale < 0.02, rmse / scale


LIT = """camera pin_hole pos 0 1 3.5 look 0 0.8 0 up 0 1 0
scene
  object simple rectangle 6 6 lambertian color 0.7 0.7 0.7
  object transformed object simple rectangle 1 1 diffuse_light color 8 8 8
    transform { rotate_x -3.14159 translate 0 2 0 }
  object transformed object simple sphere 0.5 metal 0.2 color 0.9 0.8 0.6 transform { translate -0.6 0.5 0 }
  object transformed object simple aabox 0.6 0.6 0.6 lambertian color 0.2 0.4 0.8
    transform { rotate_y 0.4 translate 0.7 0.3 0.2 }
  light point color 4 4 4 transform { rotate_z 0.3 translate 1 1.5 1.5 }
  light background_infinite color 0.1 0.1 0.1 color 0.2 0.3 0.5
  light uniform_infinite color 0.05 0.05 0.05
"""


def lit_described():
    """LIT built through SceneDesc in the loader's order: objects reversed, the emissive rectangle
    registered as an area light as it is added, then the declared lights in file order."""
    xf = Transform.from_steps
    d = SceneDesc()
    m = d.add_material("lambertian", (0.2, 0.4, 0.8))
    d.add_object("aabox", m, param=(0.6, 0.6, 0.6), transform=xf([["rotate_y", 0.4], ["translate", 0.7, 0.3, 0.2]]))
    m = d.add_material("metal", (0.9, 0.8, 0.6), fuzz=0.2)
    d.add_object("sphere", m, param=(0.5,), transform=xf([["translate", -0.6, 0.5, 0.0]]))
    m = d.add_material("diffuse_light", (8.0, 8.0, 8.0))
    panel = d.add_object("rectangle", m, param=(1.0, 1.0),
                         transform=xf([["rotate_x", -3.14159], ["translate", 0.0, 2.0, 0.0]]))
    d.add_area_light(panel, (8.0, 8.0, 8.0))
    m = d.add_material("lambertian", (0.7, 0.7, 0.7))
    d.add_object("rectangle", m, param=(6.0, 6.0))
    p, _ = xf([["rotate_z", 0.3], ["translate", 1.0, 1.5, 1.5]]).flat()
    d.add_point_light((p[3], p[7], p[11]), (4.0, 4.0, 4.0))
    d.add_background_light((0.1, 0.1, 0.1), (0.2, 0.3, 0.5))
    d.add_background_light((0.05, 0.05, 0.05), (0.05, 0.05, 0.05))
    return d


def test_declared_lights_replace_the_loaders_lights(gpu_ctx, tmp_path):
    """A `.stage` file with `light` productions: the loaded scene equals lit_described() (scene
    info, renders bit for bit with and without MIS) and its render matches the FP32 oracle; the
    same file without its light lines gets the loader's point and background lights and no area
    light, so its film differs."""
    print("stage with declared lights", flush=True)
    path = tmp_path / "lit.stage"
    path.write_text(LIT)
    loaded, view = pbrt_amd.Scene.load_stage(gpu_ctx, str(path), str(tmp_path))
    desc = lit_described()
    built = pbrt_amd.Scene(gpu_ctx, desc)
    assert loaded.info() == built.info()
    kw = dict(fov=view["fov"], up=view["up"], seed=5)
    for mis in (False, True):
        fa, _ = loaded.render(48, 36, 8, 6, view["eye"], view["look"], mis=mis, **kw)
        fb, _ = built.render(48, 36, 8, 6, view["eye"], view["look"], mis=mis, **kw)
        assert np.isfinite(fa).all() and fa.max() > 0
        np.testing.assert_array_equal(fa, fb)
    w = h = 64
    film, _ = loaded.render(w, h, 8, 6, view["eye"], view["look"], fov=view["fov"], seed=3)
    f32, _ = po.Scene(desc, precision=32, **pbrt_amd.DEVICE_BVH).render(w, h, view["fov"], view["eye"], view["look"],
                                                                      8, 6, 3)
    scale = float(f32.mean())
    assert scale > 0 and float(np.sqrt(np.mean((film - f32) ** 2))) / scale < 0.02
    unlit = tmp_path / "unlit.stage"
    unlit.write_text("".join(l + "\n" for l in LIT.splitlines() if not l.lstrip().startswith("light")))
    plain, _ = pbrt_amd.Scene.load_stage(gpu_ctx, str(unlit), str(tmp_path))
    fp, _ = plain.render(48, 36, 8, 6, view["eye"], view["look"], **kw)
    fl, _ = loaded.render(48, 36, 8, 6, view["eye"], view["look"], **kw)
    assert not np.array_equal(fp, fl)
