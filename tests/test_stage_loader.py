"""The `.stage` and PLY loaders of the native library, on the host (no GPU).

PLY (load_ply_mesh, src/loader/mesh_loader.rs; ASCII only, src/loader/ply/ply_loader.rs:251-255):
the reference's own meshes (tests/golden/test_files, copied from its test_files/) read by
pbrt_ply_read against the FP32 fixtures written by tests/golden/make_ply_fixtures.py (strtof per
coordinate, the reference's `parse::<f32>`), bit for bit; `reverse` reverses each face.

`.stage` (Loader::load_scene, src/loader.rs:29-46): every committed scene parses and its meshes
load (the call then stops at the missing device context, before any device work); syntax errors and
the shapes and cameras with no device counterpart are refused with the file line in the message.
The scenes themselves are checked on the device against an independent parse in
tests/test_gpu_stage.py.
"""
import ctypes
import os

import numpy as np
import pytest

import pbrt_amd
from conftest import ROOT, load_mesh
from pbrt_amd import _native

FILES = os.path.join(ROOT, "tests", "golden", "test_files")
STAGES = ["default", "cube_mesh", "cornell_box", "cornell_box_canonical", "csg_union_spheres", "many_spheres",
          "many_meshes"]


def load_stage_without_device(path, base_dir=os.path.join(ROOT, "tests", "golden")):
    scene = ctypes.c_void_p()
    cfg = _native.RenderConfig()
    _native.call("pbrt_scene_load_stage", None, os.fsencode(str(path)), os.fsencode(str(base_dir)), None,
                 ctypes.byref(scene), ctypes.byref(cfg))


@pytest.mark.parametrize("name", ["cube", "bun_zipper_res4"])
def test_ply_matches_the_fp32_fixture(name):
    v, f = pbrt_amd.read_ply(os.path.join(FILES, name + ".ply"))
    rv, rf = load_mesh(name)
    assert v.dtype == np.float32 and f.dtype == np.uint32
    np.testing.assert_array_equal(v, rv)
    np.testing.assert_array_equal(f, rf)


def test_ply_reverse_flips_every_face():
    v, f = pbrt_amd.read_ply(os.path.join(FILES, "bun_zipper_res4.ply"))
    v2, f2 = pbrt_amd.read_ply(os.path.join(FILES, "bun_zipper_res4.ply"), reverse=True)
    np.testing.assert_array_equal(v2, v)
    np.testing.assert_array_equal(f2, f[:, ::-1])


def test_ply_with_only_positions():
    v, f = pbrt_amd.read_ply(os.path.join(FILES, "cube_2.ply"))
    corners = np.array([[x, y, z] for z in (-0.5, 0.5) for y in (-0.5, 0.5) for x in (-0.5, 0.5)], np.float32)
    assert v.shape == (8, 3) and f.shape == (12, 3)
    assert {tuple(p) for p in v} == {tuple(p) for p in corners}
    assert list(f[0]) == [0, 1, 2] and list(f[-1]) == [3, 6, 7]


HEADER = "ply\nformat {fmt} 1.0\nelement vertex {nv}\nproperty float x\nproperty float y\nproperty float z\n" \
         "element face {nf}\nproperty list uchar int vertex_indices\nend_header\n"


@pytest.mark.parametrize("body,fmt,what", [
    ("0 0 0\n1 0 0\n0 1 0\n3 0 1 2\n", "binary_little_endian", "ASCII"),
    ("0 0 0\n1 0 0\n0 1 0\n4 0 1 2 0\n", "ascii", "triangle"),
    ("0 0 0\n1 0 0\n0 1 0\n3 0 1 3\n", "ascii", "out of range"),
    ("0 0 0\n1 0 0\n", "ascii", "truncated"),
])
def test_ply_errors(tmp_path, body, fmt, what):
    p = tmp_path / "m.ply"
    p.write_text(HEADER.format(fmt=fmt, nv=3, nf=1) + body)
    with pytest.raises(pbrt_amd.PbrtError, match=what):
        pbrt_amd.read_ply(p)
    with pytest.raises(pbrt_amd.PbrtError, match="cannot open"):
        pbrt_amd.read_ply(tmp_path / "missing.ply")


@pytest.mark.parametrize("name", STAGES)
def test_stage_files_parse_and_their_meshes_load(name):
    with pytest.raises(pbrt_amd.PbrtError, match="null context"):
        load_stage_without_device(os.path.join(FILES, name + ".stage"))


def test_stage_mesh_files_resolve_against_base_dir(tmp_path):
    # bunny_mesh.stage names test_files/bunny.ply, which is not among the fixtures
    with pytest.raises(pbrt_amd.PbrtError, match="cannot open PLY file .*test_files/bunny.ply"):
        load_stage_without_device(os.path.join(FILES, "bunny_mesh.stage"))


GOOD = """camera pin_hole
  pos 0 0 2. look 0 0 0 up 0 1 0
scene
  # a comment { with braces }
  object simple sphere 0.5 lambertian color 0.5 0.5 0.5
"""


@pytest.mark.parametrize("text,what", [
    (GOOD, "null context"),
    (GOOD.replace("pin_hole", "thin_lens"), "line 1: the thin-lens camera"),
    (GOOD.replace("sphere 0.5", "cylinder 0.5 1"), "line 5: the cylinder shape is not supported"),
    (GOOD.replace("sphere 0.5", "plane"), "line 5: the plane shape is not supported"),
    (GOOD.replace("sphere 0.5", "torus 0.5"), "line 5: unknown shape"),
    (GOOD.replace("lambertian", "glossy"), "line 5: lambertian, dielectric"),
    (GOOD.replace("color", "colour"), "line 5: color or checkerboard expected"),
    (GOOD.replace("0.5 0.5 0.5", "0.5 .5 0.5"), "line 5: unexpected character"),
    (GOOD.replace("pos 0 0 2.", "pos 0 - 2."), "line 2: malformed number"),
    (GOOD + "  object transformed object simple sphere 1 metal 0.1 color 1 1 1 transform { scale 2 }\n",
     "line 6: translate or rotate_x/y/z expected"),
    (GOOD + "  object simple mesh file \"test_files/cube.ply\n", "unterminated string"),
    (GOOD + "  object simple csg union { elem rectangle 1 1 transform { } } lambertian color 1 1 1\n",
     "line 6: CSG elements are spheres and boxes"),
    (GOOD + "  object compound { object simple sphere 1 lambertian color 1 1 1\n", "'}' or object expected"),
    (GOOD + "  light point color 15 15 15 transform { translate 0 2 1 }\n"
     "  light background_infinite color 1 1 1 color 0.5 0.7 1\n  light uniform_infinite color 0.5 0.5 0.5\n",
     "null context"),
    (GOOD + "  light spot color 1 1 1\n", "line 6: point, uniform_infinite or background_infinite expected"),
    (GOOD + "  light area color 1 1 1\n", "line 6: an area light is a diffuse_light material"),
    (GOOD + "  light background_infinite color 1 1 1\n", "line 7: 'color' expected"),
    (GOOD + "  light point 1 1 1\n", "line 6: 'color' expected"),
    (GOOD + "  light point color 1 1 1 transform { scale 2 }\n", "line 6: translate or rotate_x/y/z expected"),
    (GOOD + "  camera pin_hole\n", "line 6: object or light expected"),
])
def test_stage_syntax(tmp_path, text, what):
    p = tmp_path / "s.stage"
    p.write_text(text)
    with pytest.raises(pbrt_amd.PbrtError, match=what):
        load_stage_without_device(p, tmp_path)
