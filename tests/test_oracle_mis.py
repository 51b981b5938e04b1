"""Multiple importance sampling in the oracle's path integrator (the reference's plan after the
area light, ideas/area_light.md:158-165; an option, off by default).

Pins: MIS and the reference's estimator are both unbiased, so under the same random numbers they
give the same image mean on the room lit by its ceiling panel; a light density off by a factor 2
shifts that mean by 0.35 %, against 0.001 % here, so the bound catches a wrong density or weight.
And MIS does what it is for: on a glossy floor (GGX roughness 0.05) under a large panel, next-event
estimation alone is noisy and the BSDF sample with MIS is not.
"""
import math

import numpy as np

from oracle import pyoracle as po
from pbrt_amd import scenes
from pbrt_amd.scene import SceneDesc, Transform


def glossy_floor():
    d = SceneDesc()
    floor = d.add_material("ggx", (0.9, 0.9, 0.9), roughness=0.05)
    light = d.add_material("diffuse_light", (4.0, 4.0, 4.0))
    d.add_object("rectangle", floor, (6.0, 6.0))
    panel = d.add_object("rectangle", light, (2.0, 2.0),
                         transform=Transform.translation(0.0, 2.0, -1.0) * Transform.rotation("x", math.pi))
    d.add_area_light(panel, (4.0, 4.0, 4.0))
    return d, dict(eye=(0.0, 1.0, 3.0), look=(0.0, 0.0, -1.0), fov=1.0)


def test_mis_and_the_reference_estimator_agree_on_the_room():
    desc, cam = scenes.room(subdiv=3, ggx=True)
    sc = po.Scene(desc, precision=64)
    args = (32, 32, cam["fov"], cam["eye"], cam["look"], 256, 8, 5)
    plain, _ = sc.render(*args, mis=False)
    mis, _ = sc.render(*args, mis=True)
    rel = (mis.mean(axis=(0, 1)) - plain.mean(axis=(0, 1))) / plain.mean(axis=(0, 1))
    assert np.all(np.abs(rel) < 5e-4), rel
    assert np.isfinite(mis).all()
    assert not np.array_equal(mis, plain)  # the estimator did change


def test_mis_lowers_the_error_on_a_glossy_floor():
    desc, cam = glossy_floor()
    sc = po.Scene(desc, precision=32)
    ref, _ = sc.render(32, 32, cam["fov"], cam["eye"], cam["look"], 2048, 4, 99, mis=True)
    err = {}
    for mis in (False, True):
        e = [np.sqrt(np.mean((sc.render(32, 32, cam["fov"], cam["eye"], cam["look"], 16, 4, s, mis=mis)[0] - ref) ** 2))
             for s in range(3)]
        err[mis] = float(np.mean(e))
    assert err[True] < 0.5 * err[False], err
