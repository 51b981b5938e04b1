"""Scene description for the device scene and renderer.

Mirrors how the reference assembles a scene (src/scene.rs:80-115): shapes wrapped with a material in
a `Simple` object (src/objects/simple.rs), optionally under a `Transformed` wrapper
(src/objects/transformed.rs), and lights added beside them. A `SceneDesc` records that construction
as plain arrays; `pbrt_amd.Scene` uploads it to the device through the C ABI.

`Transform` restates src/geom/transform.rs: matrices and inverses are composed alongside, the
inverse of a rotation is its transpose and a translation inverts exactly, and a `.stage` transform
block composes its steps as translation . rotation (src/loader/visitors/scene_builder.rs:195-207).
All arithmetic in Python floats (IEEE f64, no fused operations), as the reference's f64.
"""
import math

import numpy as np

SHAPE = {"mesh": 0, "sphere": 1, "aabox": 2, "box": 2, "rectangle": 3, "rect": 3, "csg": 4}
CSG_OP = {"union": 0, "intersection": 1, "subtraction": 2, "substraction": 2}
MATERIAL = {"lambertian": 0, "dielectric": 1, "diffuse_light": 2, "ggx": 3, "metal": 4}
LIGHT = {"point": 0, "area": 1, "background": 2}
TEXTURE = {"color": 0, "checkerboard": 1}

OBJECT_DTYPE = np.dtype([("shape", "<i4"), ("mesh", "<i4"), ("material", "<i4"), ("has_xform", "<i4"),
                         ("param", "<f8", 3), ("m", "<f8", 16), ("minv", "<f8", 16)])
MATERIAL_DTYPE = np.dtype([("kind", "<i4"), ("tex", "<i4"), ("rgb", "<f8", 3), ("ior", "<f8"), ("roughness", "<f8"),
                           ("rgb2", "<f8", 3), ("scale", "<f8")])
LIGHT_DTYPE = np.dtype([("kind", "<i4"), ("object", "<i4"), ("pos", "<f8", 3), ("rgb", "<f8", 3)])
CSG_ELEM_DTYPE = np.dtype([("shape", "<i4"), ("pad", "<i4"), ("param", "<f8", 3), ("m", "<f8", 16),
                           ("minv", "<f8", 16)])


def _mat_mul(a, b):
    """src/geom/matrix4.rs:216-232: each entry summed from 0.0 over k = 0..3."""
    r = [[0.0] * 4 for _ in range(4)]
    for i in range(4):
        for j in range(4):
            s = 0.0
            for k in range(4):
                s += a[i][k] * b[k][j]
            r[i][j] = s
    return r


def _identity():
    return [[1.0, 0.0, 0.0, 0.0], [0.0, 1.0, 0.0, 0.0], [0.0, 0.0, 1.0, 0.0], [0.0, 0.0, 0.0, 1.0]]


def _transpose(m):
    return [[m[j][i] for j in range(4)] for i in range(4)]


class Transform:
    """A matrix and its inverse (src/geom/transform.rs:7-10)."""

    def __init__(self, m=None, minv=None):
        self.m = m if m is not None else _identity()
        self.minv = minv if minv is not None else _identity()

    @staticmethod
    def translation(x, y, z):  # transform.rs:37-42
        t = _identity()
        t[0][3], t[1][3], t[2][3] = x, y, z
        ti = _identity()
        ti[0][3], ti[1][3], ti[2][3] = -x, -y, -z
        return Transform(t, ti)

    @staticmethod
    def rotation(axis, theta):  # transform.rs:44-60, matrix4.rs:46-68
        s, c = math.sin(theta), math.cos(theta)
        m = _identity()
        if axis == "x":
            m[1][1], m[1][2], m[2][1], m[2][2] = c, -s, s, c
        elif axis == "y":
            m[0][0], m[0][2], m[2][0], m[2][2] = c, s, -s, c
        else:
            m[0][0], m[0][1], m[1][0], m[1][1] = c, -s, s, c
        return Transform(m, _transpose(m))

    @staticmethod
    def scale(x, y, z):
        m = _identity()
        m[0][0], m[1][1], m[2][2] = x, y, z
        mi = _identity()
        mi[0][0], mi[1][1], mi[2][2] = 1.0 / x, 1.0 / y, 1.0 / z
        return Transform(m, mi)

    def __mul__(self, o):  # transform.rs:119-127
        return Transform(_mat_mul(self.m, o.m), _mat_mul(o.minv, self.minv))

    @staticmethod
    def from_steps(steps):
        """A `.stage` transform block: steps in listed order, composed as scene_builder.rs:195-207
        does (the identity times the steps taken from the last to the first)."""
        ts = []
        for st in steps:
            if st[0] == "translate":
                ts.append(Transform.translation(*st[1:4]))
            else:
                ts.append(Transform.rotation(st[0][-1], st[1]))
        acc = Transform()
        for t in reversed(ts):
            acc = acc * t
        return acc

    def flat(self):
        return [x for row in self.m for x in row], [x for row in self.minv for x in row]


class SceneDesc:
    """Objects, materials, meshes and lights, in the order they are added."""

    def __init__(self):
        self.meshes = []      # (vertices float32 (nv,3), faces uint32 (nt,3))
        self.materials = []
        self.objects = []
        self.lights = []
        self.csg = []         # (shape, param, transform) of every CSG element, objects' elements contiguous

    def add_mesh(self, vertices, faces):
        self.meshes.append((np.ascontiguousarray(vertices, np.float32).reshape(-1, 3),
                            np.ascontiguousarray(faces, np.uint32).reshape(-1, 3)))
        return len(self.meshes) - 1

    def add_material(self, kind, rgb, ior=1.5, roughness=0.0, fuzz=None, checker=None):
        """`rgb`: the colour (texture `color`), or the first colour of a checkerboard given as
        checker=(rgb2, scale) (textures/checker_board.rs); `fuzz`: Metal's (materials/metal.rs),
        stored in the roughness slot."""
        if fuzz is not None:
            roughness = fuzz
        tex, rgb2, scale = (0, (0.0, 0.0, 0.0), 0.0) if checker is None else (1, tuple(map(float, checker[0])),
                                                                                float(checker[1]))
        self.materials.append((MATERIAL[kind], tuple(float(c) for c in rgb), float(ior), float(roughness), tex, rgb2,
                               scale))
        return len(self.materials) - 1

    def add_object(self, shape, material, param=(0.0, 0.0, 0.0), mesh=-1, transform=None):
        p = list(param) + [0.0] * (3 - len(param))
        self.objects.append((SHAPE[shape], int(mesh), int(material), transform, tuple(float(x) for x in p)))
        return len(self.objects) - 1

    def add_csg(self, op, elements, material, transform=None):
        """A CSG object (shapes/csg/*.rs): `op` is "union", "intersection" or "subtraction";
        `elements` lists (shape, param, transform) with shape "sphere" or "aabox", in the order the
        operation's constructor receives them (for a subtraction the first is the base)."""
        first = len(self.csg)
        for shape, param, xf in elements:
            if SHAPE[shape] not in (SHAPE["sphere"], SHAPE["aabox"]):
                raise ValueError("CSG elements are spheres and boxes")
            p = list(param) + [0.0] * (3 - len(param))
            self.csg.append((SHAPE[shape], tuple(float(x) for x in p), xf))
        self.objects.append((SHAPE["csg"], first, int(material), transform,
                             (float(CSG_OP[op]), float(len(elements)), 0.0)))
        return len(self.objects) - 1

    def add_point_light(self, pos, rgb):
        self.lights.append((LIGHT["point"], -1, tuple(map(float, pos)), tuple(map(float, rgb))))

    def add_area_light(self, obj, rgb):
        """An emissive rectangle object that also illuminates (ideas/area_light.md)."""
        assert self.objects[obj][0] == SHAPE["rectangle"]
        self.lights.append((LIGHT["area"], int(obj), (0.0, 0.0, 0.0), tuple(map(float, rgb))))

    def add_background_light(self, down=(1.0, 1.0, 1.0), up=(0.5, 0.7, 1.0)):
        """BackgroundInfiniteLight (lights/background_infinite_light.rs): radiance blended from
        `down` (direction y = -1) to `up` (y = +1)."""
        self.lights.append((LIGHT["background"], -1, tuple(map(float, down)), tuple(map(float, up))))

    # -- packed arrays (the layout the native code reads) --------------------------------------
    def object_array(self):
        a = np.zeros(len(self.objects), OBJECT_DTYPE)
        for i, (shape, mesh, mat, xf, param) in enumerate(self.objects):
            a[i]["shape"], a[i]["mesh"], a[i]["material"] = shape, mesh, mat
            a[i]["param"] = param
            if xf is not None:
                m, mi = xf.flat()
                a[i]["has_xform"] = 1
                a[i]["m"], a[i]["minv"] = m, mi
            else:
                a[i]["m"] = np.eye(4).ravel()
                a[i]["minv"] = np.eye(4).ravel()
        return a

    def csg_array(self):
        a = np.zeros(max(len(self.csg), 1), CSG_ELEM_DTYPE)
        for i, (shape, param, xf) in enumerate(self.csg):
            a[i]["shape"], a[i]["param"] = shape, param
            m, mi = xf.flat() if xf is not None else (np.eye(4).ravel(), np.eye(4).ravel())
            a[i]["m"], a[i]["minv"] = m, mi
        return a

    def material_array(self):
        a = np.zeros(len(self.materials), MATERIAL_DTYPE)
        for i, (kind, rgb, ior, rough, tex, rgb2, scale) in enumerate(self.materials):
            a[i]["kind"], a[i]["rgb"], a[i]["ior"], a[i]["roughness"] = kind, rgb, ior, rough
            a[i]["tex"], a[i]["rgb2"], a[i]["scale"] = tex, rgb2, scale
        return a

    def light_array(self):
        a = np.zeros(len(self.lights), LIGHT_DTYPE)
        for i, (kind, obj, pos, rgb) in enumerate(self.lights):
            a[i]["kind"], a[i]["object"], a[i]["pos"], a[i]["rgb"] = kind, obj, pos, rgb
        return a

    def mesh_arrays(self):
        """Concatenated vertices / faces with per-mesh offsets (in vertices and in triangles)."""
        voff = np.zeros(len(self.meshes) + 1, np.uint64)
        foff = np.zeros(len(self.meshes) + 1, np.uint64)
        for i, (v, f) in enumerate(self.meshes):
            voff[i + 1] = voff[i] + len(v)
            foff[i + 1] = foff[i] + len(f)
        verts = np.concatenate([v for v, _ in self.meshes]) if self.meshes else np.zeros((0, 3), np.float32)
        faces = np.concatenate([f for _, f in self.meshes]) if self.meshes else np.zeros((0, 3), np.uint32)
        return np.ascontiguousarray(verts, np.float32), voff, np.ascontiguousarray(faces, np.uint32), foff


def from_stage_json(scene, meshes):
    """A scene from a parsed `.stage` description (camera + objects in file order). The reference's
    builder adds objects to the scene in reverse file order (scene_builder.rs:53-57), and every
    `.stage` scene gets a point light at (0, 2, 1) of intensity 15 (src/loader.rs:36-42).
    `meshes` maps a PLY file name to (vertices, faces)."""
    d = SceneDesc()
    mesh_ids = {}
    for o in reversed(scene["objects"]):
        m = o["material"]
        mat = d.add_material(m["kind"] if m["kind"] in MATERIAL else "lambertian", m["color"], ior=m.get("ior", 1.5))
        xf = Transform.from_steps(o["transform"]) if o.get("transform") else None
        if o["shape"] == "csg":
            # the builder pops the elements off a stack: the constructor receives them reversed
            # (scene_builder.rs:117-145)
            els = [(e["shape"], e["param"], Transform.from_steps(e["transform"]) if e.get("transform") else None)
                   for e in reversed(o["elems"])]
            d.add_csg(o["op"], els, mat, transform=xf)
        elif o["shape"] == "mesh":
            key = (o["file"], o.get("reverse", False))
            if key not in mesh_ids:
                v, f = meshes[o["file"]]
                if o.get("reverse"):
                    f = f[:, ::-1]
                mesh_ids[key] = d.add_mesh(v, f)
            d.add_object("mesh", mat, mesh=mesh_ids[key], transform=xf)
        else:
            d.add_object(o["shape"], mat, param=o["param"], transform=xf)
    d.add_point_light((0.0, 2.0, 1.0), (15.0, 15.0, 15.0))
    return d
