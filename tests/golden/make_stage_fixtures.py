"""Writes the reference's measured `.stage` scenes as JSON fixtures (tests/golden/stages/*.json).

The four scenes of the reference's scene-mode known answers (docs/mesures_bvh.md:729-738):
cornell_box_canonical, bunny_mesh, many_spheres, many_meshes; and the reference's CSG scene,
csg_union_spheres (a union of nine spheres over a floor rectangle). Only the data the geometry needs is
kept: the camera, and per object its shape, its parameters, its material kind and colour, and its
transform steps in the order listed. Objects are stored in file order; the reference's scene builder
adds them to the scene in reverse (src/loader/visitors/scene_builder.rs:53-57), which the test
harness reproduces.

Run once, where the reference is mounted:
    python tests/golden/make_stage_fixtures.py /root/reference/test_files
"""
import json
import os
import sys

SCENES = ["cornell_box_canonical", "bunny_mesh", "many_spheres", "many_meshes", "csg_union_spheres"]


def tokens(text):
    out = []
    for line in text.splitlines():
        line = line.split("#", 1)[0]
        out += line.replace("{", " { ").replace("}", " } ").split()
    return out


class Parser:
    def __init__(self, toks):
        self.t = toks
        self.i = 0

    def next(self):
        tok = self.t[self.i]
        self.i += 1
        return tok

    def num(self):
        return float(self.next())

    def vec(self, n=3):
        return [self.num() for _ in range(n)]

    def expect(self, word):
        tok = self.next()
        assert tok == word, (tok, word)

    def stage(self):
        self.expect("camera")
        self.expect("pin_hole")
        cam = {}
        for key in ("pos", "look", "up"):
            self.expect(key)
            cam[key] = self.vec()
        self.expect("scene")
        objs = []
        while self.i < len(self.t):
            objs.append(self.object())
        return {"camera": cam, "objects": objs}

    def object(self):
        self.expect("object")
        kind = self.next()
        if kind == "transformed":
            inner = self.object()
            self.expect("transform")
            self.expect("{")
            steps = []
            while self.t[self.i] != "}":
                op = self.next()
                steps.append([op] + (self.vec() if op == "translate" else [self.num()]))
            self.expect("}")
            inner["transform"] = steps + inner.get("transform", [])
            inner.setdefault("nesting", 0)
            inner["nesting"] += 1
            return inner
        assert kind == "simple", kind
        shape = self.shape()
        shape["material"] = self.material()
        return shape

    def steps(self):
        self.expect("transform")
        self.expect("{")
        out = []
        while self.t[self.i] != "}":
            op = self.next()
            out.append([op] + (self.vec() if op == "translate" else [self.num()]))
        self.expect("}")
        return out

    def shape(self):
        kind = self.next()
        if kind == "csg":
            op = self.next()
            self.expect("{")
            elems = []
            while self.t[self.i] != "}":
                self.expect("elem")
                e = self.shape()
                e["transform"] = self.steps()
                elems.append(e)
            self.expect("}")
            return {"shape": "csg", "op": op, "elems": elems}
        if kind == "rectangle":
            return {"shape": "rectangle", "param": self.vec(2)}
        if kind == "aabox":
            return {"shape": "aabox", "param": self.vec(3)}
        if kind == "sphere":
            return {"shape": "sphere", "param": [self.num()]}
        if kind == "mesh":
            self.expect("file")
            path = self.next().strip('"')
            reverse = False
            if self.t[self.i] == "reverse":
                self.next()
                reverse = True
            return {"shape": "mesh", "file": os.path.basename(path), "reverse": reverse}
        raise ValueError(kind)

    def material(self):
        kind = self.next()
        m = {"kind": kind}
        if kind == "dielectric":
            m["ior"] = self.num()
        if kind == "metal":
            m["fuzz"] = self.num()
        tex = self.next()
        if tex == "color":
            m["color"] = self.vec()
        else:
            assert tex == "checkerboard"
            m["checker"] = self.vec(7)
            m["color"] = m["checker"][3:6]
        return m


def main(src):
    out_dir = os.path.join(os.path.dirname(os.path.abspath(__file__)), "stages")
    os.makedirs(out_dir, exist_ok=True)
    for name in SCENES:
        with open(os.path.join(src, name + ".stage")) as f:
            scene = Parser(tokens(f.read())).stage()
        for o in scene["objects"]:
            assert o.get("nesting", 0) <= 1, "nested transforms are not used by these scenes"
        with open(os.path.join(out_dir, name + ".json"), "w") as f:
            json.dump(scene, f, indent=None, separators=(",", ":"))
        print(name, len(scene["objects"]), "objects")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/test_files")
