"""Writes the reference's test meshes as compact fixtures (tests/golden/meshes/*.npz).

Reads the ASCII PLY files shipped with the reference (test_files/*.ply) and keeps exactly what the
reference's loader keeps: the x, y, z vertex properties parsed as FP32 (src/loader/ply/
ply_loader.rs:130-160, `parse::<f32>()`, correctly rounded) and the face index lists
(src/loader/mesh_loader.rs:53-86). Decimal -> FP32 goes through libc `strtof`, which is correctly
rounded, rather than through a double (which could round twice).

Run once, in a container where the reference is mounted:
    python tests/golden/make_ply_fixtures.py /root/reference/test_files
"""
import ctypes
import os
import sys

import numpy as np

MESHES = ["cube.ply", "bun_zipper_res4.ply", "dragon_vrip_res4.ply", "dragon_vrip_res3.ply", "bunny.ply"]

_libc = ctypes.CDLL(None)
_libc.strtof.restype = ctypes.c_float
_libc.strtof.argtypes = [ctypes.c_char_p, ctypes.c_void_p]


def parse_ply(path):
    with open(path, "rb") as f:
        lines = f.read().decode("ascii").splitlines()
    i = 0
    assert lines[0].strip() == "ply"
    elements = []
    while True:
        i += 1
        tok = lines[i].split()
        if not tok:
            continue
        if tok[0] == "format":
            assert tok[1] == "ascii", "only ASCII PLY is read (ply_loader.rs:251-255)"
        elif tok[0] == "element":
            elements.append([tok[1], int(tok[2]), []])
        elif tok[0] == "property":
            elements[-1][2].append(tok[1:])
        elif tok[0] == "end_header":
            i += 1
            break
    verts, faces = None, None
    for name, count, props in elements:
        rows = lines[i:i + count]
        i += count
        if name == "vertex":
            names = [p[-1] for p in props]
            ix = [names.index(c) for c in ("x", "y", "z")]
            verts = np.empty((count, 3), np.float32)
            for r, line in enumerate(rows):
                seg = line.split()
                for c in range(3):
                    verts[r, c] = _libc.strtof(seg[ix[c]].encode(), None)
        elif name == "face":
            assert len(props) == 1 and props[0][0] == "list"
            faces = np.empty((count, 3), np.uint32)
            for r, line in enumerate(rows):
                seg = line.split()
                assert int(seg[0]) == 3, "non-triangle face"
                faces[r] = [int(s) for s in seg[1:4]]
    return verts, faces


def main(src):
    out_dir = os.path.join(os.path.dirname(os.path.abspath(__file__)), "meshes")
    os.makedirs(out_dir, exist_ok=True)
    for name in MESHES:
        v, f = parse_ply(os.path.join(src, name))
        out = os.path.join(out_dir, name.replace(".ply", ".npz"))
        np.savez_compressed(out, vertices=v, faces=f)
        print(f"{name}: {len(v)} vertices, {len(f)} faces -> {out}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/test_files")
