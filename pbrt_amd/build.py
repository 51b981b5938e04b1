"""Builds the native library in-tree: pbrt_amd/libpbrt_gpu.so (HIP for gfx950 + host C++).

Flags that matter for parity with the CPU path:
  -ffp-contract=off      no fused multiply-add, host and device, so FP32 traversal arithmetic is the
                         same sequence of roundings as the FP32 CPU restatement;
  IEEE FP32 division and square root (the HIP default; no fast-math flag anywhere);
  -fno-gpu-flush-denormals-to-zero  FP32 denormals preserved, as on the host.
Packed FP32 instructions are disabled (-packed-fp32-ops), by policy for this hardware pool.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libpbrt_gpu.so")

SOURCES = ["api.cpp", "bvh_build.cpp", "scenegen.cpp", "scene_api.cpp", "nrc_api.cpp", "film.cpp", "stage_loader.cpp", "traverse.hip", "bvh_gpu.hip", "scene.hip",
           "nrc.hip", "nrc_render.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared",
    "--offload-arch=gfx950",
    "-ffp-contract=off",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
]


def build(verbose=False):
    srcs = [os.path.join(SRC, s) for s in SOURCES]
    cmd = [HIPCC] + FLAGS + ["-x", "hip"] + srcs + ["-o", OUT + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed")
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
