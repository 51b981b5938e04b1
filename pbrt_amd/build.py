"""Builds the native library in-tree: pbrt_amd/libpbrt_gpu.so (HIP for gfx950 + host C++).

Each source compiles to its own object (in parallel, under pbrt_amd/_build/), then one hipcc link
makes the shared library. Device code is not relocatable (no -fgpu-rdc): every translation unit
carries its own code object, so the units are independent.

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
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libpbrt_gpu.so")
OBJ = os.path.join(HERE, "_build")

SOURCES = ["api.cpp", "bvh_build.cpp", "scenegen.cpp", "scene_api.cpp", "nrc_api.cpp", "film.cpp", "stage_loader.cpp",
           "traverse.hip", "traverse_work.hip", "bvh_gpu.hip", "scene.hip", "nrc.hip", "nrc_render.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC",
    "--offload-arch=gfx950",
    "-ffp-contract=off",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
]


def _jobs():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n, len(SOURCES)))


def _compile(src, verbose, obj_dir, extra):
    obj = os.path.join(obj_dir, os.path.basename(src) + ".o")
    cmd = [HIPCC] + FLAGS + extra + ["-x", "hip", "-c", src, "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    return obj, r


# Host-side AddressSanitizer and UndefinedBehaviorSanitizer (device code unchanged): each flag
# applies to the host compilation only.
SANITIZE_HOST = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined", "-Xarch_host",
                 "-fno-omit-frame-pointer", "-Xarch_host", "-fno-sanitize-recover=undefined"]


def build(verbose=False, sanitize=False):
    """Compiles and links the library; sanitize=True builds libpbrt_gpu_asan.so with the host code
    under ASan + UBSan instead (loaded with PBRT_AMD_LIB_VARIANT=asan)."""
    obj_dir = OBJ + ("_asan" if sanitize else "")
    out = OUT.replace("libpbrt_gpu.so", "libpbrt_gpu_asan.so") if sanitize else OUT
    os.makedirs(obj_dir, exist_ok=True)
    srcs = [os.path.join(SRC, s) for s in SOURCES]
    extra = SANITIZE_HOST if sanitize else []
    with ThreadPoolExecutor(_jobs()) as ex:
        results = list(ex.map(lambda s: _compile(s, verbose, obj_dir, extra), srcs))
    failed = [r for _, r in results if r.returncode != 0]
    for r in results:
        if r[1].returncode != 0 or (verbose and (r[1].stdout or r[1].stderr)):
            sys.stderr.write(r[1].stdout + r[1].stderr)
    if failed:
        raise RuntimeError("hipcc failed")
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950"] + extra + [o for o, _ in results] + ["-o", out + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc link failed")
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv, sanitize="--sanitize" in sys.argv))
