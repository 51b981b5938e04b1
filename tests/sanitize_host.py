"""Host-side sanitiser run of the CPU test suite (not collected by pytest; run it by hand):

    python tests/sanitize_host.py [pytest arguments]

Builds pbrt_amd/libpbrt_gpu_asan.so (every host translation unit under AddressSanitizer and
UndefinedBehaviorSanitizer, device code unchanged) and oracle/liboracle_asan.so (the same
sanitisers, the same clang), then runs `pytest -m "not gpu"` with both loaded in their place and
clang's sanitizer runtime preloaded into the interpreter. Leak detection is off (the interpreter
itself is not instrumented); any other report fails the run.
"""
import glob
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def main(args):
    from pbrt_amd import build
    build.build(sanitize=True)
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "-B", "asan"], check=True)
    # both libraries must carry the sanitisers' hooks, or the run checks nothing
    for lib in (os.path.join(ROOT, "pbrt_amd", "libpbrt_gpu_asan.so"), os.path.join(ROOT, "oracle", "liboracle_asan.so")):
        syms = subprocess.run(["nm", "-D", lib], capture_output=True, text=True, check=True).stdout
        if "__asan_report" not in syms or "__ubsan_handle" not in syms:
            raise RuntimeError(f"{lib} is not instrumented")
    runtime = sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))[-1]
    env = dict(os.environ, LD_PRELOAD=runtime, PBRT_AMD_LIB_VARIANT="asan", ORACLE_LIB_VARIANT="asan",
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    cmd = [sys.executable, "-m", "pytest", os.path.join(ROOT, "tests"), "-q", "-m", "not gpu", "-p", "no:cacheprovider"]
    return subprocess.run(cmd + args, env=env, cwd=ROOT).returncode


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
