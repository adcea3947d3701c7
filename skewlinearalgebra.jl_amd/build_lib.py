"""Builds libskewhip.so (HIP for gfx950) in place: skewlinearalgebra/libskewhip.so.

    python build_lib.py            # from this directory, or call build() from Python
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "skewlinearalgebra", "libskewhip.so")
SOURCES = ["skh_kernels.hip", "skh_compact.hip", "skh_complex.hip", "skh_zpfaff.hip", "skh_chol.hip", "skh_rot.hip", "skh_fun.hip", "skh_noalloc.hip", "skh_grad.hip", "skh_api.cpp", "host_tridiag.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # keep the compiler from forming packed FP32 instructions (not used by this FP64 code anyway)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]


def build(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = OUT + ".tmp"
    cmd = [HIPCC] + FLAGS + ["-I", os.path.join(HERE, "..", "include"), "-o", tmp] + srcs + [
        "-L/opt/rocm/lib", "-lrocblas", "-lrocsolver", "-lrocprofiler-sdk-roctx", "-Wl,-rpath,/opt/rocm/lib"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
