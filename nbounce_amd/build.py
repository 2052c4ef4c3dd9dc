"""Builds libnbounce.so (HIP kernels for gfx950 + the C ABI + the host scene builder) in-tree.

    python nbounce_amd/build.py

Every source is compiled by hipcc with binary32 semantics fixed: no FMA contraction,
correctly rounded divide and square root, denormals kept, packed-FP32 instructions disabled.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "nbounce_amd", "csrc")
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(ROOT, "nbounce_amd", "libnbounce.so")
SOURCES = ["trace.hip", "order.hip", "serve.hip", "mlp.hip", "cache_render.hip", "display.hip", "abi.cpp", "server.cpp", "cache.cpp", "scene_build.cpp", "sobol.cpp", "gltf.cpp", "image_io.cpp", "dds.cpp"]
ARCH = os.environ.get("NB_OFFLOAD_ARCH", "gfx950")


def hipcc():
    for cand in (os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "bin", "hipcc"), shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build libnbounce")


def common_flags():
    return [
        f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
        "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-gpu-flush-denormals-to-zero",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall", "-Wno-unused-function", "-Wno-unused-value", "-Wno-unused-result",
        # a value read before it is written is an error in host and device code alike
        "-Werror=uninitialized", "-Werror=sometimes-uninitialized", "-Werror=conditional-uninitialized",
        "-I", os.path.join(ROOT, "include"), "-I", CSRC,
    ]


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    cc = hipcc()
    objs = []

    def compile_one(src):
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        cmd = [cc, "-x", "hip"] + common_flags() + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"compiling {src} failed:\n{r.stdout}\n{r.stderr}")
        return obj

    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"linking libnbounce.so failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
