"""Build libsdgso.so (HIP kernels for gfx950 + the C-ABI) and the lightdock
driver in-tree with hipcc.

    python -m sparkledock_amd.build          # or sparkledock_amd.build.build()

Objects go to sparkledock_amd/_build/; the library to
sparkledock_amd/libsdgso.so; the driver to sparkledock_amd/lightdock.
"""
import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsdgso.so")
DRIVER = os.path.join(HERE, "lightdock")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wno-unused-result",
          "-I" + os.path.join(ROOT, "include")]
HIP_FLAGS = ["-x", "hip", "--offload-arch=gfx950",
             # no packed FP32 instructions anywhere in device code
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [HIPCC] + COMMON + (HIP_FLAGS if src.endswith(".hip") else []) + ["-c", src, "-o", obj]
    _run(cmd)
    return obj


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))
    jobs = min(len(srcs), int(os.environ.get("MAX_JOBS", "8")))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs +
         ["-L/opt/rocm/lib", "-lrocsolver", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, LIB)
    host_srcs = sorted(glob.glob(os.path.join(HOST, "*.cpp")))
    if host_srcs:
        tmp = DRIVER + ".tmp"
        _run(["g++", "-O2", "-std=c++17", "-Wall", "-I" + os.path.join(ROOT, "include")] + host_srcs +
             ["-o", tmp, "-L" + HERE, "-lsdgso", "-Wl,-rpath,$ORIGIN"])
        os.replace(tmp, DRIVER)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
