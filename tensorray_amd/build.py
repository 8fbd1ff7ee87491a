"""Build recipe for the native library `libtrc.so` (HIP kernels for gfx950 +
host C++ + C ABI) and the `TensorRay` Python extension module.

Usage: python -m tensorray_amd.build   (or tensorray_amd.build.build()).

Every HIP source is compiled with
  --offload-arch=gfx950 -ffp-contract=off -Xclang -target-feature -Xclang -packed-fp32-ops
so that float arithmetic is evaluated exactly as written (reproducible
between host and device) and no packed FP32 instructions are emitted.
Host sources are compiled with g++ (-ffp-contract=off) against the HIP
headers.  Outputs land next to this file; intermediates in _build/.
"""
import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

LIB = os.path.join(HERE, "libtrc.so")

HIP_SOURCES = ["kernels/render.hip", "kernels/boundary.hip", "kernels/optim.hip", "kernels/update.hip"]
HOST_SOURCES = ["host/scene_host.cpp", "host/bvh.cpp", "capi.cpp"]

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fhip-fp32-correctly-rounded-divide-sqrt", "-Wno-unused-result",
]
HOST_FLAGS = [
    "-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__",
    "-I" + os.path.join(ROCM, "include"), "-Wall", "-Wno-unused-function", "-Wno-unknown-pragmas",
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, src.replace("/", "_") + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC] + HIP_FLAGS + ["-c", path, "-o", obj]
    else:
        cmd = ["g++"] + HOST_FLAGS + ["-c", path, "-o", obj]
    _run(cmd)
    return obj


def build_lib(jobs=8):
    os.makedirs(BUILD, exist_ok=True)
    srcs = HIP_SOURCES + HOST_SOURCES
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    # --no-undefined: a missing definition fails the build, not the first dlopen
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-Wl,--no-undefined", "-o", tmp] + objs +
         ["-lpthread"])
    os.replace(tmp, LIB)
    return LIB


def build_pymodule():
    """pybind11 module `TensorRay` (C++ host classes over the C ABI)."""
    src = os.path.join(CSRC, "pybind", "tensorray_module.cpp")
    if not os.path.exists(src):
        return None
    srcs = [src, os.path.join(CSRC, "pybind", "host_scene.cpp")]
    import pybind11
    suffix = sysconfig.get_config_var("EXT_SUFFIX")
    out = os.path.join(HERE, "TensorRay" + suffix)
    tmp = out + ".tmp"
    cmd = ["g++"] + HOST_FLAGS + [
        "-shared", "-I" + pybind11.get_include(), "-I" + sysconfig.get_paths()["include"],
        "-fvisibility=hidden", "-Wno-deprecated-declarations",
    ] + srcs + [
        "-o", tmp, "-L" + HERE, "-ltrc", "-Wl,-rpath,$ORIGIN",
        "-L" + os.path.join(ROCM, "lib"), "-lamdhip64", "-Wl,-rpath," + os.path.join(ROCM, "lib"),
    ]
    _run(cmd)
    os.replace(tmp, out)
    return out


def build():
    build_lib()
    build_pymodule()


if __name__ == "__main__":
    build()
    print("built", LIB, file=sys.stderr)
