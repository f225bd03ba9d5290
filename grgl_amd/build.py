"""Build the native parts of grgl_amd in-tree.

    python grgl_amd/build.py          # libgrgl_hip.so (gfx950) + _grgl_amd extension

libgrgl_hip.so holds the gfx950 kernels and the C ABI of include/grgl_hip.h (hipcc);
_grgl_amd*.so is the pybind11 module with the pygrgl surface (g++), linked against
libgrgl_hip.so through an $ORIGIN rpath.  Both land next to this file.
"""

import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
ARCH = os.environ.get("GRGL_AMD_ARCH", "gfx950")

HIP_SOURCES = ["grgl_hip.hip", "server.cpp", "levelize.cpp", "grouping.cpp", "host_graph.cpp", "stacking.cpp", "stream_ceiling.hip",
               "noalloc.hip"]
PY_SOURCES = ["pygrgl_module.cpp", "synth.cpp", "host_graph.cpp", "levelize.cpp", "grouping.cpp", "stacking.cpp", "grg_io.cpp",
              "construct.cpp", "fast_build.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-fvisibility=hidden",
    # Packed FP32 instructions are disabled for all device code.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]


def _run(cmd):
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout)
        raise RuntimeError("build step failed: " + " ".join(cmd[:3]) + " ...")
    return proc.stdout


def _compile_hip_object(src, out_dir):
    obj = os.path.join(out_dir, os.path.basename(src) + ".o")
    _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    return obj


def build_hip_library(verbose=False):
    out_dir = os.path.join(HERE, "_build")
    os.makedirs(out_dir, exist_ok=True)
    with ThreadPoolExecutor(max_workers=4) as ex:
        objs = list(ex.map(lambda s: _compile_hip_object(s, out_dir), HIP_SOURCES))
    lib = os.path.join(HERE, "libgrgl_hip.so")
    tmp = lib + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"])
    os.replace(tmp, lib)
    if verbose:
        print("built", lib)
    return lib


def build_python_module(verbose=False):
    import pybind11

    ext = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    target = os.path.join(HERE, "_grgl_amd" + ext)
    tmp = target + ".tmp"
    cmd = [
        CXX,
        "-O2",
        "-std=c++17",
        "-fPIC",
        "-shared",
        "-fvisibility=hidden",
        f"-I{pybind11.get_include()}",
        f"-I{sysconfig.get_paths()['include']}",
        f"-I{INCLUDE}",
        f"-I{CSRC}",
        *[os.path.join(CSRC, s) for s in PY_SOURCES],
        "-o",
        tmp,
        f"-L{HERE}",
        "-lgrgl_hip",
        "-Wl,-rpath,$ORIGIN",
        "-lpthread",
    ]
    _run(cmd)
    os.replace(tmp, target)
    if verbose:
        print("built", target)
    return target


def build(verbose=False):
    build_hip_library(verbose)
    build_python_module(verbose)


if __name__ == "__main__":
    build(verbose=True)
