"""Build the gfx950 shared libraries of gp5_amd in-tree.

    python -m gp5_amd.build          # libgp5.so + gpCpp extension

libgp5.so holds the HIP kernels and the C ABI of include/gp5.h; the gpCpp
extension is the pybind11 module (same class and method names as the
reference's src/gpCpp.cpp:327-370) layered on the C++ facade gp5pp.hpp.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "libgp5.so")
ARCH = os.environ.get("GP5_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["gp5_kernels.hip", "gp5_fast.hip", "gp5_capi.hip", "gp5_batch.hip", "gp5_nn.hip", "gp5_ccl.hip", "gp5_resample.hip", "gp5_h5.cpp"]
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 ops stay off (hipcc's host pass warns that it does not know
    # the feature; harmless)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build failed ({r.returncode}):\n{' '.join(cmd)}\n{r.stdout}")
    # keep warnings visible, but drop the known packed-fp32 host-pass notice
    lines = [l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l and l.strip()]
    if verbose and lines:
        print("\n".join(lines))


def build_lib(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES]
    tmp = LIB + ".tmp"
    _run([_hipcc(), *HIP_FLAGS, "-I", INCLUDE, "-I", CSRC, *srcs, "-ldl", "-o", tmp], verbose)
    os.replace(tmp, LIB)
    return LIB


def ext_path() -> str:
    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    return os.path.join(HERE, "gpCpp" + suffix)


def build_ext(verbose: bool = False) -> str:
    import pybind11

    out = ext_path()
    tmp = out + ".tmp"
    py_inc = sysconfig.get_paths()["include"]
    cmd = [
        os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-fPIC", "-shared", "-fvisibility=hidden",
        "-I", INCLUDE, "-I", CSRC, "-I", py_inc, "-I", pybind11.get_include(),
        os.path.join(CSRC, "gpCpp.cpp"),
        "-L", HERE, "-lgp5", "-Wl,-rpath,$ORIGIN", "-o", tmp,
    ]
    _run(cmd, verbose)
    os.replace(tmp, out)
    return out


def build(verbose: bool = False) -> None:
    build_lib(verbose)
    if os.path.exists(os.path.join(CSRC, "gpCpp.cpp")):
        build_ext(verbose)


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
