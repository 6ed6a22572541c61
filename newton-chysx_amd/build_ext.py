"""Builds the in-tree native libraries of the `chysx` package.

  chysx/libchysx_hip.so      gfx950 HIP kernels + the C ABI (include/chysx_hip.h)
  chysx/_chysx_native*.so    pybind11 module over the C++ host layer, linked
                             against libchysx_hip.so (rpath $ORIGIN)

Run `python newton-chysx_amd/build_ext.py` (or `__graft_entry__.build()`).
Object files go to newton-chysx_amd/_build/.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(CSRC, "host")
PKG = os.path.join(HERE, "chysx")
BUILD = os.path.join(HERE, "_build")
INCLUDE = os.path.join(REPO, "include")

ARCH = os.environ.get("CHYSX_HIP_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

# Packed FP32 instructions are disabled for all device code.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-I", INCLUDE, "-I", CSRC,
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-I", INCLUDE, "-I", HOST]

HIP_LIB = os.path.join(PKG, "libchysx_hip.so")
EXT_SUFFIX = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
NATIVE_LIB = os.path.join(PKG, "_chysx_native" + EXT_SUFFIX)


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("command failed: " + " ".join(cmd) + "\n" + out)
    return out


def _compile_hip(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    _run([HIPCC, *HIP_FLAGS, "-c", src, "-o", obj])
    return obj


def _compile_host(src, extra):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    _run([CXX, *HOST_FLAGS, *extra, "-c", src, "-o", obj])
    return obj


def build(verbose: bool = True) -> None:
    os.makedirs(BUILD, exist_ok=True)
    import pybind11

    py_inc = ["-I", sysconfig.get_paths()["include"], "-I", pybind11.get_include()]
    hip_srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    host_srcs = sorted(glob.glob(os.path.join(HOST, "*.cpp")))
    jobs = min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        hip_objs = list(ex.map(_compile_hip, hip_srcs))
        host_objs = list(ex.map(lambda s: _compile_host(s, py_inc if s.endswith("bindings.cpp") else []), host_srcs))
    tmp = HIP_LIB + ".tmp"
    # roctx (host-side profiler ranges) from the ROCm install
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *hip_objs, f"-L{ROCM}/lib",
          "-lrocprofiler-sdk-roctx", f"-Wl,-rpath,{ROCM}/lib"])
    os.replace(tmp, HIP_LIB)
    tmp = NATIVE_LIB + ".tmp"
    _run([CXX, "-shared", "-fPIC", "-o", tmp, *host_objs, "-L", PKG, "-lchysx_hip",
          "-Wl,-rpath,$ORIGIN", "-Wl,--no-undefined", f"-L{sys.base_prefix}/lib"] + _python_link())
    os.replace(tmp, NATIVE_LIB)
    if verbose:
        print(f"built {os.path.relpath(HIP_LIB, REPO)} and {os.path.relpath(NATIVE_LIB, REPO)}")


def _python_link():
    # Extension modules resolve Python symbols from the interpreter at load
    # time; --no-undefined therefore needs libpython on the link line only
    # when it exists as a shared library.
    libdir = sysconfig.get_config_var("LIBDIR") or ""
    ldlib = sysconfig.get_config_var("LDLIBRARY") or ""
    if ldlib.endswith(".so") or ".so." in ldlib:
        name = ldlib[3:].split(".so")[0]
        return [f"-L{libdir}", f"-l{name}"]
    return ["-Wl,--unresolved-symbols=ignore-in-object-files"]


if __name__ == "__main__":
    build()
