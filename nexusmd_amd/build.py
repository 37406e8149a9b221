"""Build the engine in-tree (no JIT cache):

  libnexusmd_hip.so    HIP kernels + the C-ABI of include/nexusmd.h (hipcc, gfx950)
  libnexusmd_host.so   C++ host side: XML readers, reporters, TaskDispatcher (g++)
  NexusMD<ext>.so      Python module with the reference's wrapper names (pybind11)

    python -m nexusmd_amd.build
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libnexusmd_hip.so")
SOURCES = ["nl.hip", "nonbonded.hip", "pme.hip", "bonded.hip", "integrate.hip", "engine.hip", "capi.hip",
           "serve.hip"]
ARCH = os.environ.get("NEXUSMD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CFLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-I" + CSRC, "-I" + os.path.join(os.path.dirname(HERE), "include")]


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, src.replace(".hip", ".o"))
    cmd = [HIPCC, *CFLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s" % (src, r.stderr[-4000:]))
    return obj


HOST_DIR = os.path.join(HERE, "host")
HOST_SOURCES = ["xml_reader.cpp", "parsers.cpp", "reporter.cpp", "dispatcher.cpp"]
HOST_LIB = os.path.join(HERE, "libnexusmd_host.so")
CXX = os.environ.get("CXX", "g++")
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wextra", "-I" + HOST_DIR,
              "-I" + os.path.join(os.path.dirname(HERE), "include")]


def module_path() -> str:
    return os.path.join(HERE, "NexusMD" + sysconfig.get_config_var("EXT_SUFFIX"))


def _run(cmd, what):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("%s failed:\n%s" % (what, r.stderr[-4000:]))


def build_host(verbose: bool = False) -> str:
    """The C++ host library and the Python module, linked against libnexusmd_hip.so."""
    import pybind11
    srcs = [os.path.join(HOST_DIR, f) for f in HOST_SOURCES]
    link = ["-L" + HERE, "-l:libnexusmd_hip.so", "-Wl,-rpath,$ORIGIN"]
    tmp = HOST_LIB + ".tmp"
    _run([CXX, *HOST_FLAGS, "-shared", "-o", tmp, *srcs, *link], "host library")
    os.replace(tmp, HOST_LIB)
    mod = module_path()
    tmp = mod + ".tmp"
    # the module finds the libraries beside it (in-tree) or in nexusmd_amd/ beside
    # it (installed from the wheel, where NexusMD sits at the top level)
    _run([CXX, *HOST_FLAGS, "-fvisibility=hidden", "-I" + pybind11.get_include(),
          "-I" + sysconfig.get_paths()["include"], "-shared", "-o", tmp, os.path.join(HOST_DIR, "pymodule.cpp"),
          "-L" + HERE, "-l:libnexusmd_host.so", *link, "-Wl,-rpath,$ORIGIN/nexusmd_amd"], "Python module")
    os.replace(tmp, mod)
    if verbose:
        print("built", HOST_LIB)
        print("built", mod)
    return mod


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr[-4000:])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_host(verbose)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
