"""Build recipe of the native parts of snapy_linfel_amd.

libsnapx.so   HIP kernels + C ABI (hipcc, gfx950), no torch dependency
_snapy.so     C++ host modules (Hydro, Reconstruct, ...) over libtorch, bound
              to Python with pybind11 (g++)

Everything is built in-tree so the shared objects travel with the source
tree.  ``build_all`` always recompiles.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sysconfig
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(ROOT, "_build")
LIBSNAPX = os.path.join(PKG, "libsnapx.so")
HOSTMOD = os.path.join(PKG, "_snapy" + (sysconfig.get_config_var("EXT_SUFFIX") or ".so"))

ARCH = os.environ.get("SNAPX_OFFLOAD_ARCH", "gfx950")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

# translation units and their extra flags; the plain FP64 path keeps the
# reference's operation order (no FMA contraction)
HIP_UNITS = {
    "abi.hip": [],
    "plain.hip": ["-ffp-contract=off"],
    "vic.hip": ["-ffp-contract=off"],
    # FP32 fused path: hardware reciprocal / square root (1-2 ulp) instead of
    # the correctly rounded division and sqrt sequences, and FP32 denormals
    # flushed to zero, so that a division is one reciprocal and one multiply
    # (without the exponent scaling that keeps denormal results); FP64 keeps
    # its denormals
    "fused.hip": ["-fno-hip-fp32-correctly-rounded-divide-sqrt", "-fgpu-flush-denormals-to-zero"],
    "adjoint.hip": ["-ffp-contract=off"],
    "control.hip": ["-ffp-contract=off"],
}

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"-I{INCLUDE}",
    # packed FP32 instructions stay disabled on this platform
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd):
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + p.stdout + p.stderr)
    return p


def build_libsnapx(jobs: int = 8) -> str:
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    cmds = []
    for unit, extra in HIP_UNITS.items():
        obj = os.path.join(BUILD, unit.replace(".hip", ".o"))
        objs.append(obj)
        cmds.append([HIPCC, *HIP_FLAGS, *extra, "-c", os.path.join(CSRC, unit), "-o", obj])
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    tmp = LIBSNAPX + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs,
          "-L/opt/rocm/lib", "-lrocprofiler-sdk-roctx"])
    os.replace(tmp, LIBSNAPX)
    return LIBSNAPX


def _torch_flags():
    import torch
    from torch.utils import cpp_extension as ce
    inc = []
    for p in ce.include_paths():
        inc.append(f"-I{p}")
    import pybind11
    inc.append(f"-I{pybind11.get_include()}")
    inc.append(f"-I{sysconfig.get_paths()['include']}")
    libdir = ce.library_paths()[0]
    abi = int(torch._C._GLIBCXX_USE_CXX11_ABI)
    cflags = ["-O2", "-std=c++17", "-fPIC", f"-D_GLIBCXX_USE_CXX11_ABI={abi}",
              "-DTORCH_API_INCLUDE_EXTENSION_H", "-DTORCH_EXTENSION_NAME=_snapy",
              "-D__HIP_PLATFORM_AMD__=1", "-DUSE_ROCM=1", "-I/opt/rocm/include", *inc]
    ldflags = [f"-L{libdir}", f"-Wl,-rpath,{libdir}", "-lc10", "-ltorch", "-ltorch_cpu",
               "-ltorch_python", "-lc10_hip", "-ltorch_hip"]
    return cflags, ldflags


def build_host(jobs: int = 8) -> str:
    """C++ host modules over libtorch (g++), linked against libsnapx.so."""
    srcs = sorted(f for f in os.listdir(HOST) if f.endswith(".cpp"))
    if not srcs:
        return ""
    os.makedirs(BUILD, exist_ok=True)
    cflags, ldflags = _torch_flags()
    cflags = cflags + [f"-I{INCLUDE}", f"-I{HOST}"]
    objs, cmds = [], []
    for s in srcs:
        obj = os.path.join(BUILD, "host_" + s.replace(".cpp", ".o"))
        objs.append(obj)
        cmds.append(["g++", *cflags, "-c", os.path.join(HOST, s), "-o", obj])
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    tmp = HOSTMOD + ".tmp"
    _run(["g++", "-shared", "-o", tmp, *objs, f"-L{PKG}", "-lsnapx",
          "-Wl,-rpath,$ORIGIN", *ldflags])
    os.replace(tmp, HOSTMOD)
    return HOSTMOD


def build_all(jobs: int = 8):
    build_libsnapx(jobs)
    build_host(jobs)
