"""Build recipe for the gfx950 NEGF library and its Fortran shims.

Produces, in-tree:
  nonequilibriumgreensfunction_amd/lib/libnegf_hip.so   -- HIP kernels + C ABI (include/negf_capi.h)
  nonequilibriumgreensfunction_amd/lib/libnegf_fortran.so -- iso_c_binding shims (when amdflang exists)
  nonequilibriumgreensfunction_amd/lib/mod/*.mod         -- Fortran module files for callers

Usage: python -m nonequilibriumgreensfunction_amd.build
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
FSRC = os.path.join(PKG, "fortran")
LIBDIR = os.path.join(PKG, "lib")
OBJDIR = os.path.join(PKG, "lib", "obj")
INCLUDE = os.path.join(ROOT, "include")

HIP_SOURCES = ["zgemm_mfma.hip", "zgemm_x3.hip", "gj_inverse.hip", "negf_kernels.hip", "negf_engine.hip"]
# Packed-FP32 instructions are disabled for every device compile.
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I" + INCLUDE, "-I" + CSRC]


def _tool(name: str, fallback: str) -> str:
    p = shutil.which(name)
    return p if p else fallback


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def build_hip(jobs: int = 4) -> str:
    hipcc = _tool("hipcc", "/opt/rocm/bin/hipcc")
    os.makedirs(OBJDIR, exist_ok=True)
    objs = []
    cmds = []
    for src in HIP_SOURCES:
        obj = os.path.join(OBJDIR, src.replace(".hip", ".o"))
        objs.append(obj)
        cmds.append([hipcc] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    out = os.path.join(LIBDIR, "libnegf_hip.so")
    tmp = out + ".tmp"
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, out)
    return out


def build_fortran() -> str | None:
    fc = shutil.which("amdflang") or ("/opt/rocm/bin/amdflang" if os.path.exists("/opt/rocm/bin/amdflang") else None)
    if fc is None or not os.path.isdir(FSRC):
        return None
    moddir = os.path.join(LIBDIR, "mod")
    os.makedirs(moddir, exist_ok=True)
    objs = []
    for src in ("negf_capi.f90", "negf_state.f90", "negf_shims.F90"):
        path = os.path.join(FSRC, src)
        obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
        _run([fc, "-O2", "-fPIC", "-module-dir", moddir, "-I" + moddir, "-c", path, "-o", obj])
        objs.append(obj)
    out = os.path.join(LIBDIR, "libnegf_fortran.so")
    _run([fc, "-shared", "-fPIC", "-o", out] + objs +
         ["-L" + LIBDIR, "-Wl,-rpath,$ORIGIN", "-lnegf_hip"])
    # host driver used by the tests: a Fortran program calling the shims
    drv = os.path.join(ROOT, "tests", "fortran", "negf_fdriver.f90")
    if os.path.exists(drv):
        _run([fc, "-O2", "-module-dir", moddir, "-I" + moddir, drv, "-o",
              os.path.join(LIBDIR, "negf_fdriver"), "-L" + LIBDIR, "-Wl,-rpath,$ORIGIN",
              "-lnegf_fortran", "-lnegf_hip"])
    return out


def build_all(jobs: int = 4) -> dict:
    out = {"hip": build_hip(jobs)}
    out["fortran"] = build_fortran()
    return out


if __name__ == "__main__":
    res = build_all(int(os.environ.get("MAX_JOBS", "4")))
    for k, v in res.items():
        print(k, v)
    sys.exit(0)
