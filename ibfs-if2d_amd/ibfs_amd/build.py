"""Builds the HIP library (gfx950) and the Fortran 90 host in-tree.

    libibfs_hip.so   hipcc, from csrc/ibfs_capi.hip (kernels in the headers)
    host/ibfs_gpu    amdflang, the Fortran host linked against libibfs_hip.so
    host/ibfs_post   amdflang, the post-processing program (Tecplot files)
"""
from __future__ import annotations

import os
import shutil
import subprocess

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))  # .../ibfs-if2d_amd
LIB = os.path.join(PKG, "libibfs_hip.so")
HOST_DIR = os.path.join(PKG, "host")
HOST_EXE = os.path.join(HOST_DIR, "ibfs_gpu")
POST_EXE = os.path.join(HOST_DIR, "ibfs_post")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FC = os.environ.get("FC_AMD", "/opt/rocm/bin/amdflang")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # no packed FP32 instructions anywhere in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]

HOST_SOURCES = ["parameters.f90", "grid.f90", "ibfs_c.f90", "variables.f90", "operators.f90", "main.f90"]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return r


def build_lib() -> str:
    src = os.path.join(PKG, "csrc", "ibfs_capi.hip")
    tmp = LIB + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-o", tmp, src])
    os.replace(tmp, LIB)
    return LIB


def build_host() -> str | None:
    if not all(os.path.exists(os.path.join(HOST_DIR, s)) for s in HOST_SOURCES + ["post.f90"]):
        return None
    if not os.path.exists(FC):
        return None
    bdir = os.path.join(HOST_DIR, "_build")
    os.makedirs(bdir, exist_ok=True)
    objs = []
    for s in HOST_SOURCES:
        o = os.path.join(bdir, s.replace(".f90", ".o"))
        _run([FC, "-O2", "-c", os.path.join(HOST_DIR, s), "-o", o, "-module-dir", bdir], cwd=bdir)
        objs.append(o)
    _run([FC, "-O2", *objs, "-o", HOST_EXE, "-L" + PKG, "-libfs_hip", "-Wl,-rpath,$ORIGIN/.."])
    # the post-processing program: the same modules with post.f90 as the driver
    po = os.path.join(bdir, "post.o")
    _run([FC, "-O2", "-c", os.path.join(HOST_DIR, "post.f90"), "-o", po, "-module-dir", bdir], cwd=bdir)
    mods = [o for o in objs if not o.endswith("main.o")]
    _run([FC, "-O2", *mods, po, "-o", POST_EXE, "-L" + PKG, "-libfs_hip", "-Wl,-rpath,$ORIGIN/.."])
    shutil.rmtree(bdir, ignore_errors=True)
    return HOST_EXE


def build_all() -> None:
    build_lib()
    build_host()


if __name__ == "__main__":
    build_all()
    print("built", LIB)
