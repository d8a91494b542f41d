"""Builds libermgpu.so (gfx950) in-tree with hipcc, and the Fortran binding
module with amdflang when it is available.

    python -m ermod_amd.build
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FC = "/opt/rocm/bin/amdflang"
SOURCES = [os.path.join(HERE, "csrc", "ermgpu.hip")]
HEADERS = [os.path.join(HERE, "csrc", f) for f in ("ermgpu_dev.h", "ermgpu_kern.h", "ermgpu_grad.h",
                                                         "ermgpu_traj.h")] + [
    os.path.join(ROOT, "include", "ermgpu.h")]
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++20", "-fPIC", "-shared",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I" + os.path.join(ROOT, "include")]


def build_lib(verbose: bool = False) -> str:
    out = os.path.join(HERE, "libermgpu.so")
    tmp = out + ".tmp"
    cmd = [HIPCC, *FLAGS, *SOURCES, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    errs = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed:\n{errs}")
    os.replace(tmp, out)
    return out


def build_fortran(verbose: bool = False):
    """The ISO_C_BINDING module and the example host program."""
    if not os.path.exists(FC):
        return None
    fdir = os.path.join(HERE, "fortran")
    bdir = os.path.join(fdir, "_build")
    os.makedirs(bdir, exist_ok=True)
    mod = os.path.join(fdir, "ermgpu.F90")
    ex = os.path.join(fdir, "ermgpu_host.F90")
    exe = os.path.join(fdir, "ermgpu_host")
    lib = os.path.join(HERE, "libermgpu.so")
    cmds = [[FC, "-O2", "-c", mod, "-J", bdir, "-o", os.path.join(bdir, "ermgpu.o")],
            [FC, "-O2", ex, os.path.join(bdir, "ermgpu.o"), "-J", bdir, "-I", bdir, "-o", exe,
             lib, "-Wl,-rpath,$ORIGIN/.."]]
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"amdflang failed:\n{r.stderr}")
    return exe


def main():
    v = "-v" in sys.argv
    print(build_lib(v))
    if os.path.exists(os.path.join(HERE, "fortran", "ermgpu.F90")):
        print(build_fortran(v))


if __name__ == "__main__":
    main()
