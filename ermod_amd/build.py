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
MAIN = os.path.join(HERE, "csrc", "ermgpu.hip")
PME = os.path.join(HERE, "csrc", "ermgpu_pme.hip")
# peak-rate microbenchmarks for the roofline (libermpeak.so)
PEAK = os.path.join(HERE, "csrc", "ermpeak.hip")
# register-FFT grid lengths per precision (the run-time dispatch in
# ermgpu.hip covers exactly these)
PME_UNITS = [("double", k) for k in (16, 24, 32, 48, 64)] + [("float", k) for k in (16, 24, 32, 48, 64, 96)]
HEADERS = [os.path.join(HERE, "csrc", f) for f in ("ermgpu_dev.h", "ermgpu_kern.h", "ermgpu_grad.h", "ermgpu_erfc.h",
                                                         "ermgpu_traj.h", "ermgpu_ctx.h")] + [
    os.path.join(ROOT, "include", "ermgpu.h")]
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++20", "-fPIC",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I" + os.path.join(ROOT, "include")]
BUILD_DIR = os.path.join(HERE, "_build")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    errs = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed: {' '.join(cmd[-3:])}\n{errs}")


def build_lib(verbose: bool = False, jobs: int = 0) -> str:
    """Compiles the host/launch unit and one unit per (precision, grid
    length) of the register-FFT kernels in parallel, then links them; the
    peak-rate microbenchmarks build into libermpeak.so alongside."""
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(BUILD_DIR, exist_ok=True)
    units = [([MAIN], os.path.join(BUILD_DIR, "ermgpu.o"))]
    for t, k in PME_UNITS:
        units.append(([f"-DERM_T={t}", f"-DERM_K={k}", PME], os.path.join(BUILD_DIR, f"pme_{t[0]}{k}.o")))
    jobs = jobs or min(len(units) + 1, max(1, os.cpu_count() or 1), 16)
    cmds = [[HIPCC, *FLAGS, "-c", *src, "-o", obj] for src, obj in units]
    peak = os.path.join(HERE, "libermpeak.so")
    cmds.append([HIPCC, *FLAGS, "-shared", PEAK, "-o", peak + ".tmp"])
    with ThreadPoolExecutor(jobs) as ex:
        for f in [ex.submit(_run, c, verbose) for c in cmds]:
            f.result()
    os.replace(peak + ".tmp", peak)
    out = os.path.join(HERE, "libermgpu.so")
    tmp = out + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *[o for _, o in units], "-o", tmp], verbose)
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
