"""Build libcosahip.so (gfx950) in-tree with hipcc.

    python -m cosa_amd.build            # library only
Flags: -O3, -ffp-contract=off (no fused multiply-add, so FP64 results keep
the reference's rounding), packed-FP32 instructions disabled;
k_resid_contracted.hip (the fast FP64 residual) with -ffp-contract=fast
-freciprocal-math.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libcosahip.so")
SOURCES = ["api.hip", "k_core.hip", "k_flux.hip", "k_resid.hip", "k_resid_contracted.hip", "k_resid_visc.hip",
           "k_turb.hip", "k_force.hip", "k_grid.hip"]
# sources built with multiply-add contraction (the fast FP64 residual)
CONTRACTED = {"k_resid_contracted.hip"}
HIPFLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-result"]


def hipcc() -> str:
    return shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


def build(verbose: bool = False, jobs: int = 4) -> str:
    os.makedirs(os.path.join(HERE, "_build"), exist_ok=True)
    objs = []
    procs = []
    for src in SOURCES:
        obj = os.path.join(HERE, "_build", src.replace(".hip", ".o"))
        flags = list(HIPFLAGS)
        if src in CONTRACTED:   # fused multiply-adds, and x / y as x * (1 / y) where 1 / y is shared
            flags = [f if f != "-ffp-contract=off" else "-ffp-contract=fast" for f in flags] + ["-freciprocal-math"]
        cmd = [hipcc()] + flags + ["-I", os.path.join(ROOT, "include"), "-I", CSRC, "-c",
                                   os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd))
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(obj)
        if len(procs) >= jobs:
            _wait(procs)
    _wait(procs)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd))
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    return LIB


def _wait(procs):
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out.decode(errors="replace"))
            raise RuntimeError(f"hipcc failed on {src}")
        msg = out.decode(errors="replace").strip()
        if msg:
            lines = [l for l in msg.splitlines() if "packed-fp32-ops" not in l and l.strip()]
            if lines:
                sys.stderr.write("\n".join(lines) + "\n")
    procs.clear()


def build_fortran(verbose: bool = False) -> str:
    """Compile the iso_c_binding module and the Fortran demo driver against
    libcosahip.so (amdflang); returns the demo executable path."""
    fc = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    fdir = os.path.join(HERE, "fortran")
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    exe = os.path.join(fdir, "binding_demo")
    cmd = [fc, "-O2", "-module-dir", bdir, "-I", bdir, os.path.join(fdir, "cosahip_binding.f90"),
           os.path.join(fdir, "binding_demo.f90"), "-L", HERE, "-lcosahip", "-Wl,-rpath,$ORIGIN/..", "-o", exe]
    if verbose:
        print(" ".join(cmd))
    subprocess.check_call(cmd)
    return exe


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
