#!/usr/bin/env python3
"""Build oracle/_ref/libcosaref.so: COSA's own hot-path routines, compiled
from the sources where they lie under /root/reference/src, plus the bind(C)
driver oracle/ref_shim.f90.  TEST INFRASTRUCTURE ONLY.

Recipe (the reference's own Makefile is not used):
  * cosa_precision, common_variables, cosa_variables, util, flux, src and
    update .f90 compile unmodified with amdflang;
  * bc.f90, cosa.f90, output.f90, input.f90 and parallelutils.f90 do not
    compile as whole files in this image: bc.f90's periodic-cut routines and
    cutman_q need the parallelutils module (MPI plumbing), cosa.f90's main
    program and the readers and
    writers of input.f90 / output.f90 need the un-vendored CGNS library.  The subroutines this path
    needs are self-contained; they are compiled verbatim, selected by name
    into a temporary file under oracle/_ref/ that is deleted right after
    compilation (nothing from the reference is kept or committed);
  * flags: -cpp -fdefault-real-8 -fdefault-double-8 -O2 (the reference's
    real*8 promotion, src/Makefile:10-71) without -ffast-math, and
    -ffp-contract=off so no fused multiply-add changes the rounding.

Usage: python3 oracle/build_ref.py [--ref /root/reference] [--quiet]
Exit status 0 when built, 2 when the reference or amdflang is absent.
"""
import argparse
import os
import re
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_ref")

WHOLE = ["cosa_precision", "common_variables", "cosa_variables", "util", "flux", "src", "update"]
BY_NAME = {
    "bc.f90": ["bbc", "bc_far0", "bc_far1", "bc_insub", "bc_outsub", "bc_walinv_14", "bc_walvis_15",
               "bc_extrapol", "bc_sym", "cut_q"],
    # the block-ownership queries cut_q calls (external subroutines after the
    # parallelutils module in the same file; they need only cosa_variables)
    "parallelutils.f90": ["myblock", "getownerid"],
    "cosa.f90": ["bresid", "lambdab", "eig_eul", "eig_ns", "eig_eul_p", "eig_ns_p", "bdeltat", "cirs_b",
                 "tridi", "ccirs_b", "scalef_b", "scaleb_b", "resg_brms", "resg_brms_hb", "rest1_b", "rest2_ho_b", "rest2_lo_b", "prol_lin_b",
                 "prol_bln_b", "rest4_b", "coarsen_b"],
    "output.f90": ["force_b", "write_brest"],
    # grid motion and metric set-up: the producers of the metric inputs
    "grid.f90": ["imetrix_b", "bvols", "pitch", "cordex_b", "volex_b", "vol_bedges", "vmetrix_b", "cut_v",
                 "cut_x"],
    "input.f90": ["read_brest"],
}
FFLAGS = ["-cpp", "-fdefault-real-8", "-fdefault-double-8", "-O2", "-ffp-contract=off", "-fPIC"]
# the CPU baseline's build: the reference's optimisation level (-O3) with the
# compiler's default floating-point contraction, without -ffast-math
FFLAGS_TIMING = ["-cpp", "-fdefault-real-8", "-fdefault-double-8", "-O3", "-fPIC"]


def extract(path, names):
    """Return the text of the named external subroutines of `path`, each from
    its `subroutine NAME` line up to the line before the next `subroutine`
    statement (or end of file)."""
    with open(path) as f:
        lines = f.readlines()
    start = re.compile(r"^\s*subroutine\s+(\w+)", re.I)
    heads = [(i, m.group(1).lower()) for i, l in enumerate(lines) for m in [start.match(l)] if m]
    out = []
    for name in names:
        hit = [k for k, (i, n) in enumerate(heads) if n == name]
        if len(hit) != 1:
            raise SystemExit(f"build_ref: subroutine {name} not found once in {path}")
        k = hit[0]
        i0 = heads[k][0]
        i1 = heads[k + 1][0] if k + 1 < len(heads) else len(lines)
        out.extend(lines[i0:i1])
    return "".join(out)


def run(cmd, quiet):
    if not quiet:
        print(" ".join(cmd))
    subprocess.check_call(cmd)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    ap.add_argument("--quiet", action="store_true")
    a = ap.parse_args()
    src = os.path.join(a.ref, "src")
    fc = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    if not os.path.isdir(src) or not os.path.exists(fc):
        print("build_ref: reference sources or amdflang absent; oracle/_ref not built")
        return 2
    build_lib(src, fc, FFLAGS, "libcosaref.so", a.quiet)
    build_lib(src, fc, FFLAGS_TIMING, "libcosaref_o3.so", a.quiet)
    return 0


def build_lib(src, fc, flags, name, quiet):
    """Compile the reference routines and the shim with `flags` into
    oracle/_ref/<name>."""
    os.makedirs(OUT, exist_ok=True)
    objs = []
    for m in WHOLE:
        o = os.path.join(OUT, m + ".o")
        run([fc] + flags + ["-module-dir", OUT, "-c", os.path.join(src, m + ".f90"), "-o", o], quiet)
        objs.append(o)
    tmp = os.path.join(OUT, "tmp_src")
    os.makedirs(tmp, exist_ok=True)
    try:
        for fname, names in BY_NAME.items():
            part = os.path.join(tmp, "sel_" + fname)
            with open(part, "w") as f:
                f.write(extract(os.path.join(src, fname), names))
            o = os.path.join(OUT, "sel_" + fname.replace(".f90", ".o"))
            run([fc] + flags + ["-I", OUT, "-c", part, "-o", o], quiet)
            objs.append(o)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    o = os.path.join(OUT, "ref_shim.o")
    run([fc] + flags + ["-I", OUT, "-module-dir", OUT, "-c", os.path.join(HERE, "ref_shim.f90"), "-o", o],
        quiet)
    objs.append(o)
    lib = os.path.join(OUT, name)
    run([fc, "-shared", "-fPIC", "-o", lib] + objs, quiet)
    for f in os.listdir(OUT):
        if f.endswith(".o") or f.endswith(".mod"):
            os.remove(os.path.join(OUT, f))
    if not quiet:
        print("built", lib)


if __name__ == "__main__":
    sys.exit(main())
