"""Fixture for the hist.dat line format: a Fortran program of this script's
own (not reference code) prints values with the reference's format 102,
(i7,2x,i7,5(1x,e13.5)) (src/output.f90:5002), compiled with amdflang -- the
compiler the reference builds with here -- and the lines are stored with
the values' bit patterns in tests/golden/hist_format102.json.

    python tests/golden/make_hist_golden.py
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def values():
    rng = np.random.default_rng(102)
    v = [0.0, -0.0, 1.0, -1.0, 0.5, 0.1, 9.999995, 9.999994999, 0.999995, 1.2345649999999999e-5, 0.123455,
         0.1234550000000001, 123456.7, 1e-99, 1e-100, 9.99995e-100, 1e300, -2.5e-310, 5e-324, 1.7976931348623157e308,
         2.220446049250313e-16, 3.3e-8, 1e-10]
    v += list(10.0 ** rng.uniform(-40, 5, 60) * rng.choice([-1, 1], 60))
    v += list(np.round(rng.uniform(0, 1, 20), 6))   # near-ties in the sixth digit
    while len(v) % 5:
        v.append(0.25)
    return [float(x) for x in v]


def main():
    fc = "/opt/rocm/bin/amdflang"
    vals = values()
    bits = [int(np.float64(x).view(np.int64)) for x in vals]
    rows = len(vals) // 5
    src = ["program hist102", "  integer(8) :: b(%d)" % len(vals), "  real(8) :: a(%d)" % len(vals),
           "  integer :: r"]
    for i, x in enumerate(bits):
        src.append("  b(%d) = %d_8" % (i + 1, x))
    src += ["  a = transfer(b, a)", "  do r = 0, %d" % (rows - 1),
            "    write(*,102) r * 37, 1 + r * 997, a(5 * r + 1:5 * r + 5)", "  end do",
            "  write(*,102) 0, 1, a(1:2)",
            "102 format(i7,2x,i7,5(1x,e13.5))", "end program hist102"]
    with tempfile.TemporaryDirectory() as d:
        f = os.path.join(d, "h.f90")
        open(f, "w").write("\n".join(src) + "\n")
        exe = os.path.join(d, "h")
        subprocess.check_call([fc, "-O0", f, "-o", exe])
        out = subprocess.check_output([exe]).decode().splitlines()
    json.dump({"bits": bits, "lines": out, "format": "(i7,2x,i7,5(1x,e13.5))",
               "rows": "itime = 37 r, it = 1 + 997 r, values 5r+1..5r+5; last line: 0, 1, values 1..2"},
              open(os.path.join(HERE, "hist_format102.json"), "w"), indent=0)
    print(f"{len(out)} lines", file=sys.stderr)


if __name__ == "__main__":
    main()
