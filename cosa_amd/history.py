"""Convergence history files in the reference's text format.

COSA's ``output`` (src/output.f90:4889-5005) writes, for pseudo-time
iteration ``it`` of physical step ``itime``:

* ``hist.dat`` (unit 37): ``itime, it, resrms(1:5)`` when ``it == 1``, ``it``
  is a multiple of ``iprint``, the solve has converged (``done``) or
  ``it == ncycle`` (and ``it != 0``);
* ``histur.dat`` (unit 41, k-omega only): the same lines with
  ``resrms(6:npde)``;
* ``histg.dat`` (unit 39): one line at convergence or at ``ncycle``.

Lines use format 102, ``(i7,2x,i7,5(1x,e13.5))`` (src/output.f90:5002);
the files are opened with status 'replace' (src/cosa.f90:107-113).  The
Fortran E13.5 edit descriptor prints a mantissa in [0.1, 1) with five
digits, ``E`` and a signed two-digit exponent (``0.12346E-04``); a three-
digit exponent drops the ``E`` (``0.12346-100``), and a value that does not
fit prints 13 asterisks.
"""
from __future__ import annotations

import math
import os
from typing import Optional, Sequence


def fortran_e(x: float, w: int = 13, d: int = 5) -> str:
    """Fortran Ew.d output editing of one real (default rounding: nearest)."""
    x = float(x)
    if math.isnan(x):
        return "NaN".rjust(w)
    if math.isinf(x):
        s = "-Infinity" if x < 0 else "Infinity"
        return (s if len(s) <= w else s[:-5]).rjust(w) if len(s) <= w or w >= 3 else "*" * w
    neg = math.copysign(1.0, x) < 0
    ax = abs(x)
    if ax == 0.0:
        mant, expo = "0" * d, 0
    else:
        # d significant digits, correctly rounded: m.mmmm e+XX -> 0.mmmmm e+(XX+1)
        s = f"{ax:.{d - 1}e}"
        digits, e = s.split("e")
        mant = digits.replace(".", "")
        expo = int(e) + 1
    if abs(expo) <= 99:
        tail = f"E{'-' if expo < 0 else '+'}{abs(expo):02d}"
    elif abs(expo) <= 999:
        tail = f"{'-' if expo < 0 else '+'}{abs(expo):03d}"
    else:
        return "*" * w
    body = ("-" if neg else "") + "0." + mant + tail
    if len(body) > w:
        body = body.replace("0.", ".", 1)
        if len(body) > w:
            return "*" * w
    return body.rjust(w)


def fortran_i(n: int, w: int) -> str:
    s = str(int(n))
    return s.rjust(w) if len(s) <= w else "*" * w


def line_102(itime: int, it: int, values: Sequence[float]) -> str:
    """One line of format 102: (i7,2x,i7,5(1x,e13.5)); fewer than five
    values end the line early, as the Fortran list does."""
    out = fortran_i(itime, 7) + "  " + fortran_i(it, 7)
    for v in list(values)[:5]:
        out += " " + fortran_e(v)
    return out


def line_100(itime: int, values: Sequence[float]) -> str:
    """One line of format 100: (i8,4(1x,e13.5)) (the time-stepping print of
    src/cosa.f90:4169)."""
    out = fortran_i(itime, 8)
    for v in list(values)[:4]:
        out += " " + fortran_e(v)
    return out


class HistoryWriter:
    """hist.dat / histur.dat / histg.dat of one solve, written as the
    reference's ``output`` writes them.  ``record(it, resrms, done)`` takes
    resrms(1:8) of pseudo-time iteration ``it`` (1-based) and flushes every
    line it writes, as the reference flushes its units."""

    def __init__(self, directory: str, npde: int, iprint: int = 1, ncycle: int = 100, kom: bool = False,
                 itime: int = 0):
        if iprint < 1:
            raise ValueError("iprint must be >= 1")
        os.makedirs(directory, exist_ok=True)
        self.npde, self.iprint, self.ncycle, self.kom, self.itime = npde, iprint, ncycle, kom, itime
        self.hist = open(os.path.join(directory, "hist.dat"), "w")
        self.histg = open(os.path.join(directory, "histg.dat"), "w")
        self.histur: Optional[object] = open(os.path.join(directory, "histur.dat"), "w") if kom else None

    def record(self, it: int, resrms: Sequence[float], done: bool = False):
        if it == 0:
            return
        r = list(resrms)
        if it == 1 or it % self.iprint == 0 or done or it == self.ncycle:
            self.hist.write(line_102(self.itime, it, r[0:5]) + "\n")
            self.hist.flush()
            if self.histur is not None:
                self.histur.write(line_102(self.itime, it, r[5:self.npde]) + "\n")
                self.histur.flush()
        if done or it == self.ncycle:
            self.histg.write(line_102(self.itime, it, r[0:5]) + "\n")
            self.histg.flush()

    def close(self):
        for f in (self.hist, self.histg, self.histur):
            if f is not None and not f.closed:
                f.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
