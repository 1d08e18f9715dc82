"""hist.dat / histur.dat / histg.dat in the reference's format: the line
format against amdflang's own output of format 102 (fixture written by
tests/golden/make_hist_golden.py), and the reference's print conditions
(src/output.f90:4946-4969)."""
import json
import os

import numpy as np

from cosa_amd.history import HistoryWriter, fortran_e, line_102

HERE = os.path.dirname(os.path.abspath(__file__))


def test_format_102_matches_fortran():
    g = json.load(open(os.path.join(HERE, "golden", "hist_format102.json")))
    vals = np.array(g["bits"], np.int64).view(np.float64)
    rows = len(vals) // 5
    assert len(g["lines"]) == rows + 1
    for r in range(rows):
        assert line_102(37 * r, 1 + 997 * r, vals[5 * r:5 * r + 5]) == g["lines"][r], r
    assert line_102(0, 1, vals[:2]) == g["lines"][-1]


def test_fortran_e_edge_cases():
    assert fortran_e(0.0) == "  0.00000E+00"
    assert fortran_e(-0.0) == " -0.00000E+00"
    assert fortran_e(9.999995) == "  0.10000E+02"       # rounds up into the next decade
    assert fortran_e(1e-100) == "  0.10000E-99"
    assert fortran_e(1e300) == "  0.10000+301"          # three-digit exponent drops the E
    assert len(fortran_e(-1.7976931348623157e308)) == 13


def test_history_print_conditions(tmp_path):
    d = str(tmp_path)
    r = [1e-3, 2e-3, 3e-3, 4e-3, 5e-3, 6e-3, 7e-3, 8e-3]
    with HistoryWriter(d, npde=7, iprint=3, ncycle=8, kom=True) as h:
        for it in range(1, 9):
            h.record(it, r, done=False)
    hist = open(os.path.join(d, "hist.dat")).read().splitlines()
    its = [int(l.split()[1]) for l in hist]
    assert its == [1, 3, 6, 8]                  # it = 1, multiples of iprint, it = ncycle
    assert hist[0] == line_102(0, 1, r[:5])
    tur = open(os.path.join(d, "histur.dat")).read().splitlines()
    assert tur[0] == line_102(0, 1, r[5:7]) and len(tur) == 4
    g = open(os.path.join(d, "histg.dat")).read().splitlines()
    assert g == [line_102(0, 8, r[:5])]


def test_history_converged_line(tmp_path):
    d = str(tmp_path)
    with HistoryWriter(d, npde=5, iprint=10, ncycle=100) as h:
        for it in range(1, 5):
            h.record(it, [float(it)] * 8, done=it == 4)
    its = [int(l.split()[1]) for l in open(os.path.join(d, "hist.dat"))]
    assert its == [1, 4]
    assert not os.path.exists(os.path.join(d, "histur.dat"))
    assert [int(l.split()[1]) for l in open(os.path.join(d, "histg.dat"))] == [4]
