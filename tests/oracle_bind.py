"""ctypes bindings of the checkers (test infrastructure):

* ``Oracle``  -- oracle/liboracle.so, the C restatement of the reference path
* ``Ref``     -- oracle/_ref/libcosaref.so, the reference's own compiled
                 routines driven through oracle/ref_shim.f90 (built only
                 where /root/reference exists)
* ``OrCase``  -- the or_case struct over numpy arrays of a cosa_amd Case
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from cosa_amd.config import CosahipConfig, komega, to_struct

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libcosaref.so")

DP = ctypes.POINTER(ctypes.c_double)
IP = ctypes.POINTER(ctypes.c_int32)


def dp(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(DP)


def ip(a):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(IP)


def _build_oracle():
    srcs = [os.path.join(ROOT, "oracle", n) for n in ("cosa_oracle.c", "grid_oracle.c", "cosa_oracle.h")]
    if (not os.path.exists(ORACLE_SO)) or os.path.getmtime(ORACLE_SO) < max(os.path.getmtime(f) for f in srcs):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "liboracle.so"])


class OrLevel(ctypes.Structure):
    _fields_ = [("nblocks", ctypes.c_int32), ("ncuts", ctypes.c_int32),
                ("imax", IP), ("jmax", IP), ("kmax", IP), ("nbcs", IP), ("bctopo", IP), ("cutopo", IP)] + [
        (n, DP) for n in ("q", "qold", "qp", "qm", "dq", "dqp", "dqm", "flux", "res", "rhs", "work1", "si", "sj",
                          "sk", "vol", "dtvol", "betaxi", "betaeta", "betazeta", "xdot", "ydot", "zdot", "pr")] + [
        (n, DP) for n in ("mut", "dist", "prd", "divvel", "delpls", "dtvolt", "tbetaxi", "tbetaeta", "tbetazeta")] + [
        ("xd", DP * 9), ("ttrm", DP), ("q1", DP), ("q2", DP)]


class OrCaseS(ctypes.Structure):
    _fields_ = [("nlevel", ctypes.c_int32), ("pad", ctypes.c_int32), ("lev", OrLevel * 4)]


class Oracle:
    def __init__(self):
        _build_oracle()
        self.lib = ctypes.CDLL(ORACLE_SO)
        L = self.lib
        L.or_get_resrms.restype = ctypes.c_double
        L.or_set_resrms.argtypes = [ctypes.c_int, ctypes.c_double]
        for n in ("or_smooth", "or_mg", "or_prtest", "or_bprtest"):
            getattr(L, n).restype = ctypes.c_int

    def set_config(self, cfg: dict, dhb=None):
        self._cfg = to_struct(cfg)
        self.lib.or_set_config(ctypes.byref(self._cfg))
        if dhb is not None:
            d = np.ascontiguousarray(np.asarray(dhb, np.float64).flatten(order="F"))
            self.lib.or_set_dhb(ctypes.c_int(dhb.shape[0]), dp(d))

    def __getattr__(self, name):
        return getattr(self.lib, name)


_REF = None


def ref_available():
    return os.path.exists(REF_SO)


class Ref:
    def __init__(self):
        self.lib = ctypes.CDLL(REF_SO, mode=os.RTLD_LAZY)
        assert self.lib.ref_cfg_size() == ctypes.sizeof(CosahipConfig)

    def set_config(self, cfg: dict, dhb=None):
        self._cfg = to_struct(cfg)
        self.lib.ref_set_config(ctypes.byref(self._cfg))
        if dhb is not None:
            d = np.ascontiguousarray(np.asarray(dhb, np.float64).flatten(order="F"))
            self.lib.ref_set_dhb(ctypes.c_int(dhb.shape[0]), dp(d))

    def __getattr__(self, name):
        return getattr(self.lib, name)


class OrCase:
    """Host arrays of every level of a Case, wired into an or_case struct."""

    FIELDS = ("q", "qold", "qp", "qm", "dq", "dqp", "dqm", "flux", "res", "rhs", "work1", "si", "sj", "sk", "vol",
              "dtvol", "betaxi", "betaeta", "betazeta", "xdot", "ydot", "zdot", "pr")

    def __init__(self, case, q1=None):
        self.case = case
        self.s = OrCaseS()
        self.s.nlevel = case.config["nlevel"]
        self.arrays = []
        self.keep = []
        for li, L in enumerate(case.levels):
            nl = li + 1
            g = case.geom[li]
            a = {}
            for n in ("q", "qold", "qp", "qm", "dq", "dqp", "dqm", "res", "rhs", "work1"):
                a[n] = np.zeros(case.size(nl, "p3"))
            a["flux"] = np.zeros(case.size(nl, "flux"))
            a["pr"] = np.zeros(case.size(nl, "p3s"))
            for n in ("si", "sj", "sk"):
                a[n] = g[n].copy()
            a["vol"] = g["vol"].copy()
            for n in ("dtvol", "betaxi", "betaeta", "betazeta"):
                a[n] = np.zeros(case.size(nl, "p1"))
            for n in ("xdot", "ydot", "zdot"):
                a[n] = g[n].copy() if n in g else np.zeros(case.size(nl, "x2"))
            a["q"][:] = case.freestream(nl)
            if nl == 1 and q1 is not None:
                a["q"][:] = q1
            lev = self.s.lev[li]
            lev.nblocks = L.nblocks
            lev.ncuts = L.ncuts
            ints = {n: np.ascontiguousarray(getattr(L, n), np.int32) for n in
                    ("imax", "jmax", "kmax", "nbcs", "bctopo", "cutopo")}
            if ints["cutopo"].size == 0:
                ints["cutopo"] = np.zeros(21, np.int32)
            self.keep.append(ints)
            for n, v in ints.items():
                setattr(lev, n, ip(v))
            for n in self.FIELDS:
                setattr(lev, n, dp(a[n]))
            c = case.config
            if c["viscous"] or komega(c):
                for n in ("mut", "prd", "divvel", "delpls"):
                    a[n] = g[n].copy() if n in g else np.zeros(case.size(nl, "p3s"))
                if komega(c) and nl == 1:
                    a["mut"] = case.turbulent_viscosity(a["q"], 1)
                for n in ("dtvolt", "tbetaxi", "tbetaeta", "tbetazeta"):
                    a[n] = np.zeros(case.size(nl, "p1"))
                a["dist"] = g["dist"].copy()
                a["xd"] = [x.copy() for x in g["xd"]]
                for n in ("mut", "dist", "prd", "divvel", "delpls", "dtvolt", "tbetaxi", "tbetaeta", "tbetazeta"):
                    setattr(lev, n, dp(a[n]))
                for m in range(9):
                    lev.xd[m] = dp(a["xd"][m])
                a["ttrm"] = np.zeros(case.size(nl, "p3s") * 3)
                lev.ttrm = dp(a["ttrm"])
            if nl == 1:   # dual time stepping: q^n, q^(n-1)
                for n in ("q1", "q2"):
                    a[n] = np.zeros(case.size(nl, "p3"))
                    setattr(lev, n, dp(a[n]))
            self.arrays.append(a)

    def ptr(self):
        return ctypes.byref(self.s)

    def __getitem__(self, nl):
        return self.arrays[nl - 1]


def or_force(o: "Oracle", case, q: np.ndarray, nsurface: int = 1, inv14: bool = True, lngth=(1.0, 1.0, 1.0),
             xyzmom=(0.25, 0.0, 0.0), nl: int = 1) -> np.ndarray:
    """force (src/output.f90:4290-4330) on the oracle: or_force_b per block
    of level nl, the block coefficients summed in block order.  q is the
    level's conservative state; returns (nsurface, nsnap, 12)."""
    L = case.levels[nl - 1]
    g = case.geom[nl - 1]
    c = case.config
    nsn, nsm, npde, lmet = case.nsnap, case.nsnap_met, case.npde, c["lmet"]
    o3, o2, o0 = L.offsets(3), L.offsets(2), L.offsets(0)
    p3, p2, p0 = L.points(3), L.points(2), L.points(0)
    zero2 = np.zeros(case.size(nl, "x2"))
    zero0 = np.zeros(case.size(nl, "d0"))
    xd = g.get("xd")
    der = (xd[0], xd[4], xd[8]) if xd is not None else (zero0, zero0, zero0)
    vel = (g.get("xdot", zero2), g.get("ydot", zero2), g.get("zdot", zero2))
    momr = np.ascontiguousarray(case.moment_centres(xyzmom), np.float64)
    lg = np.ascontiguousarray(lngth, np.float64)
    total = np.zeros((nsurface, nsn, 12))
    bt_off = 0
    for b in range(L.nblocks):
        def sl(a, off, size):
            return np.ascontiguousarray(a[off:off + size])
        qb = sl(q, o3[b] * npde * nsn, p3[b] * npde * nsn)
        xyz = [sl(g[k], o2[b] * nsm, p2[b] * nsm) for k in ("x", "y", "z")]
        v = [sl(a, o2[b] * nsm, p2[b] * nsm) for a in vel]
        m = [sl(g[k], o2[b] * lmet * nsm, p2[b] * lmet * nsm) for k in ("si", "sj", "sk")]
        d = [sl(a, o0[b] * 3 * nsm, p0[b] * 3 * nsm) for a in der]
        nb = int(L.nbcs[b])
        bt = np.ascontiguousarray(L.bctopo[bt_off:bt_off + 10 * nb], np.int32)
        bt_off += 10 * nb
        cb = np.zeros(12 * nsn * nsurface)
        o.or_force_b(dp(qb), *[dp(a) for a in xyz], *[dp(a) for a in v], *[dp(a) for a in m],
                     *[dp(a) for a in d], nb, ip(bt), nsurface, int(inv14), dp(lg), dp(momr), dp(cb),
                     int(L.imax[b]), int(L.jmax[b]), int(L.kmax[b]))
        total = total + cb.reshape(nsurface, nsn, 12)
    return total
