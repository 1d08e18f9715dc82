"""Harmonic balance on the GPU beyond the explicit Euler stage, against the
oracle through the C ABI: the implicit stage rkim (per-cell LU of the
(2N+1)x(2N+1) spectral system, update_bhb src/update.f90:1410-1841) without
IRS on one level and with IRS and multigrid, Wilcox k-omega with harmonic
balance (explicit and rkim, lim_res_ome_bhb at the restriction stage).

The LU is the same sequence of operations on both sides, so the stage
update itself is bit-identical; whole cycles pass through the device pow
(farfield, Sutherland) and are held to the FP64 bar.
"""
import ctypes

import numpy as np
import pytest

from cosa_amd import _lib
from cosa_amd.config import komega
from hb_cases import HB_CASES, hb_case
from oracle_bind import dp
from parity_util import interior_mask, make_pair, parity

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("name", ["euler_rkim_plain", "euler_rkim_mg", "rans_hb_rkim", "rans_hb_explicit"])
def test_rkim_stage_routines(name):
    """One stage routine by routine (src/cosa.f90:1719-2016) up to the
    update: the HB term kept for rkim, the restriction-stage limiter of
    k-omega with harmonic balance, CIRS, the implicit (or explicit) update."""
    c = hb_case(name)
    q0 = c.perturbed(amp=0.01, seed=11)
    kom = komega(c.config)
    plain = c.config["irsop"] == 0 and c.config["nlevel"] == 1
    _say(f"stage {name} ncell={c.ncell()} nsnap={c.nsnap} npde={c.npde} rkimp={c.config['rkimp']}")
    o, oc, g = make_pair(c, q0)
    try:
        A = oc[1]
        cp = oc.ptr()
        m = interior_mask(c)
        o.or_copy_array(cp, 1, dp(A["qold"]), dp(A["q"]))
        g.call("copy_array", 1, _lib.QOLD, _lib.Q)
        o.or_cons2prim(cp, 1, dp(A["q"]))
        g.call("cons2prim", 1, _lib.Q)
        o.or_bc(cp, 1, dp(A["q"]))
        g.call("bc", 1, _lib.Q)
        o.or_cutman_q(cp, 1, dp(A["q"]))
        g.call("cutman_q", 1, _lib.Q)
        if kom:
            o.or_tur_vis(cp, 1)
            g.call("tur_vis", 1, _lib.Q)
        o.or_lambda_deltat(cp, 1)
        g.call("lambda_deltat", 1)
        o.or_residual(cp, 1)
        g.call("residual", 1)
        o.or_prim2cons(cp, 1, dp(A["q"]))
        g.call("prim2cons", 1, _lib.Q)
        if not plain:
            o.or_src_hb(cp, 1)
            g.call("src_hb", 1)
        parity(g.download("res")[m], A["res"][m], what="res (with the HB term unless rkim without IRS/MG)")
        if kom:
            o.or_turb_limits(cp, 1, ctypes.c_double(0.25))
            g.call("turb_limits", 1, ctypes.c_double(0.25))
            parity(g.download("res")[m], A["res"][m], what="lim_res_ome_bhb")
        if c.config["irsop"]:
            o.or_cirs(cp, 1)
            g.call("cirs", 1)
            parity(g.download("res")[m], A["res"][m], what="cirs")
        o.or_update(cp, 1, ctypes.c_double(0.5))
        g.call("update", 1, ctypes.c_double(0.5))
        parity(g.download("q")[m], A["q"][m], what="update (rkim)" if c.config["rkimp"] else "update")
    finally:
        g.close()


@pytest.mark.parametrize("name", HB_CASES)
def test_hb_implicit_and_komega_cycles(name):
    c = hb_case(name)
    q0 = c.perturbed(amp=0.01, seed=5)
    _say(f"hb case {name} levels={c.config['nlevel']} ncell={c.ncell()} nsnap={c.nsnap} npde={c.npde} "
         f"rkimp={c.config['rkimp']} irsop={c.config['irsop']}")
    o, oc, g = make_pair(c, q0)
    try:
        ncyc = c.config["ncycle"]
        hist_o = np.zeros(8 * ncyc)
        where = (ctypes.c_int32 * 8)()
        n_o = o.or_mg(oc.ptr(), 1, 1, 1, 1, dp(hist_o), where)
        assert n_o == ncyc, list(where)
        n_g = g.mg(ncyc)
        assert n_g == ncyc
        npde = c.npde
        hist_g = np.array(g.history)
        hist_o = hist_o.reshape(ncyc, 8)
        parity(hist_g[:, :npde + 1].ravel(), hist_o[:, :npde + 1].ravel(), rtol=1e-10, what="resrms history")
        for nl in range(1, c.config["nlevel"] + 1):
            parity(g.download("q", nl), oc[nl]["q"], rtol=1e-10, what=f"q level {nl}")
        _say(f"  resrms(1) {hist_g[0, 0]:.4e} -> {hist_g[-1, 0]:.4e}")
    finally:
        g.close()
