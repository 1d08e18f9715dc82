"""Parity of libcosahip (HIP, gfx950) with the oracle, through the C ABI.

Every reference routine of the stage is run on the device and on the CPU
restatement from the same state, and compared at the FP64 bar
|gpu - cpu| <= 1e-12 (|cpu| + 1e-14 max|cpu|); then whole smooth() calls and
multigrid cycles.  The kernels keep the reference's operation order and
are built without FMA contraction, so most results are bit-identical; the
bar absorbs the device pow (farfield BC, CIRS coefficients).
"""
import ctypes

import numpy as np
import pytest

from cosa_amd import case as cases
from cosa_amd import _lib
from oracle_bind import dp
from parity_util import interior_mask, make_pair, parity, small_cases

pytestmark = pytest.mark.gpu

CASES = small_cases()


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(params=CASES, ids=[c.name for c in CASES])
def pair(request):
    c = request.param
    q0 = c.perturbed(amp=0.02, seed=11)
    _say(f"case {c.name} ncell={c.ncell()} nsnap={c.nsnap}")
    o, oc, g = make_pair(c, q0)
    yield c, o, oc, g
    g.close()


def test_stage_routines(pair):
    """One RK stage, routine by routine (src/cosa.f90:1719-2001)."""
    c, o, oc, g = pair
    A = oc[1]
    cp = oc.ptr()
    # cons2prim, bc, cutman_q
    o.or_cons2prim(cp, 1, dp(A["q"]))
    g.call("cons2prim", 1, _lib.Q)
    parity(g.download("q"), A["q"], what="cons2prim")
    o.or_bc(cp, 1, dp(A["q"]))
    g.call("bc", 1, _lib.Q)
    parity(g.download("q"), A["q"], what="bc")
    o.or_cutman_q(cp, 1, dp(A["q"]))
    g.call("cutman_q", 1, _lib.Q)
    parity(g.download("q"), A["q"], what="cutman_q")
    g.call("prtest", 1, 1)
    assert o.or_prtest(cp, 1, None) == 0
    # lambda + deltat
    o.or_lambda_deltat(cp, 1)
    g.call("lambda_deltat", 1)
    parity(g.download("dtvol"), A["dtvol"], what="dtvol")
    if c.config["irsop"] >= 2:
        for f in ("betaxi", "betaeta", "betazeta"):
            parity(g.download(f), A[f], what=f)
    # residual: muscl + roflux + resid, three directions
    o.or_residual(cp, 1)
    g.call("residual", 1)
    m = interior_mask(c)
    rg, rc = g.download("res"), A["res"]
    parity(rg[m], rc[m], what="residual")
    assert np.all(rg[~m] == 0) and np.all(rc[~m] == 0)
    o.or_prim2cons(cp, 1, dp(A["q"]))
    g.call("prim2cons", 1, _lib.Q)
    parity(g.download("q"), A["q"], what="prim2cons")
    if c.config["harbal"]:
        o.or_src_hb(cp, 1)
        g.call("src_hb", 1)
        parity(g.download("res")[m], A["res"][m], what="src_hb")
    r_o = np.zeros(8)
    if c.config["harbal"]:
        o.or_copy_array(cp, 1, dp(A["work1"]), dp(A["res"]))
    o.or_resg_rms(cp, 1, dp(r_o))
    r_g = g.resg_rms(1)
    parity(r_g[:6], r_o[:6], rtol=1e-13, what="resg_rms")
    o.or_cirs(cp, 1)
    g.call("cirs", 1)
    parity(g.download("res")[m], A["res"][m], what="cirs")
    o.or_copy_array(cp, 1, dp(A["qold"]), dp(A["q"]))
    g.call("copy_array", 1, _lib.QOLD, _lib.Q)
    o.or_update(cp, 1, ctypes.c_double(0.333333))
    g.call("update", 1, ctypes.c_double(0.333333))
    parity(g.download("q"), A["q"], what="update")


@pytest.mark.parametrize("niter", [1, 3])
def test_smooth_single_level(pair, niter):
    c, o, oc, g = pair
    done = (ctypes.c_int32 * 1)()
    where = (ctypes.c_int32 * 8)()
    rc = o.or_smooth(oc.ptr(), 1, 0, 0, niter, 1, done, where)
    assert rc == 0
    d, r = g.smooth(1, False, False, niter, 1)
    parity(g.download("q"), oc[1]["q"], rtol=1e-11, what="smooth q")
    r_o = np.array([o.or_get_resrms(i) for i in range(1, 9)])
    parity(r[:6], r_o[:6], rtol=1e-12, what="smooth resrms")


def _mg_case(name):
    if name == "naca3":
        return cases.naca_case(65, 17, 9, nlevel=3, name="naca3", irsop=1, cfl_irs=3.5, cfl=2.0, nl_crs=2,
                               ncycle=6)
    if name == "blocks3_lin_ho":
        return cases.two_block_case(ni=17, nj=9, nk=9, nlevel=3, distort=0.1, name="blocks3", irsop=2,
                                    cfl_irs=3.0, cfl=2.0, ncycle=5, bln_prol=0, ho_rest=1)
    if name == "naca3_pitch_hb2":
        return cases.naca_case(65, 17, 9, nlevel=3, name="naca3_pitch", irsop=2, cfl_irs=3.0, cfl=2.0, nl_crs=2,
                               ncycle=5, nharms=2, harbal=1, omega=0.6, pitch=dict(dtheta0=0.05, xh=0.25, yh=0.0))
    if name == "blocks2_ccirs":      # CIRS with cut boundaries (cutcirs = 1), irsop 2
        return cases.two_block_case(ni=17, nj=9, nk=9, nlevel=2, distort=0.1, name="blocks2_ccirs", irsop=2,
                                    cfl_irs=3.0, cfl=2.0, ncycle=5, cutcirs=1)
    if name == "naca2_ccirs":        # ccirs with irsop 1 and the O-grid's seam cut
        return cases.naca_case(65, 17, 9, nlevel=2, name="naca2_ccirs", irsop=1, cfl_irs=3.5, cfl=2.0, nl_crs=2,
                               ncycle=5, cutcirs=1)
    return cases.two_block_case(ni=17, nj=9, nk=9, nlevel=2, distort=0.1, name="blocks2", irsop=3, psirs=0.125,
                                cfl_irs=3.0, cfl=2.0, ncycle=5, nharms=1, harbal=1, omega=0.8)


@pytest.mark.parametrize("name", ["naca3", "blocks3_lin_ho", "blocks2_hb1", "naca3_pitch_hb2", "blocks2_ccirs",
                                  "naca2_ccirs"])
def test_multigrid_cycles(name):
    """Whole V-cycles (mg, src/cosa.f90:806-930): resrms history per cycle
    and the final state."""
    c = _mg_case(name)
    q0 = c.perturbed(amp=0.01, seed=5)
    _say(f"mg case {name} levels={c.config['nlevel']} ncell={c.ncell()}")
    o, oc, g = make_pair(c, q0)
    try:
        ncyc = c.config["ncycle"]
        hist_o = np.zeros(8 * ncyc)
        where = (ctypes.c_int32 * 8)()
        n_o = o.or_mg(oc.ptr(), 1, 1, 1, 1, dp(hist_o), where)
        assert n_o == ncyc
        n_g = g.mg(ncyc)
        assert n_g == ncyc
        hist_g = np.array(g.history)
        hist_o = hist_o.reshape(ncyc, 8)
        parity(hist_g[:, :6].ravel(), hist_o[:, :6].ravel(), rtol=1e-10, what="resrms history")
        for nl in range(1, c.config["nlevel"] + 1):
            parity(g.download("q", nl), oc[nl]["q"], rtol=1e-10, what=f"q level {nl}")
    finally:
        g.close()


def test_edges_corners_restrict_prolong():
    c = cases.two_block_case(ni=17, nj=9, nk=9, nlevel=2, distort=0.1, name="mgt", bln_prol=1)
    q0 = c.perturbed(amp=0.05, seed=9)
    o, oc, g = make_pair(c, q0)
    try:
        A1, A2 = oc[1], oc[2]
        cp = oc.ptr()
        o.or_edges_corners(cp, 1, dp(A1["q"]))
        g.call("edges_corners", 1, _lib.Q)
        parity(g.download("q"), A1["q"], what="edges_corners")
        res = np.random.default_rng(3).uniform(-1, 1, A1["res"].size)
        A1["res"][:] = res
        g.upload("res", res, 1)
        o.or_restrict(cp, 1)
        g.restrict(1)
        parity(g.download("q", 2), A2["q"], what="rest1")
        parity(g.download("rhs", 2), A2["rhs"], what="rest2")
        q2 = A2["q"] * (1 + np.random.default_rng(4).uniform(-0.01, 0.01, A2["q"].size))
        A2["q"][:] = q2
        g.upload("q", q2, 2)
        o.or_prolong(cp, 1)
        g.prolong(1)
        parity(g.download("q", 2), A2["q"], what="prolong coarse")
        parity(g.download("qold", 2), A2["qold"], what="prolong qold")
        parity(g.download("q", 1), A1["q"], what="prolong fine")
    finally:
        g.close()


def test_positivity_failure_reported():
    """prtest: a negative pressure aborts smooth with the reference's
    location; the oracle finds the same cell (src/cosa.f90:960-1120)."""
    c = cases.two_block_case(ni=9, nj=9, nk=5, nlevel=1, name="neg")
    q0 = c.freestream()
    L = c.levels[0]
    n3 = int(L.points(3)[0])
    # block 1, cell (4,3,2): E so small that p < 0 after cons2prim
    idx = (4 + 1) + (L.imax[0] + 3) * ((3 + 1) + (L.jmax[0] + 3) * (2 + 1))
    q0[4 * n3 + idx] = 0.01
    o, oc, g = make_pair(c, q0)
    try:
        where = (ctypes.c_int32 * 8)()
        done = (ctypes.c_int32 * 1)()
        rc = o.or_smooth(oc.ptr(), 1, 0, 0, 1, 1, done, where)
        assert rc == 5
        with pytest.raises(_lib.CosahipError) as e:
            g.smooth(1, False, False, 1, 1)
        assert e.value.code == 5
        info = e.value.info
        assert info[:5] == [1, 1, 4, 3, 2]
        assert info[6] == 1
        assert list(where)[:5] == [1, 1, 4, 3, 2]
    finally:
        g.close()


def test_freestream_preserved():
    """Uniform flow on a closed-cell grid gives a zero residual (face
    vectors of every cell sum to zero)."""
    c = cases.two_block_case(ni=13, nj=9, nk=9, distort=0.3, wall=30, alpha_deg=0.0, name="fs")
    o, oc, g = make_pair(c, c.freestream())
    try:
        g.call("cons2prim", 1, _lib.Q)
        g.call("bc", 1, _lib.Q)
        g.call("cutman_q", 1, _lib.Q)
        g.call("residual", 1)
        r = g.download("res")
        assert np.max(np.abs(r)) < 1e-13
    finally:
        g.close()


def test_unsupported_options_fail_loudly():
    c = cases.two_block_case(ni=9, nj=9, nk=5, nlevel=1, name="bad")
    c.config["npde"] = 7
    from cosa_amd.solver import Solve
    with pytest.raises(_lib.CosahipError):
        Solve(c)
