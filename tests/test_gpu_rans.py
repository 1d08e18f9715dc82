"""Parity of the viscous / Wilcox k-omega path of libcosahip (HIP, gfx950)
with the oracle, through the C ABI: every routine of a RANS stage in turn,
then whole smooth() calls and multigrid cycles.  Same FP64 bar as
test_gpu_parity.py, |gpu - cpu| <= 1e-12 (|cpu| + 1e-14 max|cpu|); the
Sutherland law's t**1.5 and the wall density's (pw/p1)**(1/gamma) go
through the device pow, everything else keeps the reference's operation
order."""
import ctypes

import numpy as np
import pytest

from cosa_amd import _lib
from cosa_amd import case as cases
from cosa_amd.config import komega, kw_turbulence
from oracle_bind import dp
from parity_util import TURB, interior_mask, make_pair, parity, rans_cases

pytestmark = pytest.mark.gpu

CASES = rans_cases()


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(params=CASES, ids=[c.name for c in CASES])
def pair(request):
    c = request.param
    q0 = c.perturbed(amp=0.02, seed=11)
    _say(f"case {c.name} ncell={c.ncell()} npde={c.npde}")
    o, oc, g = make_pair(c, q0)
    yield c, o, oc, g
    g.close()


def test_rans_stage_routines(pair):
    """One RK stage of the viscous / k-omega path, routine by routine
    (src/cosa.f90:1719-2016)."""
    c, o, oc, g = pair
    kom = komega(c.config)
    A = oc[1]
    cp = oc.ptr()
    o.or_cons2prim(cp, 1, dp(A["q"]))
    g.call("cons2prim", 1, _lib.Q)
    parity(g.download("q"), A["q"], what="cons2prim")
    o.or_bc(cp, 1, dp(A["q"]))
    g.call("bc", 1, _lib.Q)
    parity(g.download("q"), A["q"], what="bc (viscous walls)")
    o.or_cutman_q(cp, 1, dp(A["q"]))
    g.call("cutman_q", 1, _lib.Q)
    parity(g.download("q"), A["q"], what="cutman_q")
    if kom:
        o.or_tur_vis(cp, 1)
        g.call("tur_vis", 1, _lib.Q)
        parity(g.download("mut"), A["mut"], what="tur_vis (primitive q)")
    o.or_lambda_deltat(cp, 1)
    g.call("lambda_deltat", 1)
    parity(g.download("dtvol"), A["dtvol"], what="dtvol (viscous radii)")
    if kom:
        parity(g.download("dtvolt"), A["dtvolt"], what="dtvolt")
    if c.config["irsop"] >= 2:
        for f in ("betaxi", "betaeta", "betazeta"):
            parity(g.download(f), A[f], what=f)
    o.or_residual(cp, 1)
    g.call("residual", 1)
    m = interior_mask(c)
    parity(g.download("res")[m], A["res"][m], what="residual (inviscid + viscous + sources)")
    if kom:
        ms = interior_mask(c, ncomp=1)
        for f in ("prd", "divvel", "delpls"):
            parity(g.download(f)[ms], A[f][ms], what=f)
    o.or_prim2cons(cp, 1, dp(A["q"]))
    g.call("prim2cons", 1, _lib.Q)
    parity(g.download("q"), A["q"], what="prim2cons")
    if kom:
        o.or_turb_limits(cp, 1, ctypes.c_double(0.25))
        g.call("turb_limits", 1, ctypes.c_double(0.25))
        parity(g.download("res")[m], A["res"][m], what="lim_res_ome")
    o.or_cirs(cp, 1)
    g.call("cirs", 1)
    parity(g.download("res")[m], A["res"][m], what="cirs (k-omega coefficients)")
    o.or_copy_array(cp, 1, dp(A["qold"]), dp(A["q"]))
    g.call("copy_array", 1, _lib.QOLD, _lib.Q)
    o.or_update(cp, 1, ctypes.c_double(0.5))
    g.call("update", 1, ctypes.c_double(0.5))
    parity(g.download("q"), A["q"], what="update (point-implicit k-omega)")
    if kom:
        o.or_fixq(cp, 1)
        g.call("fixq", 1)
        parity(g.download("q"), A["q"], what="fixq")
        o.or_tur_vis(cp, 1)
        g.call("tur_vis", 1, _lib.Q)
        parity(g.download("mut"), A["mut"], what="tur_vis (conservative q)")


@pytest.mark.parametrize("niter", [1, 3])
def test_rans_smooth_single_level(pair, niter):
    c, o, oc, g = pair
    done = (ctypes.c_int32 * 1)()
    where = (ctypes.c_int32 * 8)()
    assert o.or_smooth(oc.ptr(), 1, 0, 0, niter, 1, done, where) == 0
    d, r = g.smooth(1, False, False, niter, 1)
    parity(g.download("q"), oc[1]["q"], rtol=1e-11, what="smooth q")
    npde = c.npde
    r_o = np.array([o.or_get_resrms(i) for i in range(1, 9)])
    parity(r[:npde + 1], r_o[:npde + 1], rtol=1e-12, what="smooth resrms")
    if komega(c.config):
        parity(g.download("mut"), oc[1]["mut"], rtol=1e-11, what="smooth mut")


def _mg_case(name):
    if name == "rans_blocks2":
        return cases.two_block_case(ni=17, nj=9, nk=9, nlevel=2, distort=0.1, name="rans_blocks2", wall=1500,
                                    irsop=1, cfl_irs=2.5, cfl=1.5, ncycle=4, nl_crs=2,
                                    **kw_turbulence(0.5, 1e5, tu=1e-3, mutfar=0.1, **TURB))
    if name == "sst_blocks2":    # SST mut on the coarse level's first visit
        return cases.two_block_case(ni=17, nj=9, nk=9, nlevel=2, distort=0.1, name="sst_blocks2", wall=1500,
                                    irsop=1, cfl_irs=2.5, cfl=1.5, ncycle=4, nl_crs=2,
                                    **kw_turbulence(0.5, 1e5, tu=1e-3, mutfar=0.1, model="sst", **TURB))
    if name == "bsl_hb_rkim2":   # BSL with harmonic balance and the implicit stage
        return cases.two_block_case(ni=17, nj=9, nk=9, nlevel=2, distort=0.1, name="bsl_hb_rkim2", wall=1500,
                                    irsop=1, cfl_irs=2.5, cfl=1.5, ncycle=3, nl_crs=2, nharms=1, harbal=1,
                                    omega=0.8, rkimp=1,
                                    **kw_turbulence(0.5, 1e5, tu=1e-3, mutfar=0.1, model="bsl", **TURB))
    if name == "rans_ramp2":     # CFL ramping (ramping2) across the cycles
        return cases.two_block_case(ni=17, nj=9, nk=9, nlevel=2, distort=0.1, name="rans_ramp2", wall=1500,
                                    irsop=1, cfl_irs=2.5, cfl=1.5, ncycle=6, nl_crs=2, ramping=2,
                                    ncycr=(1, 2, 4, 5), vcfli=(0.8, 1.6, 2.5), vcflit=(0.5, 1.0, 1.5),
                                    **kw_turbulence(0.5, 1e5, tu=1e-3, mutfar=0.1, **TURB))
    if name == "rans_ccirs2":    # k-omega with CIRS on cut boundaries, irsop 2 (tbeta for k, omega)
        return cases.two_block_case(ni=17, nj=9, nk=9, nlevel=2, distort=0.1, name="rans_ccirs2", wall=1500,
                                    irsop=2, cfl_irs=2.5, cfl=1.5, ncycle=4, nl_crs=2, cutcirs=1,
                                    **kw_turbulence(0.5, 1e5, tu=1e-3, mutfar=0.1, **TURB))
    if name == "laminar_blocks3":
        return cases.two_block_case(ni=17, nj=9, nk=9, nlevel=3, distort=0.1, name="laminar_blocks3", wall=1500,
                                    irsop=2, cfl_irs=2.5, cfl=1.5, ncycle=4, viscous=1, reyno=500.0, stemp=0.383)
    return cases.naca_case(65, 17, 9, nlevel=3, name="rans_naca3", wall=1500, irsop=2, cfl_irs=2.0, cfl=1.0,
                           ncycle=4, nl_crs=2, ds_wall=2e-3,
                           **kw_turbulence(0.5, 1e5, tu=1e-3, mutfar=0.1, **TURB))


@pytest.mark.parametrize("name", ["rans_blocks2", "laminar_blocks3", "rans_naca3", "sst_blocks2", "bsl_hb_rkim2",
                                  "rans_ramp2", "rans_ccirs2"])
def test_rans_multigrid_cycles(name):
    """Whole V-cycles with the k-omega production restricted to the coarse
    levels (rest4) and the restriction-stage limiters."""
    c = _mg_case(name)
    q0 = c.perturbed(amp=0.01, seed=5)
    _say(f"mg case {name} levels={c.config['nlevel']} ncell={c.ncell()} npde={c.npde}")
    o, oc, g = make_pair(c, q0)
    try:
        ncyc = c.config["ncycle"]
        hist_o = np.zeros(8 * ncyc)
        where = (ctypes.c_int32 * 8)()
        n_o = o.or_mg(oc.ptr(), 1, 1, 1, 1, dp(hist_o), where)
        assert n_o == ncyc
        n_g = g.mg(ncyc)
        assert n_g == ncyc
        npde = c.npde
        hist_g = np.array(g.history)
        hist_o = hist_o.reshape(ncyc, 8)
        parity(hist_g[:, :npde + 1].ravel(), hist_o[:, :npde + 1].ravel(), rtol=1e-10, what="resrms history")
        for nl in range(1, c.config["nlevel"] + 1):
            parity(g.download("q", nl), oc[nl]["q"], rtol=1e-10, what=f"q level {nl}")
    finally:
        g.close()


def test_viscous_kernel_variants_agree(monkeypatch):
    """The three arrangements of the viscous face work give bit-identical
    residuals (same operands, same accumulation order): face fluxes stored
    then differenced (default), flux and difference in one marching kernel
    (COSAHIP_VISCOUS=march), fused into the inviscid line march
    (COSAHIP_VISCOUS=fused)."""
    from cosa_amd.solver import Solve
    c = CASES[0]
    q0 = c.perturbed(amp=0.02, seed=13)
    out = []
    for mode in ("default", "fused", "march", "tile"):
        if mode == "default":
            monkeypatch.delenv("COSAHIP_VISCOUS", raising=False)
        else:
            monkeypatch.setenv("COSAHIP_VISCOUS", mode)
        g = Solve(c, device=0, q=q0)
        try:
            for f in ("cons2prim", "bc", "cutman_q"):
                g.call(f, 1, _lib.Q)
            g.call("lambda_deltat", 1)
            g.call("residual", 1)
            out.append(g.download("res"))
        finally:
            g.close()
    m = interior_mask(c)
    assert np.array_equal(out[0][m], out[1][m])
    assert np.array_equal(out[0][m], out[2][m])
    assert np.array_equal(out[0][m], out[3][m])


def _tile_pass_cases():
    return CASES + [_mg_case("bsl_hb_rkim2"), _mg_case("rans_naca3")]


@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "fp32res"])
@pytest.mark.parametrize("c", _tile_pass_cases(), ids=lambda c: c.name)
def test_viscous_tile_pass_matches(c, prec, monkeypatch):
    """The tile arrangement (COSAHIP_VISCOUS=tile: viscous face fluxes of
    the three directions stored, then one 31 x 7 tile pass forming the
    inviscid fluxes of every direction and summing them with the viscous
    differences in the reference's order) leaves the same bits as the
    default per-direction kernels after two smoothing iterations, in FP64
    and in the FP32-residual mode (laminar, Wilcox, Menter BSL / SST,
    harmonic balance, the NACA O-grid with tiles that overhang the block)."""
    from cosa_amd.solver import Solve
    q0 = c.perturbed(amp=0.02, seed=23)
    out = []
    for mode in (None, "tile"):
        if mode is None:
            monkeypatch.delenv("COSAHIP_VISCOUS", raising=False)
        else:
            monkeypatch.setenv("COSAHIP_VISCOUS", mode)
        g = Solve(c, device=0, q=q0)
        try:
            if prec:
                g.set_precision(prec)
            g.smooth(1, False, False, 2, 1)
            out.append((g.download("q"), g.download("res")))
        finally:
            g.close()
    _say(f"tile pass {c.name} prec {prec}: compared {out[0][0].size} state values")
    assert np.array_equal(out[0][0], out[1][0])
    m = interior_mask(c)
    assert np.array_equal(out[0][1][m], out[1][1][m])


def test_cirs_i_variants_agree(monkeypatch):
    """CIRS i-sweep: one thread per (line, variable) with k and omega in the
    same launch (COSAHIP_CIRS_I=iv, the default on small levels) and one
    thread per line with LDS-staged chunks (COSAHIP_CIRS_I=lds) give
    bit-identical results."""
    from cosa_amd.solver import Solve
    for c in (CASES[0], CASES[2]):
        q0 = c.perturbed(amp=0.02, seed=17)
        res = []
        for mode in ("iv", "lds"):
            monkeypatch.setenv("COSAHIP_CIRS_I", mode)
            g = Solve(c, device=0, q=q0)
            try:
                g.smooth(1, False, False, 2, 1)
                res.append(g.download("q"))
            finally:
                g.close()
        assert np.array_equal(res[0], res[1]), c.name


def _irsop1_cases():
    by_name = {c.name: c for c in CASES}
    return [by_name["bsl_blocks"], by_name["laminar_blocks"],
            cases.two_block_case(ni=17, nj=9, nk=9, nlevel=2, distort=0.1, name="euler_hb1_blocks2", irsop=1,
                                 cfl_irs=3.0, cfl=2.0, nharms=1, harbal=1, omega=0.8),
            cases.naca_case(65, 17, 9, nlevel=3, name="rans_naca3_irs1", wall=1500, irsop=1, cfl_irs=2.0, cfl=1.0,
                            nl_crs=2, **kw_turbulence(0.5, 1e5, tu=1e-3, mutfar=0.1, **TURB))]


@pytest.mark.parametrize("c", _irsop1_cases(), ids=lambda c: c.name)
def test_cirs_i_line_resident_matches(c, monkeypatch):
    """CIRS i-sweep with whole lines resident in LDS (COSAHIP_CIRS_I=line,
    irsop 1) against one thread per (line, variable) (iv): two multigrid
    cycles give bit-identical states on every level and identical norms,
    mean flow and k-omega, one and three snapshots, one and three levels."""
    from cosa_amd.solver import Solve
    q0 = c.perturbed(amp=0.02, seed=23)
    _say(f"case {c.name} ncell={c.ncell()} nsnap={c.nsnap} npde={c.npde} levels={c.config['nlevel']}")
    out = []
    for mode in ("iv", "line"):
        monkeypatch.setenv("COSAHIP_CIRS_I", mode)
        g = Solve(c, device=0, q=q0)
        try:
            for mgit in (1, 2):
                g.cycle(mgit)
            out.append(([g.download("q", nl) for nl in range(1, c.config["nlevel"] + 1)], list(g.history)))
        finally:
            g.close()
    # bitwise: signed zeros must agree too
    for nl, (qa, qb) in enumerate(zip(out[0][0], out[1][0]), 1):
        assert np.array_equal(qa.view(np.uint64), qb.view(np.uint64)), f"level {nl}: max |d| {np.abs(qa - qb).max()}"
    assert np.array_equal(np.array(out[0][1]).view(np.uint64), np.array(out[1][1]).view(np.uint64))


@pytest.mark.parametrize("irsop", [5, 6])
def test_viscous_cirs_irsop(irsop):
    """cirs_v1v / cirs_v2v (irsop 5 / 6): viscous radii enter the smoothing
    coefficients instead of the time step; three smooth iterations against
    the oracle."""
    c = cases.two_block_case(ni=13, nj=9, nk=9, nlevel=1, distort=0.1, name=f"rans_irsop{irsop}", wall=1500,
                             irsop=irsop, psirs=0.125, cfl_irs=2.5, cfl=1.5, ratt=1.4, cflt=1.2,
                             **kw_turbulence(0.5, 1e5, tu=1e-3, mutfar=0.1, **TURB))
    q0 = c.perturbed(amp=0.02, seed=19)
    _say(f"case {c.name} ncell={c.ncell()}")
    o, oc, g = make_pair(c, q0)
    try:
        done = (ctypes.c_int32 * 1)()
        where = (ctypes.c_int32 * 8)()
        assert o.or_smooth(oc.ptr(), 1, 0, 0, 3, 1, done, where) == 0
        _, r = g.smooth(1, False, False, 3, 1)
        parity(g.download("q"), oc[1]["q"], rtol=1e-11, what="smooth q")
        r_o = np.array([o.or_get_resrms(i) for i in range(1, 9)])
        parity(r[:8], r_o[:8], rtol=1e-12, what="smooth resrms")
        parity(g.download("betaxi"), oc[1]["betaxi"], what="betaxi")
    finally:
        g.close()


@pytest.mark.parametrize("name", ["rans_blocks2", "laminar_blocks3"])
def test_autotune_keeps_results(name):
    """cosahip_autotune times the interchangeable arrangements (CIRS
    i-sweep per level, viscous face fluxes) on copies and keeps the fastest:
    the state it leaves and every later cycle are bit-identical to an
    untuned solve."""
    from cosa_amd.solver import Solve
    c = _mg_case(name)
    q0 = c.perturbed(amp=0.01, seed=9)
    a = Solve(c, device=0, q=q0)
    b = Solve(c, device=0, q=q0)
    try:
        a.cycle(1)
        b.cycle(1)
        tuned = b.autotune(2)
        _say(f"autotune {name}: {tuned}")
        for nl in range(1, c.config["nlevel"] + 1):
            assert np.array_equal(a.download("q", nl), b.download("q", nl))
        for mgit in (2, 3):
            a.cycle(mgit)
            b.cycle(mgit)
        for nl in range(1, c.config["nlevel"] + 1):
            assert np.array_equal(a.download("q", nl), b.download("q", nl)), nl
        assert np.array_equal(np.array(a.history), np.array(b.history))
    finally:
        a.close()
        b.close()
