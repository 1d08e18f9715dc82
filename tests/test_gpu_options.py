"""The reference's scheme options that the default configurations leave
off, on the GPU against the oracle through the C ABI: the limiters 0-3
of limit_i/j/k (src/flux.f90:3472-3888: unlimited, smooth, minmod with
compression, Venkatakrishnan-type), the Harten-Yee entropy fix in its
three weightings and the eigenvalue floor of roflux_b
(src/flux.f90:140-210), for the mean flow and (k-omega) the turbulence
equations.  Each option must change the residual (the branch ran) and
match the oracle at the FP64 bar, routine by routine and over three
smooth iterations."""
import ctypes

import numpy as np
import pytest

from cosa_amd import _lib
from cosa_amd import case as cases
from cosa_amd.config import kw_turbulence
from oracle_bind import dp
from parity_util import TURB, interior_mask, make_pair, parity

pytestmark = pytest.mark.gpu

EFX = dict(efx_setup_f=1, cntrpy_f=0.3, entfxctf_f=0.1)
VARIANTS = {
    "limiter0": dict(ilimf=0),
    "limiter1": dict(ilimf=1, epslimf=0.05),
    "limiter2": dict(ilimf=2, epslimf=0.05),
    "limiter3": dict(ilimf=3, epslimf=0.05),
    "entropy_fix_w1": dict(EFX, efx_all_f=1, entfx_f_w=(1, 0, 0)),
    "entropy_fix_w2": dict(EFX, efx_all_f=1, entfx_f_w=(0, 1, 0)),
    # acoustic waves only (efx_all_f off); a larger width so |un - a| <= rdelta occurs
    "entropy_fix_w3_acoustic_only": dict(EFX, cntrpy_f=3.0, efx_all_f=0, entfx_f_w=(0, 0, 1)),
    "eigen_floor": dict(eig_limit_f=1, eig_lim_all_f=0, eig_cutoff_f=0.15),
    "eigen_floor_all_waves": dict(eig_limit_f=1, eig_lim_all_f=1, eig_cutoff_f=0.15),
    "rans_limiter2_turb_efx_eig": dict(ilimt=2, epslimt=0.05, efx_setup_t=1, efx_all_t=1, cntrpy_t=0.3,
                                       entfx_t_w=(0, 1, 0), eig_limit_t=1, eig_lim_all_t=1, eig_cutoff_t=0.15),
}


def _case(name, base=False):
    kw = {} if base else dict(VARIANTS[name])
    if name.startswith("rans"):
        return cases.two_block_case(ni=13, nj=9, nk=9, nlevel=1, distort=0.1, name=name, wall=1500, irsop=1,
                                    cfl_irs=2.5, cfl=1.5, **kw_turbulence(0.5, 1e5, tu=1e-3, mutfar=0.1, **TURB),
                                    **kw)
    return cases.two_block_case(ni=13, nj=9, nk=9, nlevel=1, distort=0.15, name=name, irsop=2, cfl_irs=4.0,
                                cfl=2.0, **kw)


def _residual(o, oc, g, c):
    A = oc[1]
    cp = oc.ptr()
    for f, arg in (("cons2prim", _lib.Q), ("bc", _lib.Q), ("cutman_q", _lib.Q)):
        getattr(o, "or_" + f)(cp, 1, dp(A["q"]))
        g.call(f, 1, arg)
    if c.config["kom"]:
        o.or_tur_vis(cp, 1)
        g.call("tur_vis", 1, _lib.Q)
    o.or_lambda_deltat(cp, 1)
    g.call("lambda_deltat", 1)
    o.or_residual(cp, 1)
    g.call("residual", 1)
    m = interior_mask(c)
    return g.download("res")[m], A["res"][m]


@pytest.mark.parametrize("name", list(VARIANTS))
def test_option_matches_oracle(name):
    c = _case(name)
    q0 = c.perturbed(amp=0.03, seed=31)
    print(f"case {name} ncell={c.ncell()} npde={c.npde}", flush=True)
    o, oc, g = make_pair(c, q0)
    try:
        rg, ro = _residual(o, oc, g, c)
        parity(rg, ro, what=f"{name} residual")
    finally:
        g.close()
    # the option changed the residual: the branch was taken
    cb = _case(name, base=True)
    ob, ocb, gb = make_pair(cb, q0)
    try:
        _, rb = _residual(ob, ocb, gb, cb)
    finally:
        gb.close()
    nchg = int(np.count_nonzero(ro != rb))
    print(f"  {nchg} of {ro.size} residual values changed by the option", flush=True)
    assert nchg > 0
    # three smooth iterations
    o, oc, g = make_pair(c, q0)
    try:
        done = (ctypes.c_int32 * 1)()
        where = (ctypes.c_int32 * 8)()
        assert o.or_smooth(oc.ptr(), 1, 0, 0, 3, 1, done, where) == 0
        _, r = g.smooth(1, False, False, 3, 1)
        parity(g.download("q"), oc[1]["q"], rtol=1e-11, what=f"{name} smooth q")
        r_o = np.array([o.or_get_resrms(i) for i in range(1, 9)])
        parity(r[:c.npde + 1], r_o[:c.npde + 1], rtol=1e-12, what=f"{name} smooth resrms")
    finally:
        g.close()
