"""Harmonic-balance parity cases beyond the explicit Euler ones: the
implicit stage rkim (update_bhb, src/update.f90:1410-1841) with and without
IRS / multigrid, and Wilcox k-omega with harmonic balance (lim_res_ome_bhb
at the restriction stage, explicit and rkim)."""
from __future__ import annotations

from cosa_amd import case as cases
from cosa_amd.config import kw_turbulence
from parity_util import TURB

PITCH = dict(dtheta0=0.05, xh=0.25, yh=0.0)


def hb_case(name: str):
    if name == "euler_rkim_plain":     # no IRS, one level: src_hb left out of res, rms adds it
        return cases.naca_case(33, 13, 5, nlevel=1, name=name, irsop=0, cfl=1.0, nharms=2, harbal=1, omega=0.6,
                               rkimp=1, pitch=PITCH, ncycle=3)
    if name == "euler_rkim_mg":        # IRS and two levels: the HB term taken out of the smoothed residual
        return cases.naca_case(65, 17, 9, nlevel=2, name=name, irsop=2, cfl_irs=3.0, cfl=2.0, nl_crs=2,
                               nharms=2, harbal=1, omega=0.6, rkimp=1, pitch=PITCH, ncycle=3)
    if name == "rans_hb_explicit":
        return cases.two_block_case(ni=17, nj=9, nk=9, nlevel=2, distort=0.1, name=name, wall=1500, irsop=1,
                                    cfl_irs=2.5, cfl=1.5, nl_crs=2, nharms=1, harbal=1, omega=0.8, ncycle=3,
                                    **kw_turbulence(0.5, 1e5, tu=1e-3, mutfar=0.1, **TURB))
    if name == "rans_hb_rkim":
        return cases.two_block_case(ni=17, nj=9, nk=9, nlevel=2, distort=0.1, name=name, wall=1500, irsop=1,
                                    cfl_irs=2.5, cfl=1.5, nl_crs=2, nharms=2, harbal=1, omega=0.8, rkimp=1, ncycle=3,
                                    **kw_turbulence(0.5, 1e5, tu=1e-3, mutfar=0.1, **TURB))
    if name == "rans_hb_rkim_plain":
        return cases.two_block_case(ni=13, nj=9, nk=9, nlevel=1, distort=0.1, name=name, wall=1500, irsop=0,
                                    cfl=0.8, nharms=1, harbal=1, omega=0.8, rkimp=1, ncycle=3,
                                    **kw_turbulence(0.5, 1e5, tu=1e-3, mutfar=0.1, **TURB))
    raise KeyError(name)


HB_CASES = ["euler_rkim_plain", "euler_rkim_mg", "rans_hb_explicit", "rans_hb_rkim", "rans_hb_rkim_plain"]
