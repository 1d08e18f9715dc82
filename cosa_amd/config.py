"""Solver configuration: the ctypes mirror of ``cosahip_config``
(include/cosahip.h) and the reference's defaults.

Field meanings follow common_variables (src/common_variables.f90:19-65);
defaults follow ``input`` (src/input.f90:2400-2520): gamma 1.4, rho_inf =
a_inf = 1, p_inf = 1/gamma, qmin, RK coefficients ar4 = (0.25, 0.333333, 0.5,
1.0) as the decimal literals of src/input.f90:2403.
"""
from __future__ import annotations

import ctypes
import math
from typing import Any, Dict

_INT_FIELDS = [
    "npde", "nharms", "hbmove", "lmet", "nlevel", "nl_crs", "nstage", "irsop", "cutcirs", "ilimf", "ilimt",
    "iupdt", "iprint", "ncycle", "kom", "kom_bsl", "kom_sst", "viscous", "moving", "harbal", "foturb",
    "efx_setup_f", "efx_all_f", "eig_limit_f", "eig_lim_all_f", "efx_setup_t", "efx_all_t", "eig_limit_t",
    "eig_lim_all_t",
]
_DBL_FIELDS = [
    "gamma", "machfs", "alpha", "beta", "rhoinf", "pinf", "ainf", "cfl", "cflt", "rat", "ratt", "psirs", "rkap",
    "epslimf", "epslimt", "cntrpy_f", "cntrpy_t", "entfxctf_f", "entfxctf_t", "eig_cutoff_f", "eig_cutoff_t",
    "omega", "toler",
]


# viscous / k-omega block appended after omefar
_TURB_INT_FIELDS = ["turcmp", "lim_prodtke", "lim_prodomega", "wallwilc", "ifixq", "prd_all_lvls"]
_TURB_DBL_FIELDS = ["reyno", "pranl", "prant", "stemp", "cdff", "epst", "cmu", "sigk", "sigw", "bkw", "bstrkw",
                    "gkw", "prdlim", "roughk"]


class CosahipConfig(ctypes.Structure):
    _fields_ = ([(n, ctypes.c_int32) for n in _INT_FIELDS]
                + [("entfx_f_w", ctypes.c_int32 * 3), ("entfx_t_w", ctypes.c_int32 * 3),
                   ("bln_prol", ctypes.c_int32), ("ho_rest", ctypes.c_int32), ("rkimp", ctypes.c_int32)]
                + [(n, ctypes.c_double) for n in _DBL_FIELDS]
                + [("qmin", ctypes.c_double * 7), ("alphas", ctypes.c_double * 5),
                   ("tkefar", ctypes.c_double), ("omefar", ctypes.c_double)]
                + [(n, ctypes.c_int32) for n in _TURB_INT_FIELDS]
                + [("lim_corr", ctypes.c_int32 * 3), ("pad1", ctypes.c_int32)]
                + [(n, ctypes.c_double) for n in _TURB_DBL_FIELDS]
                + [("ramping", ctypes.c_int32), ("ncycr", ctypes.c_int32 * 4), ("pad2", ctypes.c_int32),
                   ("vcfli", ctypes.c_double * 3), ("vcflit", ctypes.c_double * 3)])


def default_config(**kw: Any) -> Dict[str, Any]:
    gamma = kw.get("gamma", 1.4)
    c: Dict[str, Any] = dict(
        npde=5, nharms=0, hbmove=0, lmet=4, nlevel=1, nl_crs=4, nstage=4, irsop=1, cutcirs=0, ilimf=4, ilimt=4,
        iupdt=1, iprint=1, ncycle=100, kom=0, kom_bsl=0, kom_sst=0, viscous=0, moving=0, harbal=0, foturb=0,
        efx_setup_f=0, efx_all_f=0, eig_limit_f=0, eig_lim_all_f=0, efx_setup_t=0, efx_all_t=0, eig_limit_t=0,
        eig_lim_all_t=0, entfx_f_w=(0, 0, 0), entfx_t_w=(0, 0, 0), bln_prol=1, ho_rest=0, rkimp=0,
        gamma=gamma, machfs=0.5, alpha=0.0, beta=0.0, rhoinf=1.0, pinf=1.0 / gamma, ainf=1.0,
        cfl=2.0, cflt=2.0, rat=1.0, ratt=1.0, psirs=0.0, rkap=0.333333, epslimf=1e-6, epslimt=1e-6,
        cntrpy_f=0.0, cntrpy_t=0.0, entfxctf_f=0.0, entfxctf_t=0.0, eig_cutoff_f=0.0, eig_cutoff_t=0.0,
        omega=0.0, toler=0.0,
        qmin=(1e-8, 1.0, 1.0, 1.0, 1e-8, 1e-12, 1e-8), alphas=(0.25, 0.333333, 0.5, 1.0, 0.0),
        tkefar=0.0, omefar=0.0,
        # viscous / Wilcox k-omega (src/input.f90:2501-2525; kappa = 0.41)
        turcmp=0, lim_prodtke=0, lim_prodomega=0, wallwilc=1, ifixq=0, prd_all_lvls=0, lim_corr=(0, 0, 0),
        reyno=1.0e6, pranl=0.72, prant=0.9, stemp=110.4 / 288.15, cdff=1.0, epst=1e-12, cmu=1.0, sigk=0.5,
        sigw=0.5, bkw=3.0 / 40.0, bstrkw=0.09, gkw=(3.0 / 40.0) / 0.09 - 0.5 * 0.41 ** 2 / math.sqrt(0.09),
        prdlim=20.0, roughk=0.0,
        # CFL ramping (k-omega): off
        ramping=0, ncycr=(0, 0, 0, 0), vcfli=(0.0, 0.0, 0.0), vcflit=(0.0, 0.0, 0.0),
    )
    for k, v in kw.items():
        if k not in c and k not in ("alpha_deg", "beta_deg", "cfl_irs"):
            raise KeyError(f"unknown config field {k}")
        c[k] = v
    if "alpha_deg" in kw:
        c["alpha"] = kw["alpha_deg"] * math.pi / 180.0
    if "beta_deg" in kw:
        c["beta"] = kw["beta_deg"] * math.pi / 180.0
    if "cfl_irs" in kw:
        # input.f90:4206-4216: with IRS, cfl = vcfli(3) and rat = vcfli(3)/vcfl(3)
        c["rat"] = kw["cfl_irs"] / c["cfl"]
        c["cfl"] = kw["cfl_irs"]
    c.pop("alpha_deg", None)
    c.pop("beta_deg", None)
    c.pop("cfl_irs", None)
    return c


def komega(c: Dict[str, Any]) -> bool:
    """Any of the k-omega models (Wilcox, BSL, SST): npde = 7."""
    return bool(c["kom"] or c["kom_bsl"] or c["kom_sst"])


def to_struct(c: Dict[str, Any]) -> CosahipConfig:
    s = CosahipConfig()
    for n in _INT_FIELDS:
        setattr(s, n, int(c[n]))
    for n in _DBL_FIELDS:
        setattr(s, n, float(c[n]))
    for i in range(3):
        s.entfx_f_w[i] = int(c["entfx_f_w"][i])
        s.entfx_t_w[i] = int(c["entfx_t_w"][i])
    s.bln_prol = int(c["bln_prol"])
    s.ho_rest = int(c["ho_rest"])
    s.rkimp = int(c["rkimp"])
    for i in range(7):
        s.qmin[i] = float(c["qmin"][i])
    for i in range(5):
        s.alphas[i] = float(c["alphas"][i])
    s.tkefar = float(c["tkefar"])
    s.omefar = float(c["omefar"])
    for n in _TURB_INT_FIELDS:
        setattr(s, n, int(c[n]))
    for i in range(3):
        s.lim_corr[i] = int(c["lim_corr"][i])
    for n in _TURB_DBL_FIELDS:
        setattr(s, n, float(c[n]))
    s.ramping = int(c["ramping"])
    for i in range(4):
        s.ncycr[i] = int(c["ncycr"][i])
    for i in range(3):
        s.vcfli[i] = float(c["vcfli"][i])
        s.vcflit[i] = float(c["vcflit"][i])
    return s


def kw_turbulence(machfs: float, reyno: float, tu: float = 1e-4, mutfar: float = 1e-2, model: str = "komega",
                  **kw) -> Dict[str, Any]:
    """k-omega settings of a viscous run (the 'komega', 'bsl' and 'sst'
    branches of input, src/input.f90:2580-2610, 2695-2720): farfield k from
    the turbulence intensity squared tu, omega from mut/mu at the farfield
    mutfar."""
    tkefar = tu * machfs ** 2
    rhoinf = kw.get("rhoinf", 1.0)
    cmu = kw.get("cmu", 1.0)
    out = dict(npde=7, kom=int(model == "komega"), kom_bsl=int(model == "bsl"), kom_sst=int(model == "sst"),
               viscous=1, machfs=machfs, reyno=reyno, tkefar=tkefar,
               omefar=cmu * rhoinf * tkefar / mutfar * reyno / machfs)
    out.update(kw)
    return out


def hb_operator(nharms: int, omega: float):
    """dhb of mathb (src/input.f90:5206-5282) for N harmonics, as a
    (2N+1)x(2N+1) numpy array (row i, column j); zeit(n) = n*2*pi/omega/(2N+1)
    (src/input.f90:3185-3191)."""
    import numpy as np
    n = 2 * nharms + 1
    dt = 2 * math.pi / omega / n
    zeit = [i * dt for i in range(n)]
    ahb = np.zeros((n, n))
    for i in range(1, nharms + 1):
        ahb[2 * (i - 1) + 1, 2 * (i - 1) + 2] = i
        ahb[2 * (i - 1) + 2, 2 * (i - 1) + 1] = -i
    eihb = np.zeros((n, n))
    eihb[:, 0] = 1.0
    for j in range(1, nharms + 1):
        for i in range(n):
            eihb[i, 2 * (j - 1) + 1] = math.cos(j * omega * zeit[i])
            eihb[i, 2 * (j - 1) + 2] = math.sin(j * omega * zeit[i])
    fact = 2.0 / n
    ehb = np.zeros((n, n))
    ehb[0, :] = fact / 2
    for i in range(1, nharms + 1):
        for j in range(n):
            ehb[2 * (i - 1) + 1, j] = fact * math.cos(i * omega * zeit[j])
            ehb[2 * (i - 1) + 2, j] = fact * math.sin(i * omega * zeit[j])
    d1 = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            s = 0.0
            for k in range(n):
                s = s + ahb[i, k] * ehb[k, j]
            d1[i, j] = s
    d = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            s = 0.0
            for k in range(n):
                s = s + eihb[i, k] * d1[k, j]
            d[i, j] = s
    return d
