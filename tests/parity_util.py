"""Shared helpers of the GPU parity tests: build a case on both the oracle
and libcosahip from the same inputs and compare with the FP64 bar

    |gpu - cpu| <= 1e-12 |cpu| + 1e-14 max|cpu|      (per value)

i.e. 1e-12 relative, with an absolute floor at 1e-14 of the field's largest
magnitude: a residual value is a cancellation of fluxes of that size, and the
device pow (farfield BC, CIRS coefficients) rounds differently from glibc's
in the last bits.  Every comparison also reports how many values are
bit-identical."""
from __future__ import annotations

import ctypes

import numpy as np

from cosa_amd import case as cases
from cosa_amd.config import kw_turbulence
from oracle_bind import Oracle, OrCase, dp

RTOL = 1e-12     # per-value relative bar (FP64 GPU mode)
FLOOR = 1e-14    # absolute floor as a fraction of the field's max magnitude


def parity(gpu: np.ndarray, cpu: np.ndarray, rtol: float = RTOL, floor: float = FLOOR, what: str = "") -> float:
    """Assert the FP64 parity bar; return the worst normalised error.  NaN
    in either array fails (it would pass every comparison silently)."""
    assert gpu.shape == cpu.shape, what
    assert not np.isnan(cpu).any(), f"{what}: NaN in the oracle's values"
    assert not np.isnan(gpu).any(), f"{what}: NaN in the GPU values ({int(np.isnan(gpu).sum())})"
    ndiff = int(np.count_nonzero(gpu != cpu))
    if ndiff:
        ulp = np.abs(gpu.view(np.int64) - cpu.view(np.int64))
        print(f"  [{what}] {ndiff}/{gpu.size} values differ, max {int(ulp.max())} ulp", flush=True)
    else:
        print(f"  [{what}] bit-identical ({gpu.size} values)", flush=True)
    scale = float(np.max(np.abs(cpu))) if cpu.size else 0.0
    err = np.abs(gpu - cpu)
    tol = rtol * np.abs(cpu) + floor * scale
    bad = err > tol
    if bad.any():
        i = int(np.argmax(err - tol))
        raise AssertionError(f"{what}: {int(bad.sum())} values outside the bar; worst at {i}: "
                             f"gpu={gpu[i]!r} cpu={cpu[i]!r} |d|={err[i]:.3e} tol={tol[i]:.3e}")
    with np.errstate(invalid="ignore", divide="ignore"):
        rel = np.where(tol > 0, err / np.maximum(tol, 1e-300), 0.0)
    return float(rel.max()) if rel.size else 0.0


def interior_mask(case, nl=1, ncomp=None):
    """Boolean mask of interior cells over a whole-level P3 array."""
    L = case.levels[nl - 1]
    ncomp = ncomp or case.npde
    parts = []
    for b in range(L.nblocks):
        im, jm, km = int(L.imax[b]), int(L.jmax[b]), int(L.kmax[b])
        m = np.zeros((km + 3, jm + 3, im + 3), bool)
        m[2:km + 1, 2:jm + 1, 2:im + 1] = True
        parts.append(np.tile(m.reshape(-1), ncomp * case.nsnap))
    return np.concatenate(parts)


def make_pair(case, q0, device=0):
    """An oracle case and a GPU solve, both holding state q0 on level 1."""
    from cosa_amd.solver import Solve
    o = Oracle()
    o.set_config(case.config, case.hb_matrix())
    oc = OrCase(case, q1=q0.copy())
    g = Solve(case, device=device, q=q0)
    return o, oc, g


def small_cases():
    """Parity cases: distorted two-block with a cut and four BC types;
    NACA O-grid with a self-cut; HB on the two-block case; HB on the NACA
    O-grid pitching about the quarter chord (moving grid, lmet = 7)."""
    out = []
    c = cases.two_block_case(ni=13, nj=9, nk=9, nlevel=1, distort=0.15, name="two_block", irsop=2,
                             cfl_irs=4.0, cfl=2.0, ilimf=4, epslimf=0.01)
    out.append(c)
    c = cases.naca_case(33, 13, 5, nlevel=1, name="naca_small", irsop=1, cfl_irs=3.0, cfl=2.0)
    out.append(c)
    c = cases.two_block_case(ni=9, nj=9, nk=5, nlevel=1, distort=0.1, name="two_block_hb2", irsop=3, psirs=0.125,
                             cfl_irs=3.0, cfl=2.0, nharms=2, harbal=1, omega=0.5)
    out.append(c)
    c = cases.naca_case(33, 13, 5, nlevel=1, name="naca_pitch_hb1", irsop=2, cfl_irs=3.0, cfl=2.0, nharms=1,
                        harbal=1, omega=0.6, pitch=dict(dtheta0=0.05, xh=0.25, yh=0.0))
    out.append(c)
    return out


TURB = dict(stemp=0.383, ifixq=1, lim_prodtke=1, prdlim=20.0, turcmp=1)


def rans_cases():
    """Viscous parity cases: Wilcox k-omega on the distorted two-block case
    (viscous walls 1500), laminar Navier-Stokes on the same grid, k-omega on
    the NACA O-grid with Menter's wall omega, Menter BSL on the two-block
    case and SST on the O-grid."""
    out = []
    out.append(cases.two_block_case(ni=13, nj=9, nk=9, nlevel=1, distort=0.1, name="rans_blocks", wall=1500,
                                    irsop=2, cfl_irs=2.5, cfl=1.5, ratt=1.4, cflt=1.2,
                                    **kw_turbulence(0.5, 1e5, tu=1e-3, mutfar=0.1, **TURB)))
    out.append(cases.two_block_case(ni=13, nj=9, nk=9, nlevel=1, distort=0.1, name="laminar_blocks", wall=1500,
                                    irsop=1, cfl_irs=2.5, cfl=1.5, viscous=1, reyno=500.0, stemp=0.383))
    out.append(cases.naca_case(33, 13, 9, nlevel=1, name="rans_naca_menter", wall=1500, irsop=3, psirs=0.125,
                               cfl_irs=2.0, cfl=1.0, ds_wall=2e-3, wallwilc=0,
                               **kw_turbulence(0.5, 1e5, tu=1e-3, mutfar=0.1, **TURB)))
    # Menter BSL and SST (blended coefficients, cross diffusion, SST eddy viscosity)
    out.append(cases.two_block_case(ni=13, nj=9, nk=9, nlevel=1, distort=0.1, name="bsl_blocks", wall=1500,
                                    irsop=1, cfl_irs=2.5, cfl=1.5,
                                    **kw_turbulence(0.5, 1e5, tu=1e-3, mutfar=0.1, model="bsl", **TURB)))
    out.append(cases.naca_case(33, 13, 9, nlevel=1, name="sst_naca", wall=1500, irsop=2, cfl_irs=2.0, cfl=1.0,
                               ds_wall=2e-3, **kw_turbulence(0.5, 1e5, tu=1e-3, mutfar=0.1, model="sst", **TURB)))
    return out
