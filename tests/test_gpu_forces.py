"""Force integration on the GPU (cosahip_force, force_b + force's block
sum) against the oracle on the same state, through the C ABI.

The device keeps the reference's summation order (per-face terms, then a
sequential sum per block, snapshot, surface and component, then the block
sum), so pressure coefficients are bit-identical; the viscous ones pass
through the Sutherland law's t**1.5, where the device pow and glibc's may
differ in the last bit: bar |gpu - cpu| <= 1e-12 |cpu| + 1e-14 max|cpu|.
"""
import numpy as np
import pytest

from cosa_amd import case as cases
from oracle_bind import Oracle, or_force
from parity_util import parity, rans_cases, small_cases

pytestmark = pytest.mark.gpu


def _cases():
    sc = {c.name: c for c in small_cases()}
    rc = {c.name: c for c in rans_cases()}
    return [sc["naca_small"], sc["naca_pitch_hb1"], rc["laminar_blocks"], rc["rans_naca_menter"]]


CASES = _cases()


@pytest.mark.parametrize("c", CASES, ids=[c.name for c in CASES])
def test_force_matches_oracle(c):
    from cosa_amd.solver import Solve
    print(f"case {c.name} ncell={c.ncell()} viscous={c.config['viscous']} nsnap={c.nsnap}", flush=True)
    o = Oracle()
    o.set_config(c.config, c.hb_matrix())
    g = Solve(c, device=0, q=c.perturbed(amp=0.01, seed=21))
    try:
        g.smooth(1, False, False, 2, 1)   # realistic halos: BC-filled, conservative
        q = g.download("q")
        lngth, mom = (1.0, 0.5, 1.0), (0.25, 0.02, 0.0)
        for nsurf, inv in ((1, True), (2, False)):
            fg = g.force(nsurface=nsurf, include_inviscid=inv, lngth=lngth, xyzmom=mom)
            fo = or_force(o, c, q, nsurface=nsurf, inv14=inv, lngth=lngth, xyzmom=mom)
            assert fg.shape == fo.shape == (nsurf, c.nsnap, 12)
            assert np.array_equal(fg[..., :6], fo[..., :6]), (fg[..., :6], fo[..., :6])
            parity(fg[..., 6:].ravel(), fo[..., 6:].ravel(), what=f"viscous coefficients nsurface={nsurf}")
            print(f"  nsurface={nsurf} inviscid walls={inv}: CL_p {fg[0, :, 0]} CD_p {fg[0, :, 1]} "
                  f"CL_v {fg[0, :, 6]} CD_v {fg[0, :, 7]}", flush=True)
            walls = {int(t) for t in c.levels[0].bctopo[0::10]}
            if inv and 14 in walls or 1500 in walls:
                assert np.any(fg[0, :, :6] != 0)
            if c.config["viscous"] and 1500 in walls:
                assert np.any(fg[0, :, 6:] != 0)
    finally:
        g.close()


def test_force_argument_errors():
    from cosa_amd._lib import CosahipError
    from cosa_amd.solver import Solve
    c = cases.naca_case(33, 13, 5, nlevel=1, name="naca_small", irsop=1, cfl_irs=3.0, cfl=2.0)
    g = Solve(c, device=0, q=c.freestream())
    try:
        with pytest.raises(CosahipError):
            g.force(nsurface=0)
        f = g.force(nsurface=1, include_inviscid=False)
        assert np.all(f == 0)   # the reference's force_b: no viscous-wall surface, no force
    finally:
        g.close()
