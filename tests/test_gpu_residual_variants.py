"""The inviscid residual kernels are interchangeable: the specialised
k_residual_tile instances (31 x 7 tiles with the k-window from LDS / L2 or
in registers, 31 x 15 tiles, the 128-VGPR build) and the k_residual_sva
instances (one van Albada slope per cell shared by its two faces, as
limit_i/j/k compute it, src/flux.f90:3584-3604) give the same bits as the
generic fused kernel, which matches the oracle (muscl_b* + roflux_b + resid,
src/flux.f90:1221-1812, 45-260; src/cosa.f90:1149-1211).  Cases cover the
specialised van Albada path at second and first order, the run-time paths
(other limiters, entropy fix), harmonic balance on a moving grid, blocks
whose sizes leave partial tiles in i, j and k, and the FP32 residual."""
import numpy as np
import pytest

from cosa_amd import _lib
from cosa_amd import case as cases
from parity_util import interior_mask, make_pair, parity

pytestmark = pytest.mark.gpu

VARIANTS = ["fused", "tile7", "tile7reg", "tile15", "tile7w4", "sva7", "sva15", "sva7w4", "tile32w4", "sva32"]


def _cases():
    out = [cases.two_block_case(ni=13, nj=9, nk=9, nlevel=1, distort=0.15, name="blocks_vanalbada", irsop=2,
                                cfl_irs=4.0, cfl=2.0, ilimf=4, epslimf=0.01),
           cases.naca_case(33, 13, 5, nlevel=1, name="naca_first_order", irsop=1, cfl_irs=3.0, cfl=2.0, nl_crs=1),
           cases.naca_case(33, 13, 5, nlevel=1, name="naca_pitch_hb1", irsop=2, cfl_irs=3.0, cfl=2.0, nharms=1,
                           harbal=1, omega=0.6, pitch=dict(dtheta0=0.05, xh=0.25, yh=0.0)),
           cases.two_block_case(ni=13, nj=9, nk=9, nlevel=1, distort=0.15, name="blocks_limiter1", irsop=1,
                                cfl_irs=3.0, cfl=2.0, ilimf=1, epslimf=0.05),
           cases.two_block_case(ni=13, nj=9, nk=9, nlevel=1, distort=0.15, name="blocks_entropy_fix", irsop=1,
                                cfl_irs=3.0, cfl=2.0, efx_setup_f=1, cntrpy_f=0.3, entfxctf_f=0.1, efx_all_f=1,
                                entfx_f_w=(0, 1, 0)),
           cases.channel_case([(17, 33, 17), (33, 17, 65), (65, 17, 9)], nlevel=1, name="channels_partial_tiles",
                              machfs=0.7, irsop=1, cfl=2.0, cfl_irs=4.0, ilimf=4, epslimf=1e-6)]
    return out


CASES = _cases()


def _residual(c, q0, variant, monkeypatch, fp32=False):
    from cosa_amd.solver import Solve
    monkeypatch.setenv("COSAHIP_RESIDUAL", variant)
    g = Solve(c, device=0, q=q0)
    monkeypatch.delenv("COSAHIP_RESIDUAL")
    try:
        if fp32:
            g.set_precision(1)
        for f in ("cons2prim", "bc", "cutman_q"):
            g.call(f, 1, _lib.Q)
        g.call("residual", 1)
        return g.download("res")
    finally:
        g.close()


@pytest.mark.parametrize("c", CASES, ids=[c.name for c in CASES])
def test_residual_variants_same_bits(c, monkeypatch):
    q0 = c.perturbed(amp=0.02, seed=13)
    print(f"case {c.name}: {c.ncell()} cells x {c.nsnap}", flush=True)
    # the generic kernel against the oracle
    o, oc, g = make_pair(c, q0)
    g.close()
    cp = oc.ptr()
    from oracle_bind import dp
    for f in ("cons2prim", "bc", "cutman_q"):
        getattr(o, "or_" + f)(cp, 1, dp(oc[1]["q"]))
    o.or_residual(cp, 1)
    m = interior_mask(c)
    ref = _residual(c, q0, "fused", monkeypatch)
    parity(ref[m], oc[1]["res"][m], what=f"{c.name} generic fused vs oracle")
    for v in VARIANTS[1:]:
        r = _residual(c, q0, v, monkeypatch)
        same = np.array_equal(r.view(np.uint64), ref.view(np.uint64))
        print(f"  {v}: {'same bits' if same else 'DIFFERENT'}", flush=True)
        assert same, (c.name, v, int(np.count_nonzero(r != ref)))


def test_residual_variants_fp32_same_bits(monkeypatch):
    c = CASES[0]
    q0 = c.perturbed(amp=0.02, seed=14)
    ref = _residual(c, q0, "fused", monkeypatch, fp32=True)
    for v in VARIANTS[1:]:
        r = _residual(c, q0, v, monkeypatch, fp32=True)
        assert np.array_equal(r.view(np.uint64), ref.view(np.uint64)), v
