"""The fast FP64 mode (cosahip_set_precision mode 3): the one-pass Euler
residual kernels built with multiply-add contraction
(k_resid_contracted.hip).  Every operation stays binary64, a contracted
a*b+c rounds once, so the residual is no longer the reference's bits; it
must meet the FP64 parity bar against the oracle (muscl_b* + roflux_b +
resid, src/flux.f90:1221-1812, 45-260; src/cosa.f90:1149-1211) for every
kernel variant and option the exact mode covers, and whole multigrid cycles
must stay within the multigrid bar of test_gpu_parity.py."""
import ctypes

import numpy as np
import pytest

from cosa_amd import _lib
from cosa_amd import case as cases
from oracle_bind import dp
from parity_util import interior_mask, make_pair, parity
from test_gpu_residual_variants import CASES, VARIANTS

pytestmark = pytest.mark.gpu


def _residual(c, q0, variant, mode, monkeypatch):
    from cosa_amd.solver import Solve
    monkeypatch.setenv("COSAHIP_RESIDUAL", variant)
    g = Solve(c, device=0, q=q0)
    monkeypatch.delenv("COSAHIP_RESIDUAL")
    try:
        g.set_precision(mode)
        for f in ("cons2prim", "bc", "cutman_q"):
            g.call(f, 1, _lib.Q)
        g.call("residual", 1)
        return g.download("res")
    finally:
        g.close()


@pytest.mark.parametrize("c", CASES, ids=[c.name for c in CASES])
def test_contracted_residual_meets_fp64_bar(c, monkeypatch):
    q0 = c.perturbed(amp=0.02, seed=13)
    print(f"case {c.name}: {c.ncell()} cells x {c.nsnap}", flush=True)
    o, oc, g = make_pair(c, q0)
    g.close()
    cp = oc.ptr()
    for f in ("cons2prim", "bc", "cutman_q"):
        getattr(o, "or_" + f)(cp, 1, dp(oc[1]["q"]))
    o.or_residual(cp, 1)
    m = interior_mask(c)
    ref = oc[1]["res"][m]
    exact = _residual(c, q0, "fused", 0, monkeypatch)[m]
    changed = 0
    for v in VARIANTS:
        r = _residual(c, q0, v, 3, monkeypatch)[m]
        parity(r, ref, what=f"{c.name} {v} contracted vs oracle")
        changed += int(np.count_nonzero(r != exact))
    # contraction is in effect: some values are not the exact mode's bits
    assert changed > 0


def test_contracted_cycles_within_bar():
    c = cases.channel_case([(33, 17, 9), (17, 33, 17)], nlevel=2, name="channels2", machfs=0.7, irsop=1, cfl=2.0,
                           cfl_irs=4.0, ilimf=4, epslimf=1e-6, nl_crs=2, ncycle=5)
    q0 = c.perturbed(amp=0.01, seed=5)
    o, oc, g = make_pair(c, q0)
    try:
        g.set_precision(3)
        ncyc = c.config["ncycle"]
        hist_o = np.zeros(8 * ncyc)
        where = (ctypes.c_int32 * 8)()
        assert o.or_mg(oc.ptr(), 1, 1, 1, 1, dp(hist_o), where) == ncyc
        assert g.mg(ncyc) == ncyc
        hist_g = np.array(g.history)
        parity(hist_g[:, :6].ravel(), hist_o.reshape(ncyc, 8)[:, :6].ravel(), rtol=1e-10, what="resrms history")
        for nl in (1, 2):
            parity(g.download("q", nl), oc[nl]["q"], rtol=1e-10, what=f"q level {nl}")
        assert g.precision_in_use() == 64
    finally:
        g.close()


def test_contracted_mode_refused_for_viscous():
    from cosa_amd.solver import Solve
    from parity_util import rans_cases
    c = rans_cases()[1]
    g = Solve(c, device=0, q=c.freestream())
    try:
        with pytest.raises(RuntimeError, match="mode 3"):
            g.set_precision(3)
    finally:
        g.close()
