"""The CIRS j/k sweeps with the forward pass kept in chunks (k_cirs_jkc,
COSAHIP_CIRS_JK=chunk4|chunk8: only the last forward value of each chunk
stored, the others recomputed in the backward pass with the same
operations) leave the same bits as the sweeps that store every forward
value (tridi, src/cosa.f90:2418-2443): whole V-cycles on uneven channel
blocks (lines of 8 to 64 cells, so chunks that end inside and at the end
of a line), the fused RK update and cons2prim of the k sweep included, and
smoothing iterations of a Wilcox k-omega case (k and omega in the same
sweep)."""
import numpy as np
import pytest

from cosa_amd import case as cases
from cosa_amd.config import kw_turbulence
from parity_util import TURB

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _uneven():
    dims = cases.uneven_dims(96, seed=0)[:6]
    return cases.channel_case(dims, nlevel=3, name="c4_first6", machfs=0.7, irsop=1, cfl=2.0, cfl_irs=4.0,
                              ilimf=4, epslimf=1e-6, nl_crs=2, ncycle=100000)


def _rans():
    return cases.naca_case(65, 17, 9, nlevel=1, name="rans_naca_irs1", wall=1500, irsop=1, cfl_irs=2.0, cfl=1.0,
                           **kw_turbulence(0.5, 1e5, tu=1e-3, mutfar=0.1, **TURB))


def _run(c, mode, monkeypatch, q0, rans):
    from cosa_amd.solver import Solve
    monkeypatch.setenv("COSAHIP_CIRS_JK", mode)
    g = Solve(c, device=0, q=q0)
    try:
        if rans:
            g.smooth(1, False, False, 3, 1)
            return [g.download("q")]
        for mgit in (1, 2):
            g.cycle(mgit)
        return [g.download("q", nl) for nl in range(1, c.config["nlevel"] + 1)] + [np.array(g.history)]
    finally:
        g.close()


@pytest.mark.parametrize("mode", ["chunk4", "chunk8"])
@pytest.mark.parametrize("which", ["uneven", "rans"])
def test_cirs_jk_chunked_same_bits(which, mode, monkeypatch):
    rans = which == "rans"
    c = _rans() if rans else _uneven()
    q0 = c.perturbed(amp=0.01, seed=5)
    _say(f"{c.name}: {c.ncell()} cells, {mode} against the stored forward pass")
    ref = _run(c, "plain", monkeypatch, q0, rans)
    got = _run(c, mode, monkeypatch, q0, rans)
    for a, b in zip(ref, got):
        assert np.array_equal(a, b)
