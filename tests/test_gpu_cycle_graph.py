"""Asynchronous V-cycles (cosahip_cycle_enqueue / cosahip_cycle_wait) and
their HIP-graph replay against the synchronous path (smooth, restrict,
prolong called one by one, as mg does): every cycle, the state of every
level and the resrms line must be bit-identical.  The first cycle runs
eagerly, the second captures the graph, later cycles replay it; between the
replays the test does its own eager device work (the synchronous solve) and
synchronous downloads, and checks each replay."""
import numpy as np
import pytest

from cosa_amd import case as cases
from cosa_amd.config import kw_turbulence
from parity_util import TURB

pytestmark = pytest.mark.gpu


def _case(name):
    if name == "naca3":
        return cases.naca_case(65, 17, 9, nlevel=3, name="naca3", irsop=1, cfl_irs=3.5, cfl=2.0, nl_crs=2,
                               ncycle=100)
    if name == "blocks2_hb1":
        return cases.two_block_case(ni=17, nj=9, nk=9, nlevel=2, distort=0.1, name="blocks2", irsop=3,
                                    psirs=0.125, cfl_irs=3.0, cfl=2.0, ncycle=100, nharms=1, harbal=1, omega=0.8)
    if name == "naca3_pitch_hb2":
        return cases.naca_case(65, 17, 9, nlevel=3, name="naca3_pitch", irsop=2, cfl_irs=3.0, cfl=2.0, nl_crs=2,
                               ncycle=100, nharms=2, harbal=1, omega=0.6, pitch=dict(dtheta0=0.05, xh=0.25, yh=0.0))
    if name == "rans_blocks2":
        return cases.two_block_case(ni=17, nj=9, nk=9, nlevel=2, distort=0.1, name="rans_blocks2", wall=1500,
                                    irsop=1, cfl_irs=2.5, cfl=1.5, ncycle=100, nl_crs=2,
                                    **kw_turbulence(0.5, 1e5, tu=1e-3, mutfar=0.1, **TURB))
    if name == "single_level":
        return cases.naca_case(33, 13, 5, nlevel=1, name="naca_small", irsop=1, cfl_irs=3.0, cfl=2.0, ncycle=100)
    raise KeyError(name)


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("name", ["naca3", "blocks2_hb1", "naca3_pitch_hb2", "rans_blocks2", "single_level"])
def test_async_graph_cycles_match_sync(name, precision):
    from cosa_amd.solver import Solve
    c = _case(name)
    q0 = c.perturbed(amp=0.01, seed=5)
    print(f"case {name} precision mode {precision} ncell={c.ncell()}", flush=True)
    a = Solve(c, device=0, q=q0)
    b = Solve(c, device=0, q=q0)
    try:
        a.set_precision(precision)
        b.set_precision(precision)
        npde = c.npde
        for mgit in range(1, 7):
            a.cycle(mgit)
            b.cycle_async(mgit)
            for nl in range(1, c.config["nlevel"] + 1):
                qa, qb = a.download("q", nl), b.download("q", nl)
                assert np.array_equal(qa, qb), f"cycle {mgit} level {nl}: max |d| {np.abs(qa - qb).max()}"
            assert np.array_equal(a.history[-1][:npde + 1], b.history[-1][:npde + 1]), mgit
        print(f"  6 cycles identical; resrms(1) {a.history[0][0]:.3e} -> {a.history[-1][0]:.3e}", flush=True)
    finally:
        a.close()
        b.close()


def test_async_cycles_overlap_on_separate_streams():
    """Two solves enqueued back to back, then waited for: each result equals
    the same solve run alone."""
    from cosa_amd.solver import Solve
    c1, c2 = _case("naca3"), _case("blocks2_hb1")
    q1, q2 = c1.perturbed(amp=0.01, seed=6), c2.perturbed(amp=0.01, seed=7)
    alone = []
    for c, q in ((c1, q1), (c2, q2)):
        g = Solve(c, device=0, q=q)
        try:
            for mgit in range(1, 5):
                g.cycle_async(mgit)
            alone.append(g.download("q"))
        finally:
            g.close()
    g1, g2 = Solve(c1, device=0, q=q1), Solve(c2, device=0, q=q2)
    try:
        for mgit in range(1, 5):
            g1.cycle_enqueue(mgit)
            g2.cycle_enqueue(mgit)
            g1.cycle_wait()
            g2.cycle_wait()
        assert np.array_equal(g1.download("q"), alone[0])
        assert np.array_equal(g2.download("q"), alone[1])
    finally:
        g1.close()
        g2.close()
