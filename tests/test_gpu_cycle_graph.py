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


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("name", ["naca3", "rans_blocks2"])
def test_metric_upload_recaptures_graph(name, precision, monkeypatch):
    """A metric upload after the cycle graph was captured (cycle 2) drops the
    graph: cycle 3 is recaptured with the new metrics (and, FP32, their new
    single-precision copies; k-omega, the new cell-centre averages) and
    matches a solve that never uses graphs, bit for bit."""
    from cosa_amd.solver import Solve
    c = _case(name)
    q0 = c.perturbed(amp=0.01, seed=5)
    print(f"case {name} precision mode {precision} ncell={c.ncell()}", flush=True)
    monkeypatch.setenv("COSAHIP_GRAPH", "0")
    a = Solve(c, device=0, q=q0)
    monkeypatch.delenv("COSAHIP_GRAPH")
    b = Solve(c, device=0, q=q0)
    try:
        a.set_precision(precision)
        b.set_precision(precision)
        for mgit in (1, 2):
            a.cycle_async(mgit)
            b.cycle_async(mgit)
        # new metrics on every level: the face vectors scaled (a valid, if
        # non-closed, input) and, viscous, the metric derivatives too
        g = c.geom
        for L in c.levels:
            for f in ("si", "sj", "sk"):
                for s in (a, b):
                    s.upload(f, g[L.nl - 1][f] * 1.01, L.nl)
            if "xd" in g[L.nl - 1]:
                from cosa_amd import _lib
                for nm, arr in zip(_lib.XD_NAMES, g[L.nl - 1]["xd"]):
                    for s in (a, b):
                        s.upload(nm, arr * 0.99, L.nl)
        for mgit in (3, 4):
            a.cycle_async(mgit)
            b.cycle_async(mgit)
            for nl in range(1, c.config["nlevel"] + 1):
                qa, qb = a.download("q", nl), b.download("q", nl)
                assert np.array_equal(qa.view(np.uint64), qb.view(np.uint64)), f"cycle {mgit} level {nl}"
            assert np.array_equal(a.history[-1], b.history[-1]), mgit
    finally:
        a.close()
        b.close()


def test_define_level_refused_while_cycle_in_flight():
    """Redefining a level frees the buffers a captured graph points to: it is
    refused while a cycle is in flight, and afterwards the solve runs on the
    new level (the graph is dropped and recaptured)."""
    import ctypes
    from cosa_amd import _lib
    from cosa_amd.solver import Solve
    c = _case("single_level")
    q0 = c.perturbed(amp=0.01, seed=5)
    a = Solve(c, device=0, q=q0)
    ref = Solve(c, device=0, q=q0)
    try:
        for mgit in (1, 2):
            a.cycle_async(mgit)
        a.cycle_enqueue(3)
        L = c.levels[0]
        keep = [np.ascontiguousarray(getattr(L, n), np.int32) for n in ("imax", "jmax", "kmax", "nbcs", "bctopo",
                                                                         "cutopo")]
        rc = a.lib.cosahip_define_level(a.h, 1, L.nblocks, *[_lib.iptr(k) for k in keep[:5]], L.ncuts,
                                        _lib.iptr(keep[5]))
        assert rc == 2, rc   # COSAHIP_ERR_STATE
        a.cycle_wait()
        # redefine (fresh, zeroed buffers), upload everything again, run on
        a._define()
        a.upload("q", q0, 1)
        a.history.clear()
        for mgit in (1, 2, 3):
            a.cycle_async(mgit)
            ref.cycle_async(mgit)
        assert np.array_equal(a.download("q"), ref.download("q"))
    finally:
        a.close()
        ref.close()


def test_block_freestream_viscous_keeps_mach():
    """Viscous scaling (M/Re) is one value per solve: a block freestream with
    another Mach number is refused on a viscous solve; another incidence at
    the same Mach is accepted."""
    from cosa_amd.solver import Solve
    from cosa_amd import _lib
    c = _case("rans_blocks2")
    g = Solve(c, device=0, q=c.freestream())
    try:
        with pytest.raises(_lib.CosahipError) as e:
            g.set_block_freestream(1, c.config["machfs"] + 0.1, 2.0)
        assert e.value.code == 6, e.value
        g.set_block_freestream(2, c.config["machfs"], 3.0)
    finally:
        g.close()
