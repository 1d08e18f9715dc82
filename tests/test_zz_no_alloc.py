"""cosahip_force_noalloc (force coefficients with the buffers in the
library's fixed device array, no device allocation) against cosahip_force,
through the C ABI: two calls in a row on two solves, the second on a
smaller grid, must give the same bits as the allocating entry point.  The
pressure part of force_b is a fixed sum order on both paths; the viscous
part runs the same kernels, so everything must be bit-identical."""
import numpy as np
import pytest

from cosa_amd import case as cases
from parity_util import rans_cases

pytestmark = pytest.mark.gpu


def _forces(c, seed):
    from cosa_amd.solver import Solve
    g = Solve(c, device=0, q=c.perturbed(amp=0.01, seed=seed))
    try:
        g.smooth(1, False, False, 2, 1)
        out = []
        for nsurf, inv in ((1, True), (2, False)):
            kw = dict(nsurface=nsurf, include_inviscid=inv, lngth=(1.0, 0.5, 1.0), xyzmom=(0.25, 0.02, 0.0))
            ref = g.force(**kw)
            mine = g.force(noalloc=True, **kw)
            out.append((ref, mine, inv or bool(c.config["viscous"])))
        return out
    finally:
        g.close()


def test_force_noalloc_matches_force_twice():
    big = cases.naca_case(65, 17, 5, nlevel=1, name="naca_65", irsop=1, cfl_irs=3.0, cfl=2.0)
    small = cases.naca_case(33, 13, 5, nlevel=1, name="naca_33", irsop=1, cfl_irs=3.0, cfl=2.0)
    visc = {c.name: c for c in rans_cases()}["laminar_blocks"]
    for c, seed in ((big, 3), (small, 4), (visc, 5)):
        print(f"force_noalloc vs force: {c.name} ({c.ncell()} cells, viscous={c.config['viscous']})", flush=True)
        for ref, mine, walls in _forces(c, seed):
            assert ref.shape == mine.shape
            assert np.array_equal(ref.view(np.uint64), mine.view(np.uint64)), (ref, mine)
            # a surface with walls has forces (without them force_b reports zeros)
            assert np.any(ref != 0) == walls


def test_force_noalloc_rejects_oversize():
    """A request beyond the fixed array (here 1024 x 128 wall faces: 12
    terms each, 1.57 M doubles against 1,048,576) is refused on the host
    with an error before any copy or launch; a bad surface count too."""
    from cosa_amd._lib import CosahipError
    from cosa_amd.solver import Solve
    c = cases.naca_case(1025, 5, 129, nlevel=1, name="naca_wide", irsop=1, cfl_irs=3.0, cfl=2.0)
    g = Solve(c, device=0, q=c.freestream())
    try:
        with pytest.raises(CosahipError):
            g.force(nsurface=0, noalloc=True)
        with pytest.raises(CosahipError, match="scratch"):
            g.force(nsurface=1, include_inviscid=True, noalloc=True)
        f = g.force(nsurface=1, include_inviscid=True)   # the allocating entry point still serves it
        assert f.shape == (1, 1, 12)
    finally:
        g.close()
