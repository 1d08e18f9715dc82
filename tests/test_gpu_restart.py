"""Restart through a resident solve (writerest / readrest layout,
src/output.f90:3481-3566, src/input.f90:4620-4714): a checkpoint taken
asynchronously while later cycles are enqueued holds the state as of the
moment it was requested, and a new solve started from the file continues
with the same bits as the uninterrupted one.  The hist.dat lines of the
uninterrupted run are written as the reference's output writes them."""
import os

import numpy as np
import pytest

from cosa_amd import case as cases
from cosa_amd import restart
from cosa_amd.config import kw_turbulence
from cosa_amd.history import HistoryWriter, line_102
from parity_util import TURB

pytestmark = pytest.mark.gpu


def _cases():
    return [cases.naca_case(65, 17, 9, nlevel=2, name="restart_euler_mg", irsop=1, cfl=2.0, cfl_irs=3.0, nl_crs=2,
                            iprint=2, ncycle=5),
            cases.two_block_case(ni=13, nj=9, nk=9, nlevel=1, distort=0.1, name="restart_rans", wall=1500, irsop=1,
                                 cfl_irs=2.5, cfl=1.5, iprint=2, ncycle=5,
                                 **kw_turbulence(0.5, 1e5, tu=1e-3, mutfar=0.1, **TURB))]


@pytest.mark.parametrize("c", _cases(), ids=lambda c: c.name)
def test_async_checkpoint_and_restart(c, tmp_path):
    from cosa_amd.solver import Solve
    kom = bool(c.config["kom"])
    q0 = c.perturbed(amp=0.01, seed=21)
    a = Solve(c, device=0, q=q0)
    try:
        hdir = str(tmp_path / "hist")
        with HistoryWriter(hdir, npde=c.npde, iprint=2, ncycle=5, kom=kom) as h:
            for m in (1, 2, 3):
                a.cycle(m)
                h.record(m, a.history[-1], False)
            q3 = a.download("q")
            mut3 = a.download("mut") if kom else None
            for m in (4, 5):
                a.cycle(m)
                h.record(m, a.history[-1], False)
        q5 = a.download("q")
        hist = open(os.path.join(hdir, "hist.dat")).read().splitlines()
        assert [int(l.split()[1]) for l in hist] == [1, 2, 4, 5]
        assert hist[-1] == line_102(0, 5, a.history[4][:5])
        hist_a = [r.copy() for r in a.history]
    finally:
        a.close()
    # the same run, checkpointed after cycle 3 while cycles 4 and 5 are enqueued
    b = Solve(c, device=0, q=q0)
    try:
        for m in (1, 2, 3):
            b.cycle_async(m)
        path = str(tmp_path / "restart.bin")
        ck = b.checkpoint_async(path)
        b.cycle_enqueue(4)
        ck.wait()
        b.cycle_wait()
        b.cycle_async(5)
        assert np.array_equal(b.download("q").view(np.uint64), q5.view(np.uint64))
    finally:
        b.close()
    qf, mutf = restart.read_restart(path, c)
    assert np.array_equal(qf.view(np.uint64), q3.view(np.uint64))
    if kom:
        assert np.array_equal(mutf.view(np.uint64), mut3.view(np.uint64))
    # a new solve from the file continues with the same bits
    r = Solve(c, device=0)
    try:
        r.read_restart(path)
        for m in (4, 5):
            r.cycle(m)
        assert np.array_equal(r.download("q").view(np.uint64), q5.view(np.uint64))
        for k, m in enumerate((4, 5)):
            assert np.array_equal(r.history[k], hist_a[m - 1]), m
    finally:
        r.close()
