"""Restart files (cosa_amd/restart.py) against the reference's own
write_brest / read_brest (src/output.f90:3531-3566, src/input.f90:4665-4714,
compiled into oracle/_ref): identical bytes for the same arrays, and the
reference's file read back to the same arrays.  Where the reference is
absent, against the file it wrote, frozen in tests/golden/ref_kernels.npz."""
import os

import numpy as np
import pytest

from cosa_amd import case as cases
from cosa_amd import restart
from cosa_amd.config import komega, kw_turbulence
from oracle_bind import Ref, ref_available

HERE = os.path.dirname(os.path.abspath(__file__))


def _case(kom, nharms):
    kw = dict(nharms=nharms, harbal=1, omega=0.5) if nharms else {}
    if kom:
        kw.update(kw_turbulence(0.5, 1e5, tu=1e-3, mutfar=0.1))
    return cases.two_block_case(ni=7, nj=5, nk=5, nlevel=1, distort=0.1, name="rst", irsop=1,
                                wall=1500 if kom else 14, **kw)


def _arrays(c, seed):
    rng = np.random.default_rng(seed)
    q = rng.uniform(-2, 2, c.size(1, "p3"))
    mut = rng.uniform(0, 5, c.size(1, "p3s")) if komega(c.config) else None
    return q, mut


def _ref_file(r, c, path, q, mut, mode):
    L = c.levels[0]
    im, jm, km = (np.ascontiguousarray(getattr(L, n), np.int32) for n in ("imax", "jmax", "kmax"))
    m = mut if mut is not None else np.zeros(1)
    b = path.encode()
    from oracle_bind import dp, ip
    r.ref_restart(mode, b, len(b), dp(q), dp(m), L.nblocks, ip(im), ip(jm), ip(km))


@pytest.mark.ref
@pytest.mark.skipif(not ref_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("kom,nharms", [(False, 0), (True, 0), (True, 1), (False, 2)])
def test_restart_bytes_match_reference(tmp_path, kom, nharms):
    c = _case(kom, nharms)
    r = Ref()
    r.set_config(c.config, c.hb_matrix())
    q, mut = _arrays(c, 3 + nharms + 10 * kom)
    ours, theirs = str(tmp_path / "ours"), str(tmp_path / "theirs")
    restart.write_restart(ours, c, q, mut)
    _ref_file(r, c, theirs, q, mut, 0)
    assert open(ours, "rb").read() == open(theirs, "rb").read()
    q2, m2 = restart.read_restart(theirs, c)
    assert np.array_equal(q2, q)
    if kom:
        assert np.array_equal(m2, mut)
    # and the reference reads ours
    q3, m3 = np.zeros_like(q), (np.zeros_like(mut) if kom else None)
    _ref_file(r, c, ours, q3, m3, 1)
    assert np.array_equal(q3, q)
    if kom:
        assert np.array_equal(m3, mut)


def test_restart_matches_frozen_reference_file(tmp_path):
    data = np.load(os.path.join(HERE, "golden", "ref_kernels.npz"), allow_pickle=False)
    c = _case(True, 1)
    q, mut = _arrays(c, 14)
    ours = str(tmp_path / "ours")
    restart.write_restart(ours, c, q, mut)
    assert open(ours, "rb").read() == data["restart_kom_hb1/out/bytes"].tobytes()


def test_restart_rejects_wrong_sizes(tmp_path):
    c = _case(False, 0)
    q, _ = _arrays(c, 1)
    p = str(tmp_path / "r")
    restart.write_restart(p, c, q)
    with pytest.raises(ValueError):
        restart.read_restart(p, _case(False, 1))
