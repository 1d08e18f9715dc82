"""Pin the oracle's dual time stepping orchestration (src/cosa.f90:437-672:
q1 = q, rhs_uns, mg(), q2 = q1; inside smooth src_uns, the finest residual
minus rhs and the rkimp norm, src/cosa.f90:1910-1941) by what the
mathematics fixes.  The per-block routines (src_uns_b, rhs_uns_b, the rkimp
update and limits) are pinned bit for bit against COSA's compiled ones in
test_oracle_vs_ref_dual.py; here the physical step as a whole:

* once the pseudo-time iterations have converged, the state solves the
  physical-time scheme: backward Euler on the first step of a first-order
  start, R(q1) + (q1 - q0) vol / dt = 0, BDF2 afterwards,
  R(q2) + (3 q2 / 2 - 2 q1 + q0 / 2) vol / dt = 0, with R the spatial
  residual of the pinned routines;
* that solution does not depend on how the pseudo-time iteration gets there:
  explicit or implicit (rkimp) physical-time term, with or without IRS, one
  or two grid levels (a wrong sign, a missing subtraction of rhs, a source
  added twice or a coarse-grid forcing taken wrongly changes it or stops
  convergence);
* a uniform state (the freestream of a wall-free box) stays uniform.
"""
import ctypes

import numpy as np
import pytest

from cosa_amd import case as cases
from oracle_bind import Oracle, OrCase, dp

DT = 0.6


def box(nlevel=1, **kw):
    """A bump channel (walls on the j faces), second-order van Albada on the
    fine level."""
    cfg = dict(irsop=0, cfl=1.5, nl_crs=2, ilimf=4, epslimf=1e-2, toler=1e-13, machfs=0.5)
    cfg.update(kw)
    return cases.channel_case([(17, 9, 5)], nlevel=nlevel, name="channel", **cfg)


def interior(c, nl=1):
    """(P3 interior mask over the blocks of level nl, vol index of each P3 interior point)."""
    L = c.levels[nl - 1]
    m3s, p1s, o1 = [], [], 0
    for b in range(L.nblocks):
        im, jm, km = int(L.imax[b]), int(L.jmax[b]), int(L.kmax[b])
        m = np.zeros((c.npde, km + 3, jm + 3, im + 3), bool)
        m[:, 2:km + 1, 2:jm + 1, 2:im + 1] = True
        m3s.append(m.reshape(-1))
        p1 = np.arange((km + 1) * (jm + 1) * (im + 1)).reshape(km + 1, jm + 1, im + 1)[1:km, 1:jm, 1:im]
        p1s.append(np.tile(p1.reshape(-1) + o1, c.npde))
        o1 += (km + 1) * (jm + 1) * (im + 1)
    return np.concatenate(m3s), np.concatenate(p1s)


def run_steps(c, q0, nsteps, rkimp=0, ncycle=4000):
    """nsteps physical steps from q0 (backward-Euler start); the states."""
    c.config["rkimp"] = rkimp
    c.config["ncycle"] = ncycle
    o = Oracle()
    o.set_config(c.config)
    oc = OrCase(c, q1=q0.copy())
    hist = np.zeros(8 * ncycle)
    where = (ctypes.c_int32 * 8)()
    out, cycles = [q0.copy()], []
    for it in range(1, nsteps + 1):
        o.or_set_dual(1, ctypes.c_double(DT), 0, it)
        o.or_dual_begin(oc.ptr(), it)
        n = o.or_mg(oc.ptr(), 1, 1, 1, 1, dp(hist), where)
        assert 0 < n < ncycle, f"step {it}: no convergence ({n})"
        o.or_dual_end(oc.ptr())
        out.append(oc[1]["q"].copy())
        cycles.append(n)
    o.or_set_dual(0, ctypes.c_double(DT), 0, 1)
    return out, cycles


def spatial_residual(c, q):
    """R(q): the pinned routines of one residual evaluation (cons2prim, bc,
    cut, lambda, residual), q conservative in, residual out."""
    o = Oracle()
    cfg = dict(c.config, rkimp=0)
    o.set_config(cfg)
    oc = OrCase(c, q1=q.copy())
    cp = oc.ptr()
    a = oc[1]
    o.or_cons2prim(cp, 1, dp(a["q"]))
    o.or_bc(cp, 1, dp(a["q"]))
    o.or_cutman_q(cp, 1, dp(a["q"]))
    o.or_lambda_deltat(cp, 1)
    o.or_residual(cp, 1)
    return a["res"].copy()


@pytest.fixture(scope="module")
def reference_steps():
    c = box()
    q0 = c.perturbed(amp=0.03, seed=5)
    qs, cycles = run_steps(c, q0, 2)
    return c, q0, qs, cycles


def test_converged_step_solves_backward_euler_then_bdf2(reference_steps):
    c, q0, qs, cycles = reference_steps
    m3, p1 = interior(c)
    vol = c.geom[0]["vol"]
    v = vol[p1]
    for step, (terms, name) in enumerate([((1.0, -1.0, 0.0), "backward Euler"), ((1.5, -2.0, 0.5), "BDF2")]):
        qn = qs[step + 1]
        r = spatial_residual(c, qn)[m3]
        a, b, d = terms
        qm2 = qs[step - 1][m3] if step >= 1 else np.zeros(m3.sum())
        dqdt = (a * qn[m3] + b * qs[step][m3] + d * qm2) * v / DT
        e = np.abs(r + dqdt).max()
        scale = np.abs(dqdt).max()
        print(f"{name}: {cycles[step]} cycles, max |R + dq/dt vol| = {e:.2e} (|dq/dt vol| up to {scale:.2e})")
        assert scale > 1e-4
        assert e < 1e-9 * max(scale, 1.0)


@pytest.mark.parametrize("variant", ["irsop1", "two_levels", "rkimp_plain", "rkimp_irs", "rkimp_two_levels"])
def test_step_independent_of_pseudo_time_scheme(reference_steps, variant):
    c0, q0, qs, _ = reference_steps
    kw, rkimp = {}, 0
    if variant in ("irsop1", "rkimp_irs"):
        kw = dict(irsop=1, cfl=3.0, cfl_irs=4.0)
    if variant in ("two_levels", "rkimp_two_levels"):
        kw["nlevel"] = 2
    if variant.startswith("rkimp"):
        rkimp = 1
    c = box(**kw)
    got, cycles = run_steps(c, q0, 2, rkimp=rkimp)
    m3, _ = interior(c)
    scale = np.abs(qs[2][m3]).max()
    for s in (1, 2):
        d = np.abs(got[s][m3] - qs[s][m3]).max() / scale
        moved = np.abs(qs[s][m3] - qs[s - 1][m3]).max() / scale
        print(f"{variant} step {s}: {cycles[s - 1]} cycles, max |q - q_ref| / max|q| = {d:.2e} (step moved {moved:.2e})")
        assert moved > 1e-5
        assert d < 1e-10


def test_uniform_state_stays():
    c = cases.two_block_case(ni=9, nj=9, nk=5, nlevel=2, wall=30, name="free", irsop=1, cfl=3.0, cfl_irs=4.0,
                             toler=1e-13)
    q0 = c.freestream(1).copy()
    for rkimp in (0, 1):
        qs, cycles = run_steps(c, q0, 2, rkimp=rkimp, ncycle=50)
        m3, _ = interior(c)
        d = np.abs(qs[2][m3] - q0[m3]).max()
        assert d < 1e-13, (rkimp, d)
