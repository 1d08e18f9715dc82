"""Pin the force integration of the C restatement (or_force_b,
oracle/cosa_oracle.c) against COSA's own compiled force_b
(src/output.f90:4388-4886, built into oracle/_ref by oracle/build_ref.py):
identical random inputs, identical binary64 coefficients required.

The type-14 (inviscid wall) extension has no counterpart in the reference,
whose force_b integrates only the viscous-wall surfaces 1500 + s - 1 (an
Euler run reports zero forces).  It is pinned by closed forms instead: a
uniform pressure on a closed box of walls gives no net force, and on one
flat wall of area A with unit normal n gives the force -p A n.
"""
import ctypes
import math

import numpy as np
import pytest

from blockgen import Patch, bctopo_of, n3, prim_to_cons, rand_prim
from cosa_amd.config import default_config, hb_operator
from oracle_bind import Oracle, Ref, dp, ip, ref_available

DIMS = (7, 6, 5)
PATCHES = [Patch((1, 1, 1), (1, 6, 5), 1500), Patch((7, 1, 1), (7, 6, 5), 1501),
           Patch((1, 1, 1), (7, 1, 5), 1500), Patch((1, 6, 1), (7, 6, 5), 14),
           Patch((1, 1, 1), (7, 6, 1), 1501), Patch((1, 1, 5), (7, 6, 5), 80)]
LNGTH = np.array([1.3, 0.9, 0.7])
XYZMOM = np.array([0.25, 0.1, -0.05])


def n2(d):
    return (d[0] + 2) * (d[1] + 2) * (d[2] + 2)


def n0(d):
    return d[0] * d[1] * d[2]


def momr_of(dth, xh, yh, pitching):
    """Moment centre carried with the pitching body, as force_b computes it
    (src/output.f90:4476-4486)."""
    out = []
    for t in dth:
        if pitching:
            d1 = (XYZMOM[0] - xh) * (math.cos(t) - 1) - (XYZMOM[1] - yh) * math.sin(t)
            d2 = (XYZMOM[0] - xh) * math.sin(t) + (XYZMOM[1] - yh) * (math.cos(t) - 1)
            out += [XYZMOM[0] + d1, XYZMOM[1] + d2, XYZMOM[2]]
        else:
            out += list(XYZMOM)
    return np.array(out)


def inputs(rng, d, nsnap, nmet, lmet):
    q = prim_to_cons(rand_prim(rng, *d, nsnap=nsnap), *d, nsnap=nsnap)
    xyz = [rng.uniform(-1, 1, n2(d) * nmet) for _ in range(3)]
    vel = [rng.uniform(-0.05, 0.05, n2(d) * nmet) for _ in range(3)]
    mets = []
    for _ in range(3):
        s = rng.uniform(-1, 1, (nmet * n2(d), lmet))
        s[:, 3] = rng.uniform(0.01, 0.2, nmet * n2(d))
        mets.append(s.reshape(-1).copy())
    der = [rng.uniform(-3, 3, n0(d) * 3 * nmet) for _ in range(3)]
    return q, xyz, vel, mets, der


@pytest.mark.ref
@pytest.mark.skipif(not ref_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("viscous", [0, 1])
@pytest.mark.parametrize("motion", ["static", "pitching_hb1"])
def test_force_b_matches_reference(viscous, motion):
    o, r = Oracle(), Ref()
    rng = np.random.default_rng(7 + viscous + 2 * (motion != "static"))
    hb = motion != "static"
    kw = dict(viscous=viscous, reyno=2.0e5, stemp=0.383, machfs=0.45, alpha_deg=3.0, beta_deg=1.5)
    if hb:
        kw.update(nharms=1, harbal=1, omega=0.6, moving=1, hbmove=1, lmet=7)
    c = default_config(**kw)
    dhb = hb_operator(1, c["omega"]) if hb else None
    o.set_config(c, dhb)
    r.set_config(c, dhb)
    nsnap = 3 if hb else 1
    q, (x, y, z), (xd, yd, zd), (si, sj, sk), (xdi, edj, zdk) = inputs(rng, DIMS, nsnap, nsnap, c["lmet"])
    bt = bctopo_of(PATCHES)
    dth = np.array([0.05, -0.03, 0.02])[:nsnap].copy()
    xh, yh = 0.3, 0.05
    cr = np.zeros(12 * nsnap * 2)
    r.ref_force_b(dp(q), dp(x), dp(y), dp(z), dp(xd), dp(yd), dp(zd), dp(si), dp(sj), dp(sk), dp(xdi), dp(edj),
                  dp(zdk), len(PATCHES), ip(bt), 2, dp(LNGTH), dp(XYZMOM), int(hb), ctypes.c_double(xh),
                  ctypes.c_double(yh), dp(dth), dp(cr), *DIMS)
    co = np.zeros_like(cr)
    momr = momr_of(dth, xh, yh, hb)
    o.or_force_b(dp(q), dp(x), dp(y), dp(z), dp(xd), dp(yd), dp(zd), dp(si), dp(sj), dp(sk), dp(xdi), dp(edj),
                 dp(zdk), len(PATCHES), ip(bt), 2, 0, dp(LNGTH), dp(momr), dp(co), *DIMS)
    assert np.count_nonzero(cr) > (6 if not viscous else 12)
    np.testing.assert_array_equal(co, cr)


def _box_case(d, p0, types):
    """Cartesian box of unit cells scaled by h: exact metrics (unit normals,
    face areas), uniform pressure p0 and zero velocity everywhere."""
    h = (0.2, 0.15, 0.1)
    sm = []
    for dirn in range(3):
        s = np.zeros((n2(d), 4))
        s[:, dirn] = 1.0
        s[:, 3] = h[0] * h[1] * h[2] / h[dirn]
        sm.append(s.reshape(-1).copy())
    k, j, i = np.meshgrid(*(np.arange(m + 2) for m in d[::-1]), indexing="ij")
    xyz = [np.ascontiguousarray((a * hh).reshape(-1), np.float64) for a, hh in zip((i, j, k), h)]
    prim = np.zeros((5, n3(*d)))
    prim[0] = 1.0
    prim[4] = p0
    q = prim_to_cons(prim.reshape(-1).copy(), *d)
    pats = [Patch((1, 1, 1), (1, d[1], d[2]), types[0]), Patch((d[0], 1, 1), (d[0], d[1], d[2]), types[1]),
            Patch((1, 1, 1), (d[0], 1, d[2]), types[2]), Patch((1, d[1], 1), (d[0], d[1], d[2]), types[3]),
            Patch((1, 1, 1), (d[0], d[1], 1), types[4]), Patch((1, 1, d[2]), (d[0], d[1], d[2]), types[5])]
    return q, xyz, sm, bctopo_of(pats), h


def _force(o, d, q, xyz, sm, bt, nbcs, inv14):
    z2 = np.zeros(n2(d))
    z0 = np.zeros(n0(d) * 3)
    co = np.zeros(12)
    o.or_force_b(dp(q), dp(xyz[0]), dp(xyz[1]), dp(xyz[2]), dp(z2), dp(z2), dp(z2), dp(sm[0]), dp(sm[1]),
                 dp(sm[2]), dp(z0), dp(z0), dp(z0), nbcs, ip(bt), 1, inv14, dp(LNGTH), dp(XYZMOM), dp(co), *d)
    return co


def test_inviscid_walls_closed_box_has_no_net_force():
    o = Oracle()
    c = default_config(machfs=0.5, alpha_deg=4.0)
    o.set_config(c)
    d = (9, 7, 6)
    q, xyz, sm, bt, h = _box_case(d, 0.8, (14,) * 6)
    co = _force(o, d, q, xyz, sm, bt, 6, 1)
    one_face = 0.8 * (d[0] - 1) * (d[2] - 1) * h[0] * h[2] * 2 / (LNGTH[0] * LNGTH[1] * 0.25)
    assert np.all(np.abs(co[:3]) <= 1e-13 * one_face), co[:3]
    # the reference's own selection (viscous-wall surfaces only) sees no surface
    assert np.all(_force(o, d, q, xyz, sm, bt, 6, 0) == 0)


def test_inviscid_wall_plate_force_closed_form():
    """Pressure p0 on the j = 1 wall (area A, outward normal of the body
    pointing into the flow along +y): force -p0 A (0, 1, 0); CL, CD, CZ are
    its projections on the lift, drag and side directions of force_b
    (src/output.f90:4451-4462) scaled by 2 / (l1 l2 M^2)."""
    o = Oracle()
    alpha, beta, mach = math.radians(4.0), math.radians(1.5), 0.5
    c = default_config(machfs=mach, alpha=alpha, beta=beta)
    o.set_config(c)
    d = (9, 7, 6)
    p0 = 0.8
    q, xyz, sm, bt, h = _box_case(d, p0, (30, 30, 14, 30, 80, 80))
    co = _force(o, d, q, xyz, sm, bt, 6, 1)
    area = (d[0] - 1) * (d[2] - 1) * h[0] * h[2]
    f = np.array([0.0, -p0 * area, 0.0])
    den = math.sqrt(math.sin(beta) ** 2 + math.cos(alpha) ** 2 * math.cos(beta) ** 2)
    lift = np.array([-math.sin(alpha) * math.cos(alpha) * math.cos(beta) ** 2 / den, den,
                     math.sin(alpha) * math.sin(beta) * math.cos(beta) / den])
    drag = np.array([math.cos(alpha) * math.cos(beta), math.sin(alpha) * math.cos(beta), math.sin(beta)])
    side = np.array([-math.sin(beta) / den, 0.0, math.cos(alpha) * math.cos(beta) / den])
    scale = 2 / (LNGTH[0] * LNGTH[1] * mach ** 2)
    want = [f @ lift * scale, f @ drag * scale, f @ side * scale]
    np.testing.assert_allclose(co[:3], want, rtol=1e-13, atol=1e-15)
    assert np.all(co[6:] == 0)   # no viscous part for inviscid walls
