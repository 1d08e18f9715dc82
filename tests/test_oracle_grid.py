"""Pin the oracle's grid motion and metric set-up (oracle/grid_oracle.c):

* against COSA's own compiled routines (oracle/_ref): imetrix_b, bvols,
  volex_b, vol_bedges, cordex_b, vmetrix_b (src/grid.f90), coarsen_b
  (src/cosa.f90), pitch, and the cut copies cut_v / cut_x -- every output
  array bit-identical, on distorted blocks, static and moving (lmet 7,
  several snapshots), both pitch branches, every cut orientation;
* by what the mathematics fixes, where the reference is absent: the face
  vectors of every cell close (sum of outward area vectors = 0), the
  volume of an affinely mapped cell is det(A) times the unit cell's, the
  metric derivatives of an affine map are the rows of A^-1 exactly (up to
  rounding), coarsening keeps every second node, a rigid pitch keeps every
  distance and its node velocity is omega x r."""
import ctypes
import itertools

import numpy as np
import pytest

from cosa_amd.config import default_config
from oracle_bind import Oracle, Ref, dp, ip, ref_available
from test_oracle_cut import cut_record

needs_ref = pytest.mark.skipif(not ref_available(), reason="oracle/_ref not built")
D = ctypes.c_double


def n2(d):
    return (d[0] + 2) * (d[1] + 2) * (d[2] + 2)


def n1(d):
    return (d[0] + 1) * (d[1] + 1) * (d[2] + 1)


def n0(d):
    return d[0] * d[1] * d[2]


def nodes(d, rng, amp=0.15, nsm=1, affine=None):
    """Node coordinates x, y, z (P2 layout, nsm snapshots) of a distorted
    stretched box with every node 1..imax etc. set and the halo nodes filled
    as cordex_b does (copies of the boundary nodes), or of an affine map."""
    im, jm, km = d
    i, j, k = np.meshgrid(np.arange(im + 2), np.arange(jm + 2), np.arange(km + 2), indexing="ij")
    ii, jj, kk = (np.clip(a, 1, m) - 1.0 for a, m in ((i, im), (j, jm), (k, km)))
    out = []
    for n in range(nsm):
        if affine is not None:
            A, c = affine
            P = np.stack([ii, jj, kk], -1) @ A.T + c
        else:
            P = np.stack([ii * 1.1 + 0.3 * np.sin(jj), jj ** 1.3 + 0.2 * ii, kk * 0.7 + 0.1 * ii * jj / (im * jm)], -1)
            P = P + amp * rng.uniform(-1, 1, P.shape) * (1 + 0.1 * n)
            # halo nodes: copies of the boundary nodes (cordex_b)
            P = P[np.clip(i, 1, im), np.clip(j, 1, jm), np.clip(k, 1, km)]
        out.append([np.ascontiguousarray(P[..., c].transpose(2, 1, 0)).reshape(-1) for c in range(3)])
    return [np.ascontiguousarray(np.concatenate([o[c] for o in out])) for c in range(3)]


def cfg(**kw):
    return default_config(**kw)


def _pair(**kw):
    o, c = Oracle(), cfg(**kw)
    o.set_config(c)
    r = None
    if ref_available():
        r = Ref()
        r.set_config(c)
    return o, r


def bits_equal(a, b, what):
    np.testing.assert_array_equal(a.view(np.uint64), b.view(np.uint64), err_msg=what)


# ---- against the compiled reference ---------------------------------------

@needs_ref
@pytest.mark.parametrize("moving,nharms", [(0, 0), (1, 1), (1, 2)])
def test_imetrix_bvols_match_reference(moving, nharms):
    lmet = 7 if moving else 4
    kw = dict(lmet=lmet, moving=moving, nharms=nharms, harbal=int(nharms > 0), hbmove=int(moving), omega=0.4)
    o, r = _pair(**kw)
    rng = np.random.default_rng(3 + nharms)
    nsm = 2 * nharms * moving + 1
    for d in ((7, 6, 5), (9, 4, 6), (5, 8, 3)):
        x, y, z = nodes(d, rng, nsm=nsm)
        vel = [rng.uniform(-1, 1, x.size) for _ in range(3)]
        outs = []
        for lib, fn in ((o, "or_imetrix_b"), (r, "ref_imetrix_b")):
            s = [np.full(n2(d) * lmet * nsm, 7.0) for _ in range(3)]
            if fn == "or_imetrix_b":
                assert lib.or_imetrix_b(*map(dp, (x, y, z, *vel, *s)), *d, lmet, nsm, moving) == 0
            else:
                lib.ref_imetrix_b(*map(dp, (x, y, z, *vel, *s)), *d)
            outs.append(s)
        for m in range(3):
            bits_equal(outs[0][m], outs[1][m], f"imetrix_b {d} face family {m + 1}")
        va, vb = np.full(n1(d), 3.0), np.full(n1(d), 3.0)
        assert o.or_bvols(dp(x), dp(y), dp(z), dp(va), *d) == 0
        r.ref_bvols(dp(x), dp(y), dp(z), dp(vb), *d)
        bits_equal(va, vb, f"bvols {d}")


@needs_ref
def test_volume_halos_and_node_halos_match_reference():
    o, r = _pair()
    rng = np.random.default_rng(5)
    for d in ((7, 6, 5), (4, 9, 6)):
        x, y, z = nodes(d, rng)
        v = np.zeros(n1(d))
        o.or_bvols(dp(x), dp(y), dp(z), dp(v), *d)
        a, b = v.copy(), v.copy()
        o.or_volex_b(dp(a), *d)
        r.ref_volex_b(dp(b), *d)
        bits_equal(a, b, "volex_b")
        a2, b2 = a.copy(), a.copy()
        o.or_vol_bedges(dp(a2), *d)
        r.ref_vol_bedges(dp(b2), *d)
        assert not np.array_equal(a2, a)
        bits_equal(a2, b2, "vol_bedges")
        # cordex_b on a grid whose halo nodes hold garbage
        xs = [rng.uniform(-5, 5, x.size) for _ in range(3)]
        for m, src in enumerate((x, y, z)):
            im, jm, km = d
            inner = src.reshape(km + 2, jm + 2, im + 2)[1:-1, 1:-1, 1:-1]
            xs[m].reshape(km + 2, jm + 2, im + 2)[1:-1, 1:-1, 1:-1] = inner
        ao, ar = [u.copy() for u in xs], [u.copy() for u in xs]
        o.or_cordex_b(*map(dp, ao), *d, 1)
        r.ref_cordex_b(*map(dp, ar), *d)
        for m in range(3):
            bits_equal(ao[m], ar[m], "cordex_b")


@needs_ref
@pytest.mark.parametrize("moving,nharms", [(0, 0), (1, 2)])
def test_vmetrix_matches_reference(moving, nharms):
    kw = dict(lmet=7 if moving else 4, moving=moving, nharms=nharms, harbal=int(nharms > 0), hbmove=int(moving),
              omega=0.4, viscous=1, reyno=1e5)
    o, r = _pair(**kw)
    rng = np.random.default_rng(11)
    nsm = 2 * nharms * moving + 1
    for d in ((7, 6, 5), (6, 9, 4)):
        x, y, z = nodes(d, rng, nsm=nsm)
        # a collapsed side: the halo node equals the boundary node (the
        # doubled one-sided difference of vmetrix_b)
        v = np.zeros(n1(d))
        o.or_bvols(dp(x), dp(y), dp(z), dp(v), *d)
        o.or_volex_b(dp(v), *d)
        outs = []
        for side in ("o", "r"):
            dd = [np.full(n0(d) * 3 * nsm, 5.0) for _ in range(9)]
            if side == "o":
                arr = (ctypes.POINTER(ctypes.c_double) * 9)(*[dp(a) for a in dd])
                o.or_vmetrix_b(dp(x), dp(y), dp(z), arr, dp(v), *d, nsm)
            else:
                r.ref_vmetrix_b(dp(x), dp(y), dp(z), *[dp(a) for a in dd], dp(v), *d)
            outs.append(dd)
        for m in range(9):
            bits_equal(outs[0][m], outs[1][m], f"vmetrix_b {d} array {m}")


@needs_ref
def test_coarsen_and_pitch_match_reference():
    o, r = _pair(moving=1, lmet=7)
    rng = np.random.default_rng(13)
    d, dc = (9, 5, 7), (5, 3, 4)
    x, y, z = nodes(d, rng)
    outs = []
    for lib, fn in ((o, "or_coarsen_b"), (r, "ref_coarsen_b")):
        c = [np.full(n2(dc), -1.0) for _ in range(3)]
        getattr(lib, fn)(dp(x), dp(y), dp(z), *d, *map(dp, c), *dc)
        outs.append(c)
    for m in range(3):
        bits_equal(outs[0][m], outs[1][m], "coarsen_b")
    # pitch: unit displacement scales (the first branch) and scales that
    # take the second branch at some nodes
    ones = np.ones(n1(d))
    mixed = np.where(rng.uniform(size=n1(d)) < 0.3, 1e-5, rng.uniform(0.5, 2.0, n1(d)))
    for scale in (ones, mixed):
        for zeit, dtheta in ((0.0, 0.0), (1.3, 0.07), (4.0, -0.08)):
            dtheta0, omegas, xh, yh = 0.0873, 0.2, 0.25, 0.01
            po = [np.zeros(n2(d)) for _ in range(6)]
            pr = [np.zeros(n2(d)) for _ in range(6)]
            o.or_pitch(D(np.cos(dtheta)), D(np.sin(dtheta)), D(np.cos(omegas * zeit)), D(dtheta0), D(omegas), D(xh),
                       D(yh), dp(x), dp(y), dp(z), dp(scale), dp(scale), *map(dp, po), *d)
            r.ref_pitch(D(zeit), D(dtheta), D(dtheta0), D(omegas), D(xh), D(yh), dp(x), dp(y), dp(z), dp(scale),
                        dp(scale), dp(scale), *map(dp, pr), *d)
            for m in range(6):
                bits_equal(po[m], pr[m], f"pitch component {m} zeit {zeit}")


@needs_ref
def test_cut_v_cut_x_every_orientation():
    o, r = _pair(moving=1, lmet=7, nharms=1, harbal=1, hbmove=1, omega=0.3)
    rng = np.random.default_rng(17)
    d1, d2 = (7, 6, 5), (6, 8, 7)
    nsm = 3
    n = 0
    for idir1, side1, idir2, side2, swap in itertools.product((1, 2, 3), (0, 1), (1, 2, 3), (0, 1), (0, 1)):
        for sg1 in itertools.product((1, -1), repeat=2):
            for sg2 in ((1, 1), (-1, -1)):
                rec = cut_record(rng, d1, d2, idir1, side1, idir2, side2, swap, sg1, sg2)
                v2 = rng.uniform(size=n1(d2))
                va, vb = rng.uniform(size=n1(d1)), None
                vb = va.copy()
                o.or_cut_layer(0, 1, *d1, dp(va), *d2, dp(v2), ip(rec))
                r.ref_cut_v(*d1, dp(vb), *d2, dp(v2), ip(rec))
                bits_equal(va, vb, f"cut_v {rec.tolist()}")
                x2 = [rng.uniform(size=n2(d2) * nsm) for _ in range(3)]
                xa = [rng.uniform(size=n2(d1) * nsm) for _ in range(3)]
                xb = [a.copy() for a in xa]
                for m in range(3):
                    o.or_cut_layer(1, nsm, *d1, dp(xa[m]), *d2, dp(x2[m]), ip(rec))
                r.ref_cut_x(*d1, *map(dp, xb), *d2, *map(dp, x2), ip(rec))
                for m in range(3):
                    bits_equal(xa[m], xb[m], f"cut_x {rec.tolist()}")
                n += 1
    assert n == 3 * 2 * 3 * 2 * 2 * 8


# ---- closed forms (no reference needed) ------------------------------------

def _unpack_faces(s, d, lmet):
    im, jm, km = d
    return s.reshape(km + 2, jm + 2, im + 2, lmet)


def test_face_vectors_close_every_cell():
    """Sum over a cell's six faces of (outward unit normal x area) = 0: the
    face vectors of imetrix_b are halved diagonal cross products, so each
    face vector is exact in the edge vectors and they cancel around a
    closed cell up to rounding."""
    o = Oracle()
    o.set_config(cfg())
    rng = np.random.default_rng(23)
    d = (9, 7, 6)
    x, y, z = nodes(d, rng, amp=0.2)
    s = [np.zeros(n2(d) * 4) for _ in range(3)]
    assert o.or_imetrix_b(dp(x), dp(y), dp(z), None, None, None, *map(dp, s), *d, 4, 1, 0) == 0
    si, sj, sk = (_unpack_faces(a, d, 4) for a in s)
    im, jm, km = d
    vi = si[..., :3] * si[..., 3:4]
    vj = sj[..., :3] * sj[..., 3:4]
    vk = sk[..., :3] * sk[..., 3:4]
    c = (slice(1, km), slice(1, jm), slice(1, im))
    tot = (vi[1:km, 1:jm, 2:im + 1] - vi[c] + vj[1:km, 2:jm + 1, 1:im] - vj[c] + vk[2:km + 1, 1:jm, 1:im] - vk[c])
    assert np.abs(tot).max() < 1e-13 * np.abs(vi).max()


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_affine_volumes_and_metric_derivatives(seed):
    """An affine map P = A (i-1, j-1, k-1) + c: every cell volume is
    det(A), every face area vector is the cofactor of the two edge vectors,
    and the metric derivatives on every face are the rows of A^-1 -- the
    inverse Jacobian of the map -- to rounding."""
    rng = np.random.default_rng(seed)
    A = np.eye(3) + 0.3 * rng.uniform(-1, 1, (3, 3))
    if np.linalg.det(A) < 0:
        A[:, 0] = -A[:, 0]
    c0 = rng.uniform(-1, 1, 3)
    o = Oracle()
    o.set_config(cfg(viscous=1, reyno=1e5))
    d = (6, 5, 7)
    x, y, z = nodes(d, rng, affine=(A, c0))
    v = np.zeros(n1(d))
    assert o.or_bvols(dp(x), dp(y), dp(z), dp(v), *d) == 0
    im, jm, km = d
    vv = v.reshape(km + 1, jm + 1, im + 1)[1:km, 1:jm, 1:im]
    det = np.linalg.det(A)
    assert np.abs(vv - det).max() < 1e-13 * det
    o.or_volex_b(dp(v), *d)
    # the halo nodes of an affine map extended linearly keep the Jacobian
    # exact at the block faces too
    xe, ye, ze = nodes(d, rng, affine=(A, c0))
    i, j, k = np.meshgrid(np.arange(im + 2), np.arange(jm + 2), np.arange(km + 2), indexing="ij")
    P = np.stack([i - 1.0, j - 1.0, k - 1.0], -1) @ A.T + c0
    xe, ye, ze = (np.ascontiguousarray(P[..., m].transpose(2, 1, 0)).reshape(-1) for m in range(3))
    dd = [np.zeros(n0(d) * 3) for _ in range(9)]
    arr = (ctypes.POINTER(ctypes.c_double) * 9)(*[dp(a) for a in dd])
    o.or_vmetrix_b(dp(xe), dp(ye), dp(ze), arr, dp(v), *d, 1)
    Ainv = np.linalg.inv(A)
    fams = [(im, jm - 1, km - 1), (im - 1, jm, km - 1), (im - 1, jm - 1, km)]
    for f, (ni, nj, nk) in enumerate(fams):
        for row in range(3):
            a = dd[3 * f + row].reshape(km, jm, im, 3)[:nk, :nj, :ni]
            assert np.abs(a - Ainv[row]).max() < 1e-12 * np.abs(Ainv).max(), (f, row)


def test_coarsen_keeps_every_second_node_and_pitch_is_rigid():
    o = Oracle()
    o.set_config(cfg(moving=1, lmet=7))
    rng = np.random.default_rng(29)
    d, dc = (9, 7, 5), (5, 4, 3)
    x, y, z = nodes(d, rng)
    c = [np.zeros(n2(dc)) for _ in range(3)]
    o.or_coarsen_b(dp(x), dp(y), dp(z), *d, *map(dp, c), *dc)
    for m, a in enumerate((x, y, z)):
        fine = a.reshape(d[2] + 2, d[1] + 2, d[0] + 2)[1:-1:2, 1:-1:2, 1:-1:2]
        coarse = c[m].reshape(dc[2] + 2, dc[1] + 2, dc[0] + 2)[1:-1, 1:-1, 1:-1]
        assert np.array_equal(fine, coarse)
    th, om, dth0, xh, yh, zeit = 0.11, 0.3, 0.2, 0.25, -0.1, 0.7
    p = [np.zeros(n2(d)) for _ in range(6)]
    ones = np.ones(n1(d))
    o.or_pitch(D(np.cos(th)), D(np.sin(th)), D(np.cos(om * zeit)), D(dth0), D(om), D(xh), D(yh), dp(x), dp(y),
               dp(z), dp(ones), dp(ones), *map(dp, p), *d)
    sh = (d[2] + 2, d[1] + 2, d[0] + 2)
    inner = (slice(1, -1),) * 3
    X0, Y0 = x.reshape(sh)[inner] - xh, y.reshape(sh)[inner] - yh
    X, Y = p[0].reshape(sh)[inner] - xh, p[1].reshape(sh)[inner] - yh
    assert np.abs(np.hypot(X, Y) - np.hypot(X0, Y0)).max() < 1e-13
    ang = (np.arctan2(Y, X) - np.arctan2(Y0, X0) - th + np.pi) % (2 * np.pi) - np.pi
    assert np.abs(ang).max() < 1e-12
    rate = om * dth0 * np.cos(om * zeit)        # d theta / dt
    assert np.abs(p[3].reshape(sh)[inner] - (-rate * Y)).max() < 1e-13
    assert np.abs(p[4].reshape(sh)[inner] - rate * X).max() < 1e-13
