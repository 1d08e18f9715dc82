"""Pin the C restatement (oracle/cosa_oracle.c) against COSA's own compiled
routines (oracle/_ref, built from /root/reference by oracle/build_ref.py).

Each test feeds both the same random block inputs and requires identical
binary64 results (both builds keep every operation separately rounded).
Skipped where the reference is not available (e.g. on the GPU host); the
same comparisons are frozen into tests/golden/ fixtures by
tests/golden/make_golden.py and re-checked by test_oracle_golden.py.
"""
import ctypes

import numpy as np
import pytest

from blockgen import all_faces_patches, bctopo_of, box_block, box_block_moving, n3, prim_to_cons, rand_prim
from cosa_amd.config import default_config, hb_operator
from oracle_bind import Oracle, Ref, dp, ip, ref_available

pytestmark = [pytest.mark.ref, pytest.mark.skipif(not ref_available(), reason="oracle/_ref not built")]

DIMS = [(9, 7, 5), (6, 8, 4), (5, 5, 6)]


@pytest.fixture(scope="module")
def libs():
    return Oracle(), Ref()


def setcfg(libs, nharms=0, **kw):
    o, r = libs
    if nharms:
        kw.setdefault("omega", 0.7)
    c = default_config(nharms=nharms, harbal=int(nharms > 0), **kw)
    d = hb_operator(nharms, c["omega"]) if nharms else None
    o.set_config(c, d)
    r.set_config(c, d)
    return c


def same(a, b, what=""):
    np.testing.assert_array_equal(a, b, err_msg=what)


@pytest.mark.parametrize("dims", DIMS)
@pytest.mark.parametrize("nh", [0, 1])
def test_cons_prim(libs, dims, nh):
    o, r = libs
    setcfg(libs, nharms=nh)
    rng = np.random.default_rng(1)
    ns = 2 * nh + 1
    q = prim_to_cons(rand_prim(rng, *dims, nsnap=ns), *dims, nsnap=ns)
    a, b = q.copy(), q.copy()
    w = np.zeros_like(q)
    o.or_cons2prim_b(dp(a), *dims, 0)
    r.ref_cons2prim_b(dp(b), dp(w), *dims, 0)
    same(a, b, "cons2prim")
    o.or_prim2cons_b(dp(a), *dims)
    r.ref_prim2cons_b(dp(b), dp(w), *dims)
    same(a, b, "prim2cons")


@pytest.mark.parametrize("types", [(30, 14, 30, 14, 80, 80), (4, 30, 14, 80, 80, 4), (14, 14, 80, 30, 30, 14)])
@pytest.mark.parametrize("alpha_deg", [1.25, 181.0])
def test_bbc(libs, types, alpha_deg):
    o, r = libs
    dims = (9, 7, 6)
    setcfg(libs, machfs=0.5, alpha_deg=alpha_deg)
    rng = np.random.default_rng(2)
    si, sj, sk, vol = box_block(*dims, seed=4)
    q = rand_prim(rng, *dims)
    bt = bctopo_of(all_faces_patches(*dims, types))
    a, b = q.copy(), q.copy()
    dist = np.zeros(vol.size)
    o.or_bbc(dp(a), dp(si), dp(sj), dp(sk), ip(bt), 6, *dims)
    r.ref_bbc(dp(b), dp(si), dp(sj), dp(sk), dp(dist), 1, ip(bt), 6, *dims)
    same(a, b, "bbc")


def test_bbc_supersonic(libs):
    """Supersonic in/outflow branches of the farfield condition."""
    o, r = libs
    dims = (7, 6, 5)
    setcfg(libs, machfs=1.6, alpha_deg=10.0)
    rng = np.random.default_rng(3)
    si, sj, sk, vol = box_block(*dims, seed=5)
    q = rand_prim(rng, *dims)
    q.reshape(5, -1)[1] += 1.2
    bt = bctopo_of(all_faces_patches(*dims, (30, 30, 30, 30, 30, 30)))
    a, b = q.copy(), q.copy()
    o.or_bbc(dp(a), dp(si), dp(sj), dp(sk), ip(bt), 6, *dims)
    r.ref_bbc(dp(b), dp(si), dp(sj), dp(sk), dp(np.zeros(vol.size)), 1, ip(bt), 6, *dims)
    same(a, b, "bbc supersonic")


MOVING = dict(moving=1, hbmove=1, lmet=7)


@pytest.mark.parametrize("nh", [1, 2])
def test_bbc_moving(libs, nh):
    """Inviscid wall on a pitching grid: the normal velocity is taken
    relative to the face velocity (metric components 5-7 of each snapshot)."""
    o, r = libs
    dims = (8, 7, 5)
    setcfg(libs, nharms=nh, machfs=0.5, alpha_deg=2.0, **MOVING)
    ns = 2 * nh + 1
    (si, sj, sk), vol, _ = box_block_moving(*dims, ns, seed=8)
    q = rand_prim(np.random.default_rng(31), *dims, nsnap=ns)
    bt = bctopo_of(all_faces_patches(*dims, (14, 30, 14, 14, 80, 14)))
    a, b = q.copy(), q.copy()
    o.or_bbc(dp(a), dp(si), dp(sj), dp(sk), ip(bt), 6, *dims)
    r.ref_bbc(dp(b), dp(si), dp(sj), dp(sk), dp(np.zeros(vol.size)), 1, ip(bt), 6, *dims)
    same(a, b, "bbc moving")
    assert not np.array_equal(a, q)


@pytest.mark.parametrize("nh", [1, 2])
def test_lambda_moving(libs, nh):
    """Spectral radii with the node velocity of each snapshot subtracted."""
    o, r = libs
    dims = (8, 7, 5)
    setcfg(libs, nharms=nh, irsop=2, cfl=3.0, rat=1.7, **MOVING)
    ns = 2 * nh + 1
    (si, sj, sk), vol, (xd, yd, zd) = box_block_moving(*dims, ns, seed=9)
    q = rand_prim(np.random.default_rng(32), *dims, nsnap=ns)
    N3 = n3(*dims)
    n1 = (dims[0] + 1) * (dims[1] + 1) * (dims[2] + 1)
    lam_o, lam_r = np.zeros(N3 * 5 * ns), np.zeros(N3 * 5 * ns)
    o.or_lambdab(dp(q), dp(si), dp(sj), dp(sk), dp(xd), dp(yd), dp(zd), dp(lam_o), *dims)
    cut = np.zeros(n1 * ns)
    r.ref_lambdab(dp(q), dp(si), dp(sj), dp(sk), dp(xd), dp(yd), dp(zd), dp(vol), dp(np.zeros(N3 * ns)),
                  dp(np.zeros(N3 * 5 * ns)), dp(lam_r), dp(cut), dp(cut), *dims)
    same(lam_o, lam_r, "lambda moving")
    lam_s = np.zeros_like(lam_o)
    o.or_lambdab(dp(q), dp(si), dp(sj), dp(sk), dp(np.zeros_like(xd)), dp(np.zeros_like(yd)), dp(np.zeros_like(zd)),
                 dp(lam_s), *dims)
    assert not np.array_equal(lam_s, lam_o)


@pytest.mark.parametrize("idir", [1, 2, 3])
def test_roflux_moving(libs, idir):
    """Roe flux with the face velocity ubn (lmet = 7), three snapshots."""
    o, r = libs
    dims = (8, 6, 5)
    setcfg(libs, nharms=1, efx_setup_f=1, efx_all_f=1, cntrpy_f=0.3, entfx_f_w=(0, 0, 1), **MOVING)
    (si, sj, sk), vol, _ = box_block_moving(*dims, 3, seed=10)
    s = (si, sj, sk)[idir - 1]
    rng = np.random.default_rng(40 + idir)
    qr = rand_prim(rng, *dims, nsnap=3)
    ql = rand_prim(rng, *dims, nsnap=3)
    nf = dims[0] * dims[1] * dims[2] * 5 * 3
    fo, fr = np.zeros(nf), np.zeros(nf)
    o.or_roflux_b(idir, dp(qr), dp(ql), dp(fo), dp(s), *dims)
    scratch = [np.zeros_like(qr) for _ in range(2)]
    r.ref_roflux_b(idir, dp(qr), dp(ql), dp(scratch[0]), dp(scratch[1]), dp(fr), dp(s), *dims)
    same(fo, fr, "roflux moving")


@pytest.mark.parametrize("irsop", [1, 2, 3])
@pytest.mark.parametrize("nh", [0, 2])
def test_lambda_deltat(libs, irsop, nh):
    o, r = libs
    dims = (8, 7, 5)
    setcfg(libs, nharms=nh, irsop=irsop, cfl=3.0, rat=1.7, psirs=0.125)
    ns = 2 * nh + 1
    rng = np.random.default_rng(4)
    si, sj, sk, vol = box_block(*dims, seed=6)
    q = rand_prim(rng, *dims, nsnap=ns)
    N3 = n3(*dims)
    n1 = (dims[0] + 1) * (dims[1] + 1) * (dims[2] + 1)
    n2 = (dims[0] + 2) * (dims[1] + 2) * (dims[2] + 2)
    lam_o, lam_r = np.zeros(N3 * 5 * ns), np.zeros(N3 * 5 * ns)
    xd = np.zeros(n2)
    o.or_lambdab(dp(q), dp(si), dp(sj), dp(sk), dp(xd), dp(xd), dp(xd), dp(lam_o), *dims)
    work = np.zeros(N3 * 5 * ns)
    mut = np.zeros(N3 * ns)
    cut = np.zeros(n1 * ns)
    r.ref_lambdab(dp(q), dp(si), dp(sj), dp(sk), dp(xd), dp(xd), dp(xd), dp(vol), dp(mut), dp(work), dp(lam_r),
                  dp(cut), dp(cut), *dims)
    same(lam_o, lam_r, "lambda")
    out_o = [np.zeros(n1 * ns) for _ in range(4)]
    out_r = [np.zeros(n1 * ns) for _ in range(8)]
    o.or_bdeltat(dp(out_o[0]), dp(lam_o), dp(out_o[1]), dp(out_o[2]), dp(out_o[3]), *dims)
    r.ref_bdeltat(dp(out_r[0]), dp(out_r[4]), dp(lam_r), dp(np.zeros_like(lam_r)), dp(out_r[1]), dp(out_r[2]),
                  dp(out_r[3]), dp(out_r[5]), dp(out_r[6]), dp(out_r[7]), *dims)
    for m in range(4):
        same(out_o[m], out_r[m], f"deltat {m}")


@pytest.mark.parametrize("idir", [1, 2, 3])
@pytest.mark.parametrize("ilim", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("nl,nl_crs", [(1, 4), (2, 2)])
def test_muscl(libs, idir, ilim, nl, nl_crs):
    o, r = libs
    dims = (9, 7, 6)
    setcfg(libs, ilimf=ilim, epslimf=0.05, rkap=0.333333, nl_crs=nl_crs)
    rng = np.random.default_rng(10 + idir + 3 * ilim)
    q = rand_prim(rng, *dims)
    bt = bctopo_of(all_faces_patches(*dims, (14, 30, 14, 14, 80, 14)))
    ao = [np.zeros_like(q) for _ in range(5)]
    ar = [np.zeros_like(q) for _ in range(5)]
    o.or_muscl_b(idir, nl, ip(bt), 6, dp(q), *[dp(a) for a in ao], *dims)
    r.ref_muscl_b(idir, nl, ip(bt), 6, dp(q), *[dp(a) for a in ar], *dims)
    for m, (a, b) in enumerate(zip(ao, ar)):
        same(a, b, f"muscl array {m}")


@pytest.mark.parametrize("idir", [1, 2, 3])
@pytest.mark.parametrize("efx", [0, 1])
def test_roflux_resid(libs, idir, efx):
    o, r = libs
    dims = (8, 6, 5)
    kw = {}
    if efx:
        kw = dict(efx_setup_f=1, efx_all_f=1, cntrpy_f=0.3, entfx_f_w=(0, 1, 0), eig_limit_f=1, eig_lim_all_f=1,
                  eig_cutoff_f=0.05)
    setcfg(libs, **kw)
    rng = np.random.default_rng(20 + idir)
    si, sj, sk, vol = box_block(*dims, seed=7)
    s = (si, sj, sk)[idir - 1]
    qr = rand_prim(rng, *dims)
    ql = rand_prim(rng, *dims)
    nf = dims[0] * dims[1] * dims[2] * 5
    fo, fr = np.zeros(nf), np.zeros(nf)
    o.or_roflux_b(idir, dp(qr), dp(ql), dp(fo), dp(s), *dims)
    scratch = [np.zeros_like(qr) for _ in range(2)]
    r.ref_roflux_b(idir, dp(qr), dp(ql), dp(scratch[0]), dp(scratch[1]), dp(fr), dp(s), *dims)
    same(fo, fr, "roflux")
    res0 = rng.uniform(-1, 1, qr.size)
    a, b = res0.copy(), res0.copy()
    o.or_bresid(idir, dp(fo), dp(a), *dims)
    r.ref_bresid(idir, dp(fr), dp(b), *dims)
    same(a, b, "resid")


@pytest.mark.parametrize("nh", [1, 2, 4])
def test_src_hb(libs, nh):
    o, r = libs
    dims = (6, 5, 4)
    setcfg(libs, nharms=nh, omega=0.314)
    ns = 2 * nh + 1
    rng = np.random.default_rng(30 + nh)
    q = prim_to_cons(rand_prim(rng, *dims, nsnap=ns), *dims, nsnap=ns)
    vol = rng.uniform(0.5, 1.5, (dims[0] + 1) * (dims[1] + 1) * (dims[2] + 1))
    res = rng.uniform(-1, 1, q.size)
    a, b = res.copy(), res.copy()
    ho, hr = np.zeros_like(q), np.zeros_like(q)
    o.or_src_bhb(dp(q), dp(vol), dp(a), dp(ho), *dims)
    r.ref_src_bhb(dp(q), dp(vol), dp(b), dp(hr), *dims)
    same(a, b, "src_hb res")
    same(ho, hr, "src_hb qhb")


@pytest.mark.parametrize("nh", [0, 1])
def test_resg_brms(libs, nh):
    o, r = libs
    dims = (7, 6, 5)
    setcfg(libs, nharms=nh)
    ns = 2 * nh + 1
    rng = np.random.default_rng(40)
    N = n3(*dims) * 5 * ns
    res, rhs = rng.uniform(-1, 1, N), rng.uniform(-1, 1, N)
    bo, br = np.zeros(8), np.zeros(8)
    vol = np.ones((dims[0] + 1) * (dims[1] + 1) * (dims[2] + 1))
    if nh == 0:
        o.or_resg_brms(dp(res), dp(rhs), dp(bo), *dims)
        r.ref_resg_brms(dp(res), dp(res), dp(rhs), dp(vol), dp(br), *dims)
    else:
        o.or_resg_brms_hb(dp(res), dp(bo), *dims)
        r.ref_resg_brms_hb(dp(res), dp(res.copy()), dp(vol), dp(br), *dims)
    same(bo[:6], br[:6], "rms partial sums")


@pytest.mark.parametrize("irsop", [1, 2, 3])
@pytest.mark.parametrize("dims", [(9, 7, 6), (6, 5, 3)])
def test_cirs(libs, irsop, dims):
    o, r = libs
    setcfg(libs, irsop=irsop, rat=1.8)
    rng = np.random.default_rng(50 + irsop)
    N = n3(*dims) * 5
    n1 = (dims[0] + 1) * (dims[1] + 1) * (dims[2] + 1)
    res = rng.uniform(-1, 1, N)
    dtv = rng.uniform(0.5, 2.0, n1)
    betas = [rng.uniform(0.0, 1.5, n1) for _ in range(3)]
    a, b = res.copy(), res.copy()
    o.or_cirs_b(dp(a), dp(dtv), *[dp(x) for x in betas], *dims)
    work = np.zeros(max(dims) * 5 * 4)
    r.ref_cirs_b(dp(b), dp(dtv), *[dp(x) for x in betas], dp(dtv), *[dp(x) for x in betas], dp(work), *dims)
    same(a, b, "cirs")


def test_tridi(libs):
    o, r = libs
    rng = np.random.default_rng(60)
    for n in (2, 3, 7, 40):
        a, b, c = rng.uniform(-1, 0, n), rng.uniform(2.5, 3, n), rng.uniform(-1, 0, n)
        a[0] = 0
        c[-1] = 0
        x = rng.uniform(-1, 1, n)
        xo, xr = x.copy(), x.copy()
        o.or_tridi(dp(a), dp(b), dp(c), dp(xo), n)
        r.ref_tridi(dp(a), dp(b), dp(c), dp(xr), n, n)
        same(xo, xr, f"tridi n={n}")


@pytest.mark.parametrize("nh", [0, 1])
def test_update(libs, nh):
    o, r = libs
    dims = (7, 6, 5)
    setcfg(libs, nharms=nh)
    ns = 2 * nh + 1
    rng = np.random.default_rng(70)
    N = n3(*dims) * 5 * ns
    n1 = (dims[0] + 1) * (dims[1] + 1) * (dims[2] + 1)
    res, qold = rng.uniform(-1, 1, N), rng.uniform(0.5, 1.5, N)
    dtv = rng.uniform(0.1, 1.0, n1 * ns)
    q0 = rng.uniform(0.5, 1.5, N)
    a, b = q0.copy(), q0.copy()
    o.or_update_b(dp(res), dp(a), dp(qold), dp(dtv), ctypes.c_double(0.333333), *dims)
    z = np.zeros(N * 3)
    r.ref_update(dp(res.copy()), dp(b), dp(qold), dp(np.ones(n1)), dp(dtv), dp(dtv), dp(z), dp(z), dp(z), dp(z),
                 dp(z), ctypes.c_double(0.333333), *dims)
    same(a, b, "update")


@pytest.mark.parametrize("ho", [0, 1])
def test_restrict(libs, ho):
    o, r = libs
    setcfg(libs)
    f = (9, 9, 5)
    c = (5, 5, 3)
    rng = np.random.default_rng(80)
    q1 = rng.uniform(0.5, 1.5, n3(*f) * 5)
    v1 = rng.uniform(0.5, 1.5, (f[0] + 1) * (f[1] + 1) * (f[2] + 1))
    v2 = rng.uniform(3.5, 9.5, (c[0] + 1) * (c[1] + 1) * (c[2] + 1))
    qo, qr = np.zeros(n3(*c) * 5), np.zeros(n3(*c) * 5)
    o.or_rest1_b(*f, *c, dp(q1), dp(qo), dp(v1), dp(v2))
    r.ref_rest1_b(*f, *c, dp(q1), dp(qr), dp(v1), dp(v2))
    same(qo, qr, "rest1")
    res1 = rng.uniform(-1, 1, q1.size)
    ro, rr = np.zeros(n3(*c) * 5), np.zeros(n3(*c) * 5)
    o.or_rest2_b(ho, *f, *c, dp(ro), dp(res1))
    r.ref_rest2_b(ho, *f, *c, dp(np.zeros_like(res1)), dp(rr), dp(res1))
    same(ro, rr, "rest2")


@pytest.mark.parametrize("bln", [0, 1])
def test_prolong_kernel(libs, bln):
    o, r = libs
    setcfg(libs, bln_prol=bln)
    f = (9, 9, 5)
    c = (5, 5, 3)
    rng = np.random.default_rng(90)
    q2 = rng.uniform(-0.1, 0.1, n3(*c) * 5)
    v1 = rng.uniform(0.5, 1.5, (f[0] + 1) * (f[1] + 1) * (f[2] + 1))
    q1 = rng.uniform(0.5, 1.5, n3(*f) * 5)
    a, b = q1.copy(), q1.copy()
    o.or_prol_b(bln, *f, *c, dp(a), dp(q2), dp(v1))
    r.ref_prol_b(bln, *f, *c, dp(b), dp(q2), dp(v1))
    same(a, b, "prolong")


def test_edges_corners(libs):
    o, r = libs
    setcfg(libs)
    dims = (7, 6, 5)
    rng = np.random.default_rng(95)
    q = rng.uniform(0.5, 1.5, n3(*dims) * 5)
    a, b = q.copy(), q.copy()
    mut = np.zeros(n3(*dims))
    o.or_q_bedges(dp(a), *dims)
    r.ref_q_bedges(dp(b), dp(mut), *dims)
    same(a, b, "edges")
    o.or_q_bcorns(dp(a), *dims)
    r.ref_q_bcorns(dp(b), dp(mut), *dims)
    same(a, b, "corners")
