"""Pin the viscous / Wilcox k-omega part of the C restatement
(oracle/cosa_oracle.c) against COSA's own compiled routines
(oracle/_ref, built from /root/reference by oracle/build_ref.py).

Same contract as test_oracle_vs_ref.py: identical random inputs, identical
binary64 outputs required, except where a routine calls the math library's
pow (Sutherland's law t**1.5, the isentropic wall density) -- there glibc's
pow (oracle) and flang's runtime pow (reference) are both correctly rounded
in practice, and the tests still require equality.
"""
import numpy as np
import pytest

from blockgen import all_faces_patches, bctopo_of, box_block, n3, prim_to_cons, rand_prim
from cosa_amd.config import default_config, hb_operator, kw_turbulence
from oracle_bind import Oracle, Ref, dp, ip, ref_available

pytestmark = [pytest.mark.ref, pytest.mark.skipif(not ref_available(), reason="oracle/_ref not built")]

DIMS = (8, 7, 6)


@pytest.fixture(scope="module")
def libs():
    return Oracle(), Ref()


def setcfg(libs, nharms=0, kom=True, **kw):
    o, r = libs
    base = kw_turbulence(0.4, 2.0e5, tu=1e-3, mutfar=0.1) if kom else dict(viscous=1)
    base.update(kw)
    if nharms:
        base.setdefault("omega", 0.7)
        base["harbal"] = 1
    c = default_config(nharms=nharms, stemp=0.383, **base)
    d = hb_operator(nharms, c["omega"]) if nharms else None
    o.set_config(c, d)
    r.set_config(c, d)
    return c


def same(a, b, what=""):
    np.testing.assert_array_equal(a, b, err_msg=what)


def prim7(rng, dims, nsnap=1):
    q = rand_prim(rng, *dims, npde=7, nsnap=nsnap)
    v = q.reshape(nsnap, 7, -1)
    v[:, 5] = rng.uniform(1e-4, 1e-2, v[:, 5].shape)      # k
    v[:, 6] = rng.uniform(5.0, 50.0, v[:, 6].shape)       # omega
    return q


def n1(dims):
    return (dims[0] + 1) * (dims[1] + 1) * (dims[2] + 1)


def n0(dims):
    return dims[0] * dims[1] * dims[2]


def mder(rng, dims, nsnap=1):
    """Nine metric-derivative arrays (3, imax, jmax, kmax) near identity."""
    out = []
    for m in range(9):
        a = rng.uniform(-0.2, 0.2, (nsnap, n0(dims), 3))
        a[:, :, m % 3] += 3.0 if m // 3 == m % 3 else 0.0
        out.append(a.reshape(-1).copy())
    return out


@pytest.mark.parametrize("nh", [0, 1])
def test_cons_prim_kom(libs, nh):
    o, r = libs
    setcfg(libs, nharms=nh)
    ns = 2 * nh + 1
    q = prim_to_cons(prim7(np.random.default_rng(1), DIMS, ns), *DIMS, npde=7, nsnap=ns)
    a, b = q.copy(), q.copy()
    w = np.zeros_like(q)
    o.or_cons2prim_b(dp(a), *DIMS, 0)
    r.ref_cons2prim_b(dp(b), dp(w), *DIMS, 0)
    same(a, b, "cons2prim")
    o.or_prim2cons_b(dp(a), *DIMS)
    r.ref_prim2cons_b(dp(b), dp(w), *DIMS)
    same(a, b, "prim2cons")


@pytest.mark.parametrize("wallwilc,roughk", [(1, 0.0), (1, 3e-4), (0, 0.0)])
@pytest.mark.parametrize("types", [(30, 1500, 1500, 30, 80, 80), (1500, 4, 30, 1501, 1500, 80)])
def test_bbc_kom(libs, wallwilc, roughk, types):
    o, r = libs
    setcfg(libs, machfs=0.4, alpha_deg=3.0, wallwilc=wallwilc, roughk=roughk)
    rng = np.random.default_rng(2)
    si, sj, sk, vol = box_block(*DIMS, seed=4)
    q = prim7(rng, DIMS)
    dist = rng.uniform(1e-4, 1e-2, n1(DIMS))
    bt = bctopo_of(all_faces_patches(*DIMS, types))
    a, b = q.copy(), q.copy()
    o.or_bbc_dist(dp(a), dp(si), dp(sj), dp(sk), dp(dist), ip(bt), 6, *DIMS)
    r.ref_bbc(dp(b), dp(si), dp(sj), dp(sk), dp(dist), 1, ip(bt), 6, *DIMS)
    same(a, b, "bbc kom")
    # the viscous wall put k = 0 and a positive omega on its halo layers
    k = a.reshape(7, -1)[5]
    assert np.count_nonzero(k == 0.0) > 0 and np.all(np.isfinite(a))


@pytest.mark.parametrize("kom", [True, False])
@pytest.mark.parametrize("irsop", [1, 2, 3, 5, 6])
def test_lambda_deltat_viscous(libs, kom, irsop):
    o, r = libs
    setcfg(libs, kom=kom, irsop=irsop, cfl=3.0, cflt=1.5, rat=1.7, ratt=1.3, psirs=0.125, cdff=2.0)
    rng = np.random.default_rng(4)
    si, sj, sk, vol = box_block(*DIMS, seed=6)
    npde = 7 if kom else 5
    q = prim7(rng, DIMS) if kom else rand_prim(rng, *DIMS)
    N3 = n3(*DIMS)
    mut = rng.uniform(0.0, 50.0, N3)
    xd = np.zeros((DIMS[0] + 2) * (DIMS[1] + 2) * (DIMS[2] + 2))
    lam_o, lam_r = np.zeros(N3 * npde), np.zeros(N3 * npde)
    lv_o, lv_r = np.zeros(N3 * npde), np.zeros(N3 * npde)
    o.or_lambdab_v(dp(q), dp(si), dp(sj), dp(sk), dp(xd), dp(xd), dp(xd), dp(vol), dp(mut), dp(lam_o), dp(lv_o),
                   *DIMS)
    cut = np.zeros(n1(DIMS))
    r.ref_lambdab(dp(q), dp(si), dp(sj), dp(sk), dp(xd), dp(xd), dp(xd), dp(vol), dp(mut), dp(lv_r), dp(lam_r),
                  dp(cut), dp(cut), *DIMS)
    m = np.zeros((npde, N3), bool)
    m[:3] = True
    same(lam_o.reshape(npde, -1)[:3], lam_r.reshape(npde, -1)[:3], "lambdas")
    same(lv_o.reshape(npde, -1)[:3], lv_r.reshape(npde, -1)[:3], "viscous lambdas")
    oo = [np.zeros(n1(DIMS)) for _ in range(8)]
    rr = [np.zeros(n1(DIMS)) for _ in range(8)]
    o.or_bdeltat_v(*[dp(x) for x in oo[:2]], dp(lam_o), dp(lv_o), *[dp(x) for x in oo[2:]], *DIMS)
    r.ref_bdeltat(dp(rr[0]), dp(rr[1]), dp(lam_r), dp(lv_r), *[dp(x) for x in rr[2:]], *DIMS)
    for k in range(8):
        same(oo[k], rr[k], f"deltat array {k}")
    if irsop in (5, 6):   # the viscous floor of the smoothing is active somewhere
        assert np.any(oo[2] > 0)


@pytest.mark.parametrize("idir", [1, 2, 3])
@pytest.mark.parametrize("foturb,nl", [(0, 1), (1, 1), (0, 2)])
def test_muscl_kom(libs, idir, foturb, nl):
    o, r = libs
    setcfg(libs, ilimf=4, ilimt=2, epslimf=0.05, epslimt=0.02, foturb=foturb, nl_crs=2)
    q = prim7(np.random.default_rng(10 + idir), DIMS)
    bt = bctopo_of(all_faces_patches(*DIMS, (1500, 30, 14, 1500, 80, 30)))
    ao = [np.zeros_like(q) for _ in range(5)]
    ar = [np.zeros_like(q) for _ in range(5)]
    o.or_muscl_b(idir, nl, ip(bt), 6, dp(q), *[dp(a) for a in ao], *DIMS)
    r.ref_muscl_b(idir, nl, ip(bt), 6, dp(q), *[dp(a) for a in ar], *DIMS)
    for m, (a, b) in enumerate(zip(ao, ar)):
        same(a, b, f"muscl array {m}")


@pytest.mark.parametrize("idir", [1, 2, 3])
@pytest.mark.parametrize("efx", [0, 1])
def test_roflux_kom(libs, idir, efx):
    o, r = libs
    kw = {}
    if efx:
        kw = dict(efx_setup_t=1, efx_all_t=1, cntrpy_t=0.2, entfx_t_w=(0, 1, 0), eig_limit_t=1, eig_cutoff_t=0.05,
                  efx_setup_f=1, cntrpy_f=0.1, entfx_f_w=(0, 0, 1))
    setcfg(libs, nharms=1, **kw)
    rng = np.random.default_rng(20 + idir)
    si, sj, sk, vol = box_block(*DIMS, seed=7)
    s = (si, sj, sk)[idir - 1]
    qr, ql = prim7(rng, DIMS, 3), prim7(rng, DIMS, 3)
    fo, fr = np.zeros(n0(DIMS) * 7 * 3), np.zeros(n0(DIMS) * 7 * 3)
    o.or_roflux_b(idir, dp(qr), dp(ql), dp(fo), dp(s), *DIMS)
    sc = [np.zeros_like(qr) for _ in range(2)]
    r.ref_roflux_b(idir, dp(qr), dp(ql), dp(sc[0]), dp(sc[1]), dp(fr), dp(s), *DIMS)
    same(fo, fr, "roflux kom")
    assert np.count_nonzero(fo.reshape(3, 7, -1)[:, 5:7]) > 0.5 * fo.reshape(3, 7, -1)[:, 5:7].size


VTYPES = (1500, 30, 1500, 14, 80, 1501)


@pytest.mark.parametrize("idir", [1, 2, 3])
@pytest.mark.parametrize("calmut", [0, 1])
def test_q_face(libs, idir, calmut):
    o, r = libs
    setcfg(libs)
    rng = np.random.default_rng(30 + idir)
    q = prim7(rng, DIMS)
    mut = rng.uniform(0, 20, n3(*DIMS))
    bt = bctopo_of(all_faces_patches(*DIMS, VTYPES))
    qo, qr_ = np.zeros_like(q), np.zeros_like(q)
    fo, fr = np.full(n0(DIMS) * 7, 7.0), np.full(n0(DIMS) * 7, 7.0)
    o.or_q_face_b(idir, ip(bt), 6, dp(q), dp(qo), dp(mut), dp(fo), calmut, *DIMS)
    r.ref_q_face_b(idir, ip(bt), 6, dp(q), dp(qr_), dp(mut), dp(fr), calmut, *DIMS)
    same(qo, qr_, "qf")
    same(fo, fr, "mutf")


@pytest.mark.parametrize("idir", [1, 2, 3])
def test_q_der(libs, idir):
    o, r = libs
    setcfg(libs)
    rng = np.random.default_rng(40 + idir)
    q = prim7(rng, DIMS)
    bt = bctopo_of(all_faces_patches(*DIMS, VTYPES))
    ao = [np.zeros_like(q) for _ in range(3)]
    ar = [np.zeros_like(q) for _ in range(3)]
    o.or_q_der_b(idir, ip(bt), 6, dp(q), *[dp(a) for a in ao], *DIMS)
    r.ref_q_der_b(idir, ip(bt), 6, dp(q), *[dp(a) for a in ar], *DIMS)
    for m in range(3):
        same(ao[m], ar[m], f"q_der component {m}")
        assert np.count_nonzero(ao[m]) > 0


@pytest.mark.parametrize("idir", [1, 2, 3])
@pytest.mark.parametrize("kom,turcmp", [(True, 0), (True, 1), (False, 0)])
def test_rtst_vflux_resid(libs, idir, kom, turcmp):
    o, r = libs
    setcfg(libs, kom=kom, turcmp=turcmp)
    npde = 7 if kom else 5
    rng = np.random.default_rng(50 + idir)
    si, sj, sk, vol = box_block(*DIMS, seed=8)
    s = (si, sj, sk)[idir - 1]
    qf = prim7(rng, DIMS) if kom else rand_prim(rng, *DIMS)
    d = [rng.uniform(-0.05, 0.05, qf.size) for _ in range(3)]
    xd = mder(rng, DIMS)[3 * (idir - 1):3 * idir]
    mutf = np.zeros(n0(DIMS) * npde)
    mutf[:n0(DIMS)] = rng.uniform(0, 30, n0(DIMS))
    to, tr = np.zeros_like(qf), np.zeros_like(qf)
    if kom:
        o.or_rtst_b(idir, dp(qf), *[dp(x) for x in d], *[dp(x) for x in xd], dp(to), dp(mutf), *DIMS)
        r.ref_rtst_b(idir, dp(qf), *[dp(x) for x in d], *[dp(x) for x in xd], dp(tr), dp(mutf), *DIMS)
        same(to, tr, "rtst")
    fo, fr = mutf.copy(), mutf.copy()
    dist = np.zeros(n0(DIMS))
    o.or_vflux_b(idir, dp(qf), *[dp(x) for x in d], dp(to), dp(fo), dp(dist), *[dp(x) for x in xd], dp(s), *DIMS)
    r.ref_vflux_b(idir, dp(qf), *[dp(x) for x in d], dp(tr), dp(fr), dp(dist), *[dp(x) for x in xd], dp(s), *DIMS)
    same(fo, fr, "vflux")
    fv = fo.reshape(npde, -1)
    assert np.count_nonzero(fv[1:5]) > 0.5 * fv[1:5].size
    res0 = rng.uniform(-1, 1, qf.size)
    a, b = res0.copy(), res0.copy()
    o.or_bresid(idir, dp(fo), dp(a), *DIMS)
    r.ref_bresid(idir, dp(fr), dp(b), *DIMS)
    same(a, b, "resid")


@pytest.mark.parametrize("idir", [1, 2, 3])
def test_q_der_c(libs, idir):
    o, r = libs
    setcfg(libs)
    rng = np.random.default_rng(60 + idir)
    qf = prim7(rng, DIMS)
    a, b = np.zeros_like(qf), np.zeros_like(qf)
    o.or_q_der_c_b(idir, dp(qf), dp(a), *DIMS)
    r.ref_q_der_c_b(idir, dp(qf), dp(b), *DIMS)
    same(a, b, "q_der_c")


@pytest.mark.parametrize("turcmp", [0, 1])
@pytest.mark.parametrize("lim", [(0, 0), (1, 0), (1, 1)])
@pytest.mark.parametrize("nl", [1, 2])
def test_turprod_srckw(libs, turcmp, lim, nl):
    o, r = libs
    setcfg(libs, turcmp=turcmp, lim_prodtke=lim[0], lim_prodomega=lim[1], prdlim=0.5)
    rng = np.random.default_rng(70 + turcmp)
    q = prim7(rng, DIMS)
    d = [rng.uniform(-0.3, 0.3, q.size) for _ in range(3)]
    xd = mder(rng, DIMS)
    N3 = n3(*DIMS)
    po = [np.zeros(N3) for _ in range(3)]
    pr_ = [np.zeros(N3) for _ in range(3)]
    XD = (ctypes_ptr_array(xd))
    o.or_turprod_b(dp(q), *[dp(x) for x in d], *[dp(x) for x in po], XD, *DIMS)
    r.ref_turprod_b(dp(q), *[dp(x) for x in d], *[dp(x) for x in pr_], *[dp(x) for x in xd], *DIMS)
    for m in range(3):
        same(po[m], pr_[m], f"turprod output {m}")
    mut = rng.uniform(0, 200, N3)
    vol = rng.uniform(0.5, 1.5, n1(DIMS))
    res0 = rng.uniform(-1, 1, q.size)
    a, b = res0.copy(), res0.copy()
    o.or_srckw_b(dp(q), dp(mut), dp(po[0]), dp(po[1]), dp(a), dp(vol), nl, *DIMS)
    r.ref_srckw_b(dp(q), dp(mut), dp(pr_[0]), dp(pr_[1]), dp(b), dp(vol), nl, *DIMS)
    same(a, b, "srckw")
    qc = prim_to_cons(q, *DIMS, npde=7)
    p1, p2 = po[0].copy(), pr_[0].copy()
    o.or_lim_prod_mg_b(dp(qc), dp(p1), dp(mut), *DIMS)
    r.ref_lim_prod_mg_b(dp(qc), dp(p2), dp(mut), *DIMS)
    same(p1, p2, "lim_prod_mg")


def ctypes_ptr_array(arrs):
    import ctypes
    t = (ctypes.POINTER(ctypes.c_double) * len(arrs))()
    for m, a in enumerate(arrs):
        t[m] = dp(a)
    return t


def test_fixq_tur_vis(libs):
    o, r = libs
    setcfg(libs, qmin=(1e-8, 1.0, 1.0, 1.0, 1e-8, 3e-3, 20.0))
    rng = np.random.default_rng(80)
    q = prim_to_cons(prim7(rng, DIMS), *DIMS, npde=7)
    a, b = q.copy(), q.copy()
    o.or_fixq_b(dp(a), *DIMS)
    r.ref_fixq_b(dp(b), *DIMS)
    same(a, b, "fixq")
    assert not np.array_equal(a, q)
    mo, mr = np.zeros(n3(*DIMS)), np.zeros(n3(*DIMS))
    o.or_tur_vis_b(dp(a), dp(mo), *DIMS)
    r.ref_tur_vis_b(dp(b), dp(mr), *DIMS)
    same(mo, mr, "tur_vis")


@pytest.mark.parametrize("nh", [0, 1])
def test_update_kom(libs, nh):
    o, r = libs
    setcfg(libs, nharms=nh)
    ns = 2 * nh + 1
    rng = np.random.default_rng(90 + nh)
    q = prim_to_cons(prim7(rng, DIMS, ns), *DIMS, npde=7, nsnap=ns)
    qold = q * (1 + rng.uniform(-0.01, 0.01, q.size))
    res = rng.uniform(-1e-2, 1e-2, q.size)
    vol = rng.uniform(0.5, 1.5, n1(DIMS))
    dtv = rng.uniform(0.1, 1.0, n1(DIMS) * ns)
    dtvt = rng.uniform(0.1, 1.0, n1(DIMS) * ns)
    N3 = n3(*DIMS)
    delpls = rng.uniform(0, 0.5, N3 * ns)
    prd = rng.uniform(0, 4.0, N3 * ns)
    # ttrm of Wilcox k-omega (update_ttrm_b kom branch)
    ttrm = np.zeros(N3 * 3 * ns)
    z = np.zeros(q.size)
    zz = np.zeros(n0(DIMS) * 3 * ns)
    r.ref_update_ttrm_b(dp(q), dp(z), dp(z), dp(z), *[dp(zz) for _ in range(9)], dp(ttrm), dp(vol), *DIMS)
    a, b = q.copy(), q.copy()
    o.or_update_kb(dp(res), dp(a), dp(qold), dp(vol), dp(dtv), dp(dtvt), dp(delpls), dp(prd), ctypes_double(0.5),
                   *DIMS, None)
    r.ref_update(dp(res), dp(b), dp(qold), dp(vol), dp(dtv), dp(dtvt), dp(z), dp(z), dp(delpls), dp(prd), dp(ttrm),
                 ctypes_double(0.5), *DIMS)
    same(a, b, "update kom")


def ctypes_double(x):
    import ctypes
    return ctypes.c_double(x)


def test_lim_res_ome(libs):
    o, r = libs
    setcfg(libs)
    rng = np.random.default_rng(95)
    q = prim_to_cons(prim7(rng, DIMS), *DIMS, npde=7)
    res = rng.uniform(-50, 50, q.size)
    vol = rng.uniform(0.5, 1.5, n1(DIMS))
    dtvt = rng.uniform(0.1, 1.0, n1(DIMS))
    N3 = n3(*DIMS)
    delpls = rng.uniform(0, 0.5, N3)
    prd = rng.uniform(0, 400.0, N3)
    ttrm = np.zeros(N3 * 3)
    z = np.zeros(q.size)
    zz = np.zeros(n0(DIMS) * 3)
    r.ref_update_ttrm_b(dp(q), dp(z), dp(z), dp(z), *[dp(zz) for _ in range(9)], dp(ttrm), dp(vol), *DIMS)
    a, b = res.copy(), res.copy()
    o.or_lim_res_ome_b(dp(a), dp(q), dp(vol), dp(dtvt), dp(delpls), dp(prd), ctypes_double(0.25), *DIMS, None)
    r.ref_lim_res_ome_b(dp(b), dp(q), dp(vol), dp(dtvt), dp(delpls), dp(prd), dp(ttrm), ctypes_double(0.25), *DIMS)
    same(a, b, "lim_res_ome")
    assert not np.array_equal(a, res)


@pytest.mark.parametrize("irsop", [1, 2, 3])
def test_cirs_kom(libs, irsop):
    o, r = libs
    setcfg(libs, irsop=irsop, rat=1.8)
    rng = np.random.default_rng(100 + irsop)
    res = rng.uniform(-1, 1, n3(*DIMS) * 7)
    arrs = [rng.uniform(0.5, 2, n1(DIMS)) if m in (0, 4) else rng.uniform(0, 1.5, n1(DIMS)) for m in range(8)]
    a, b = res.copy(), res.copy()
    o.or_cirs_kb(dp(a), *[dp(x) for x in arrs], *DIMS)
    w = np.zeros(max(DIMS) * 7 * 4)
    r.ref_cirs_b(dp(b), *[dp(x) for x in arrs], dp(w), *DIMS)
    same(a, b, "cirs kom")


def test_rest4(libs):
    o, r = libs
    setcfg(libs)
    rng = np.random.default_rng(110)
    f, c = (9, 9, 5), (5, 5, 3)
    p1 = rng.uniform(0, 1, n3(*f))
    v1 = rng.uniform(0.5, 1.5, n1(f))
    v2 = rng.uniform(3.5, 9.5, n1(c))
    a, b = np.zeros(n3(*c)), np.zeros(n3(*c))
    o.or_rest4_b(*f, *c, dp(p1), dp(a), dp(v1), dp(v2))
    r.ref_rest4_b(*f, *c, dp(p1), dp(b), dp(v1), dp(v2))
    same(a, b, "rest4")


# ---------------------------------------------------------------------------
# Menter BSL / SST (src/src.f90:306-378, src/update.f90:2375-2524, 3036-3161,
# src/flux.f90:987-1061, 2949-3131)

def setcfg_sst(libs, model, **kw):
    o, r = libs
    base = kw_turbulence(0.4, 2.0e5, tu=1e-3, mutfar=0.1)
    base.update(kom=0, kom_bsl=int(model == "bsl"), kom_sst=int(model == "sst"))
    base.update(kw)
    c = default_config(stemp=0.383, **base)
    o.set_config(c)
    r.set_config(c)
    return c


def sst_inputs(seed):
    rng = np.random.default_rng(seed)
    q = prim7(rng, DIMS)
    d = [rng.uniform(-0.3, 0.3, q.size) for _ in range(3)]
    xd = mder(rng, DIMS)
    # wall distances from near-wall to far: both F1 / F2 branches are taken
    dist = 10.0 ** rng.uniform(-5, 0, n1(DIMS))
    return rng, q, d, xd, dist


@pytest.mark.parametrize("idir", [1, 2, 3])
def test_q_der_c_sst(libs, idir):
    o, r = libs
    setcfg_sst(libs, "sst")
    rng = np.random.default_rng(160 + idir)
    qf = prim7(rng, DIMS)
    a, b = np.zeros_like(qf), np.zeros_like(qf)
    o.or_q_der_c_b(idir, dp(qf), dp(a), *DIMS)
    r.ref_q_der_c_b(idir, dp(qf), dp(b), *DIMS)
    same(a, b, "q_der_c (k, omega too)")
    assert np.count_nonzero(a.reshape(7, -1)[5:]) > 0


@pytest.mark.parametrize("model", ["bsl", "sst"])
def test_update_ttrm_srckw_sst(libs, model):
    o, r = libs
    setcfg_sst(libs, model, lim_prodtke=1, lim_prodomega=1, prdlim=0.5)
    rng, q, d, xd, dist = sst_inputs(170)
    N3 = n3(*DIMS)
    to, tr = np.zeros(N3 * 3), np.zeros(N3 * 3)
    o.or_update_ttrm_b(dp(q), *[dp(x) for x in d], ctypes_ptr_array(xd), dp(to), dp(dist), *DIMS)
    r.ref_update_ttrm_b(dp(q), *[dp(x) for x in d], *[dp(x) for x in xd], dp(tr), dp(dist), *DIMS)
    same(to, tr, f"update_ttrm {model}")
    t = to.reshape(3, -1)
    inner = np.zeros((DIMS[2] + 3, DIMS[1] + 3, DIMS[0] + 3), bool)
    inner[2:DIMS[2] + 1, 2:DIMS[1] + 1, 2:DIMS[0] + 1] = True
    g = t[1][inner.ravel()]
    # F1 spans the range: blended gamma strictly between the inner and outer values somewhere
    assert g.max() > g.min()
    prd = rng.uniform(0, 4.0, N3)
    div = rng.uniform(-0.5, 0.5, N3)
    mut = rng.uniform(0, 200, N3)
    vol = rng.uniform(0.5, 1.5, n1(DIMS))
    res0 = rng.uniform(-1, 1, q.size)
    for nl in (1, 2):
        a, b = res0.copy(), res0.copy()
        o.or_srckw_sst_b(dp(q), dp(prd), dp(div), dp(a), dp(mut), dp(to), dp(vol), nl, *DIMS)
        r.ref_srckw_sst_b(dp(q), dp(prd), dp(div), dp(b), dp(mut), dp(tr), dp(vol), nl, *DIMS)
        same(a, b, f"srckw_sst nl={nl}")
        assert not np.array_equal(a, res0)


def test_tur_vis_sst(libs):
    o, r = libs
    setcfg_sst(libs, "sst")
    rng, q, d, xd, dist = sst_inputs(180)
    qc = prim_to_cons(q, *DIMS, npde=7)
    N3 = n3(*DIMS)
    mo, mr = np.zeros(N3), np.zeros(N3)
    o.or_tur_vis_sst_b(dp(qc), dp(mo), *[dp(x) for x in d], ctypes_ptr_array(xd), dp(dist), *DIMS)
    r.ref_tur_vis_sst_b(dp(qc), dp(mr), *[dp(x) for x in d], *[dp(x) for x in xd], dp(dist), *DIMS)
    same(mo, mr, "tur_vis_sst")
    assert np.count_nonzero(mo) > 0


@pytest.mark.parametrize("idir", [1, 2, 3])
@pytest.mark.parametrize("model", ["bsl", "sst"])
def test_vflux_sst(libs, idir, model):
    o, r = libs
    setcfg_sst(libs, model)
    rng = np.random.default_rng(190 + idir)
    si, sj, sk, vol = box_block(*DIMS, seed=8)
    s = (si, sj, sk)[idir - 1]
    qf = prim7(rng, DIMS)
    d = [rng.uniform(-0.05, 0.05, qf.size) for _ in range(3)]
    xd = mder(rng, DIMS)[3 * (idir - 1):3 * idir]
    mutf = np.zeros(n0(DIMS) * 7)
    mutf[:n0(DIMS)] = rng.uniform(0, 30, n0(DIMS))
    tau = np.zeros_like(qf)
    r.ref_rtst_b(idir, dp(qf), *[dp(x) for x in d], *[dp(x) for x in xd], dp(tau), dp(mutf), *DIMS)
    fdist = 10.0 ** rng.uniform(-5, 0, n0(DIMS))
    fo, fr = mutf.copy(), mutf.copy()
    o.or_vflux_b(idir, dp(qf), *[dp(x) for x in d], dp(tau), dp(fo), dp(fdist), *[dp(x) for x in xd], dp(s), *DIMS)
    r.ref_vflux_b(idir, dp(qf), *[dp(x) for x in d], dp(tau), dp(fr), dp(fdist), *[dp(x) for x in xd], dp(s),
                  *DIMS)
    same(fo, fr, f"vflux {model}")


def test_facedist_closed_form(libs):
    """facedist_b: with a wall distance linear in the index the face values
    are the exact midpoints, the boundary faces take the halo cell."""
    o, _ = libs
    setcfg_sst(libs, "sst")
    im, jm, km = DIMS
    i = np.arange(im + 1)[None, None, :]
    j = np.arange(jm + 1)[None, :, None]
    k = np.arange(km + 1)[:, None, None]
    dist = np.ascontiguousarray((0.25 * i + 0.5 * j + 1.0 * k + 0.125).astype(np.float64).reshape(-1))
    for idir, step in ((1, 0.25), (2, 0.5), (3, 1.0)):
        fd = np.zeros(im * jm * km)
        o.or_facedist_b(idir, dp(dist), dp(fd), *DIMS)
        f = fd.reshape(km, jm, im)
        ii, jj, kk = np.meshgrid(np.arange(1, im + 1), np.arange(1, jm + 1), np.arange(1, km + 1), indexing="ij")
        ii, jj, kk = ii.transpose(2, 1, 0), jj.transpose(2, 1, 0), kk.transpose(2, 1, 0)
        a = {1: ii, 2: jj, 3: kk}[idir]
        L = {1: im, 2: jm, 3: km}[idir]
        want = 0.25 * ii + 0.5 * jj + 1.0 * kk + 0.125 - step / 2
        want = np.where(a == 1, want - step / 2, want)          # dist of cell 0
        want = np.where(a == L, want + step / 2, want)          # dist of cell L
        m = (ii < im) | (idir == 1)
        m &= (jj < jm) | (idir == 2)
        m &= (kk < km) | (idir == 3)
        np.testing.assert_array_equal(f[m], want[m])


@pytest.mark.parametrize("ramping", [1, 2])
def test_cfl_ramping(libs, ramping):
    """bdeltat with CFL ramping (src/cosa.f90:4760-4820) over a cycle
    schedule crossing every ramp segment: dtvol / dtvolt of the oracle and
    the compiled reference agree cycle by cycle, and the ramp reaches the
    end values."""
    o, r = libs
    ncycr = (3, 8, 15, 24)
    vcfli, vcflit = (0.5, 2.0, 4.0), (0.3, 1.0, 2.5)
    setcfg(libs, kom=True, irsop=1, cfl=7.0, cflt=7.0, rat=1.7, ratt=1.3, cdff=1.0, ramping=ramping, ncycr=ncycr,
           vcfli=vcfli, vcflit=vcflit)
    rng = np.random.default_rng(44)
    si, sj, sk, vol = box_block(*DIMS, seed=6)
    q = prim7(rng, DIMS)
    N3 = n3(*DIMS)
    mut = rng.uniform(0.0, 50.0, N3)
    xd = np.zeros((DIMS[0] + 2) * (DIMS[1] + 2) * (DIMS[2] + 2))
    lam, lv = np.zeros(N3 * 7), np.zeros(N3 * 7)
    o.or_lambdab_v(dp(q), dp(si), dp(sj), dp(sk), dp(xd), dp(xd), dp(xd), dp(vol), dp(mut), dp(lam), dp(lv), *DIMS)
    o.or_get_cfl.restype = __import__("ctypes").c_double
    seen = []
    for mgit in (1, 2, 3, 4, 6, 8, 9, 12, 15, 16, 20, 24, 25, 99, 100, 101):
        o.or_set_mgit(mgit)
        r.ref_set_mgit(mgit)
        oo = [np.zeros(n1(DIMS)) for _ in range(8)]
        rr = [np.zeros(n1(DIMS)) for _ in range(8)]
        o.or_bdeltat_v(*[dp(x) for x in oo[:2]], dp(lam), dp(lv), *[dp(x) for x in oo[2:]], *DIMS)
        r.ref_bdeltat(dp(rr[0]), dp(rr[1]), dp(lam), dp(lv), *[dp(x) for x in rr[2:]], *DIMS)
        same(oo[0], rr[0], f"dtvol cycle {mgit}")
        same(oo[1], rr[1], f"dtvolt cycle {mgit}")
        seen.append((mgit, o.or_get_cfl(0), o.or_get_cfl(1)))
    cfl = dict((m, c) for m, c, _ in seen)
    assert cfl[1] == vcfli[0] and cfl[25] == vcfli[2] or ramping == 1
    if ramping == 2:
        assert cfl[3] == vcfli[0] and cfl[9] == vcfli[1] and cfl[25] == vcfli[2]
        assert vcfli[0] < cfl[6] < vcfli[1] and vcfli[1] < cfl[20] < vcfli[2]
    else:   # ramping1 refreshes at cycle 1 and every 100th only
        assert cfl[1] == vcfli[0] and cfl[99] == vcfli[0] and cfl[100] == vcfli[2] and cfl[101] == vcfli[2]


@pytest.mark.parametrize("kom", [False, True])
@pytest.mark.parametrize("irsop", [1, 2, 3])
def test_ccirs(libs, kom, irsop):
    """ccirs_b (cutcirs = 1, src/cosa.f90:2067-2415) and scalef/scaleb: the
    xi sweep from ores, every line with the halo residual as Dirichlet data
    at both ends."""
    o, r = libs
    if kom:
        setcfg(libs, irsop=irsop, cutcirs=1, rat=1.8, ratt=1.4)
    else:
        setcfg(libs, kom=False, irsop=irsop, cutcirs=1, rat=1.8)
    npde = 7 if kom else 5
    rng = np.random.default_rng(300 + irsop + 10 * kom)
    N3 = n3(*DIMS)
    res = rng.uniform(-1, 1, N3 * npde)          # halos hold neighbour residuals
    ores = rng.uniform(-1, 1, N3 * npde)
    b = [rng.uniform(0.0, 2.0, n1(DIMS)) for _ in range(6)]
    dtv, dtvt = rng.uniform(0.1, 1.0, n1(DIMS)), rng.uniform(0.1, 1.0, n1(DIMS))
    a1, a2 = res.copy(), res.copy()
    o.or_ccirs_b(dp(a1), dp(ores), *[dp(x) for x in b], *DIMS)
    r.ref_ccirs_b(dp(a2), dp(ores), *[dp(x) for x in b], *DIMS)
    same(a1, a2, "ccirs")
    assert not np.array_equal(a1, res)
    for back in (0, 1):
        s1, s2 = res.copy(), res.copy()
        o.or_scale_b(dp(s1), dp(dtv), dp(dtvt), back, *DIMS)
        r.ref_scale_b(dp(s2), dp(dtv), dp(dtvt), back, *DIMS)
        same(s1, s2, f"scale back={back}")
