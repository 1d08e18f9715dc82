"""Device grid motion and metric set-up (cosahip_grid_metrics / _coarsen /
_pitch, k_grid.hip) against the oracle's restatement of the reference
routines (oracle/grid_oracle.c, itself bit-identical to COSA's compiled
imetrix_b, bvols, volex_b, cut_v, vol_bedges, cordex_b, cut_x, vmetrix_b,
coarsen_b and pitch): from the same level-1 node coordinates, every level's
node coordinates and velocities, face metrics, volumes and metric
derivatives must come out bit-identical.  Cases: a distorted two-block grid
with a cut (Euler, two levels, linear prolongation: vol_bedges), the NACA
O-grid with its seam self-cut (three levels), laminar and k-omega two-block
and O-grid cases (coordex, cut_x, vmetrix), a pitching harmonic-balance
O-grid (pitch, face velocities of every snapshot) and the first blocks of
the uneven channel set."""
import ctypes

import numpy as np
import pytest

from cosa_amd import case as cases
from cosa_amd.config import kw_turbulence
from oracle_bind import Oracle, dp, ip
from parity_util import TURB

pytestmark = pytest.mark.gpu

D = ctypes.c_double
FIELDS = ("x", "y", "z", "xdot", "ydot", "zdot", "si", "sj", "sk", "vol")


def _cases():
    out = [
        cases.two_block_case(ni=13, nj=9, nk=9, nlevel=2, distort=0.15, name="two_block_lin", irsop=1,
                             cfl_irs=3.0, cfl=2.0, bln_prol=0),
        cases.naca_case(33, 13, 9, nlevel=3, name="naca_3lev", irsop=1, cfl_irs=3.0, cfl=2.0),
        cases.two_block_case(ni=13, nj=9, nk=9, nlevel=2, distort=0.1, name="laminar_blocks_2lev", wall=1500,
                             irsop=1, cfl_irs=2.5, cfl=1.5, viscous=1, reyno=500.0, stemp=0.383),
        cases.naca_case(33, 13, 9, nlevel=2, name="rans_naca", wall=1500, irsop=1, cfl_irs=2.0, cfl=1.0,
                        ds_wall=2e-3, wallwilc=0, **kw_turbulence(0.5, 1e5, tu=1e-3, mutfar=0.1, **TURB)),
        cases.naca_case(33, 13, 5, nlevel=2, name="naca_pitch_hb2", irsop=1, cfl_irs=3.0, cfl=2.0, nharms=2,
                        harbal=1, omega=0.6, pitch=dict(dtheta0=0.05, xh=0.25, yh=0.0)),
        cases.channel_case(cases.uneven_dims(96, seed=0)[:5], nlevel=3, name="uneven_first5", machfs=0.7,
                           irsop=1, cfl=2.0, cfl_irs=4.0),
    ]
    return out


CASES = _cases()


class OracleGeometry:
    """The driver's set-up sequence (src/cosa.f90:160-166, 297-322) on the
    oracle, block by block, on copies of the device's starting arrays."""

    def __init__(self, c, start):
        self.c, self.a = c, start
        self.o = Oracle()
        self.o.set_config(c.config, c.hb_matrix())

    def _views(self, nl, b):
        c, L = self.c, self.c.levels[nl - 1]
        nsm, lmet = c.nsnap_met, c.config["lmet"]
        o2, p2 = L.offsets(2), L.points(2)
        o1, p1 = L.offsets(1), L.points(1)
        o0, p0 = L.offsets(0), L.points(0)
        A = self.a[nl - 1]
        v = {}
        for f in ("x", "y", "z", "xdot", "ydot", "zdot"):
            v[f] = A[f][o2[b] * nsm:(o2[b] + p2[b]) * nsm]
        for f in ("si", "sj", "sk"):
            v[f] = A[f][o2[b] * lmet * nsm:(o2[b] + p2[b]) * lmet * nsm]
        v["vol"] = A["vol"][o1[b]:o1[b] + p1[b]]
        if "xd" in A:
            v["xd"] = [a[o0[b] * 3 * nsm:(o0[b] + p0[b]) * 3 * nsm] for a in A["xd"]]
        return v, (int(L.imax[b]), int(L.jmax[b]), int(L.kmax[b]))

    def run(self):
        c, o = self.c, self.o
        cfg = c.config
        nsm, lmet = c.nsnap_met, cfg["lmet"]
        for nl in range(1, cfg["nlevel"] + 1):
            L = c.levels[nl - 1]
            if nl > 1:
                for b in range(L.nblocks):
                    vf, df = self._views(nl - 1, b)
                    vc, dc = self._views(nl, b)
                    o.or_coarsen_b(dp(vf["x"]), dp(vf["y"]), dp(vf["z"]), *df, dp(vc["x"]), dp(vc["y"]),
                                   dp(vc["z"]), *dc)
            if cfg["moving"]:
                m = c.motion
                pose = c.pitch_poses().reshape(-1, 3)
                for b in range(L.nblocks):
                    v, d = self._views(nl, b)
                    n2 = (d[0] + 2) * (d[1] + 2) * (d[2] + 2)
                    base = [v[f][:n2].copy() for f in ("x", "y", "z")]
                    ones = np.ones((d[0] + 1) * (d[1] + 1) * (d[2] + 1))
                    for n in range(nsm):
                        sl = slice(n * n2, (n + 1) * n2)
                        outs = [v[f][sl] for f in ("x", "y", "z", "xdot", "ydot", "zdot")]
                        o.or_pitch(D(pose[n, 0]), D(pose[n, 1]), D(pose[n, 2]), D(m["dtheta0"]), D(cfg["omega"]),
                                   D(m.get("xh", 0.25)), D(m.get("yh", 0.0)), *map(dp, base), dp(ones), dp(ones),
                                   *map(dp, outs), *d)
            for b in range(L.nblocks):
                v, d = self._views(nl, b)
                o.or_bvols(dp(v["x"]), dp(v["y"]), dp(v["z"]), dp(v["vol"]), *d)
                rc = o.or_imetrix_b(*[dp(v[f]) for f in ("x", "y", "z", "xdot", "ydot", "zdot", "si", "sj", "sk")],
                                    *d, lmet, nsm, int(cfg["moving"]))
                assert rc == 0
            cuts = np.ascontiguousarray(L.cutopo, np.int32).reshape(-1, 21) if L.ncuts else np.zeros((0, 21),
                                                                                                 np.int32)
            if cfg["viscous"] or cfg["nlevel"] > 1:
                for b in range(L.nblocks):
                    v, d = self._views(nl, b)
                    o.or_volex_b(dp(v["vol"]), *d)
                for rec in cuts:
                    v1, d1 = self._views(nl, int(rec[0]) - 1)
                    v2, d2 = self._views(nl, int(rec[9]) - 1)
                    o.or_cut_layer(0, 1, *d1, dp(v1["vol"]), *d2, dp(v2["vol"]), ip(np.ascontiguousarray(rec)))
            if nl < cfg["nlevel"] and not cfg["bln_prol"]:
                for b in range(L.nblocks):
                    v, d = self._views(nl, b)
                    o.or_vol_bedges(dp(v["vol"]), *d)
            if cfg["viscous"]:
                for b in range(L.nblocks):
                    v, d = self._views(nl, b)
                    o.or_cordex_b(dp(v["x"]), dp(v["y"]), dp(v["z"]), *d, nsm)
                for rec in cuts:
                    v1, d1 = self._views(nl, int(rec[0]) - 1)
                    v2, d2 = self._views(nl, int(rec[9]) - 1)
                    for f in ("x", "y", "z"):
                        o.or_cut_layer(1, nsm, *d1, dp(v1[f]), *d2, dp(v2[f]), ip(np.ascontiguousarray(rec)))
                for b in range(L.nblocks):
                    v, d = self._views(nl, b)
                    arr = (ctypes.POINTER(ctypes.c_double) * 9)(*[dp(a) for a in v["xd"]])
                    o.or_vmetrix_b(dp(v["x"]), dp(v["y"]), dp(v["z"]), arr, dp(v["vol"]), *d, nsm)


def _download_all(g, c):
    out = []
    for nl in range(1, c.config["nlevel"] + 1):
        A = {}
        for f in FIELDS:
            A[f] = g.download(f, nl) if g.size(f, nl) > 0 else np.zeros(c.size(nl, "x2"))
        if c.config["viscous"]:
            from cosa_amd import _lib
            A["xd"] = [g.download(n, nl) for n in _lib.XD_NAMES]
        out.append(A)
    return out


@pytest.mark.parametrize("c", CASES, ids=[c.name for c in CASES])
def test_device_geometry_matches_oracle(c):
    from cosa_amd.solver import Solve
    print(f"grid set-up {c.name}: {c.levels[0].nblocks} blocks, {c.config['nlevel']} levels, "
          f"viscous={c.config['viscous']} moving={c.config['moving']} nsnap_met={c.nsnap_met}", flush=True)
    g = Solve(c, device=0, q=c.freestream())
    try:
        start = _download_all(g, c)
        og = OracleGeometry(c, [{k: (v.copy() if k != "xd" else [a.copy() for a in v]) for k, v in A.items()}
                                for A in start])
        og.run()
        g.device_geometry()
        dev = _download_all(g, c)
    finally:
        g.close()
    for nl in range(1, c.config["nlevel"] + 1):
        for f in FIELDS:
            a, b = dev[nl - 1][f], og.a[nl - 1][f]
            assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), (c.name, nl, f, np.abs(a - b).max())
        if c.config["viscous"]:
            for m in range(9):
                a, b = dev[nl - 1]["xd"][m], og.a[nl - 1]["xd"][m]
                assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), (c.name, nl, m)
        changed = [f for f in ("si", "vol") if not np.array_equal(dev[nl - 1][f], start[nl - 1][f])]
        print(f"  level {nl}: every field bit-identical to the oracle; changed from the host geometry: {changed}",
              flush=True)


def test_grid_metrics_rejects_a_folded_face():
    """A collapsed face (zero area) is refused as imetrix_b stops on it."""
    from cosa_amd._lib import CosahipError
    from cosa_amd.solver import Solve
    c = cases.two_block_case(ni=9, nj=5, nk=5, nlevel=1, name="folded", irsop=1, cfl_irs=3.0, cfl=2.0)
    g = Solve(c, device=0, q=c.freestream())
    try:
        x = g.download("x", 1)
        y = g.download("y", 1)
        z = g.download("z", 1)
        L = c.levels[0]
        im, jm = int(L.imax[0]), int(L.jmax[0])
        # collapse the first block's node column (i = 3, j = 2) onto (i = 3, j = 1) for every k
        sh = (int(L.kmax[0]) + 2, jm + 2, im + 2)
        for a in (x, y, z):
            blk = a[:sh[0] * sh[1] * sh[2]].reshape(sh)
            blk[:, 2, 3] = blk[:, 1, 3]
            blk[:, 2, 4] = blk[:, 1, 4]
        with pytest.raises(CosahipError, match="face area"):
            g.device_geometry(x, y, z)
    finally:
        g.close()
