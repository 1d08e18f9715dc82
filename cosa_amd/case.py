"""A solve case: configuration, multiblock topology and geometric inputs of
every multigrid level, in the reference's flat array layout.

This is the host-side bookkeeping the reference does in ``input``/``alloc``
/``setoffs`` and ``grid`` (src/input.f90, src/grid.f90) before the
pseudo-time loop starts.  It produces exactly the arrays that the Fortran
driver would upload through the C ABI.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import math
import numpy as np

from . import grids
from .config import default_config, hb_operator
from .layout import Block, Cut, Level, Patch, build_levels, check_coarsenable


@dataclass
class Case:
    config: Dict
    blocks: List[Block]
    cuts: List[Cut]
    xyz: List[Tuple[np.ndarray, np.ndarray, np.ndarray]]   # fine-level nodes per block, [k, j, i]
    levels: List[Level] = field(default_factory=list)
    geom: List[Dict[str, np.ndarray]] = field(default_factory=list)
    name: str = "case"
    # rigid pitching of the whole grid for harmonic-balance moving-grid runs:
    # {"dtheta0": amplitude [rad], "xh": hinge x, "yh": hinge y}; snapshot n
    # sits at t_n = 2 pi n / ((2N+1) omega) with theta = dtheta0 sin(omega t)
    motion: Optional[Dict] = None

    def __post_init__(self):
        nlevel = self.config["nlevel"]
        check_coarsenable(self.blocks, nlevel)
        self.levels = build_levels(self.blocks, self.cuts, nlevel)
        self.geom = [self._level_geometry(L) for L in self.levels]

    # -- sizes ---------------------------------------------------------------
    @property
    def npde(self) -> int:
        return self.config["npde"]

    @property
    def nsnap(self) -> int:
        return 2 * self.config["nharms"] + 1

    @property
    def nsnap_met(self) -> int:
        return 2 * self.config["nharms"] * self.config["hbmove"] + 1

    def size(self, nl: int, kind: str) -> int:
        """Doubles in a whole-level array of a given kind."""
        L = self.levels[nl - 1]
        lmet = self.config["lmet"]
        return {
            "p3": L.total(3) * self.npde * self.nsnap,
            "p3s": L.total(3) * self.nsnap,
            "met": L.total(2) * lmet * self.nsnap_met,
            "x2": L.total(2) * self.nsnap_met,
            "p1": L.total(1) * self.nsnap,
            "vol": L.total(1),
            "flux": L.total(0) * self.npde * self.nsnap,
            "d0": L.total(0) * 3 * self.nsnap_met,
        }[kind]

    def ncell(self, nl: int = 1) -> int:
        return self.levels[nl - 1].ncell()

    # -- geometry ------------------------------------------------------------
    def snapshot_pose(self, n: int) -> Tuple[float, float]:
        """Pitch angle and rate of harmonic-balance snapshot n."""
        m = self.motion
        om = self.config["omega"]
        t = 2.0 * math.pi * n / ((2 * self.config["nharms"] + 1) * om)
        return m["dtheta0"] * math.sin(om * t), m["dtheta0"] * om * math.cos(om * t)

    def pitch_poses(self) -> np.ndarray:
        """Per harmonic-balance snapshot n: dcos(dtheta(n)), dsin(dtheta(n)),
        dcos(omegas zeit(n)) with zeit(n) = n dt, dt = 2 pi / omegas / (2N+1)
        (src/input.f90:3187-3190) and dtheta(n) = dtheta0 dsin(omegas zeit(n))
        (src/grid.f90:4540): the constants pitch uses (src/grid.f90:1921-1958)."""
        m, om = self.motion, self.config["omega"]
        ns = 2 * self.config["nharms"] + 1
        dt = 2 * math.pi / om / ns
        out = []
        for n in range(self.nsnap_met):
            zeit = n * dt
            dth = m["dtheta0"] * math.sin(om * zeit)
            out += [math.cos(dth), math.sin(dth), math.cos(om * zeit)]
        return np.array(out)

    def moment_centres(self, xyzmom=(0.25, 0.0, 0.0)) -> np.ndarray:
        """(xmomr, ymomr, zmomr) of every snapshot: the moment reference
        point carried with the pitching body as force_b computes it
        (src/output.f90:4476-4486); the point itself for a fixed body."""
        out = []
        for n in range(self.nsnap):
            xm, ym, zm = xyzmom
            if self.config["moving"] and self.motion is not None:
                th = self.snapshot_pose(n)[0]
                xh, yh = self.motion.get("xh", 0.25), self.motion.get("yh", 0.0)
                d1 = (xm - xh) * (math.cos(th) - 1) - (ym - yh) * math.sin(th)
                d2 = (xm - xh) * math.sin(th) + (ym - yh) * (math.cos(th) - 1)
                out += [xm + d1, ym + d2, zm]
            else:
                out += [xm, ym, zm]
        return np.array(out)

    def _level_geometry(self, L: Level) -> Dict[str, np.ndarray]:
        lmet = self.config["lmet"]
        nsm = self.nsnap_met
        moving = bool(self.config["moving"])
        if moving and (self.motion is None or lmet != 7):
            raise ValueError("a moving case needs lmet = 7 and a motion")
        si = np.zeros(self.size(L.nl, "met"))
        sj = np.zeros_like(si)
        sk = np.zeros_like(si)
        vol = np.zeros(self.size(L.nl, "vol"))
        vel = [np.zeros(self.size(L.nl, "x2")) for _ in range(3)] if moving else None
        xyz2 = [np.zeros(self.size(L.nl, "x2")) for _ in range(3)]
        o2 = L.offsets(2)
        o1 = L.offsets(1)
        for b in range(L.nblocks):
            x, y, z = (grids.coarsen(a, L.nl) for a in self.xyz[b])
            assert x.shape == (L.kmax[b], L.jmax[b], L.imax[b])
            n2 = int(L.points(2)[b])
            for n in range(nsm):
                if moving:
                    th, thd = self.snapshot_pose(n)
                    xn, yn, zn, *v = grids.pitch(x, y, z, th, thd, self.motion.get("xh", 0.25),
                                                  self.motion.get("yh", 0.0))
                    mets = grids.face_metrics(xn, yn, zn, lmet, vel=v)
                    for dst, src in zip(vel, v):
                        p2 = np.zeros((x.shape[0] + 2, x.shape[1] + 2, x.shape[2] + 2))
                        p2[1:-1, 1:-1, 1:-1] = src        # node (i,j,k) at P2 index (i,j,k)
                        dst[o2[b] * nsm + n * n2: o2[b] * nsm + (n + 1) * n2] = p2.reshape(-1)
                    pose = (xn, yn, zn)
                else:
                    pose = (x, y, z)
                    if n == 0:
                        mets = grids.face_metrics(x, y, z, lmet)
                for dst, src in zip(xyz2, pose):
                    dst[o2[b] * nsm + n * n2: o2[b] * nsm + (n + 1) * n2] = node_coords_p2(src).reshape(-1)
                for arr, m in zip((si, sj, sk), mets):
                    blk = arr[o2[b] * lmet * nsm: (o2[b] + n2) * lmet * nsm]
                    blk[n * n2 * lmet:(n + 1) * n2 * lmet] = m.reshape(-1)   # [k][j][i][lmet]
            v = grids.cell_volumes(x, y, z)
            vol[o1[b]: o1[b] + L.points(1)[b]] = v.reshape(-1)
        out = {"si": si, "sj": sj, "sk": sk, "vol": vol, "x": xyz2[0], "y": xyz2[1], "z": xyz2[2]}
        if moving:
            out.update(xdot=vel[0], ydot=vel[1], zdot=vel[2])
        if self.config["viscous"]:
            out.update(self._viscous_geometry(L))
        return out

    def _viscous_geometry(self, L: Level) -> Dict[str, np.ndarray]:
        """Metric derivatives (nine (3, imax, jmax, kmax) arrays, off_0
        offsets) and the cell-centre distance to the nearest viscous wall
        (types 15xx) of every block of level L; the nearest wall face centre
        stands for the wall (dist2wall, src/grid.f90:2506-2808, is the
        caller's set-up step)."""
        from scipy.spatial import cKDTree
        nsm = self.nsnap_met
        xd = [np.zeros(self.size(L.nl, "d0")) for _ in range(9)]
        dist = np.zeros(self.size(L.nl, "vol"))
        o0, o1 = L.offsets(0), L.offsets(1)
        walls = []
        coords = [tuple(grids.coarsen(a, L.nl) for a in self.xyz[b]) for b in range(L.nblocks)]
        for b, blk in enumerate(self.blocks):
            x, y, z = coords[b]
            P = np.stack([x, y, z], axis=-1)
            f = 2 ** (L.nl - 1)
            for pt in blk.patches:
                if pt.bctype // 100 != 15:
                    continue
                lo = [(v - 1) // f for v in pt.istrt]
                hi = [(v - 1) // f for v in pt.iend]
                sub = P[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1]
                # face centres of the wall patch: average over the two in-plane directions
                axes = [a for a in range(3) if sub.shape[a] > 1]
                c = sub
                for a in axes:
                    sl1 = [slice(None)] * 4
                    sl2 = [slice(None)] * 4
                    sl1[a] = slice(0, -1)
                    sl2[a] = slice(1, None)
                    c = 0.5 * (c[tuple(sl1)] + c[tuple(sl2)])
                walls.append(c.reshape(-1, 3))
        tree = cKDTree(np.concatenate(walls)) if walls else None
        for b in range(L.nblocks):
            x, y, z = coords[b]
            md = grids.metric_derivatives(x, y, z)
            n0 = int(L.points(0)[b])
            for m in range(9):
                flat = md[m].reshape(-1)
                for n in range(nsm):
                    xd[m][(o0[b] * nsm + n * n0) * 3:(o0[b] * nsm + (n + 1) * n0) * 3] = flat
            cc = grids.cell_centres(x, y, z).reshape(-1, 3)
            d = tree.query(cc, workers=8)[0] if tree is not None else np.full(len(cc), 1.0e6)
            dist[o1[b]:o1[b] + L.points(1)[b]] = d
        return {"xd": xd, "dist": dist}

    def turbulent_viscosity(self, q: np.ndarray, nl: int = 1) -> np.ndarray:
        """tur_vis_b (src/update.f90:2297-2337) of a conservative state:
        mut = cmu rho (rho k) / (rho omega + epst) Re / M over 0:imax,
        0:jmax, 0:kmax of every block, zero elsewhere -- the driver's
        initial mut after init2fs (src/cosa.f90:229-232)."""
        c = self.config
        L = self.levels[nl - 1]
        mut = np.zeros(self.size(nl, "p3s"))
        o3 = L.offsets(3)
        for b in range(L.nblocks):
            im, jm, km = int(L.imax[b]), int(L.jmax[b]), int(L.kmax[b])
            n3 = (im + 3) * (jm + 3) * (km + 3)
            for n in range(self.nsnap):
                base = o3[b] * self.npde * self.nsnap + n * self.npde * n3
                qq = q[base:base + self.npde * n3].reshape(self.npde, km + 3, jm + 3, im + 3)
                m = mut[o3[b] * self.nsnap + n * n3: o3[b] * self.nsnap + (n + 1) * n3].reshape(km + 3, jm + 3, im + 3)
                sl = (slice(1, km + 2), slice(1, jm + 2), slice(1, im + 2))
                m[sl] = c["cmu"] * qq[0][sl] * qq[5][sl] / (qq[6][sl] + c["epst"]) * c["reyno"] / c["machfs"]
        return mut

    # -- states --------------------------------------------------------------
    def freestream(self, nl: int = 1) -> np.ndarray:
        """init2fs (src/input.f90:4741-4800), conservative, halos included."""
        c = self.config
        L = self.levels[nl - 1]
        q = np.empty(self.size(nl, "p3"))
        vals = [c["rhoinf"],
                c["rhoinf"] * c["machfs"] * math.cos(c["alpha"]) * math.cos(c["beta"]),
                c["rhoinf"] * c["machfs"] * math.sin(c["alpha"]) * math.cos(c["beta"]),
                c["rhoinf"] * c["machfs"] * math.sin(c["beta"]),
                1 / c["rhoinf"] / c["gamma"] / (c["gamma"] - 1) + 0.5 * c["machfs"] ** 2]
        if self.npde == 7:
            # init2fs_b (src/input.f90:4741-4842): rho k, rho omega on the
            # interior only (halos stay 0); the energy is the Euler value
            vals += [c["rhoinf"] * c["tkefar"], c["rhoinf"] * c["omefar"]]
        o3 = L.offsets(3)
        for b in range(L.nblocks):
            n3 = int(L.points(3)[b])
            base = o3[b] * self.npde * self.nsnap
            for n in range(self.nsnap):
                for v in range(self.npde):
                    s = base + (n * self.npde + v) * n3
                    if v < 5:
                        q[s:s + n3] = vals[v]
                    else:
                        im, jm, km = int(L.imax[b]), int(L.jmax[b]), int(L.kmax[b])
                        blk = np.zeros((km + 3, jm + 3, im + 3))
                        blk[2:km + 1, 2:jm + 1, 2:im + 1] = vals[v]
                        q[s:s + n3] = blk.reshape(-1)
        return q

    def perturbed(self, amp: float = 0.01, seed: int = 1, nl: int = 1) -> np.ndarray:
        """Freestream times (1 + uniform(-amp, amp)) per value (C4 recipe)."""
        q = self.freestream(nl)
        rng = np.random.default_rng(seed)
        return q * (1.0 + rng.uniform(-amp, amp, q.shape))

    def hb_matrix(self) -> Optional[np.ndarray]:
        if self.config["nharms"] == 0:
            return None
        return hb_operator(self.config["nharms"], self.config["omega"])


# ---------------------------------------------------------------------------
# configurations

def node_coords_p2(a: np.ndarray) -> np.ndarray:
    """One coordinate of a block's nodes (kmax, jmax, imax) in the P2 layout
    (0:imax+1, 0:jmax+1, 0:kmax+1) with the face halo nodes copied from the
    boundary nodes as cordex_b does (src/grid.f90:23-78); edge and corner
    halo nodes stay zero."""
    p = np.zeros((a.shape[0] + 2, a.shape[1] + 2, a.shape[2] + 2))
    p[1:-1, 1:-1, 1:-1] = a
    p[1:-1, 0, 1:-1] = p[1:-1, 1, 1:-1]
    p[1:-1, -1, 1:-1] = p[1:-1, -2, 1:-1]
    p[1:-1, 1:-1, 0] = p[1:-1, 1:-1, 1]
    p[1:-1, 1:-1, -1] = p[1:-1, 1:-1, -2]
    p[0, 1:-1, 1:-1] = p[1, 1:-1, 1:-1]
    p[-1, 1:-1, 1:-1] = p[-2, 1:-1, 1:-1]
    return p


def naca_case(ni=193, nj=49, nk=5, nlevel=1, name="naca0012", pitch=None, wall=14, r_far=20.0, ds_wall=1e-3,
              **cfg) -> Case:
    """NACA 0012 O-grid, single block: j=1 inviscid wall (14), j=jmax
    farfield (30), k faces symmetry (80), seam i=1 <-> i=imax as a self-cut.
    pitch = {"dtheta0", "xh", "yh"} makes it a harmonic-balance pitching
    aerofoil on a moving grid (moving, hbmove = 1, lmet = 7)."""
    cfg.setdefault("machfs", 0.5)
    cfg.setdefault("alpha_deg", 1.25)
    if pitch is not None:
        cfg.update(moving=1, hbmove=1, lmet=7)
    c = default_config(nlevel=nlevel, **cfg)
    x, y, z = grids.naca0012_ogrid(ni, nj, nk, r_far=r_far, ds_wall=ds_wall)
    b = Block(ni, nj, nk, [
        Patch((1, 1, 1), (ni, 1, nk), wall),
        Patch((1, nj, 1), (ni, nj, nk), 30),
        Patch((1, 1, 1), (ni, nj, 1), 80),
        Patch((1, 1, nk), (ni, nj, nk), 80),
    ])
    cut = Cut(1, (1, 1, 1), (1, nj, nk), 1, (ni, 1, 1), (ni, nj, nk), (1, 2, 3))
    return Case(c, [b], [cut], [(x, y, z)], name=name, motion=pitch)


def wing_case(ni=257, nj=65, nk=65, nblocks=4, nlevel=3, name="wing4", semispan=2.0, ds_wall=5e-6, r_far=20.0,
              **cfg) -> Case:
    """Rectangular wing with a NACA 0012 section (chord 1, semi-span
    `semispan`: aspect ratio 2*semispan), O-H topology: every block is the
    O-grid of the section (i around, j normal with geometric stretching from
    ds_wall to r_far) extruded over 1/nblocks of the semi-span; blocks are
    stacked spanwise and joined by 1:1 cuts, each block's O-grid seam is a
    self-cut.  Viscous wall (1500) on j = 1, farfield (30) on j = jmax,
    symmetry (80) at the root and farfield at the outer spanwise plane.
    4 blocks of 257 x 65 x 65 nodes = 4,194,304 cells."""
    cfg.setdefault("machfs", 0.3)
    cfg.setdefault("alpha_deg", 4.0)
    c = default_config(nlevel=nlevel, **cfg)
    x, y, z0 = grids.naca0012_ogrid(ni, nj, nk, r_far=r_far, ds_wall=ds_wall, dz=semispan / nblocks / (nk - 1))
    blocks, xyz, cuts = [], [], []
    for b in range(nblocks):
        patches = [Patch((1, 1, 1), (ni, 1, nk), 1500), Patch((1, nj, 1), (ni, nj, nk), 30)]
        if b == 0:
            patches.append(Patch((1, 1, 1), (ni, nj, 1), 80))
        if b == nblocks - 1:
            patches.append(Patch((1, 1, nk), (ni, nj, nk), 30))
        blocks.append(Block(ni, nj, nk, patches))
        xyz.append((x.copy(), y.copy(), z0 + b * semispan / nblocks))
        cuts.append(Cut(b + 1, (1, 1, 1), (1, nj, nk), b + 1, (ni, 1, 1), (ni, nj, nk), (1, 2, 3)))
        if b + 1 < nblocks:
            cuts.append(Cut(b + 1, (1, 1, nk), (ni, nj, nk), b + 2, (1, 1, 1), (ni, nj, 1), (1, 2, 3)))
    return Case(c, blocks, cuts, xyz, name=name)


def channel_block(ni, nj, nk) -> Block:
    return Block(ni, nj, nk, [
        Patch((1, 1, 1), (1, nj, nk), 30),
        Patch((ni, 1, 1), (ni, nj, nk), 30),
        Patch((1, 1, 1), (ni, 1, nk), 14),
        Patch((1, nj, 1), (ni, nj, nk), 14),
        Patch((1, 1, 1), (ni, nj, 1), 80),
        Patch((1, 1, nk), (ni, nj, nk), 80),
    ])


def uneven_dims(nblocks=96, seed=0, sides=(17, 33, 65, 129)):
    """Block node counts of the uneven multiblock configuration."""
    rng = np.random.default_rng(seed)
    return [tuple(int(v) for v in rng.choice(sides, 3)) for _ in range(nblocks)]


def channel_case(dims: Sequence[Tuple[int, int, int]], nlevel=3, name="bump_channels", **cfg) -> Case:
    """Independent bump channels (one block each): inflow/outflow farfield
    (30) on i faces, inviscid walls (14) on j faces, symmetry (80) on k."""
    cfg.setdefault("machfs", 0.7)
    c = default_config(nlevel=nlevel, **cfg)
    blocks, xyz = [], []
    for (ni, nj, nk) in dims:
        blocks.append(channel_block(ni, nj, nk))
        xyz.append(grids.bump_channel(ni, nj, nk))
    return Case(c, blocks, [], xyz, name=name)


def two_block_case(ni=17, nj=9, nk=9, nlevel=1, distort=0.0, name="two_blocks", wall=14, **cfg) -> Case:
    """Two blocks joined by a 1:1 cut across i (block 1 i=imax <-> block 2
    i=1), farfield elsewhere except a wall and symmetry planes; used by the
    parity tests to exercise cuts between different blocks.  wall = 30
    gives a wall-free box (freestream is then an exact steady state)."""
    cfg.setdefault("machfs", 0.6)
    cfg.setdefault("alpha_deg", 3.0)
    c = default_config(nlevel=nlevel, **cfg)
    x, y, z = grids.bump_channel(2 * ni - 1, nj, nk, length=3.0, height=1.0, width=0.6)
    if distort:
        # distort: amplitude as a fraction of the smallest node spacing
        h = min(np.min(np.abs(np.diff(a, axis=ax))[np.abs(np.diff(a, axis=ax)) > 0])
                for a, ax in ((x, 2), (y, 1), (z, 0)))
        x, y, z = grids.distort(x, y, z, distort * h, seed=3)
    a = tuple(arr[:, :, :ni].copy() for arr in (x, y, z))
    b = tuple(arr[:, :, ni - 1:].copy() for arr in (x, y, z))
    b1 = Block(ni, nj, nk, [
        Patch((1, 1, 1), (1, nj, nk), 30),
        Patch((1, 1, 1), (ni, 1, nk), wall),
        Patch((1, nj, 1), (ni, nj, nk), 30),
        Patch((1, 1, 1), (ni, nj, 1), 80),
        Patch((1, 1, nk), (ni, nj, nk), 80),
    ])
    b2 = Block(ni, nj, nk, [
        Patch((ni, 1, 1), (ni, nj, nk), 4),
        Patch((1, 1, 1), (ni, 1, nk), wall),
        Patch((1, nj, 1), (ni, nj, nk), 30),
        Patch((1, 1, 1), (ni, nj, 1), 80),
        Patch((1, 1, nk), (ni, nj, nk), 80),
    ])
    cut = Cut(1, (ni, 1, 1), (ni, nj, nk), 2, (1, 1, 1), (1, nj, nk), (1, 2, 3))
    return Case(c, [b1, b2], [cut], [a, b], name=name)
