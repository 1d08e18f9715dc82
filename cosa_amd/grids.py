"""Analytic grids and geometric inputs of the hot path.

The reference reads node coordinates from mesh.dat and computes face
metrics and volumes itself (src/grid.f90); those are setup, outside the
pseudo-time loop, and reach the loop as inputs.  This module builds the
synthetic grids of the benchmark configurations and restates the metric
definition of ``imetrix_b`` (src/grid.f90:848-1253) so the inputs handed to
libcosahip have the reference's meaning:

    s(1:3) = unit normal of the face (diagonal cross product / 2, normalised)
    s(4)   = face area

Node arrays are indexed [k, j, i] (i fastest, as the reference stores them).
"""
from __future__ import annotations

import math
from typing import Tuple

import numpy as np


# ---------------------------------------------------------------------------
# grid generators

def naca0012_y(x: np.ndarray) -> np.ndarray:
    """NACA 0012 half thickness, closed trailing edge."""
    return 0.6 * (0.2969 * np.sqrt(x) - 0.1260 * x - 0.3516 * x ** 2 + 0.2843 * x ** 3 - 0.1036 * x ** 4)


def geometric_distribution(n: int, ds1: float, length: float) -> np.ndarray:
    """n points from 0 to length, first spacing ds1, constant growth ratio."""
    if n < 2:
        return np.zeros(1)
    m = n - 1
    if ds1 * m >= length:
        return np.linspace(0.0, length, n)
    lo, hi = 1.0 + 1e-12, 2.0
    while ds1 * (hi ** m - 1) / (hi - 1) < length:
        hi *= 2
    for _ in range(200):
        r = 0.5 * (lo + hi)
        if ds1 * (r ** m - 1) / (r - 1) < length:
            lo = r
        else:
            hi = r
    r = 0.5 * (lo + hi)
    s = np.concatenate([[0.0], np.cumsum(ds1 * r ** np.arange(m))])
    return s * (length / s[-1])


def naca0012_ogrid(ni: int, nj: int, nk: int, r_far: float = 20.0, ds_wall: float = 1e-3,
                   dz: float = 0.05) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """O-grid around a NACA 0012 section (chord 1).  i runs from the trailing
    edge along the lower surface to the leading edge and back along the upper
    surface (i = 1 and i = ni coincide: the seam is a self-cut); j runs from
    the wall to a circle of radius r_far about mid chord with geometric
    stretching from ds_wall; k is a uniform extrusion in z (quasi-2D)."""
    th = 2.0 * np.pi * np.arange(ni) / (ni - 1)
    xw = 0.5 * (1.0 + np.cos(th))
    yw = np.where(th <= np.pi, -naca0012_y(xw), naca0012_y(xw))
    yw[0] = 0.0
    yw[-1] = 0.0
    xo = 0.5 + r_far * np.cos(th)
    yo = -r_far * np.sin(th)
    yo[0] = 0.0
    yo[-1] = 0.0
    s = geometric_distribution(nj, ds_wall, r_far) / r_far
    x2 = xw[None, :] + s[:, None] * (xo - xw)[None, :]
    y2 = yw[None, :] + s[:, None] * (yo - yw)[None, :]
    z = np.arange(nk) * dz
    x = np.broadcast_to(x2[None], (nk, nj, ni)).copy()
    y = np.broadcast_to(y2[None], (nk, nj, ni)).copy()
    zz = np.broadcast_to(z[:, None, None], (nk, nj, ni)).copy()
    return x, y, zz


def bump_channel(ni: int, nj: int, nk: int, length: float = 3.0, height: float = 1.0, width: float = 1.0,
                 bump: float = 0.05, stretch: float = 1.5) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Cartesian-stretched channel with a sinusoidal bump of relative height
    `bump` on the lower wall between x = 1 and x = 2 (length 3)."""
    u = np.linspace(0.0, 1.0, ni)
    x1 = length * u
    v = np.linspace(0.0, 1.0, nj)
    eta = np.sinh(stretch * v) / np.sinh(stretch)          # clustering at j = 1
    w = np.linspace(0.0, 1.0, nk)
    zb = np.where((x1 > 1.0) & (x1 < 2.0), bump * height * np.sin(np.pi * (x1 - 1.0)) ** 2, 0.0)
    y2 = zb[None, :] + eta[:, None] * (height - zb)[None, :]
    x = np.broadcast_to(x1[None, None, :], (nk, nj, ni)).copy()
    y = np.broadcast_to(y2[None, :, :], (nk, nj, ni)).copy()
    z = np.broadcast_to((width * w)[:, None, None], (nk, nj, ni)).copy()
    return x, y, z


def distort(x, y, z, amp: float, seed: int = 0):
    """Smooth random distortion of interior nodes (tests: non-Cartesian cells)."""
    rng = np.random.default_rng(seed)
    nk, nj, ni = x.shape
    out = []
    for a in (x, y, z):
        d = rng.uniform(-amp, amp, a.shape)
        d[:, :, 0] = d[:, :, -1] = 0
        d[:, 0, :] = d[:, -1, :] = 0
        d[0, :, :] = d[-1, :, :] = 0
        out.append(a + d)
    return tuple(out)


def coarsen(x: np.ndarray, nl: int) -> np.ndarray:
    """Every 2**(nl-1)-th node (coarsen_b, src/cosa.f90:2609-2669)."""
    f = 2 ** (nl - 1)
    return np.ascontiguousarray(x[::f, ::f, ::f])


# ---------------------------------------------------------------------------
# metrics and volumes

def face_metrics(x: np.ndarray, y: np.ndarray, z: np.ndarray, lmet: int = 4, vel=None):
    """si, sj, sk of one block (imetrix_b, src/grid.f90:848-1253), shaped
    [k, j, i, lmet] over 0:kmax+1 etc. (the P2 family with lmet fastest).
    With lmet = 7 and vel = (xdot, ydot, zdot) at the nodes, components 5-7
    are the face-midpoint velocity, the mean of the face's four node
    velocities (src/grid.f90:1031-1064)."""
    nk, nj, ni = x.shape
    out = []
    V = np.stack(vel, axis=-1) if vel is not None else None
    for d in range(3):
        s = np.zeros((nk + 2, nj + 2, ni + 2, lmet))
        P = np.stack([x, y, z], axis=-1)  # [k, j, i, 3]
        if d == 0:   # i faces: i = 1..imax, j = 1..jmax-1, k = 1..kmax-1
            a = P[1:, 1:, :] - P[:-1, :-1, :]          # P(j+1,k+1) - P(j,k)
            b = P[1:, :-1, :] - P[:-1, 1:, :]          # P(j,k+1) - P(j+1,k)
            sl = (slice(1, nk), slice(1, nj), slice(1, ni + 1))
            sgn = 1.0
        elif d == 1:  # j faces: d1 = P(i+1,k+1)-P(i,k), d2 = P(i,k+1)-P(i+1,k)
            a = P[1:, :, 1:] - P[:-1, :, :-1]
            b = P[1:, :, :-1] - P[:-1, :, 1:]
            sl = (slice(1, nk), slice(1, nj + 1), slice(1, ni))
            sgn = -1.0
        else:         # k faces: d1 = P(i+1,j+1)-P(i,j), d2 = P(i,j+1)-P(i+1,j)
            a = P[:, 1:, 1:] - P[:, :-1, :-1]
            b = P[:, 1:, :-1] - P[:, :-1, 1:]
            sl = (slice(1, nk + 1), slice(1, nj), slice(1, ni))
            sgn = 1.0
        c0 = (a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1]) / 2
        c1 = (a[..., 2] * b[..., 0] - a[..., 0] * b[..., 2]) / 2
        c2 = (a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]) / 2
        if sgn < 0:
            c0, c1, c2 = -c0, -c1, -c2
        area = np.sqrt(c0 * c0 + c1 * c1 + c2 * c2)
        if np.any(area <= 0):
            raise ValueError("non-positive face area")
        s[sl + (0,)] = c0 / area
        s[sl + (1,)] = c1 / area
        s[sl + (2,)] = c2 / area
        s[sl + (3,)] = area
        if V is not None and lmet >= 7:
            if d == 0:
                fv = 0.25 * (V[:-1, :-1, :] + V[:-1, 1:, :] + V[1:, :-1, :] + V[1:, 1:, :])
            elif d == 1:
                fv = 0.25 * (V[:-1, :, :-1] + V[:-1, :, 1:] + V[1:, :, :-1] + V[1:, :, 1:])
            else:
                fv = 0.25 * (V[:, :-1, :-1] + V[:, :-1, 1:] + V[:, 1:, :-1] + V[:, 1:, 1:])
            s[sl + (slice(4, 7),)] = fv
        out.append(s)
    return out


def pitch(x0: np.ndarray, y0: np.ndarray, z0: np.ndarray, theta: float, theta_dot: float,
          xh: float = 0.25, yh: float = 0.0):
    """Rigid pitching about the hinge (xh, yh) in the x-y plane: node
    positions rotated by theta and node velocities for a pitch rate
    theta_dot (the rigid-body motion of pitch, src/grid.f90:1872-1960)."""
    c, s = np.cos(theta), np.sin(theta)
    dx, dy = x0 - xh, y0 - yh
    x = xh + dx * c - dy * s
    y = yh + dx * s + dy * c
    xd = -theta_dot * (dx * s + dy * c)
    yd = theta_dot * (dx * c - dy * s)
    return x, y, z0.copy(), xd, yd, np.zeros_like(z0)


def cell_volumes(x: np.ndarray, y: np.ndarray, z: np.ndarray) -> np.ndarray:
    """Cell volumes [k, j, i] over 0:kmax, 0:jmax, 0:imax (the P1 family),
    from V = 1/3 sum_f c_f . S_f with outward face vectors and face centres
    at the average of the four face nodes.  Halo values copy the adjacent
    interior cell (they are read only by the linear prolongation)."""
    nk, nj, ni = x.shape
    P = np.stack([x, y, z], axis=-1)
    si, sj, sk = face_metrics(x, y, z)

    def fc(d):
        if d == 0:
            return 0.25 * (P[:-1, :-1, :] + P[1:, :-1, :] + P[:-1, 1:, :] + P[1:, 1:, :])
        if d == 1:
            return 0.25 * (P[:-1, :, :-1] + P[1:, :, :-1] + P[:-1, :, 1:] + P[1:, :, 1:])
        return 0.25 * (P[:, :-1, :-1] + P[:, 1:, :-1] + P[:, :-1, 1:] + P[:, 1:, 1:])

    ci, cj, ck = fc(0), fc(1), fc(2)
    Si = si[1:nk, 1:nj, 1:ni + 1, :3] * si[1:nk, 1:nj, 1:ni + 1, 3:4]
    Sj = sj[1:nk, 1:nj + 1, 1:ni, :3] * sj[1:nk, 1:nj + 1, 1:ni, 3:4]
    Sk = sk[1:nk + 1, 1:nj, 1:ni, :3] * sk[1:nk + 1, 1:nj, 1:ni, 3:4]
    fi = np.sum(ci * Si, axis=-1)
    fj = np.sum(cj * Sj, axis=-1)
    fk = np.sum(ck * Sk, axis=-1)
    v = (fi[:, :, 1:] - fi[:, :, :-1] + fj[:, 1:, :] - fj[:, :-1, :] + fk[1:, :, :] - fk[:-1, :, :]) / 3.0
    if np.any(v <= 0):
        raise ValueError("non-positive cell volume")
    vol = np.zeros((nk + 1, nj + 1, ni + 1))
    vol[1:nk, 1:nj, 1:ni] = v
    vol[:, :, 0] = vol[:, :, 1]
    vol[:, :, ni] = vol[:, :, ni - 1]
    vol[:, 0, :] = vol[:, 1, :]
    vol[:, nj, :] = vol[:, nj - 1, :]
    vol[0, :, :] = vol[1, :, :]
    vol[nk, :, :] = vol[nk - 1, :, :]
    return vol


def metric_derivatives(x: np.ndarray, y: np.ndarray, z: np.ndarray):
    """Metric derivatives (xi, eta, zeta)_(x, y, z) on the i, j and k faces
    of one block: nine arrays [k, j, i, 3] over 1:kmax, 1:jmax, 1:imax (the
    (3, imax, jmax, kmax) layout of xideri ... zetaderk, src/grid.f90:
    3150-3583), zero outside each face family's range.  Computed as the
    inverse of the face Jacobian [x_xi x_eta x_zeta]: the derivative normal
    to the face from node central differences (one-sided at the block
    ends) averaged over the face's four nodes, the tangential ones from the
    face edges.  These are inputs the caller uploads; the hot path only
    reads them."""
    nk, nj, ni = x.shape
    P = np.stack([x, y, z], axis=-1)                     # [k, j, i, 3]

    def dnode(ax):
        d = np.empty_like(P)
        sl = [slice(None)] * 3
        lo, hi, mid = list(sl), list(sl), list(sl)
        n = P.shape[ax]
        a = [slice(None)] * 4
        b = [slice(None)] * 4
        c = [slice(None)] * 4
        a[ax], b[ax], c[ax] = slice(2, n), slice(0, n - 2), slice(1, n - 1)
        d[tuple(c)] = (P[tuple(a)] - P[tuple(b)]) / 2
        a[ax], b[ax], c[ax] = 1, 0, 0
        d[tuple(c)] = P[tuple(a)] - P[tuple(b)]
        a[ax], b[ax], c[ax] = n - 1, n - 2, n - 1
        d[tuple(c)] = P[tuple(a)] - P[tuple(b)]
        return d

    dxi, deta, dzeta = dnode(2), dnode(1), dnode(0)     # node derivatives along i, j, k
    out = []
    for fam in range(3):
        J = np.zeros((nk, nj, ni, 3, 3))
        if fam == 0:      # i faces: nodes (i, j..j+1, k..k+1)
            sl = (slice(0, nk - 1), slice(0, nj - 1), slice(0, ni))
            avg = lambda A: 0.25 * (A[:-1, :-1, :] + A[:-1, 1:, :] + A[1:, :-1, :] + A[1:, 1:, :])
            e_eta = 0.5 * ((P[:-1, 1:, :] - P[:-1, :-1, :]) + (P[1:, 1:, :] - P[1:, :-1, :]))
            e_zeta = 0.5 * ((P[1:, :-1, :] - P[:-1, :-1, :]) + (P[1:, 1:, :] - P[:-1, 1:, :]))
            cols = (avg(dxi), e_eta, e_zeta)
        elif fam == 1:    # j faces: nodes (i..i+1, j, k..k+1)
            sl = (slice(0, nk - 1), slice(0, nj), slice(0, ni - 1))
            avg = lambda A: 0.25 * (A[:-1, :, :-1] + A[:-1, :, 1:] + A[1:, :, :-1] + A[1:, :, 1:])
            e_xi = 0.5 * ((P[:-1, :, 1:] - P[:-1, :, :-1]) + (P[1:, :, 1:] - P[1:, :, :-1]))
            e_zeta = 0.5 * ((P[1:, :, :-1] - P[:-1, :, :-1]) + (P[1:, :, 1:] - P[:-1, :, 1:]))
            cols = (e_xi, avg(deta), e_zeta)
        else:             # k faces: nodes (i..i+1, j..j+1, k)
            sl = (slice(0, nk), slice(0, nj - 1), slice(0, ni - 1))
            avg = lambda A: 0.25 * (A[:, :-1, :-1] + A[:, :-1, 1:] + A[:, 1:, :-1] + A[:, 1:, 1:])
            e_xi = 0.5 * ((P[:, :-1, 1:] - P[:, :-1, :-1]) + (P[:, 1:, 1:] - P[:, 1:, :-1]))
            e_eta = 0.5 * ((P[:, 1:, :-1] - P[:, :-1, :-1]) + (P[:, 1:, 1:] - P[:, :-1, 1:]))
            cols = (e_xi, e_eta, avg(dzeta))
        Jf = np.stack(cols, axis=-1)                      # [..., 3 (x,y,z), 3 (xi,eta,zeta)]
        inv = np.linalg.inv(Jf)                           # rows: xi_(x,y,z), eta_(...), zeta_(...)
        for r in range(3):
            a = np.zeros((nk, nj, ni, 3))
            a[sl] = inv[..., r, :]
            out.append(a)
    # order: xideri, etaderi, zetaderi, xiderj, etaderj, zetaderj, xiderk, etaderk, zetaderk
    return out


def cell_centres(x: np.ndarray, y: np.ndarray, z: np.ndarray) -> np.ndarray:
    """Cell centres [k, j, i, 3] over 0:kmax, 0:jmax, 0:imax (P1 family);
    halo cells take the centre of the adjacent interior cell."""
    P = np.stack([x, y, z], axis=-1)
    c = 0.125 * (P[:-1, :-1, :-1] + P[:-1, :-1, 1:] + P[:-1, 1:, :-1] + P[:-1, 1:, 1:] +
                 P[1:, :-1, :-1] + P[1:, :-1, 1:] + P[1:, 1:, :-1] + P[1:, 1:, 1:])
    nk, nj, ni = x.shape
    out = np.zeros((nk + 1, nj + 1, ni + 1, 3))
    out[1:nk, 1:nj, 1:ni] = c
    out[:, :, 0] = out[:, :, 1]
    out[:, :, ni] = out[:, :, ni - 1]
    out[:, 0, :] = out[:, 1, :]
    out[:, nj, :] = out[:, nj - 1, :]
    out[0, :, :] = out[1, :, :]
    out[nk, :, :] = out[nk - 1, :, :]
    return out
