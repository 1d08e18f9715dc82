"""Multiblock layout of COSA, restated for the host side of libcosahip.

Mirrors the reference's setup bookkeeping so that arrays handed to the
library have exactly the reference's flat layout:

* level node counts  ``(n-1)/2**(nl-1)+1``            (src/input.f90:4393-4434)
* per-block offsets of the P3/P2/P1/0 families         (src/input.f90:5502-5582)
* BC records ``bcdata``/``bctopo`` from node indices   (src/input.f90:3831-3903,
  5155-5203)
* cut records ``cutdata``/``cutopo``                   (src/input.f90:3905-4044,
  5285-5344)

Blocks are numbered from 1 in every record, as in the reference.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Sequence, Tuple

import numpy as np

# cyc(idir, m) of src/input.f90:2411 (0-based rows, 1-based values)
CYC = ((1, 2, 3), (2, 3, 1), (3, 1, 2))


def level_nodes(n: int, nl: int) -> int:
    """Node count of a direction on level nl (src/input.f90:4426-4431)."""
    return (n - 1) // 2 ** (nl - 1) + 1


@dataclass
class Patch:
    """One BC patch as written in input.dat: node index ranges and type."""
    istrt: Tuple[int, int, int]
    iend: Tuple[int, int, int]
    bctype: int


@dataclass
class Cut:
    """One 1:1 cut as written in input.dat (two records)."""
    blk1: int
    istrt1: Tuple[int, int, int]
    iend1: Tuple[int, int, int]
    blk2: int
    istrt2: Tuple[int, int, int]
    iend2: Tuple[int, int, int]
    iord: Tuple[int, int, int] = (1, 2, 3)


@dataclass
class Block:
    imax: int
    jmax: int
    kmax: int
    patches: List[Patch] = field(default_factory=list)


def bcdata_of(p: Patch) -> List[int]:
    """bcdata(1:9) of one patch (src/input.f90:3844-3868)."""
    istrt = list(p.istrt)
    iend = list(p.iend)
    idir = 0
    for i in range(3):
        if istrt[i] == iend[i]:
            idir = i + 1
    if idir == 0:
        raise ValueError("BC patch has no constant index")
    for i in range(3):
        if i + 1 != idir:
            iend[i] -= 1
    return [p.bctype, idir, istrt[idir - 1], istrt[0], iend[0], istrt[1], iend[1], istrt[2], iend[2]]


def bctopo_level(bcd: Sequence[int], nl: int) -> List[int]:
    """bctopo(1:10) on level nl (src/input.f90:5166-5186); row 10 unused."""
    t = list(bcd[:2])
    if nl == 1:
        t += list(bcd[2:9])
    else:
        f = 2 ** (nl - 1)
        t.append(bcd[2] // f + 1)
        t += [(v - 1) // f + 1 for v in bcd[3:9]]
    return t + [0]


def _cut_face(istrt, iend):
    istrt = list(istrt)
    iend = list(iend)
    idir = 0
    for l in range(3):
        if istrt[l] == iend[l]:
            idir = l + 1
    if idir == 0:
        raise ValueError("cut face has no constant index")
    for l in range(3):
        if l + 1 != idir:
            if iend[l] > istrt[l]:
                iend[l] -= 1
            else:
                istrt[l] -= 1
    return idir, istrt, iend


def cutdata_of(c: Cut) -> List[List[int]]:
    """The two directional cutdata(1:21) records of one cut
    (src/input.f90:3909-4043)."""
    d1, s1, e1 = _cut_face(c.istrt1, c.iend1)
    d2, s2, e2 = _cut_face(c.istrt2, c.iend2)
    a = [0] * 21
    b = [0] * 21
    a[0:9] = [c.blk1, d1, s1[d1 - 1], s1[0], e1[0], s1[1], e1[1], s1[2], e1[2]]
    b[9:18] = [c.blk1, d1, s1[d1 - 1], s1[0], e1[0], s1[1], e1[1], s1[2], e1[2]]
    a[9:18] = [c.blk2, d2, s2[d2 - 1], s2[0], e2[0], s2[1], e2[1], s2[2], e2[2]]
    a[18:21] = list(c.iord)
    b[0:9] = [c.blk2, d2, s2[d2 - 1], s2[0], e2[0], s2[1], e2[1], s2[2], e2[2]]
    io = c.iord
    if io[0] == 2 and io[1] == 3:
        col = 2
    elif io[0] == 3 and io[1] == 1:
        col = 3
    else:
        col = 1
    b[18:21] = [CYC[io[m] - 1][col - 1] for m in range(3)]
    return [a, b]


def cutopo_level(cd: Sequence[int], nl: int) -> List[int]:
    """cutopo(1:21) on level nl (src/input.f90:5294-5340)."""
    t = list(cd)
    if nl > 1:
        f = 2 ** (nl - 1)
        t[2] = cd[2] // f + 1
        t[11] = cd[11] // f + 1
        for i in list(range(3, 9)) + list(range(12, 18)):
            t[i] = (cd[i] - 1) // f + 1
    return t


@dataclass
class Level:
    """Sizes and offsets of one multigrid level (setoffs)."""
    nl: int
    imax: np.ndarray
    jmax: np.ndarray
    kmax: np.ndarray
    nbcs: np.ndarray
    bctopo: np.ndarray
    cutopo: np.ndarray

    @property
    def nblocks(self) -> int:
        return len(self.imax)

    @property
    def ncuts(self) -> int:
        return len(self.cutopo) // 21

    def points(self, fam: int) -> np.ndarray:
        """Points per block of family fam: 3 (-1:n+1), 2 (0:n+1), 1 (0:n), 0 (1:n)."""
        e = fam if fam > 0 else 0
        return ((self.imax.astype(np.int64) + e) * (self.jmax + e) * (self.kmax + e)).astype(np.int64)

    def offsets(self, fam: int) -> np.ndarray:
        p = self.points(fam)
        return np.concatenate([[0], np.cumsum(p)[:-1]]).astype(np.int64)

    def total(self, fam: int) -> int:
        return int(self.points(fam).sum())

    def ncell(self) -> int:
        return int(((self.imax - 1).astype(np.int64) * (self.jmax - 1) * (self.kmax - 1)).sum())


def build_levels(blocks: Sequence[Block], cuts: Sequence[Cut], nlevel: int) -> List[Level]:
    bcd = [[bcdata_of(p) for p in b.patches] for b in blocks]
    cd = [r for c in cuts for r in cutdata_of(c)]
    levels = []
    for nl in range(1, nlevel + 1):
        im = np.array([level_nodes(b.imax, nl) for b in blocks], np.int32)
        jm = np.array([level_nodes(b.jmax, nl) for b in blocks], np.int32)
        km = np.array([level_nodes(b.kmax, nl) for b in blocks], np.int32)
        bt = [v for blk in bcd for p in blk for v in bctopo_level(p, nl)]
        ct = [v for r in cd for v in cutopo_level(r, nl)]
        levels.append(Level(nl, im, jm, km, np.array([len(b.patches) for b in blocks], np.int32),
                            np.array(bt, np.int32), np.array(ct, np.int32)))
    return levels


def check_coarsenable(blocks: Sequence[Block], nlevel: int) -> None:
    """Every level must keep whole cells: (n-1) divisible by 2**(nlevel-1)."""
    f = 2 ** (nlevel - 1)
    for ib, b in enumerate(blocks):
        for n in (b.imax, b.jmax, b.kmax):
            if (n - 1) % f:
                raise ValueError(f"block {ib + 1}: {n} nodes cannot be coarsened {nlevel - 1} times")


# ---- index helpers (0-based flat index into one block's P3 array) ----------

def p3_index(imax: int, jmax: int, kmax: int, i, j, k):
    return (np.asarray(i) + 1) + (imax + 3) * ((np.asarray(j) + 1) + (jmax + 3) * (np.asarray(k) + 1))


def p3_view(arr: np.ndarray, off: int, imax: int, jmax: int, kmax: int, ncomp: int, nsnap: int) -> np.ndarray:
    """View of one block's P3 array as [n, comp, k, j, i] with i,j,k
    shifted by +1 (so index 0 is the reference's -1)."""
    n3 = (imax + 3) * (jmax + 3) * (kmax + 3)
    return arr[off * ncomp * nsnap: off * ncomp * nsnap + n3 * ncomp * nsnap].reshape(
        nsnap, ncomp, kmax + 3, jmax + 3, imax + 3)
