"""Restart files in the reference's byte layout: writerest / write_brest
(src/output.f90:3481-3566) and readrest / read_brest (src/input.f90:
4620-4714), serial form.

A file holds, for every block of the level in order, one Fortran
unformatted sequential record of q(-1:imax+1,-1:jmax+1,-1:kmax+1,npde,0:2N)
(conservative, binary64) and, for k-omega, one record of
mut(-1:imax+1,-1:jmax+1,-1:kmax+1,0:2N).  A record is its byte count as a
4-byte integer, the data, and the byte count again (the record markers of
flang and gfortran); records of 2 GiB or more would need the compilers'
subrecord continuation and are refused.  Byte order is the host's (little
endian here), as the reference's unformatted writes are.

Deviation: writerest passes mut at the offset of q (`mut(iq)`,
src/output.f90:3520), so the reference's files hold the wrong eddy
viscosity for every block after the first; this writer stores each block's
own mut, which readrest reads back correctly.
"""
from __future__ import annotations

import os
from typing import Optional, Tuple

import numpy as np

_MAX_RECORD = 2 ** 31 - 1


def _blocks(case, nl: int):
    L = case.levels[nl - 1]
    o3, p3 = L.offsets(3), L.points(3)
    return [(int(o3[b]), int(p3[b])) for b in range(L.nblocks)]


def _write_record(f, a: np.ndarray):
    data = np.ascontiguousarray(a, dtype="<f8").tobytes()
    if len(data) > _MAX_RECORD:
        raise ValueError("restart record of 2 GiB or more (needs subrecords)")
    marker = np.int32(len(data)).tobytes()
    f.write(marker)
    f.write(data)
    f.write(marker)


def _read_record(f, n: int) -> np.ndarray:
    head = f.read(4)
    if len(head) != 4:
        raise ValueError("restart file ends before a record")
    nbytes = int(np.frombuffer(head, "<i4")[0])
    if nbytes != 8 * n:
        raise ValueError(f"restart record of {nbytes} bytes, expected {8 * n}")
    data = np.frombuffer(f.read(nbytes), "<f8")
    tail = f.read(4)
    if len(data) != n or len(tail) != 4 or int(np.frombuffer(tail, "<i4")[0]) != nbytes:
        raise ValueError("restart record truncated or its markers disagree")
    return data.astype(np.float64)


def write_restart(path: str, case, q: np.ndarray, mut: Optional[np.ndarray] = None, nl: int = 1):
    """writerest: q (and mut for k-omega) of level nl, whole-level arrays in
    the library's layout."""
    ns, npde = case.nsnap, case.npde
    kom = bool(case.config["kom"] or case.config["kom_bsl"] or case.config["kom_sst"])
    if kom and mut is None:
        raise ValueError("k-omega restart needs mut")
    tmp = path + ".part"
    with open(tmp, "wb") as f:
        for o, n3 in _blocks(case, nl):
            _write_record(f, q[o * npde * ns: (o + n3) * npde * ns])
            if kom:
                _write_record(f, mut[o * ns: (o + n3) * ns])
    os.replace(tmp, path)


def read_restart(path: str, case, nl: int = 1) -> Tuple[np.ndarray, Optional[np.ndarray]]:
    """readrest: (q, mut) of level nl (mut None without k-omega)."""
    ns, npde = case.nsnap, case.npde
    kom = bool(case.config["kom"] or case.config["kom_bsl"] or case.config["kom_sst"])
    q = np.zeros(case.size(nl, "p3"))
    mut = np.zeros(case.size(nl, "p3s")) if kom else None
    with open(path, "rb") as f:
        for o, n3 in _blocks(case, nl):
            q[o * npde * ns: (o + n3) * npde * ns] = _read_record(f, n3 * npde * ns)
            if kom:
                mut[o * ns: (o + n3) * ns] = _read_record(f, n3 * ns)
        if f.read(1):
            raise ValueError("restart file longer than the level's blocks")
    return q, mut
