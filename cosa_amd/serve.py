"""Continuous-batching serving of steady and harmonic-balance solves on one
GPU (the north star's serving form; the reference runs one solve per
process, src/cosa.f90:806-930).

Requests arrive over time, each a NACA 0012 O-grid solve at its own Mach
number, incidence, grid size and (optionally) harmonic-balance pitching.
The server keeps every admitted solve resident in HBM (its own
cosahip_solve and HIP stream), admits waiting requests while the estimated
device footprint of the resident solves stays within a budget, and
keeps one V-cycle of every resident solve in flight (graph replays on
separate streams, so their kernels overlap on the device): a finished cycle
is read and the next one enqueued at once; a solve retires when its density
residual rms has dropped by
`drop` from its first cycle (or after `max_cycles`), its force coefficients
are read and its device memory is freed.  Reported: time-to-converged per
request (from arrival), p50 / p99, converged solves per second and the
aggregate cell-update rate.

    python -m cosa_amd.serve --requests 24 --precision mixed
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from . import _lib
from . import case as cases

POSITIVITY = 5   # COSAHIP_ERR_POSITIVITY: a cell's pressure or density fell to qmin

# C5 grid sizes: O-grid nodes (i around the section, j normal); 5 k-nodes
# (4 cells) so that both multigrid levels keep at least 2 cells in k
GRIDS = {"small": (129, 33), "medium": (257, 65), "large": (513, 129)}
NK = 5


@dataclass
class Request:
    rid: int
    grid: str = "small"
    mach: float = 0.5
    alpha_deg: float = 1.0
    nharms: int = 0            # > 0: harmonic-balance pitching, dtheta0 = 1 deg
    rans: bool = False         # Wilcox k-omega RANS (viscous wall) instead of Euler
    reyno: float = 1.0e6       # Reynolds number of a RANS request
    arrival: float = 0.0       # seconds after the server starts
    drop: float = 1e-3         # retire when resrms(1) <= drop * first resrms(1)
    max_cycles: int = 400
    # retire as stalled when the best resrms(1) has not halved over this many
    # cycles (a limit cycle of the limiter / shock); 0: never
    stall_window: int = 0
    precision: int = 0         # cosahip_set_precision mode
    tail_rms_factor: float = 1e-1   # mode 2: FP64 tail below this fraction of the first rms

    def make_case(self) -> "cases.Case":
        ni, nj = GRIDS[self.grid]
        # CIRS with variable coefficients (cirs_v2), the fastest-converging
        # smoothing of these O-grids
        kw = dict(machfs=self.mach, alpha_deg=self.alpha_deg, irsop=3, psirs=0.125, cfl=2.0, cfl_irs=5.0, nl_crs=2,
                  ncycle=self.max_cycles, toler=0.0)
        pitch = None
        if self.nharms:
            kw.update(nharms=self.nharms, harbal=1, omega=0.2, irsop=1, cfl_irs=3.5)
            pitch = dict(dtheta0=math.radians(1.0), xh=0.25, yh=0.0)
        if self.rans:
            from .config import kw_turbulence
            # Wilcox k-omega, viscous wall (type 1500) resolved to y+ ~ 1 at Re 1e6
            kw.update(irsop=1, cfl=1.5, cfl_irs=3.0, ifixq=1,
                      **kw_turbulence(self.mach, self.reyno, tu=1e-4, mutfar=1e-2, stemp=110.4 / 288.15))
            return cases.naca_case(ni, nj, NK, nlevel=2, name=f"req{self.rid}", wall=1500, ds_wall=2e-5, **kw)
        return cases.naca_case(ni, nj, NK, nlevel=2, name=f"req{self.rid}", pitch=pitch, **kw)


@dataclass
class Result:
    rid: int
    cells: int
    cycles: int
    converged: bool
    t_admit: float
    t_done: float
    latency: float             # t_done - arrival
    resrms: List[float]        # resrms(1) per cycle
    forces: np.ndarray         # (nsnap, 12) of surface 1 (type-14 walls)
    updates: float = 0.0       # interior (cell, snapshot) RK-stage updates performed
    q: Optional[np.ndarray] = None
    drop_reached: float = 1.0  # lowest resrms(1) / first resrms(1)
    stalled: bool = False      # retired by the stall rule before its drop
    failed: bool = False       # retired after a positivity failure (the reference stops there)


def device_bytes_estimate(c: "cases.Case") -> int:
    """Device memory of a solve: the library's per-level fields, scratch and
    tables, from the array extents (an upper estimate)."""
    tot = 0
    for L in c.levels:
        nl = L.nl
        p3 = c.size(nl, "p3")
        tot += 8 * (6 * p3 + 6 * c.size(nl, "p3s") + 4 * c.size(nl, "met") + 9 * c.size(nl, "x2")
                    + 6 * c.size(nl, "p1") + c.size(nl, "vol") + 9 * c.size(nl, "d0") + p3)
    return tot + (8 << 20)


@dataclass
class Server:
    budget_bytes: int = 8 << 30
    max_active: int = 64
    device: int = 0
    keep_state: bool = False
    verbose: bool = False      # a line on stderr per retired request
    results: List[Result] = field(default_factory=list)

    def run(self, requests: List[Request], clock=time.perf_counter) -> List[Result]:
        """Serve the requests.  Grids and metrics are built before the clock
        starts (the client's set-up step, done on the host as the reference's
        driver does); admission, upload, the cycles, retirement and the
        force read-back are timed."""
        from .solver import Solve
        built = {r.rid: r.make_case() for r in requests}
        waiting = sorted(requests, key=lambda r: (r.arrival, r.rid))
        active: Dict[int, dict] = {}
        used = 0
        t0 = clock()
        while waiting or active:
            now = clock() - t0
            # admission: arrived requests, in order, while the footprint fits
            while waiting and waiting[0].arrival <= now and len(active) < self.max_active:
                r = waiting[0]
                c = built[r.rid]
                need = device_bytes_estimate(c)
                if active and used + need > self.budget_bytes:
                    break
                waiting.pop(0)
                s = Solve(c, device=self.device, q=c.freestream())
                if r.precision:   # mode 2: threshold set after the first cycle
                    s.set_precision(r.precision, 1e-300 if r.precision == 2 else 0.0)
                active[r.rid] = dict(req=r, case=c, solve=s, mgit=0, need=need, t_admit=clock() - t0,
                                     first=None, pending=False)
                used += need
            if not active:
                # idle until the next arrival
                time.sleep(max(0.0, min(0.001, waiting[0].arrival - now)))
                continue
            # every resident solve keeps one cycle in flight: a finished cycle
            # is read, tested and followed at once by the next (no lockstep
            # rounds, so short and long cycles of different grids overlap)
            progressed = False
            retire = []
            for rid, a in active.items():
                s, r = a["solve"], a["req"]
                if not a["pending"]:
                    a["mgit"] += 1
                    s.cycle_enqueue(a["mgit"])
                    a["pending"] = progressed = True
                    continue
                if not s.cycle_ready():
                    continue
                a["pending"] = False
                progressed = True
                try:
                    s.cycle_wait()
                except _lib.CosahipError as e:
                    if e.code != POSITIVITY:
                        raise
                    a["failed"] = True
                    retire.append((rid, False))
                    continue
                r1 = float(s.history[-1][0])
                if a["first"] is None:
                    a["first"] = r1
                    a["best"], a["best_at"] = r1, a["mgit"]
                    if r.precision == 2:   # FP64 tail once the rms is down by tail_rms_factor
                        s.set_precision(2, r.tail_rms_factor * r1)
                if r1 < 0.5 * a["best"]:
                    a["best"], a["best_at"] = r1, a["mgit"]
                a["min"] = min(a.get("min", r1), r1)
                conv = r1 <= r.drop * a["first"]
                stalled = bool(r.stall_window) and a["mgit"] - a["best_at"] >= r.stall_window
                if conv or stalled or a["mgit"] >= r.max_cycles:
                    retire.append((rid, conv))
                else:
                    a["mgit"] += 1
                    s.cycle_enqueue(a["mgit"])
                    a["pending"] = True
            if not progressed and not retire:
                time.sleep(2e-5)
            for rid, conv in retire:
                a = active.pop(rid)
                s, r, c = a["solve"], a["req"], a["case"]
                t_done = clock() - t0
                failed = a.get("failed", False)
                f = np.zeros((c.nsnap, 12)) if failed else s.force(nsurface=1, include_inviscid=True)[0]
                q = s.download("q") if self.keep_state else None
                upd = s.cell_updates_per_cycle() * a["mgit"]
                first = a["first"] if a["first"] is not None else 1.0
                self.results.append(Result(rid=rid, cells=c.ncell() * c.nsnap, cycles=a["mgit"], converged=conv,
                                           t_admit=a["t_admit"], t_done=t_done, latency=t_done - r.arrival,
                                           resrms=[float(h[0]) for h in s.history], forces=f, updates=upd, q=q,
                                           drop_reached=a.get("min", first) / first, failed=failed,
                                           stalled=not conv and not failed and a["mgit"] < r.max_cycles))
                s.close()
                used -= a["need"]
                if self.verbose:
                    res = self.results[-1]
                    print(f"retired {rid} ({r.grid}, {'rans' if r.rans else 'hb' if r.nharms else 'euler'}, "
                          f"M {r.mach:.2f}, alpha {r.alpha_deg:.1f}): {res.cycles} cycles, drop "
                          f"{res.drop_reached:.1e}, {'converged' if conv else 'failed' if res.failed else 'stalled' if res.stalled else 'max'}, "
                          f"t {t_done:.1f} s", file=sys.stderr, flush=True)
        return self.results


# slots of one batch solve per grid size (requests of one grid run as the
# blocks of one solve, so small grids share their launches)
BATCH_SLOTS = {"small": 16, "medium": 8, "large": 2}


def batch_case(req: Request, nslots: int) -> "cases.Case":
    """`nslots` copies of the request's grid as the blocks of one case, each
    with its own seam self-cut and no cut between copies: every block
    evolves as the grid solved alone."""
    one = req.make_case()
    b = one.blocks[0]
    cut = one.cuts[0]
    cuts = [cases.Cut(k + 1, cut.istrt1, cut.iend1, k + 1, cut.istrt2, cut.iend2, cut.iord) for k in range(nslots)]
    return cases.Case(dict(one.config), [b] * nslots, cuts, [one.xyz[0]] * nslots, name=f"batch_{req.grid}",
                      motion=one.motion)


def block_slice(c: "cases.Case", block: int, nl: int = 1) -> slice:
    """The P3 slice (all variables and snapshots) of block `block` (0-based)."""
    L = c.levels[nl - 1]
    o3, p3 = L.offsets(3), L.points(3)
    w = c.npde * c.nsnap
    return slice(o3[block] * w, (o3[block] + p3[block]) * w)


@dataclass
class BatchServer:
    """Continuous batching at block granularity: the requests of one grid
    size and harmonic count share a batch solve of BATCH_SLOTS[grid] blocks;
    an arriving request takes a free slot (its block gets the request's
    freestream and initial state between cycles), every batch keeps one
    cycle in flight, and a slot retires at its own residual drop (per-block
    norms) with its own force coefficients, freeing the slot for the next
    request.  FP64 or FP32-residual (precision 1) per batch."""
    device: int = 0
    keep_state: bool = False
    verbose: bool = False
    slots: Dict[str, int] = field(default_factory=lambda: dict(BATCH_SLOTS))
    results: List[Result] = field(default_factory=list)

    def run(self, requests: List[Request], clock=time.perf_counter) -> List[Result]:
        from .solver import Solve
        keyof = lambda r: (r.grid, r.nharms, r.precision)  # noqa: E731
        built = {}      # key -> batch case (built before the clock, the client's set-up)
        init = {}       # rid -> the request's initial block state
        for r in requests:
            k = keyof(r)
            if k not in built:
                built[k] = batch_case(r, self.slots[r.grid])
            one = r.make_case()
            init[r.rid] = [one.freestream(nl) for nl in range(1, one.config["nlevel"] + 1)]
        waiting = sorted(requests, key=lambda r: (r.arrival, r.rid))
        batches: Dict[tuple, dict] = {}
        t0 = clock()
        while waiting or any(any(sl is not None for sl in b["slot"]) for b in batches.values()):
            now = clock() - t0
            # admission into free slots; a batch is created on its first request
            admitted = {}
            while waiting and waiting[0].arrival <= now:
                r = waiting[0]
                k = keyof(r)
                b = batches.get(k)
                if b is None:
                    c = built[k]
                    sv = Solve(c, device=self.device, q=c.freestream())
                    sv.set_block_rms(True)
                    if r.precision == 1:
                        sv.set_precision(1)
                    b = batches[k] = dict(case=c, solve=sv, slot=[None] * c.levels[0].nblocks, pending=False,
                                          mgit=0, q=None)
                if b["pending"] or None not in b["slot"]:
                    break        # wait for the batch's cycle / a free slot (arrival order kept)
                waiting.pop(0)
                i = b["slot"].index(None)
                admitted.setdefault(k, []).append((i, r))
                b["slot"][i] = dict(req=r, first=None, cycles=0, t_admit=clock() - t0)
            for k, lst in admitted.items():
                # the slot's block on every level starts from the request's
                # freestream, as a solve of its own grid does
                b = batches[k]
                sv, c = b["solve"], b["case"]
                for i, r in lst:
                    sv.set_block_freestream(i + 1, r.mach, r.alpha_deg)
                for nl in range(1, c.config["nlevel"] + 1):
                    q = sv.download("q", nl)
                    for i, r in lst:
                        q[block_slice(c, i, nl)] = init[r.rid][nl - 1]
                    sv.upload("q", q, nl)
            progressed = False
            for k, b in batches.items():
                sv, c = b["solve"], b["case"]
                if not any(sl is not None for sl in b["slot"]):
                    continue
                if not b["pending"]:
                    b["mgit"] += 1
                    sv.cycle_enqueue(b["mgit"])
                    b["pending"] = progressed = True
                    continue
                if not sv.cycle_ready():
                    continue
                b["pending"] = False
                progressed = True
                bad = -1   # the slot whose block failed the positivity test this cycle
                try:
                    sv.cycle_wait()
                except _lib.CosahipError as e:
                    if e.code != POSITIVITY:
                        raise
                    bad = int(e.info[1]) - 1
                rms = sv.block_rms()
                done = []
                for i, sl in enumerate(b["slot"]):
                    if sl is None:
                        continue
                    sl["cycles"] += 1
                    if i == bad:
                        sl["failed"] = True
                        done.append((i, False))
                        continue
                    r1 = float(rms[i, 0])
                    if sl["first"] is None:
                        sl["first"] = r1
                        sl["best"], sl["best_at"] = r1, sl["cycles"]
                    if r1 < 0.5 * sl["best"]:
                        sl["best"], sl["best_at"] = r1, sl["cycles"]
                    sl["min"] = min(sl.get("min", r1), r1)
                    conv = r1 <= sl["req"].drop * sl["first"]
                    w = sl["req"].stall_window
                    if conv or sl["cycles"] >= sl["req"].max_cycles or (w and sl["cycles"] - sl["best_at"] >= w):
                        done.append((i, conv))
                if done:
                    f = sv.force_blocks(nsurface=1, include_inviscid=True)
                    q = sv.download("q") if self.keep_state else None
                    t_done = clock() - t0
                    nc = c.ncell() // c.levels[0].nblocks
                    for i, conv in done:
                        sl = b["slot"][i]
                        r = sl["req"]
                        failed = sl.get("failed", False)
                        first = sl["first"] if sl["first"] is not None else 1.0
                        self.results.append(Result(
                            rid=r.rid, cells=nc * c.nsnap, cycles=sl["cycles"], converged=conv, t_admit=sl["t_admit"],
                            t_done=t_done, latency=t_done - r.arrival, resrms=[],
                            forces=np.zeros_like(f[i, 0]) if failed else f[i, 0],
                            updates=sv.cell_updates_per_cycle() / c.levels[0].nblocks * sl["cycles"],
                            q=q[block_slice(c, i)].copy() if q is not None else None,
                            drop_reached=sl.get("min", first) / first, failed=failed,
                            stalled=not conv and not failed and sl["cycles"] < r.max_cycles))
                        b["slot"][i] = None
                        if failed:
                            # the failed block restarts from its freestream so the batch goes on
                            for nl in range(1, c.config["nlevel"] + 1):
                                qq = sv.download("q", nl)
                                qq[block_slice(c, i, nl)] = init[r.rid][nl - 1]
                                sv.upload("q", qq, nl)
                        if self.verbose:
                            print(f"retired {r.rid} ({r.grid}, batch slot {i}): {sl['cycles']} cycles, "
                                  f"{'converged' if conv else 'max'}, t {t_done:.1f} s", file=sys.stderr, flush=True)
            if not progressed:
                time.sleep(2e-5)
        for b in batches.values():
            b["solve"].close()
        return self.results


def summarize(results: List[Result], wall: float, cell_updates: float) -> dict:
    """Converged solves per second and p50 / p99 time-to-converged count only
    the requests that reached their drop; the ones retired at max_cycles are
    reported separately (count and their own retirement latencies)."""
    conv = [r for r in results if r.converged]
    unconv = [r for r in results if not r.converged]
    lat = np.array([r.latency for r in conv])
    out = {"requests": len(results), "converged": len(conv), "unconverged_retired": len(unconv),
           "stalled": sum(1 for r in unconv if r.stalled), "failed": sum(1 for r in unconv if r.failed),
           "wall_s": round(wall, 3), "solves_per_s": round(len(conv) / wall, 3),
           "latency_p50_s": round(float(np.percentile(lat, 50)), 4) if conv else None,
           "latency_p99_s": round(float(np.percentile(lat, 99)), 4) if conv else None,
           "mcell_updates_per_s": round(cell_updates / wall / 1e6, 2)}
    if unconv:
        ul = np.array([r.latency for r in unconv])
        out["unconverged_latency_p50_s"] = round(float(np.percentile(ul, 50)), 4)
        out["unconverged_drop_reached_median"] = float(np.median([r.drop_reached for r in unconv]))
        out["unconverged_drop_reached_best"] = float(min(r.drop_reached for r in unconv))
    return out


def request_mix(n: int, seed: int = 0, rate: float = 4.0, precision: int = 0, drop: float = 1e-8,
                grids=("small", "medium", "large"), hb_fraction: float = 0.1, rans_fraction: float = 0.45,
                max_cycles: int = 4000, stall_window: int = 400, mach=(0.2, 0.75),
                alpha_deg=(-4.0, 8.0)) -> List[Request]:
    """n requests with Poisson arrivals at `rate` per second, grid sizes
    cycling through `grids`, Mach and incidence uniform in the given ranges;
    a fraction hb_fraction harmonic-balance pitching with N = 2, a fraction
    rans_fraction steady Wilcox k-omega RANS, the rest steady Euler (the
    serving configuration's defaults: M in [0.2, 0.75], alpha in [-4, 8]
    degrees, a 1e-8 drop).  RANS requests get three times the stall window:
    their density residual stagnates for hundreds of cycles while the
    boundary layer develops."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += float(rng.exponential(1.0 / rate))
        u = rng.random()
        hb = u < hb_fraction
        rans = (not hb) and u < hb_fraction + rans_fraction
        out.append(Request(rid=i, grid=grids[i % len(grids)], mach=float(rng.uniform(*mach)),
                           alpha_deg=float(rng.uniform(*alpha_deg)), nharms=2 if hb else 0, rans=rans,
                           arrival=t, drop=drop, precision=precision, max_cycles=max_cycles,
                           stall_window=3 * stall_window if rans else stall_window))
    return out


def cell_updates(results: List[Result]) -> float:
    """Interior (cell, snapshot) RK-stage updates of the solves served."""
    return float(sum(r.updates for r in results))


def sequential(requests: List[Request], device: int = 0, verbose: bool = False) -> List[Result]:
    """The same requests solved one after another with the plain FP64
    synchronous path (smooth / restrict / prolong per routine), for
    comparison."""
    from .solver import Solve
    out = []
    built = {r.rid: r.make_case() for r in requests}
    t0 = time.perf_counter()
    for r in sorted(requests, key=lambda r: r.arrival):
        c = built[r.rid]
        s = Solve(c, device=device, q=c.freestream())
        ta = time.perf_counter() - t0
        first = None
        conv = False
        mgit = 0
        best = best_at = None
        failed = False
        while mgit < r.max_cycles:
            mgit += 1
            try:
                s.cycle(mgit)
            except _lib.CosahipError as e:
                if e.code != POSITIVITY:
                    raise
                failed = True
                break
            r1 = float(s.history[-1][0])
            first = r1 if first is None else first
            if best is None or r1 < 0.5 * best:
                best, best_at = r1, mgit
            if r1 <= r.drop * first:
                conv = True
                break
            if r.stall_window and mgit - best_at >= r.stall_window:
                break
        td = time.perf_counter() - t0
        h1 = [float(h[0]) for h in s.history] or [1.0]
        out.append(Result(rid=r.rid, cells=c.ncell() * c.nsnap, cycles=mgit, converged=conv, t_admit=ta,
                          t_done=td, latency=td - r.arrival, resrms=h1,
                          forces=np.zeros((c.nsnap, 12)) if failed else s.force(nsurface=1, include_inviscid=True)[0],
                          updates=s.cell_updates_per_cycle() * mgit, drop_reached=min(h1) / h1[0], failed=failed,
                          stalled=not conv and not failed and mgit < r.max_cycles))
        if verbose:
            print(f"sequential {r.rid}: {mgit} cycles, drop {min(h1) / h1[0]:.1e}, t {td:.1f} s", file=sys.stderr,
                  flush=True)
        s.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--requests", type=int, default=24)
    ap.add_argument("--rate", type=float, default=4.0, help="mean arrivals per second")
    ap.add_argument("--precision", choices=["fp64", "fp32", "mixed"], default="mixed")
    ap.add_argument("--drop", type=float, default=1e-8)
    ap.add_argument("--max-cycles", type=int, default=4000)
    ap.add_argument("--stall-window", type=int, default=400,
                    help="retire a request whose best resrms(1) has not halved in this many cycles (0: never)")
    ap.add_argument("--rans-fraction", type=float, default=0.45,
                    help="fraction of steady k-omega RANS requests (the batch mode serves Euler and HB only)")
    ap.add_argument("--batch", action="store_true",
                    help="serve requests of one grid as the blocks of shared batch solves (FP64 / FP32 residual)")
    ap.add_argument("--hw-queues", type=int, default=16,
                    help="GPU_MAX_HW_QUEUES of this process (HIP's default 4 lets only 4 streams run at once)")
    ap.add_argument("--budget-gib", type=float, default=64.0)
    ap.add_argument("--baseline", action="store_true", help="also run the requests sequentially in FP64")
    a = ap.parse_args()
    # before the HIP runtime starts (first solve): hardware queues for the
    # concurrent streams of the resident solves
    os.environ.setdefault("GPU_MAX_HW_QUEUES", str(min(max(a.hw_queues, 1), 32)))
    prec = {"fp64": 0, "fp32": 1, "mixed": 2}[a.precision]
    reqs = request_mix(a.requests, rate=a.rate, precision=prec, drop=a.drop, max_cycles=a.max_cycles,
                       stall_window=a.stall_window, rans_fraction=0.0 if a.batch else a.rans_fraction)
    if a.batch:
        for r in reqs:    # batches run one precision throughout: mixed is served as FP32 residual
            r.precision = 1 if prec else 0
    srv = BatchServer(verbose=True) if a.batch else Server(budget_bytes=int(a.budget_gib * (1 << 30)), verbose=True)
    t0 = time.perf_counter()
    res = srv.run(reqs)
    wall = time.perf_counter() - t0
    out = {"serving": summarize(res, wall, cell_updates(res)), "mode": "batch" if a.batch else "per-solve streams",
           "precision": a.precision,
           "drop": a.drop, "rate_per_s": a.rate,
           "mix": {"euler": sum(1 for r in reqs if not r.rans and not r.nharms),
                   "rans": sum(1 for r in reqs if r.rans), "hb_n2": sum(1 for r in reqs if r.nharms)}}
    if a.baseline:
        t0 = time.perf_counter()
        base = sequential(reqs, verbose=True)
        wb = time.perf_counter() - t0
        out["sequential_fp64"] = summarize(base, wb, cell_updates(base))
        # force coefficients (pressure + viscous) of the requests converged in both runs: the
        # served (mixed-precision) result against the plain FP64 one
        tot = lambda r, m: float(r.forces[0, m] + r.forces[0, 6 + m])  # noqa: E731
        bconv = {r.rid: r for r in base if r.converged}
        both = [r for r in res if r.converged and r.rid in bconv]
        out["converged_in_both"] = len(both)
        if both:
            out["max_abs_dCL_vs_fp64"] = max(abs(tot(r, 0) - tot(bconv[r.rid], 0)) for r in both)
            out["max_abs_dCD_vs_fp64"] = max(abs(tot(r, 1) - tot(bconv[r.rid], 1)) for r in both)
        sq = out["sequential_fp64"]["solves_per_s"]
        out["speedup_solves_per_s"] = round(out["serving"]["solves_per_s"] / sq, 2) if sq > 0 else None
        # every retired request, converged or not, per second of wall time
        out["speedup_retired_per_s"] = round((len(res) / wall) / (len(base) / wb), 2)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
