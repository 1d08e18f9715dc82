#!/usr/bin/env python3
"""Benchmark of the COSA pseudo-time residual/update path on one MI355X.

Workload (default, --workload wing4): steady RANS with the Wilcox k-omega
model around a rectangular wing with a NACA 0012 section and aspect ratio 4
-- 4 blocks of 257 x 65 x 65 nodes (4,194,304 cells), O-H topology, wall
spacing 5e-6 chord -- at M = 0.3, Re = 3e6, alpha = 4 deg: point-implicit
4-stage RK with central implicit residual smoothing (cirs_c), van Albada
MUSCL, 3-level multigrid V-cycles, FP64, freestream plus a 1% uniform
perturbation (default_rng(1)).

--workload uneven96: 96 independent bump-channel blocks whose node counts
per direction are drawn from {17, 33, 65, 129} with numpy.random
.default_rng(0) (24,285,184 cells), Euler at M = 0.7, same scheme.

One step = one multigrid cycle of mg() (src/cosa.f90:822-912) over all
blocks.  The metric is cell-updates per second: one RK stage applied to one
interior cell, every level counted.  Inputs are resident in HBM before the
timed region; the timed region is bracketed by device synchronisation.

Prints ONE JSON line with the driver's contract plus ``roofline`` (dominant
kernel family, achieved algorithmic GB/s from HIP events recorded on the
library's stream over the timed region) and ``cpu_baseline`` (the CPU
restatement timed on this host's cores on a bounded sample).
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# algorithmic bytes per cell-update of each kernel family (FP64, static grid;
# each required array read once and written once, halos not counted)
BYTES_PER_CELL_EULER = {
    "residual": 176.0,   # q 5x8 R + face metrics 3x4x8 R + res 5x8 W
    "cirs": 248.0,       # res 5x8 R+W per sweep (3 sweeps) + dtvol 8 R
    "update": 128.0,     # qold 40 R + res 40 R + dtvol 8 R + q 40 W
    "halo": None,        # surface work
    "convert": 80.0,     # q 40 R + 40 W
    "lambda": 144.0,     # q 40 R + metrics 96 R + dtvol 8 W
    "other": None,
}
# FP32-residual mode (Euler): the residual reads FP32 copies of q and the
# face metrics and writes res in FP64; cons2prim also writes the FP32 copy
BYTES_PER_CELL_FP32_RESIDUAL = {
    "residual": 108.0,   # q 5x4 R + face metrics 3x4x4 R + res 5x8 W
    "convert": 100.0,    # q 40 R + 40 W + FP32 copy 5x4 W
}
BYTES_PER_CELL_RANS_FP32_RESIDUAL = {
    # q 7x4 R + face metrics 48 R + metric derivatives 108 R + mut 4 R + vol 8 R
    # + res 7x8 W + prd, divvel, delpls 24 W
    "residual": 276.0,
    "convert": 140.0,    # q 7x8 R + W + FP32 copy 7x4 W
}
BYTES_PER_CELL_RANS = {  # Wilcox k-omega, npde = 7
    # q 7x8 R + face metrics 96 R + metric derivatives 3 faces x 9 x 8 = 216 R
    # + mut 8 R + vol 8 R + res 7x8 W + prd, divvel, delpls 24 W
    "residual": 464.0,
    "cirs": 352.0,       # res 7x8 R+W per sweep (3 sweeps) + dtvol, dtvolt 16 R
    # qold 56 R + res 56 R + q 56 R+W + dtvol, dtvolt, vol, delpls, prd 40 R
    "update": 264.0,
    "halo": None,
    "convert": 112.0,    # q 7x8 R + W
    "lambda": 176.0,     # q 5x8 R + metrics 96 R + vol 8 + mut 8 R + dtvol, dtvolt 16 W + ... betas
    "other": None,
}
# HBM bytes per cell-update of the residual pass measured with rocprofv3 --pmc
# FETCH_SIZE / WRITE_SIZE (scaled by a plain-copy calibration: FETCH x2.00,
# WRITE x1.00): profiles/residual_uneven96_fp64_hbm_traffic.md and
# profiles/residual_wing4_rans_fp64_hbm_traffic.md.  Re-measure when the
# kernels change.
MEASURED_TRAFFIC_PER_CELL = {("uneven96", "fp64", "residual"): 322.7, ("wing4", "fp64", "residual"): 2290.0}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def _wing_cfg():
    from cosa_amd.config import kw_turbulence
    return dict(irsop=1, cfl=2.0, cfl_irs=4.0, nl_crs=2, ncycle=100000, ds_wall=5e-6,
                **kw_turbulence(0.3, 3.0e6, tu=1e-4, mutfar=1e-2, stemp=110.4 / 288.15, ifixq=1, lim_prodtke=1,
                                prdlim=20.0))


WING_CFG = None


def make_case(workload: str):
    from cosa_amd import case as cases
    global WING_CFG
    WING_CFG = _wing_cfg()
    if workload == "uneven96":
        dims = cases.uneven_dims(96, seed=0)
        c = cases.channel_case(dims, nlevel=3, name="uneven96", machfs=0.7, irsop=1, cfl=2.0, cfl_irs=4.0,
                               ilimf=4, epslimf=1e-6, nl_crs=2, ncycle=100000)
        return c, dict(workload="uneven96", blocks=96, dims_seed=0, cells=c.ncell(), levels=3, physics="euler",
                       mach=0.7, rk_stages=4, irs="cirs_c", limiter="van_albada", precision="fp64")
    if workload == "wing4":
        c = cases.wing_case(257, 65, 65, nblocks=4, nlevel=3, name="wing4", **WING_CFG)
        return c, dict(workload="wing4_rans_komega", blocks=4, block_nodes=[257, 65, 65], cells=c.ncell(), levels=3,
                       physics="rans_wilcox_komega", mach=0.3, reynolds=3.0e6, alpha_deg=4.0, rk_stages=4,
                       update="point_implicit_komega", irs="cirs_c", precision="fp64",
                       grid="O-H wing, NACA0012 section, semi-span 2 (AR 4), wall spacing 5e-6 chord")
    if workload == "naca":
        c = cases.naca_case(193, 49, 9, nlevel=3, name="naca0012", irsop=1, cfl=2.0, cfl_irs=4.0, nl_crs=2,
                            ncycle=100000)
        return c, dict(workload="naca0012_193x49x9", cells=c.ncell(), levels=3, physics="euler", mach=0.5,
                       alpha_deg=1.25, rk_stages=4, irs="cirs_c", precision="fp64")
    raise SystemExit(f"unknown workload {workload}")


def _cpu_sample_case(workload: str):
    """The CPU sample of a workload: the wing at 1/64 of the cells, or the
    first uneven channel blocks up to 1.5 M cells."""
    from cosa_amd import case as cases
    if workload == "wing4":
        c = cases.wing_case(65, 17, 17, nblocks=4, nlevel=3, name="sample", **_wing_cfg())
        return c, "the wing at 1/64 of the cells: 4 blocks of 65x17x17 nodes"
    if workload == "uneven96":
        dims = cases.uneven_dims(96, seed=0)
        nb, cells = 0, 0
        for d in dims:
            nb += 1
            cells += (d[0] - 1) * (d[1] - 1) * (d[2] - 1)
            if cells >= 1_500_000:
                break
        c = cases.channel_case(dims[:nb], nlevel=3, name="sample", machfs=0.7, irsop=1, cfl=2.0, cfl_irs=4.0,
                               ilimf=4, epslimf=1e-6, nl_crs=2, ncycle=100000)
        return c, f"the first {nb} of the 96 uneven blocks"
    c, _ = make_case(workload)
    return c, "the whole case"


def _cpu_sample_run(workload: str, budget_s: float):
    """One single-threaded oracle process: whole multigrid cycles of the
    sample for about budget_s; returns (cell-updates, seconds, cycles,
    cells, description)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_bind import Oracle, OrCase
    sample, desc = _cpu_sample_case(workload)
    o = Oracle()
    o.set_config(sample.config)
    oc = OrCase(sample, q1=sample.perturbed(amp=0.01, seed=1))
    st = sample.config["nstage"]
    per_cycle = st * (sample.ncell(1) + 2 * sample.ncell(2) + sample.ncell(3))
    where = (ctypes.c_int32 * 8)()
    done = (ctypes.c_int32 * 1)()
    t0 = time.perf_counter()
    ncyc = 0
    cp = oc.ptr()
    while True:
        mgit = ncyc + 1
        rc = o.or_smooth(cp, 1, 0, 1, 1, mgit, done, where)
        o.or_restrict(cp, 1)
        rc |= o.or_smooth(cp, 2, 1, 1, 1, mgit, done, where)
        o.or_restrict(cp, 2)
        rc |= o.or_smooth(cp, 3, 1, 0, 1, mgit, done, where)
        o.or_prolong(cp, 2)
        rc |= o.or_smooth(cp, 2, 0, 0, 1, mgit, done, where)
        o.or_prolong(cp, 1)
        ncyc += 1
        if rc:
            raise RuntimeError("oracle sample failed")
        if time.perf_counter() - t0 > budget_s or ncyc >= 400:
            break
    dt = time.perf_counter() - t0
    return per_cycle * ncyc, dt, ncyc, sample.ncell(), desc


def cpu_baseline(workload: str, budget_s: float, workers: int):
    """The CPU restatement (oracle/liboracle.so, gcc -O2, one thread per
    process) timed on this host's cores: `workers` independent copies of the
    sample run at once in fresh processes (the throughput form of the
    reference's single-rank solves), aggregate cell-updates per second."""
    import multiprocessing as mp
    t0 = time.perf_counter()
    if workers > 1:
        with mp.get_context("spawn").Pool(workers) as pool:
            runs = pool.starmap(_cpu_sample_run, [(workload, budget_s)] * workers)
    else:
        runs = [_cpu_sample_run(workload, budget_s)]
    wall = time.perf_counter() - t0
    rate = sum(u / dt for u, dt, *_ in runs)
    u, dt, ncyc, cells, desc = runs[0]
    return {"value": round(rate / 1e6, 4), "unit": "Mcell-updates/s", "cores": workers, "kind": "port",
            "sample": f"{workers} concurrent single-threaded copies of {desc} ({cells} cells), {ncyc} multigrid "
                      f"cycles each in {dt:.1f} s ({wall:.0f} s wall), oracle/liboracle.so (C restatement, gcc -O2)"}


def copy_calibration():
    """A plain copy of known size: 512 MiB read + 512 MiB written per launch.
    The counters FETCH_SIZE / WRITE_SIZE read on it give the factor between
    what they report and the bytes that really crossed HBM."""
    import torch
    n = 64 << 20
    x = torch.ones(n, dtype=torch.float64, device="cuda:0")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": n * 8, "copy_calibration_bytes_written": n * 8,
                      "launches": 3}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="wing4", choices=["wing4", "uneven96", "naca"])
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default kernel arrangements")
    ap.add_argument("--cpu-workers", type=int, default=min(16, os.cpu_count() or 1),
                    help="concurrent single-threaded oracle processes for cpu_baseline (the host's cores, at most 16)")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32", "mixed"],
                    help="fp32: FP32 residual / FP64 state (cosahip_set_precision mode 1); mixed: FP32 until "
                         "resrms(1) < --tail-rms, then FP64 (mode 2)")
    ap.add_argument("--tail-rms", type=float, default=1e-6)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 512 MiB device-to-device copy three times (scale for the HBM "
                         "byte counters FETCH_SIZE / WRITE_SIZE) and exit")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs the path on one GPU only (one GPU's share of the blocks); "
              "--gpus must be 1", file=sys.stderr)
        sys.exit(2)

    if a.copy_calibration:
        copy_calibration()
        return

    from cosa_amd.config import komega
    from cosa_amd.solver import Solve
    t0 = time.time()
    case, cfg_desc = make_case(a.workload)
    cfg_desc["precision"] = {"fp64": "fp64", "fp32": "fp32 residual / fp64 state",
                             "mixed": f"fp32 residual until resrms(1) < {a.tail_rms:g}, then fp64"}[a.precision]
    log(f"case {case.name}: {case.ncell()} cells, built in {time.time() - t0:.1f} s")
    q0 = case.perturbed(amp=0.01, seed=1)
    s = Solve(case, device=0, q=q0)
    del q0
    prec_mode = {"fp64": 0, "fp32": 1, "mixed": 2}[a.precision]
    if prec_mode:
        s.set_precision(prec_mode, a.tail_rms if prec_mode == 2 else 0.0)
    log("device state resident; warmup")
    mgit = 0
    tuned = None
    for w in range(a.warmup):
        mgit += 1
        s.cycle(mgit)
        if w == 0 and not a.no_autotune:
            # after one cycle the level states are realistic: pick the kernel arrangements
            tuned = s.autotune(3)
            log(f"autotune: {tuned}")
    s.synchronize()
    s.timing_reset()
    s.timing(True)
    s.synchronize()
    log(f"timed region: {a.steps} cycles")
    t1 = time.perf_counter()
    for _ in range(a.steps):
        mgit += 1
        s.cycle(mgit)
    s.synchronize()
    t2 = time.perf_counter()
    fam = s.timing_read()
    s.timing(False)
    elapsed = t2 - t1
    per_cycle = s.cell_updates_per_cycle()
    value = per_cycle * a.steps / elapsed / 1e6
    res_hist = [float(h[0]) for h in s.history]
    log(f"elapsed {elapsed:.3f} s, {value:.1f} Mcell-updates/s, resrms(1) {res_hist[0]:.3e} -> {res_hist[-1]:.3e}")

    # roofline of the dominant family (most device time in the timed region)
    BYTES_PER_CELL = BYTES_PER_CELL_RANS if komega(case.config) else BYTES_PER_CELL_EULER
    if prec_mode and s.precision_in_use() == 32:
        BYTES_PER_CELL = dict(BYTES_PER_CELL, **(BYTES_PER_CELL_RANS_FP32_RESIDUAL if komega(case.config)
                                                 else BYTES_PER_CELL_FP32_RESIDUAL))
    fams = {}
    for name, d in fam.items():
        bpc = BYTES_PER_CELL.get(name)
        if d["ms"] > 0 and bpc:
            gbs = bpc * d["items"] / (d["ms"] / 1e3) / 1e9
            fams[name] = {"ms_total": round(d["ms"], 3), "launches": d["launches"], "bytes_per_cell": bpc,
                          "achieved_GBs": round(gbs, 1)}
        elif d["ms"] > 0:
            fams[name] = {"ms_total": round(d["ms"], 3), "launches": d["launches"]}
    dom = max((n for n in fams if BYTES_PER_CELL.get(n)), key=lambda n: fams[n]["ms_total"])
    d = fam[dom]
    achieved = BYTES_PER_CELL[dom] * d["items"] / (d["ms"] / 1e3) / 1e9
    tpc = MEASURED_TRAFFIC_PER_CELL.get((a.workload, a.precision, dom))
    traffic = round(tpc * d["items"] / max(d["launches"], 1)) if tpc else None
    roof = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "bytes_per_cell_update": BYTES_PER_CELL[dom], "avg_launch_ms": round(d["ms"] / max(d["launches"], 1), 4)}
    s.close()

    cpu = None
    if not a.no_cpu:
        log("cpu baseline (oracle, 1 thread)")
        cpu = cpu_baseline(a.workload, a.cpu_seconds, a.cpu_workers)

    out = {"metric": "Mcell-updates/s", "value": round(value, 2), "unit": "Mcell-updates/s", "n_gpus": 1,
           "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(elapsed / a.steps * 1e3, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64" if not prec_mode else "f32 residual / f64 state", "data": "synthetic",
           "config": dict(cfg_desc, autotune=tuned), "roofline": roof, "cpu_baseline": cpu, "families": fams,
           "resrms1_first_last": [res_hist[0], res_hist[-1]]}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
