#!/usr/bin/env python3
"""Benchmark of the COSA pseudo-time residual/update path on one MI355X.

Headline workload (default, --workload uneven96): 96 independent
bump-channel blocks whose node counts per direction are drawn from
{17, 33, 65, 129} with numpy.random.default_rng(0) (24,285,184 cells), Euler
at M = 0.7, freestream plus a 1% uniform perturbation (default_rng(1)),
4-stage RK, cirs_c, van Albada MUSCL, 3-level multigrid V-cycles, FP64 --
the largest configuration that fits one GPU.  The same JSON line carries the
RANS wing (4 blocks of 257 x 65 x 65 nodes, Wilcox k-omega, M 0.3, Re 3e6,
alpha 4 deg, 3 levels) under "extra_workloads" (--no-extras leaves it out).

One step = one multigrid cycle of mg() (src/cosa.f90:822-912) over all
blocks, enqueued as one launch sequence and, from the second cycle on,
replayed as one HIP graph (cosahip_cycle_enqueue / cosahip_cycle_wait).  The
metric is cell-updates per second: one RK stage applied to one interior cell,
every level counted.  Inputs are resident in HBM before the timed region;
the timed region is bracketed by device synchronisation.  A second timed
region runs the same cycles synchronously with HIP events around every
kernel family on the library's stream ("sync_path", "families") for the
roofline of the dominant kernel.

cpu_baseline: COSA's own compiled routines (oracle/_ref/libcosaref_o3.so,
amdflang -O3, no fast-math) driven in smooth()/mg() order by
oracle/ref_mg.py, one block per process as the reference's MPI ranks own
whole blocks, on this host's cores, for a bounded time.
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
# HBM bytes per cell-update of the residual pass and of the three CIRS sweeps
# measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (scaled by a
# plain-copy calibration: FETCH x2.00, WRITE x1.00) on the level-1 kernels:
# profiles/residual_uneven96_fp64_hbm_traffic.md
# and profiles/residual_wing4_rans_fp64_hbm_traffic.md.  Re-measure when the
# kernels change.
MEASURED_TRAFFIC_PER_CELL = {("uneven96", "fp64", "residual"): 312.7, ("uneven96", "fp64", "cirs"): 445.8,
                             ("wing4", "fp64", "residual"): 2112.1}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def _wing_cfg():
    from cosa_amd.config import kw_turbulence
    return dict(irsop=1, cfl=2.0, cfl_irs=4.0, nl_crs=2, ncycle=100000, ds_wall=5e-6,
                **kw_turbulence(0.3, 3.0e6, tu=1e-4, mutfar=1e-2, stemp=110.4 / 288.15, ifixq=1, lim_prodtke=1,
                                prdlim=20.0))


def _uneven_cfg():
    return dict(machfs=0.7, irsop=1, cfl=2.0, cfl_irs=4.0, ilimf=4, epslimf=1e-6, nl_crs=2, ncycle=100000)


def make_case(workload: str):
    from cosa_amd import case as cases
    if workload == "uneven96":
        dims = cases.uneven_dims(96, seed=0)
        c = cases.channel_case(dims, nlevel=3, name="uneven96", **_uneven_cfg())
        return c, dict(workload="uneven96", blocks=96, dims_seed=0, cells=c.ncell(), levels=3, physics="euler",
                       mach=0.7, rk_stages=4, irs="cirs_c", limiter="van_albada", precision="fp64",
                       grid="bump channels, node counts per direction from {17,33,65,129}")
    if workload == "wing4":
        c = cases.wing_case(257, 65, 65, nblocks=4, nlevel=3, name="wing4", **_wing_cfg())
        return c, dict(workload="wing4_rans_komega", blocks=4, block_nodes=[257, 65, 65], cells=c.ncell(), levels=3,
                       physics="rans_wilcox_komega", mach=0.3, reynolds=3.0e6, alpha_deg=4.0, rk_stages=4,
                       update="point_implicit_komega", irs="cirs_c", precision="fp64",
                       grid="O-H wing, NACA0012 section, semi-span 2 (AR 4), wall spacing 5e-6 chord")
    if workload.startswith("c3_hb"):
        import math
        nh = int(workload[5:] or 4)
        c = cases.naca_case(385, 97, 5, nlevel=2, name=f"c3_pitch_hb{nh}", irsop=1, cfl=2.0, cfl_irs=3.5, nl_crs=2,
                            ncycle=100000, machfs=0.5, alpha_deg=0.0, nharms=nh, harbal=1, omega=0.1,
                            pitch=dict(dtheta0=math.radians(5.0), xh=0.25, yh=0.0))
        return c, dict(workload=f"c3_hb_pitch_naca0012_385x97_N{nh}", cells=c.ncell(), snapshots=c.nsnap, levels=2,
                       physics="euler harmonic balance, pitching 5 deg at k = 0.1", mach=0.5, rk_stages=4,
                       irs="cirs_c", precision="fp64")
    if workload == "naca":
        c = cases.naca_case(193, 49, 9, nlevel=3, name="naca0012", irsop=1, cfl=2.0, cfl_irs=4.0, nl_crs=2,
                            ncycle=100000)
        return c, dict(workload="naca0012_193x49x9", cells=c.ncell(), levels=3, physics="euler", mach=0.5,
                       alpha_deg=1.25, rk_stages=4, irs="cirs_c", precision="fp64")
    raise SystemExit(f"unknown workload {workload}")


# ---------------------------------------------------------------------------
# CPU baseline: the reference's compiled routines

def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _ref_block_run(dims, budget_s, lib):
    """One process = one reference rank owning one block: whole multigrid
    cycles of the block for about budget_s; returns (cell-updates, s, cycles)."""
    from cosa_amd import case as cases
    from oracle.ref_mg import RefMG
    c = cases.channel_case([dims], nlevel=3, name="sample", **_uneven_cfg())
    r = RefMG(c, c.perturbed(amp=0.01, seed=1), lib=lib)
    per = r.cell_updates_per_cycle()
    t0 = time.perf_counter()
    n = 0
    while True:
        n += 1
        r.cycle(n)
        if time.perf_counter() - t0 > budget_s or n >= 200:
            break
    return per * n, time.perf_counter() - t0, n


def cpu_baseline(budget_s: float, workers: int):
    """COSA's compiled routines timed on this host: `workers` processes, each
    one of the first `workers` blocks of the uneven set (the reference runs
    one MPI rank per block), aggregate cell-updates per second."""
    import multiprocessing as mp
    from cosa_amd import case as cases
    from oracle.ref_mg import TIMING_LIB
    if not os.path.exists(TIMING_LIB):
        return None
    dims = cases.uneven_dims(96, seed=0)[:workers]
    t0 = time.perf_counter()
    with mp.get_context("spawn").Pool(workers) as pool:
        runs = pool.starmap(_ref_block_run, [(d, budget_s, TIMING_LIB) for d in dims])
    wall = time.perf_counter() - t0
    rate = sum(u / dt for u, dt, _ in runs)
    cells = sum((a - 1) * (b - 1) * (c - 1) for a, b, c in dims)
    return {"value": round(rate / 1e6, 4), "unit": "Mcell-updates/s", "cores": workers, "kind": "reference",
            "sample": f"the first {workers} of the 96 uneven blocks ({cells} cells, 3 levels), one block per "
                      f"process as one reference rank, {min(n for *_, n in runs)}-{max(n for *_, n in runs)} "
                      f"multigrid cycles each in ~{budget_s:.0f} s ({wall:.0f} s wall); COSA's compiled kernels "
                      f"(amdflang -O3 -fdefault-real-8, no fast-math) in smooth/mg order (oracle/ref_mg.py); "
                      f"CPU: {_cpu_model()}"}


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


# ---------------------------------------------------------------------------
# GPU measurement

def measure(workload: str, steps: int, warmup: int, precision: str, tail_rms: float, autotune: bool,
            sync_region: bool = True):
    """Graph-replayed cycles (value), then the synchronous path with kernel
    family timing (roofline).  Returns a dict."""
    from cosa_amd.config import komega
    from cosa_amd.solver import Solve
    t0 = time.time()
    case, cfg_desc = make_case(workload)
    cfg_desc["precision"] = {"fp64": "fp64", "fp32": "fp32 residual / fp64 state",
                             "mixed": f"fp32 residual until resrms(1) < {tail_rms:g}, then fp64",
                             "fp64c": "fp64, residual with multiply-add contraction"}[precision]
    log(f"case {case.name}: {case.ncell()} cells, built in {time.time() - t0:.1f} s")
    q0 = case.perturbed(amp=0.01, seed=1)
    s = Solve(case, device=0, q=q0)
    del q0
    prec_mode = {"fp64": 0, "fp32": 1, "mixed": 2, "fp64c": 3}[precision]
    if prec_mode:
        s.set_precision(prec_mode, tail_rms if prec_mode == 2 else 0.0)
    log("device state resident; warmup")
    mgit = 0
    tuned = None
    for w in range(max(warmup, 2)):
        mgit += 1
        s.cycle_async(mgit)
        if w == 0 and autotune:
            # after one cycle the level states are realistic: pick the kernel arrangements
            tuned = s.autotune(3)
            log(f"autotune: {tuned}")
    s.synchronize()
    log(f"timed region (graph-replayed cycles): {steps} cycles")
    t1 = time.perf_counter()
    for _ in range(steps):
        mgit += 1
        s.cycle_async(mgit)
    s.synchronize()
    t2 = time.perf_counter()
    elapsed = t2 - t1
    per_cycle = s.cell_updates_per_cycle()
    value = per_cycle * steps / elapsed / 1e6
    res_hist = [float(h[0]) for h in s.history]
    log(f"elapsed {elapsed:.3f} s, {value:.1f} Mcell-updates/s, resrms(1) {res_hist[0]:.3e} -> {res_hist[-1]:.3e}")
    out = dict(value=value, ms_per_step=elapsed / steps * 1e3, config=dict(cfg_desc, autotune=tuned),
               resrms1_first_last=[res_hist[0], res_hist[-1]], cycles_run=mgit)
    if sync_region:
        # the synchronous path with HIP events around each kernel family
        s.timing_reset()
        s.timing(True)
        s.synchronize()
        log(f"timed region (synchronous, kernel families timed): {steps} cycles")
        t1 = time.perf_counter()
        for _ in range(steps):
            mgit += 1
            s.cycle(mgit)
        s.synchronize()
        t2 = time.perf_counter()
        fam = s.timing_read()
        s.timing(False)
        out["sync_path"] = {"value": round(per_cycle * steps / (t2 - t1) / 1e6, 2),
                            "ms_per_step": round((t2 - t1) / steps * 1e3, 3)}
        bpc_tab = BYTES_PER_CELL_RANS if komega(case.config) else BYTES_PER_CELL_EULER
        if prec_mode and s.precision_in_use() == 32:
            bpc_tab = dict(bpc_tab, **(BYTES_PER_CELL_RANS_FP32_RESIDUAL if komega(case.config)
                                       else BYTES_PER_CELL_FP32_RESIDUAL))
        if fam.get("update", {}).get("launches", 0) == 0 and not komega(case.config):
            # the RK update runs inside the last CIRS sweep: that sweep reads
            # qold and writes q instead of writing res for the update to read
            bpc_tab = dict(bpc_tab, cirs=BYTES_PER_CELL_EULER["cirs"] + BYTES_PER_CELL_EULER["update"] - 80.0)
        fams = {}
        for name, d in fam.items():
            bpc = bpc_tab.get(name)
            if d["ms"] > 0 and bpc:
                gbs = bpc * d["items"] / (d["ms"] / 1e3) / 1e9
                fams[name] = {"ms_total": round(d["ms"], 3), "launches": d["launches"], "bytes_per_cell": bpc,
                              "achieved_GBs": round(gbs, 1)}
            elif d["ms"] > 0:
                fams[name] = {"ms_total": round(d["ms"], 3), "launches": d["launches"]}
        dom = max((n for n in fams if bpc_tab.get(n)), key=lambda n: fams[n]["ms_total"])
        d = fam[dom]
        achieved = bpc_tab[dom] * d["items"] / (d["ms"] / 1e3) / 1e9
        tpc = MEASURED_TRAFFIC_PER_CELL.get((workload, precision, dom))
        traffic = round(tpc * d["items"] / max(d["launches"], 1)) if tpc else None
        out["roofline"] = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                           "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                           "bytes_per_cell_update": bpc_tab[dom],
                           "avg_launch_ms": round(d["ms"] / max(d["launches"], 1), 4)}
        out["families"] = fams
    s.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="uneven96", choices=["uneven96", "wing4", "naca", "c3_hb1", "c3_hb2",
                                                                "c3_hb4"])
    ap.add_argument("--no-extras", action="store_true", help="leave out the wing workload of extra_workloads")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default kernel arrangements")
    ap.add_argument("--cpu-workers", type=int, default=min(16, os.cpu_count() or 1),
                    help="concurrent single-block reference processes for cpu_baseline (at most 16)")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32", "mixed", "fp64c"],
                    help="fp32: FP32 residual / FP64 state (cosahip_set_precision mode 1); mixed: FP32 until "
                         "resrms(1) < --tail-rms, then FP64 (mode 2); fp64c: FP64 residual with multiply-add "
                         "contraction (mode 3)")
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

    head = measure(a.workload, a.steps, a.warmup, a.precision, a.tail_rms, not a.no_autotune)
    extras = {}
    if not a.no_extras and a.workload == "uneven96" and a.precision == "fp64":
        # the same workload in the fast FP64 mode (contracted residual, held to the FP64 parity bar)
        u = measure("uneven96", a.steps, a.warmup, "fp64c", a.tail_rms, not a.no_autotune)
        extras["uneven96_fp64_contracted"] = {"value": round(u["value"], 2), "unit": "Mcell-updates/s",
                                              "ms_per_step": round(u["ms_per_step"], 3),
                                              "roofline": u.get("roofline"),
                                              "residual": (u.get("families") or {}).get("residual")}
    if not a.no_extras and a.workload == "uneven96":
        w = measure("wing4", a.steps, a.warmup, a.precision, a.tail_rms, not a.no_autotune)
        extras["wing4_rans_komega"] = {"value": round(w["value"], 2), "unit": "Mcell-updates/s",
                                       "ms_per_step": round(w["ms_per_step"], 3), "config": w["config"],
                                       "sync_path": w.get("sync_path"), "roofline": w.get("roofline")}
    cpu = None
    if not a.no_cpu:
        log("cpu baseline (the reference's compiled routines)")
        cpu = cpu_baseline(a.cpu_seconds, max(1, min(16, a.cpu_workers)))

    out = {"metric": "Mcell-updates/s", "value": round(head["value"], 2), "unit": "Mcell-updates/s", "n_gpus": 1,
           "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(head["ms_per_step"], 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64" if a.precision in ("fp64", "fp64c") else "f32 residual / f64 state", "data": "synthetic",
           "config": head["config"], "roofline": head.get("roofline"), "cpu_baseline": cpu,
           "sync_path": head.get("sync_path"), "families": head.get("families"),
           "resrms1_first_last": head["resrms1_first_last"]}
    if extras:
        out["extra_workloads"] = extras
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
