"""Host driver of a libcosahip solve, mirroring COSA's own driver.

``Solve`` owns one device context: it defines every multigrid level of a
:class:`cosa_amd.case.Case`, uploads the geometric inputs and the state, and
runs the reference's multigrid cycle ``mg`` (src/cosa.f90:806-930) over the
C ABI: ``smooth`` on the fine level, ``restrict``, coarse smoothing,
``prolong``.  Method names follow the reference routines.
"""
from __future__ import annotations

import ctypes
from typing import Dict, List, Optional

import numpy as np

from . import _lib
from .case import Case
from .config import komega, to_struct


class PinnedArray:
    """float64 array in pinned host memory (cosahip_host_alloc)."""

    def __init__(self, lib, n: int):
        self.lib, self.n = lib, int(n)
        self.ptr = ctypes.c_void_p()
        rc = lib.cosahip_host_alloc(max(self.n, 1) * 8, ctypes.byref(self.ptr))
        if rc:
            raise _lib.CosahipError(rc, "cosahip_host_alloc failed")
        self.array = np.ctypeslib.as_array(ctypes.cast(self.ptr, ctypes.POINTER(ctypes.c_double)), (max(self.n, 1),))[:self.n]

    def free(self):
        if self.ptr is not None and self.ptr.value:
            self.array = None
            self.lib.cosahip_host_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Checkpoint:
    """A checkpoint in flight (Solve.checkpoint_async)."""

    def __init__(self, solve: "Solve", path: str, nl: int):
        self.solve, self.path, self.nl = solve, path, nl
        kom = komega(solve.case.config)
        self.q = PinnedArray(solve.lib, solve.size("q", nl))
        self.mut = PinnedArray(solve.lib, solve.size("mut", nl)) if kom else None
        mp = _lib.dptr(self.mut.array) if kom else None
        solve._check(solve.lib.cosahip_checkpoint_begin(solve.h, nl, _lib.dptr(self.q.array), self.q.n, mp,
                                                        self.mut.n if kom else 0), "checkpoint_begin")
        self.done = False

    def wait(self):
        """Block until the snapshot is on the host, then write the file
        (writerest layout); returns the path."""
        from . import restart
        if not self.done:
            self.solve._check(self.solve.lib.cosahip_checkpoint_wait(self.solve.h), "checkpoint_wait")
            restart.write_restart(self.path, self.solve.case, self.q.array,
                                  self.mut.array if self.mut is not None else None, self.nl)
            self.q.free()
            if self.mut is not None:
                self.mut.free()
            self.done = True
        return self.path


class Solve:
    def __init__(self, case: Case, device: int = 0, q: Optional[np.ndarray] = None):
        self.case = case
        self.lib = _lib.load()
        self._cfg = to_struct(case.config)
        h = ctypes.c_void_p()
        rc = self.lib.cosahip_create(ctypes.byref(self._cfg), device, ctypes.byref(h))
        if rc:
            raise _lib.CosahipError(rc, "cosahip_create failed (see stderr)")
        self.h = h
        self.history: List[np.ndarray] = []
        try:
            self._define()
            if q is not None:
                self.upload("q", q, 1)
                if komega(case.config):
                    self.upload("mut", case.turbulent_viscosity(q, 1), 1)
        except Exception:
            self.close()
            raise

    # -- plumbing ------------------------------------------------------------
    def _check(self, rc: int, what: str):
        if rc:
            msg = self.lib.cosahip_last_error(self.h).decode()
            info = (ctypes.c_int32 * 8)()
            self.lib.cosahip_error_info(self.h, info)
            raise _lib.CosahipError(rc, f"{what}: {msg}", list(info))

    def _define(self):
        c = self.case
        for L in c.levels:
            keep = [np.ascontiguousarray(getattr(L, n), np.int32) for n in
                    ("imax", "jmax", "kmax", "nbcs", "bctopo", "cutopo")]
            cut = keep[5] if keep[5].size else np.zeros(21, np.int32)
            self._check(self.lib.cosahip_define_level(self.h, L.nl, L.nblocks, _lib.iptr(keep[0]),
                                                       _lib.iptr(keep[1]), _lib.iptr(keep[2]), _lib.iptr(keep[3]),
                                                       _lib.iptr(keep[4]), L.ncuts, _lib.iptr(cut)),
                        f"define_level {L.nl}")
            g = c.geom[L.nl - 1]
            for f in ("si", "sj", "sk", "vol", "xdot", "ydot", "zdot", "dist", "x", "y", "z"):
                if f in g:
                    self.upload(f, g[f], L.nl)
            if "xd" in g:
                for name, arr in zip(_lib.XD_NAMES, g["xd"]):
                    self.upload(name, arr, L.nl)
            q0 = c.freestream(L.nl)
            self.upload("q", q0, L.nl)
            if komega(c.config):
                # the driver's mut after init2fs (src/cosa.f90:229-232); coarse
                # levels get theirs on their first visit
                self.upload("mut", c.turbulent_viscosity(q0, L.nl) if L.nl == 1 else
                            np.zeros(c.size(L.nl, "p3s")), L.nl)
        d = c.hb_matrix()
        if d is not None:
            dd = np.ascontiguousarray(d.flatten(order="F"))
            self._check(self.lib.cosahip_set_hb_operator(self.h, d.shape[0], _lib.dptr(dd)), "set_hb_operator")

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.cosahip_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def size(self, field: str, nl: int = 1) -> int:
        return int(self.lib.cosahip_field_size(self.h, nl, _lib.FIELD_IDS[field]))

    def upload(self, field: str, arr: np.ndarray, nl: int = 1):
        a = np.ascontiguousarray(arr, np.float64)
        self._check(self.lib.cosahip_upload(self.h, nl, _lib.FIELD_IDS[field], _lib.dptr(a), a.size),
                    f"upload {field}")

    def download(self, field: str, nl: int = 1) -> np.ndarray:
        n = self.size(field, nl)
        a = np.empty(n)
        self._check(self.lib.cosahip_download(self.h, nl, _lib.FIELD_IDS[field], _lib.dptr(a), n),
                    f"download {field}")
        return a

    def synchronize(self):
        self._check(self.lib.cosahip_synchronize(self.h), "synchronize")

    # -- reference routines ----------------------------------------------------
    def call(self, name: str, *args):
        self._check(getattr(self.lib, "cosahip_" + name)(self.h, *args), name)

    def resg_rms(self, nl: int = 1) -> np.ndarray:
        r = np.zeros(8)
        self._check(self.lib.cosahip_resg_rms(self.h, nl, _lib.dptr(r)), "resg_rms")
        return r

    def smooth(self, nl: int, from_fine: bool, to_coarse: bool, niter: int, mgit: int):
        done = (ctypes.c_int32 * 1)()
        r = np.zeros(8)
        self._check(self.lib.cosahip_smooth(self.h, nl, int(from_fine), int(to_coarse), niter, mgit, done,
                                            _lib.dptr(r)), f"smooth level {nl}")
        return bool(done[0]), r

    def restrict(self, nl: int):
        self._check(self.lib.cosahip_restrict(self.h, nl), "restrict")

    def prolong(self, nl: int):
        self._check(self.lib.cosahip_prolong(self.h, nl), "prolong")

    # -- grid motion and metric set-up on the device ---------------------------
    def grid_metrics(self, nl: int = 1):
        """VOL, SI/SJ/SK and (viscous) the metric derivatives of level nl
        from its uploaded node coordinates (cosahip_grid_metrics)."""
        self._check(self.lib.cosahip_grid_metrics(self.h, nl), "grid_metrics")

    def grid_coarsen(self, nl: int):
        """Level nl + 1's node coordinates from level nl's (coarsen_b)."""
        self._check(self.lib.cosahip_grid_coarsen(self.h, nl), "grid_coarsen")

    def grid_pitch(self, nl: int = 1):
        """Node positions and velocities of every harmonic-balance snapshot
        of level nl, rigid pitching about the case's hinge (movegrid with
        pitch)."""
        m, c = self.case.motion, self.case.config
        pose = np.ascontiguousarray(self.case.pitch_poses(), np.float64)
        self._check(self.lib.cosahip_grid_pitch(self.h, nl, m["dtheta0"], c["omega"], m.get("xh", 0.25),
                                                m.get("yh", 0.0), _lib.dptr(pose)), "grid_pitch")

    def device_geometry(self, x=None, y=None, z=None):
        """The whole set-up of the driver on the device (src/cosa.f90:160-166,
        297-322): level-1 node coordinates (uploaded if given), then per level
        movegrid (moving cases), the metrics, and the next level's nodes by
        coarsening."""
        for f, a in (("x", x), ("y", y), ("z", z)):
            if a is not None:
                self.upload(f, a, 1)
        nlevel = self.case.config["nlevel"]
        for nl in range(1, nlevel + 1):
            if nl > 1:
                self.grid_coarsen(nl - 1)
            if self.case.config["moving"]:
                self.grid_pitch(nl)
            self.grid_metrics(nl)

    def force(self, nsurface: int = 1, include_inviscid: bool = True, lngth=(1.0, 1.0, 1.0),
              xyzmom=(0.25, 0.0, 0.0), nl: int = 1, noalloc: bool = False) -> np.ndarray:
        """Force and moment coefficients (force_b): array (nsurface, nsnap,
        12) of cl_p, cd_p, cz_p, cmx_p, cmy_p, cm_p, cl_v, cd_v, cz_v, cmx_v,
        cmy_v, cm_v.  Q must be conservative with its halos filled (between
        smooth calls it is).  noalloc: through cosahip_force_noalloc (the
        library's fixed scratch array instead of device allocations)."""
        ns = self.case.nsnap
        momr = np.ascontiguousarray(self.case.moment_centres(xyzmom), np.float64)
        lg = np.ascontiguousarray(lngth, np.float64)
        out = np.zeros(12 * ns * nsurface)
        fn = self.lib.cosahip_force_noalloc if noalloc else self.lib.cosahip_force
        self._check(fn(self.h, nl, nsurface, int(include_inviscid), _lib.dptr(lg), _lib.dptr(momr), _lib.dptr(out)),
                    "force_noalloc" if noalloc else "force")
        return out.reshape(nsurface, ns, 12)

    def write_restart(self, path: str, nl: int = 1):
        """The reference's restart file (writerest) of the resident state."""
        from . import restart
        kom = komega(self.case.config)
        restart.write_restart(path, self.case, self.download("q", nl), self.download("mut", nl) if kom else None, nl)

    def checkpoint_async(self, path: str, nl: int = 1) -> "Checkpoint":
        """Start a restart file of the state as of every launch enqueued so
        far, without stopping the solve (cosahip_checkpoint_begin): the
        device snapshot is stream-ordered and the device-to-host copy runs
        on the library's second stream while later cycles run.  The file is
        written by ``wait()`` on the returned object."""
        return Checkpoint(self, path, nl)

    def read_restart(self, path: str, nl: int = 1):
        """Load a restart file (readrest) into the resident state."""
        from . import restart
        q, mut = restart.read_restart(path, self.case, nl)
        self.upload("q", q, nl)
        if mut is not None:
            self.upload("mut", mut, nl)

    def set_precision(self, mode: int, tail_rms: float = 0.0):
        """0 FP64; 1 FP32 residual / FP64 state; 2 FP32 until the level-1
        density residual rms drops below tail_rms, FP64 after."""
        self._check(self.lib.cosahip_set_precision(self.h, int(mode), float(tail_rms)), "set_precision")

    def precision_in_use(self) -> int:
        return int(self.lib.cosahip_precision_in_use(self.h))

    # -- multigrid driver ------------------------------------------------------
    def cycle(self, mgit: int, nstart: int = 1, npre: int = 1, npost: int = 1, ncrs: int = 1) -> bool:
        """One pass of mg's cycle loop (src/cosa.f90:822-912): fine smooth,
        then the coarse levels of the V and the prolongations.  Appends the
        fine-level resrms(1:8) to ``history``; returns the convergence flag."""
        nlevel = self.case.config["nlevel"]
        niter = 1 if nlevel == 1 else (nstart if mgit == 1 else npost)
        done, r = self.smooth(1, False, nlevel > 1, niter, mgit)
        self.history.append(r)
        if done or nlevel == 1:
            return done
        self.restrict(1)
        niter = ncrs if nlevel == 2 else npre
        self.smooth(2, True, nlevel > 2, niter, mgit)
        if nlevel >= 3:
            self.restrict(2)
            self.smooth(3, True, False, ncrs, mgit)
            self.prolong(2)
            self.smooth(2, False, False, npost, mgit)
        self.prolong(1)
        return False

    def cycle_enqueue(self, mgit: int):
        """Put one V-cycle on the device without waiting (graph replay from
        the second cycle on); pair with cycle_wait."""
        self._check(self.lib.cosahip_cycle_enqueue(self.h, mgit), "cycle_enqueue")

    def cycle_wait(self) -> bool:
        """Wait for the enqueued cycle; appends its resrms(1:8) to history and
        returns the convergence flag."""
        done = (ctypes.c_int32 * 1)()
        r = np.zeros(8)
        self._check(self.lib.cosahip_cycle_wait(self.h, done, _lib.dptr(r)), "cycle_wait")
        self.history.append(r)
        return bool(done[0])

    def autotune(self, reps: int = 3) -> Dict[str, object]:
        """Time the interchangeable kernel arrangements on the current state
        and keep the fastest (cosahip_autotune); results are unchanged."""
        ch = (ctypes.c_int32 * 8)()
        self._check(self.lib.cosahip_autotune(self.h, reps, ch), "autotune")
        names = {0: None, 1: "lds", 2: "per_variable", 3: "line_resident"}
        rnames = {-1: None, 0: "generic_fused", 1: "tile31x7", 2: "tile31x7_register_window", 3: "tile31x15",
                  4: "tile31x7_128vgpr", 5: "tile31x7_shared_slopes", 6: "tile31x15_shared_slopes",
                  7: "tile31x7_shared_slopes_128vgpr", 8: "tile32x7_128vgpr", 9: "tile32x7_shared_slopes"}
        vnames = {0: "march", 1: "stored_faces", 2: "tile_pass"}
        return {"cirs_i_sweep": [names[ch[m]] for m in range(3)], "viscous": vnames[ch[3]],
                "residual": [rnames.get(ch[4 + m]) for m in range(3)]}

    # -- blocks as independent solves (serving batches) -----------------------
    def set_block_freestream(self, block: int, machfs: float, alpha_deg: float, beta_deg: float = 0.0):
        """Freestream of block `block` (1-based) for bc_far0 and its forces;
        angles in degrees converted as the configuration converts them."""
        import math
        self._check(self.lib.cosahip_set_block_freestream(self.h, block, machfs, alpha_deg * math.pi / 180.0,
                                                          beta_deg * math.pi / 180.0), "set_block_freestream")

    def set_block_rms(self, on: bool = True):
        self._check(self.lib.cosahip_set_block_rms(self.h, int(on)), "set_block_rms")

    def block_rms(self) -> np.ndarray:
        """resrms(1:8) per block of the last level-1 norms, (nblocks, 8)."""
        nb = self.case.levels[0].nblocks
        out = np.zeros(8 * nb)
        self._check(self.lib.cosahip_block_rms(self.h, _lib.dptr(out)), "block_rms")
        return out.reshape(nb, 8)

    def force_blocks(self, nsurface: int = 1, include_inviscid: bool = True, lngth=(1.0, 1.0, 1.0),
                     xyzmom=(0.25, 0.0, 0.0), nl: int = 1) -> np.ndarray:
        """cosahip_force per block with the block's own freestream:
        (nblocks, nsurface, nsnap, 12)."""
        nb, ns = self.case.levels[nl - 1].nblocks, self.case.nsnap
        lg = np.ascontiguousarray(lngth, np.float64)
        momr = np.ascontiguousarray(self.case.moment_centres(xyzmom), np.float64)
        out = np.zeros(nb * nsurface * ns * 12)
        self._check(self.lib.cosahip_force_blocks(self.h, nl, nsurface, int(include_inviscid), _lib.dptr(lg),
                                                  _lib.dptr(momr), _lib.dptr(out)), "force_blocks")
        return out.reshape(nb, nsurface, ns, 12)

    def cycle_ready(self) -> bool:
        """True when the enqueued cycle has finished (cycle_wait will not block)."""
        r = self.lib.cosahip_cycle_ready(self.h)
        if r < 0:
            self._check(r, "cycle_ready")
        return r == 1

    def cycle_async(self, mgit: int) -> bool:
        self.cycle_enqueue(mgit)
        return self.cycle_wait()

    def mg(self, ncycle: Optional[int] = None, nstart: int = 1, npre: int = 1, npost: int = 1, ncrs: int = 1,
           on_cycle=None, hist=None) -> int:
        """The V-cycle schedule of mg (src/cosa.f90:806-930).  Records
        resrms(1:8) of every cycle in ``history`` and, given a
        cosa_amd.history.HistoryWriter, the reference's hist.dat lines;
        returns cycles run."""
        nlevel = self.case.config["nlevel"]
        if nlevel > 3:
            raise ValueError("the reference's V-cycle has at most 3 levels")
        ncycle = ncycle or self.case.config["ncycle"]
        for mgit in range(1, ncycle + 1):
            done = self.cycle(mgit, nstart, npre, npost, ncrs)
            if hist is not None:
                hist.record(mgit, self.history[-1], done)
            if on_cycle:
                on_cycle(mgit, self.history[-1])
            if done:
                return mgit
        if nlevel > 1:
            self.smooth(1, False, False, npost, ncycle)
        return ncycle

    # -- dual time stepping ----------------------------------------------------
    def set_dual_time(self, dt: float, prop_start: bool = False):
        """Switch the solve to dual time stepping (src/cosa.f90:437-672):
        physical step dt; prop_start: two-state start (upload q2 first),
        else backward Euler on the first step.  config rkimp selects the
        implicit physical-time term of the RK stages."""
        self._check(self.lib.cosahip_set_dual_time(self.h, float(dt), int(bool(prop_start))), "set_dual_time")

    def dual_step(self, itime: int, ncycle: Optional[int] = None, nstart: int = 1, npre: int = 1, npost: int = 1,
                  ncrs: int = 1, hist=None) -> int:
        """One physical step (src/cosa.f90:640-652): q1 = q and rhs =
        rhs_uns(q1, q2), the pseudo-time cycles of mg, then q2 = q1.
        Returns the cycles run."""
        self._check(self.lib.cosahip_dual_begin(self.h, int(itime)), "dual_begin")
        n = self.mg(ncycle, nstart, npre, npost, ncrs, hist=hist)
        self._check(self.lib.cosahip_dual_end(self.h), "dual_end")
        return n

    # -- instrumentation -------------------------------------------------------
    def timing(self, on: bool):
        self._check(self.lib.cosahip_timing_enable(self.h, int(on)), "timing_enable")

    def timing_reset(self):
        self._check(self.lib.cosahip_timing_reset(self.h), "timing_reset")

    def timing_read(self) -> Dict[str, Dict[str, float]]:
        out = {}
        for f, name in enumerate(_lib.FAMILIES):
            ms = ctypes.c_double()
            n = ctypes.c_int64()
            it = ctypes.c_int64()
            self._check(self.lib.cosahip_timing_read(self.h, f, ctypes.byref(ms), ctypes.byref(n), ctypes.byref(it)),
                        "timing_read")
            out[name] = {"ms": ms.value, "launches": n.value, "items": it.value}
        return out

    def cell_updates_per_cycle(self, nstart: int = 1, npre: int = 1, npost: int = 1, ncrs: int = 1,
                               first: bool = False) -> int:
        """Interior (cell, snapshot) RK-stage updates of one mg cycle, every
        level counted (the residual-only pass before restriction excluded)."""
        c = self.case
        ns = c.nsnap
        st = c.config["nstage"]
        nlevel = c.config["nlevel"]
        n1 = c.ncell(1) * ns * st
        tot = n1 * (nstart if first else (npost if nlevel > 1 else 1))
        if nlevel >= 2:
            n2 = c.ncell(2) * ns * st
            tot += n2 * (ncrs if nlevel == 2 else npre)
        if nlevel >= 3:
            tot += c.ncell(3) * ns * st * ncrs + c.ncell(2) * ns * st * npost
        return tot
