/*
 * cosahip.h -- C ABI of libcosahip, the gfx950 (MI355X) implementation of
 * COSA's pseudo-time residual/update loop.
 *
 * The caller is COSA's Fortran driver (program cosa -> mg -> smooth), bound
 * through iso_c_binding, or any host that mirrors it (the Python package
 * cosa_amd does).  Every entry point below replaces one reference routine
 * (cited as file:line into LANCASTER-CFD/COSA src/) and keeps its meaning:
 * same level numbering (nl = 1 is the finest grid), same 1-based block and
 * BC/cut topology records, same flat array layout as the reference's
 * Cray-pointer allocations (src/input.f90:4878-5131, offsets of setoffs
 * src/input.f90:5502-5582), so host<->device transfers are single copies.
 *
 * Conventions
 *   - All functions return COSAHIP_OK (0) or a positive error code; the text
 *     of the last error is available from cosahip_last_error().  The
 *     reference writes a message and stops in the same situations.
 *   - The library owns device memory.  State stays resident in HBM between
 *     calls; host arrays are touched only by upload/download.
 *   - One host thread per solve.  Different solves may share a device.
 *   - Floating point is IEEE binary64 (Fortran real(8) with the reference's
 *     -fdefault-real-8 promotion); integers are 32-bit (cosa_int).
 */
#ifndef COSAHIP_H
#define COSAHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define COSAHIP_ABI_VERSION 1

/* ---- error codes ------------------------------------------------------ */
enum {
  COSAHIP_OK = 0,
  COSAHIP_ERR_ARG = 1,         /* bad argument / unsupported option        */
  COSAHIP_ERR_STATE = 2,       /* call out of order (level not defined...) */
  COSAHIP_ERR_DEVICE = 3,      /* HIP runtime failure                      */
  COSAHIP_ERR_NOMEM = 4,       /* device allocation failed                 */
  COSAHIP_ERR_POSITIVITY = 5,  /* prtest: p <= qmin(5) or rho <= qmin(1)   */
  COSAHIP_ERR_UNSUPPORTED = 6  /* option of the reference not built here   */
};

/* ---- solver configuration ---------------------------------------------
 * Mirror of the common_variables members the hot path reads
 * (src/common_variables.f90:19-65, src/cosa_variables.f90:31-36).
 * Logicals are int32 (0/1).  Angles in radians (the reference converts the
 * degrees of input.dat on input).  Field order is fixed: the Fortran binding
 * declares the same members in a bind(C) derived type.                    */
typedef struct cosahip_config {
  int32_t npde;        /* 5 Euler/laminar, 7 k-omega                       */
  int32_t nharms;      /* harmonic-balance harmonics N (0: steady)         */
  int32_t hbmove;      /* 1: metrics carry the snapshot index              */
  int32_t lmet;        /* metric rows per face: 4 static, 7 moving         */
  int32_t nlevel;      /* multigrid levels (reference max 3)               */
  int32_t nl_crs;      /* first-order MUSCL on levels nl >= nl_crs         */
  int32_t nstage;      /* RK stages (4)                                    */
  int32_t irsop;       /* 0 no_irs, 1 cirs_c, 2 cirs_v1, 3 cirs_v2         */
  int32_t cutcirs;     /* 0 cirs, 1 ccirs: cut-coupled CIRS (src/cosa.f90:2035-2415, 5238-5773) */
  int32_t ilimf;       /* flow limiter 0..4                                */
  int32_t ilimt;       /* turbulence limiter 0..4                          */
  int32_t iupdt;       /* local time step refresh period                   */
  int32_t iprint;      /* residual print period                            */
  int32_t ncycle;      /* number of MG cycles (lmax)                       */
  int32_t kom, kom_bsl, kom_sst, viscous;
  int32_t moving, harbal, foturb;
  int32_t efx_setup_f, efx_all_f, eig_limit_f, eig_lim_all_f;
  int32_t efx_setup_t, efx_all_t, eig_limit_t, eig_lim_all_t;
  int32_t entfx_f_w[3], entfx_t_w[3];
  int32_t bln_prol;    /* 1: bilinear prolongation, 0: linear          */
  int32_t ho_rest;     /* 1: 64-weight residual restriction             */
  int32_t rkimp;       /* 1: HB implicit RK (rkim): the spectral term solved
                          per cell with an LU of (2N+1)x(2N+1) (update_bhb) */
  double gamma, machfs, alpha, beta;       /* alpha, beta in radians       */
  double rhoinf, pinf, ainf;
  double cfl, cflt, rat, ratt, psirs, rkap;
  double epslimf, epslimt;
  double cntrpy_f, cntrpy_t, entfxctf_f, entfxctf_t;
  double eig_cutoff_f, eig_cutoff_t;
  double omega;        /* HB angular frequency                             */
  double toler;        /* convergence: max(resrms(1:5)) < toler            */
  double qmin[7];
  double alphas[5];    /* RK stage coefficients (ar4 literals)             */
  double tkefar, omefar;
  /* ---- viscous flow and Wilcox k-omega (src/common_variables.f90:23-65,
   *      defaults and parsing src/input.f90:2501-2525, 2695-2770, 3700) ---- */
  int32_t turcmp;        /* 1: compressibility term in div(u) of k-omega    */
  int32_t lim_prodtke;   /* 1: limit k production (prdlim)                  */
  int32_t lim_prodomega; /* 1: limit omega production too                   */
  int32_t wallwilc;      /* type-15 wall omega: 1 Wilcox rough wall, 0 Menter */
  int32_t ifixq;         /* 1: floor k, omega at qmin(6:7) after updates    */
  int32_t prd_all_lvls;  /* 1: production on every level (0 only here)     */
  int32_t lim_corr[3];   /* k-omega prolongation limiters (all 0 here)      */
  int32_t pad1;
  double reyno;          /* Reynolds number                                  */
  double pranl, prant;   /* laminar / turbulent Prandtl numbers              */
  double stemp;          /* Sutherland temperature / reference temperature   */
  double cdff;           /* weight of the viscous spectral radii             */
  double epst, cmu, sigk, sigw, bkw, bstrkw, gkw;   /* k-omega constants     */
  double prdlim;         /* production limiter factor                       */
  double roughk;         /* wall roughness (Wilcox wall, 0: smooth)          */
  /* ---- CFL ramping of k-omega runs (bdeltat, src/cosa.f90:4760-4820;
   *      input src/input.f90:3304-3345) ---- */
  int32_t ramping;       /* 0 none, 1 ramping1 (cubic, refreshed at cycle 1 and
                            every 100th), 2 ramping2 (piecewise linear)    */
  int32_t ncycr[4];      /* ncycr(0:3): cycles where the ramp segments end   */
  int32_t pad2;
  double vcfli[3];       /* CFL at the marks (with IRS), vcfli(1:3)           */
  double vcflit[3];      /* turbulence CFL at the marks, vcflit(1:3)          */
} cosahip_config;

/* ---- device fields (per level; reference layout and extents) ----------
 *   P3 family: (-1:imax+1,-1:jmax+1,-1:kmax+1, ncomp, 0:2N)   off_p3
 *   P2 family: (lmet, 0:imax+1, 0:jmax+1, 0:kmax+1, 0:2N*hbmove) off_p2
 *   P1 family: (0:imax, 0:jmax, 0:kmax [, 0:2N])              off_p1     */
enum {
  COSAHIP_Q = 0,        /* conservative state (P3, npde)                   */
  COSAHIP_QOLD = 1,     /* state at iteration start (P3, npde)             */
  COSAHIP_RES = 2,      /* residual (P3, npde)                             */
  COSAHIP_RHS = 3,      /* FAS forcing (P3, npde)                          */
  COSAHIP_SI = 4,       /* xi-face metrics (P2, lmet)                      */
  COSAHIP_SJ = 5,
  COSAHIP_SK = 6,
  COSAHIP_VOL = 7,      /* cell volumes (P1, no snapshot index)            */
  COSAHIP_DTVOL = 8,    /* dt/vol (P1, 0:2N)                               */
  COSAHIP_BETAXI = 9,   /* CIRS coefficients (P1, 0:2N)                    */
  COSAHIP_BETAETA = 10,
  COSAHIP_BETAZETA = 11,
  COSAHIP_XDOT = 12,    /* node grid velocity (P2 scalar, 0:2N*hbmove); lambdab
                           reads node (i,j,k) for cell (i,j,k)              */
  COSAHIP_YDOT = 13,
  COSAHIP_ZDOT = 14,
  COSAHIP_MUT = 15,     /* turbulent viscosity (P3 scalar, 0:2N)           */
  COSAHIP_DIST = 16,    /* cell-centre wall distance (P1, no snapshot index) */
  COSAHIP_XIDERI = 17,  /* metric derivatives (3, imax, jmax, kmax, 0:2N*hbmove),
                           off_0 offsets: xi/eta/zeta derivatives on i, j, k
                           faces in the order xideri, etaderi, zetaderi,
                           xiderj, etaderj, zetaderj, xiderk, etaderk, zetaderk */
  COSAHIP_ETADERI = 18,
  COSAHIP_ZETADERI = 19,
  COSAHIP_XIDERJ = 20,
  COSAHIP_ETADERJ = 21,
  COSAHIP_ZETADERJ = 22,
  COSAHIP_XIDERK = 23,
  COSAHIP_ETADERK = 24,
  COSAHIP_ZETADERK = 25,
  COSAHIP_PRD = 26,     /* k-omega production P_d (P3 scalar, 0:2N)        */
  COSAHIP_DIVVEL = 27,  /* velocity divergence (P3 scalar, 0:2N)           */
  COSAHIP_DELPLS = 28,  /* 2/3 max(div u, 0) (P3 scalar, 0:2N)             */
  COSAHIP_DTVOLT = 29,  /* k-omega dt/vol (P1, 0:2N)                       */
  COSAHIP_X = 30,       /* node coordinates (P2 scalar, 0:2N*hbmove), halo nodes
                           as cordex_b fills them (src/grid.f90:23-78); read
                           only by cosahip_force                            */
  COSAHIP_Y = 31,
  COSAHIP_Z = 32,
  COSAHIP_TTRM = 33,    /* BSL/SST blended coefficients (P3, 3: cross diffusion,
                           gamma, beta; 0:2N), refreshed by every residual;
                           download only                                     */
  COSAHIP_Q1 = 34,      /* dual time stepping, finest level: q^n (P3, npde);
                           allocated by cosahip_set_dual_time (else empty)  */
  COSAHIP_Q2 = 35,      /* dual time stepping, finest level: q^(n-1)       */
  COSAHIP_NFIELDS = 36
};

typedef struct cosahip_solve cosahip_solve;

/* ---- lifetime ---------------------------------------------------------- */

/* Replaces alloc (src/input.f90:4878-5131) for the hot-path arrays: copies
 * the configuration, selects device `device` (-1: current).                */
int cosahip_create(const cosahip_config* cfg, int device, cosahip_solve** out);
int cosahip_destroy(cosahip_solve* s);
const char* cosahip_last_error(const cosahip_solve* s);
/* On COSAHIP_ERR_POSITIVITY: info = {level, block(1-based), i, j, k, iter,
 * kind (1: pressure, 2: density), snapshot}.  Mirrors the message of
 * bprtest (src/cosa.f90:960-1120).                                         */
int cosahip_error_info(const cosahip_solve* s, int32_t info[8]);

/* HB time-derivative operator dhb(0:2N,0:2N) as built by mathb
 * (src/input.f90:5206-5282); Fortran column-major, nsnap = 2N+1.          */
int cosahip_set_hb_operator(cosahip_solve* s, int nsnap, const double* dhb);

/* Dual time stepping (`unsteady dual dt ... rkex/rkim` of input.dat,
 * src/input.f90:2801-2874; the physical-time loop src/cosa.f90:437-672).
 * cosahip_set_dual_time switches the solve to it once level 1 is defined:
 * dt the physical time step, dual_prop_start 1 for the two-state start
 * (restart holding q^(n-1)), 0 for a backward-Euler first step.  It
 * allocates COSAHIP_Q1/Q2 on level 1 (zeroed; upload Q2 for a two-state
 * start).  cfg.rkimp then selects the implicit physical-time term of
 * update_b / lim_res_ome_b (src/update.f90:649-762, 2753-2794) instead of
 * harmonic balance's rkim.  Harmonic balance excludes it.
 * Per physical step, around the multigrid cycles of mg():
 *   cosahip_dual_begin(s, itime): q1 = q, rhs = rhs_uns(q1, q2) on level 1
 *     (src/cosa.f90:640-650, src/src.f90:453-525);
 *   cosahip_dual_end(s): q2 = q1 (src/cosa.f90:651-652).
 * Inside smooth the reference's dualt branches run: src_uns on every level
 * (without it on one level without IRS when rkimp), the finest residual
 * minus rhs, and the rkimp norm (src/cosa.f90:1910-1941, 4534-4549).      */
int cosahip_set_dual_time(cosahip_solve* s, double dt, int dual_prop_start);
int cosahip_dual_begin(cosahip_solve* s, int itime);
int cosahip_dual_end(cosahip_solve* s);

/* Level topology: block node counts i_imax/j_jmax/k_kmax of level nl
 * (src/input.f90:4393-4434), BC records bctopo(10,nbcs(b)) of setdim_bbc
 * (src/input.f90:5155-5203) concatenated over blocks, cut records
 * cutopo(21,ncuts) of setdimcut (src/input.f90:5285-5344).  Blocks are
 * numbered 1..nblocks as in the reference.                                 */
int cosahip_define_level(cosahip_solve* s, int nl, int nblocks,
                         const int32_t* imax, const int32_t* jmax,
                         const int32_t* kmax, const int32_t* nbcs,
                         const int32_t* bctopo, int ncuts,
                         const int32_t* cutopo);

/* Number of doubles of `field` on level nl (whole level, all blocks).     */
int64_t cosahip_field_size(const cosahip_solve* s, int nl, int field);
int cosahip_upload(cosahip_solve* s, int nl, int field, const double* host,
                   int64_t count);
int cosahip_download(cosahip_solve* s, int nl, int field, double* host,
                     int64_t count);
/* Checkpoint of the resident state while the solve goes on (the data of
 * writerest, src/output.f90:3481-3566; the asynchronous copy has no
 * counterpart in the reference): a device snapshot of level nl's q (and,
 * k-omega, mut) is taken in stream order -- after every launch enqueued so
 * far, before any enqueued later -- then copied to host_q / host_mut on a
 * second stream of the library while the solve's stream goes on.  counts
 * must equal cosahip_field_size; host_mut null and nmut 0 without
 * k-omega.  Host buffers from cosahip_host_alloc (pinned) let the copy
 * overlap the solve; any other host memory works but may serialise.
 * cosahip_checkpoint_wait blocks until the host buffers hold the snapshot;
 * one checkpoint per solve at a time.                                     */
int cosahip_checkpoint_begin(cosahip_solve* s, int nl, double* host_q, int64_t nq, double* host_mut,
                             int64_t nmut);
int cosahip_checkpoint_wait(cosahip_solve* s);
/* Pinned host memory for checkpoint buffers (hipHostMalloc / hipHostFree). */
int cosahip_host_alloc(int64_t bytes, void** out);
int cosahip_host_free(void* p);
int cosahip_synchronize(cosahip_solve* s);

/* ---- reference routines, whole level, in place on device state ------- */

/* cons2prim(nl,0,q,work): interior only (src/update.f90:2112-2188).
 * `field` is COSAHIP_Q or COSAHIP_QOLD.                                     */
int cosahip_cons2prim(cosahip_solve* s, int nl, int field);
/* prim2cons: interior + face halo strips (src/update.f90:1851-2084).      */
int cosahip_prim2cons(cosahip_solve* s, int nl, int field);
/* bc (src/bc.f90:2-121): BC patches of every block, in input order.       */
int cosahip_bc(cosahip_solve* s, int nl, int field);
/* cutman_q (src/bc.f90:1801-1891): 1:1 cut halo copies.                   */
int cosahip_cutman_q(cosahip_solve* s, int nl, int field);
/* q_edges + q_corns (src/flux.f90:3889-4263, 3132-3471) on field.        */
int cosahip_edges_corners(cosahip_solve* s, int nl, int field);
/* prtest (src/cosa.f90:933-1120) on Q: COSAHIP_ERR_POSITIVITY on failure. */
int cosahip_prtest(cosahip_solve* s, int nl, int iter);
/* lambda + deltat (src/cosa.f90:1214-1591, 4183-4237, 4701-5203).         */
int cosahip_lambda_deltat(cosahip_solve* s, int nl);
/* zero(res) on 0..imax (src/util.f90:238-369), then for idir=1..3:
 * muscl + roflux + resid (src/flux.f90:1183-1812, 2-387, 3472-3888;
 * src/cosa.f90:1123-1211) and, viscous, q_edges + q_face + q_der + rtst +
 * vflux + resid (src/flux.f90:3889-4263, 2391-2801, 1813-2390, 2802-2948,
 * 831-1182); k-omega adds the sources q_face + q_der_c + turprod + srckw
 * (src/cosa.f90:1869-1902; src/src.f90:2-272).  Q must hold primitive
 * variables with halos.                                                    */
int cosahip_residual(cosahip_solve* s, int nl);
/* src_hb (src/src.f90:530-613): res += omega*vol*D*q, q conservative.     */
int cosahip_src_hb(cosahip_solve* s, int nl);
/* src_uns (src/src.f90:381-450): res += q*vol/dt on the first step of a
 * first-order start, else 3*q*vol/2/dt; q conservative (dual time).       */
int cosahip_src_uns(cosahip_solve* s, int nl);
/* resg_rms / resg_rms_hb (src/cosa.f90:4458-4697) -> resrms(1:npde+1).    */
int cosahip_resg_rms(cosahip_solve* s, int nl, double* resrms);
/* add_array(..., off_l=-1, fact) on the interior (src/util.f90:137-236):
 * dst += fact*src, fields RES/RHS.                                        */
int cosahip_add_array(cosahip_solve* s, int nl, int dst, int src, double fact);
/* copy_array interior (src/util.f90:2-135): dst = src.                   */
int cosahip_copy_array(cosahip_solve* s, int nl, int dst, int src);
/* cirs (src/cosa.f90:5206-5623) for cutcirs = 0; for cutcirs = 1 ccirs
 * (src/cosa.f90:2035-2415) with scalef/scaleb (5626-5773): two passes with
 * the cut-exchanged halo residuals as Dirichlet data.                     */
int cosahip_cirs(cosahip_solve* s, int nl);
/* update (src/update.f90:2-100, 568-762, 1301-1408), explicit RK; with
 * k-omega the point-implicit rho*omega, rho*k update first.               */
int cosahip_update(cosahip_solve* s, int nl, double ar);
/* tur_vis + cutman_mut (src/update.f90:2272-2337; src/bc.f90:2713-2800):
 * mut of level nl from `field` (Q or QOLD, primitive or conservative as the
 * caller holds it), k-omega only.                                          */
int cosahip_tur_vis(cosahip_solve* s, int nl, int field);
/* fixq (src/update.f90:2191-2269): k, omega floors qmin(6:7) on Q.        */
int cosahip_fixq(cosahip_solve* s, int nl);
/* restriction-stage limits of k-omega (src/cosa.f90:1965-1973):
 * lim_prod_mg (src/src.f90:674-757, when nlevel > 1, nl < nlevel and
 * lim_prodtke) then lim_res_ome (src/update.f90:2527-2835) with stage
 * coefficient ar; Q conservative.                                          */
int cosahip_turb_limits(cosahip_solve* s, int nl, double ar);

/* ---- drivers --------------------------------------------------------- */

/* smooth (src/cosa.f90:1594-2032): niter pseudo-time iterations on level
 * nl (one more residual-only pass when to_coarse).  mgit is the current MG
 * cycle (controls the RMS print cycles).  On return *done is 1 when
 * max(resrms(1:5)) < toler; resrms holds the last RMS evaluated on nl=1.  */
int cosahip_smooth(cosahip_solve* s, int nl, int from_fine, int to_coarse,
                   int niter, int mgit, int32_t* done, double* resrms);
/* restrict (src/cosa.f90:2672-2754) from nl to nl+1.                     */
int cosahip_restrict(cosahip_solve* s, int nl);
/* prolong (src/cosa.f90:3167-3249) correction from nl+1 onto nl.         */
int cosahip_prolong(cosahip_solve* s, int nl);

/* ---- asynchronous cycles (no counterpart: the reference's mg loop,
 * src/cosa.f90:822-912, runs on the host one routine at a time) ---------
 * cosahip_cycle_enqueue puts one whole V-cycle of mg (nstart = npre = npost
 * = ncrs = 1: fine smooth, restrict, coarse smooths, prolongs) on the
 * solve's stream and returns without waiting.  From the second cycle on the
 * launch sequence is captured once as a HIP graph and replayed (recaptured
 * when the residual precision changes; not while timing is enabled;
 * COSAHIP_GRAPH=0 at create time disables it).  cosahip_cycle_wait waits
 * for it and returns the cycle's level-1 resrms(1:8), *done = max(resrms(1:5))
 * < toler, and COSAHIP_ERR_POSITIVITY (cosahip_error_info) when prtest
 * failed on any level.  Differences from calling smooth/restrict/prolong:
 * the norms are evaluated every cycle, the convergence test and the mode-2
 * precision switch are applied after the whole cycle.  Solves on different
 * streams run concurrently.                                              */
int cosahip_cycle_enqueue(cosahip_solve* s, int mgit);
int cosahip_cycle_wait(cosahip_solve* s, int32_t* done, double* resrms);
/* 1 when the enqueued cycle has finished on the device (cosahip_cycle_wait
 * then returns without blocking), 0 while it runs, < 0 on error; no
 * counterpart in the reference (a serving loop polls its solves with it). */
int cosahip_cycle_ready(cosahip_solve* s);

/* ---- autotuner (no counterpart in the reference) -------------------------
 * Times the interchangeable kernel arrangements on the solve's current
 * state and keeps the fastest: per level the CIRS i-sweep (one thread per
 * line with LDS-staged chunks, or one thread per line and variable), and
 * for viscous runs the viscous face fluxes (stored then differenced, flux
 * and difference in one line march, or the fluxes of all three directions
 * stored and differenced in one tile pass with the inviscid fluxes).  The
 * arrangements give
 * bit-identical results; the state is not changed (timing runs on a copy of
 * the residual).  choice[0..2]: CIRS i-sweep of levels 1..3 (1 one thread
 * per line with LDS-staged chunks, 2 one thread per line and variable,
 * 3 whole lines resident in LDS -- irsop 1 only --, 0 level absent),
 * choice[3]: 1 stored face fluxes, 0 march, 2 tile pass; choice[4..6]: inviscid
 * residual kernel of levels 1..3 (0 generic fused, 1 specialised 31x7
 * tiles with the k-window from LDS/L2, 2 the same with a register
 * k-window, 3 31x15 tiles, 4 as 1 held to 128 VGPRs, 5 / 6 31x7 / 31x15
 * tiles with the van Albada slope evaluated once per cell, 7 as 5 held to
 * 128 VGPRs, 8 as 4 on 32x7 tiles, 9 as 5 on 32x7 tiles; -1 level absent);
 * choice[7] is reserved (0).  choice holds 8 entries.
 * Environment switches at create time take precedence:
 * COSAHIP_RESIDUAL=fused|tile7|tile7reg|tile15|tile7w4|sva7|sva15|sva7w4|tile32w4|sva32|split
 * forces the
 * inviscid residual kernel, COSAHIP_CIRS_I=lds|iv|line forces the i-sweep, COSAHIP_CIRS_IL_LDS sets
 * the LDS bytes per work-group of the line-resident sweep (4096..65536,
 * default 24576), COSAHIP_VISCOUS=march|fused|tile the viscous arrangement,
 * COSAHIP_CIRS_JK=chunk4|chunk8|plain the irsop 1 j/k sweeps (default chunk4). */
int cosahip_autotune(cosahip_solve* s, int reps, int32_t* choice);

/* ---- blocks as independent solves (no counterpart in the reference) -----
 * For serving, requests of one grid run as the blocks of one solve: every
 * block has its own freestream (bc_far0's qfst, set from the Mach number and
 * angles as the configuration's), its own norms and its own force
 * coefficients.  Blocks without cuts between them evolve exactly as the
 * same grid solved alone.  Block numbers are 1-based.
 *   cosahip_set_block_freestream  between cycles
 *   cosahip_set_block_rms(on)     keep every tile's norm partials of level 1
 *   cosahip_block_rms(out)        resrms(1:8) per block of the last norm
 *                                 evaluation, out[8 * nblocks]
 *   cosahip_force_blocks(...)     cosahip_force per block with the block's
 *                                 Mach number and angles, coef[nblocks][nsurface][2N+1][12] */
int cosahip_set_block_freestream(cosahip_solve* s, int block, double machfs, double alpha, double beta);
int cosahip_set_block_rms(cosahip_solve* s, int on);
int cosahip_block_rms(cosahip_solve* s, double* out);
int cosahip_force_blocks(cosahip_solve* s, int nl, int nsurface, int include_inviscid, const double* lngth,
                         const double* momr, double* coef);

/* ---- grid motion and metric set-up on the device ---------------------
 * The producers of the metric inputs, computed from the node coordinates
 * X, Y, Z (and XDOT, YDOT, ZDOT of moving grids) already uploaded to the
 * level, with the reference routines' arithmetic (the same bits as COSA's
 * CPU set-up), so a solve can be admitted from its grid alone.
 *
 * cosahip_grid_metrics: the per-level sequence of the driver
 * (src/cosa.f90:304-322): vols/bvols (VOL interior, src/grid.f90:2162-2258),
 * imetrix/imetrix_b (SI, SJ, SK incl. face velocities when moving,
 * src/grid.f90:848-1253); when viscous or nlevel > 1 volex_b and cutman_v
 * (VOL face halos, src/grid.f90:83-141, 525-838); when nl < nlevel and
 * the prolongation is linear vol_edges (src/grid.f90:3584-3826); when
 * viscous coordex/cordex_b and cutman_x (X/Y/Z halo nodes, src/grid.f90:
 * 2-78, 144-524) and vmetrix_b (the nine metric derivatives,
 * src/grid.f90:3150-3583).  The rotating-frame face moments (hawt) are not
 * built.  COSAHIP_ERR_ARG when a face area is not positive (imetrix_b
 * stops there); a non-positive volume is not an error (bvols only reports
 * it).  Drops a captured cycle graph.                                     */
int cosahip_grid_metrics(cosahip_solve* s, int nl);
/* coarsen (src/cosa.f90:2575-2669): level nl + 1's X, Y, Z (nodes 1..imax
 * etc. of snapshot 0) as every second node of level nl's.                 */
int cosahip_grid_coarsen(cosahip_solve* s, int nl);
/* movegrid with pitch (src/grid.f90:1783-1964): X, Y, Z, XDOT, YDOT, ZDOT
 * of every metric snapshot n of level nl from the base pose (snapshot 0 of
 * X, Y, Z at the first call, kept as x0), pitching about (xh, yh) with unit
 * displacement scales dx = dy = 1 (the reference's mesh.dat for rigid
 * motion).  pose[3n + 0..2] = dcos(dtheta(n)), dsin(dtheta(n)),
 * dcos(omegas * zeit(n)) with dtheta(n) = dtheta0 dsin(omegas zeit(n))
 * (src/grid.f90:4540), evaluated by the caller as the reference does.
 * Moving configurations only.                                             */
int cosahip_grid_pitch(cosahip_solve* s, int nl, double dtheta0, double omegas, double xh, double yh,
                       const double* pose);

/* ---- forces ------------------------------------------------------------
 * force_b (src/output.f90:4388-4886) summed over the level's blocks as
 * force does (src/output.f90:4290-4330), aircraft conventions: pressure and
 * viscous force and moment coefficients of surfaces s = 1..nsurface, where
 * surface s is the union of the BC patches of type 1500 + s - 1 (viscous
 * walls; the viscous part only when the run is viscous).  With
 * include_inviscid != 0 the type-14 walls also count as surface 1 (pressure
 * part only): an extension without counterpart in the reference, whose
 * force_b skips them and so reports zero forces for Euler runs.
 * Reads Q (conservative, halos filled as after a completed stage), SI/SJ/SK,
 * X/Y/Z, XDOT/YDOT/ZDOT (moving) and XIDERI/ETADERJ/ZETADERK (viscous).
 * lngth[3]: the reference lengths lngth(1:3); momr[3 * (2N+1)]: the moment
 * centre (xmomr, ymomr, zmomr) of each snapshot as force_b derives it from
 * xmom and the motion.  coef(12, 0:2N, nsurface) receives cl_p, cd_p, cz_p,
 * cmx_p, cmy_p, cm_p, cl_v, cd_v, cz_v, cmx_v, cmy_v, cm_v.                */
int cosahip_force(cosahip_solve* s, int nl, int nsurface, int include_inviscid,
                  const double* lngth, const double* momr, double* coef);
/* cosahip_force without device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, kernels and results as cosahip_force, but the wall-face
 * runs, per-face terms, per-block sums, coefficients and geometry live in a
 * fixed device array of the library, 8 MiB (1,048,576 doubles) per GPU,
 * held until the process ends; it calls no hipMalloc / hipFree.  The array
 * needs 12 * faces * (2N+1) + 12 * blocks * (2N+1) * nsurface + 12 * (2N+1)
 * * nsurface + 13 + 3 * (2N+1) doubles plus the runs (each part rounded up
 * to 32 doubles); a larger request returns COSAHIP_ERR_ARG.  Every solve on
 * one GPU shares that GPU's array: calls that overlap on one GPU (solves on
 * other host threads) race on it.                                         */
int cosahip_force_noalloc(cosahip_solve* s, int nl, int nsurface, int include_inviscid,
                          const double* lngth, const double* momr, double* coef);

/* ---- precision (no counterpart in the reference, which computes in
 * real(kind=cosa_real), FP64, throughout: src/cosa_precision.f90:17) -----
 * mode 0: FP64 everywhere (default, the parity mode).
 * mode 1: the residual (MUSCL, limiters, Roe fluxes, viscous and
 *         turbulent diffusion fluxes, k-omega sources) is evaluated in
 *         FP32 from FP32 copies of the primitive state, face metrics,
 *         metric derivatives and mut; res and its accumulation, the
 *         conserved state, the stage update, CIRS and the residual norms
 *         stay FP64.
 * mode 2: as mode 1 until the level-1 density residual rms reported by
 *         cosahip_smooth falls below tail_rms, FP64 from then on (FP64
 *         defect-correction tail).  Calling again restarts the switch.
 * mode 3: fast FP64 -- the Euler residual in binary64 with multiply-add
 *         contraction (one rounding per a*b+c instead of two) and
 *         quotients by a shared divisor as products with its reciprocal; results
 *         differ from the reference's in the last bits and are held to
 *         the FP64 parity bar.  Euler only (COSAHIP_ERR_ARG when viscous).
 * COSAHIP_ERR_ARG for the per-direction inviscid variant
 * (COSAHIP_RESIDUAL=split), which has no FP32 or contracted form.        */
int cosahip_set_precision(cosahip_solve* s, int mode, double tail_rms);
/* 32 while the next residual is evaluated in FP32, else 64.              */
int cosahip_precision_in_use(const cosahip_solve* s);

/* ---- instrumentation (no counterpart in the reference, whose only timers
 * are whole-run wall clocks, src/cosa.f90:789-797) ----------------------
 * When enabled, every kernel launch of a family is bracketed by HIP events
 * recorded on the solve's own stream; cosahip_timing_read synchronises and
 * returns the family's summed device time, launch count and work items
 * (interior cell-snapshots it processed).  Disabled by default.          */
enum {
  COSAHIP_FAM_RESIDUAL = 0,  /* MUSCL + Roe + flux difference, 3 directions */
  COSAHIP_FAM_CIRS = 1,      /* implicit residual smoothing, 3 sweeps       */
  COSAHIP_FAM_UPDATE = 2,    /* RK update                                   */
  COSAHIP_FAM_HALO = 3,      /* BC patches and cut copies                   */
  COSAHIP_FAM_CONVERT = 4,   /* cons2prim / prim2cons                       */
  COSAHIP_FAM_LAMBDA = 5,    /* spectral radii + local time step            */
  COSAHIP_FAM_OTHER = 6,     /* copies, FAS, norms, HB, MG transfer, checks */
  COSAHIP_NFAM = 7
};
int cosahip_timing_enable(cosahip_solve* s, int on);
int cosahip_timing_read(cosahip_solve* s, int family, double* total_ms, int64_t* launches, int64_t* items);
int cosahip_timing_reset(cosahip_solve* s);

/* Library build string. */
const char* cosahip_build_info(void);

#ifdef __cplusplus
}
#endif
#endif /* COSAHIP_H */
