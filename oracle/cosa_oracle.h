/*
 * cosa_oracle.h -- CPU restatement of COSA's pseudo-time residual/update
 * path (LANCASTER-CFD/COSA, src/), used ONLY as the checker of libcosahip.
 *
 * TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library.  The product (cosa_amd,
 * libcosahip) never links, imports or calls it.
 *
 * Every function restates one reference routine (file:line in its comment)
 * on the reference's own flat array layout, one block at a time (*_b) or a
 * whole level.  Arithmetic is written in the reference's operation order so
 * that a build with -O2 -ffp-contract=off reproduces the reference's
 * binary64 results (see oracle/Makefile).
 */
#ifndef COSA_ORACLE_H
#define COSA_ORACLE_H
#include <stdint.h>
#include "cosahip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One multigrid level: block sizes, offsets (setoffs) and the arrays the
 * hot path touches.  Arrays are owned by the caller.                     */
typedef struct or_level {
  int32_t nblocks;
  int32_t ncuts;
  const int32_t* imax;        /* [nblocks] node counts                      */
  const int32_t* jmax;
  const int32_t* kmax;
  const int32_t* nbcs;        /* [nblocks]                                  */
  const int32_t* bctopo;      /* 10 per patch, concatenated over blocks     */
  const int32_t* cutopo;      /* 21 per cut                                 */
  double *q, *qold, *qp, *qm, *dq, *dqp, *dqm, *flux, *res, *rhs, *work1;
  double *si, *sj, *sk, *vol, *dtvol, *betaxi, *betaeta, *betazeta;
  double *xdot, *ydot, *zdot, *pr;
  /* viscous / k-omega (NULL otherwise) */
  double *mut, *dist, *prd, *divvel, *delpls, *dtvolt, *tbetaxi, *tbetaeta, *tbetazeta;
  double* xd[9];   /* xideri, etaderi, zetaderi, xiderj, ..., zetaderk */
  double* ttrm;    /* (P3, 3, 0:2N): cross diffusion, gamma, beta (k-omega) */
  double *q1, *q2; /* dual time stepping, finest level: q^n, q^(n-1) (P3) */
} or_level;

typedef struct or_case {
  int32_t nlevel;
  int32_t pad;
  or_level lev[4];
} or_case;

/* parameters (common_variables mirror) */
void or_set_config(const cosahip_config* cfg);
void or_set_dhb(int nsnap, const double* dhb);
double or_get_resrms(int i);            /* resrms(i), i = 1..8            */
void or_set_resrms(int i, double v);
void or_set_mgit(int m);                /* cycle counter of CFL ramping   */
void or_set_dual(int dualt, double dt, int prop_start, int itime);   /* dual time */
double or_get_cfl(int turb);            /* cfl (0) / cflt (1) in use      */

/* per-block restatements (explicit-shape dummies of the reference) */
void or_cons2prim_b(double* q, int imax, int jmax, int kmax, int xtnd);
void or_prim2cons_b(double* q, int imax, int jmax, int kmax);
int or_bprtest(const double* q, int imax, int jmax, int kmax, int32_t* where);
void or_bbc(double* q, const double* si, const double* sj, const double* sk,
            const int32_t* bctopo, int nbcs, int imax, int jmax, int kmax);
void or_cut_q(int imax1, int jmax1, int kmax1, double* q1, int imax2,
              int jmax2, int kmax2, const double* q2, const int32_t* cutopo);
void or_lambdab(const double* q, const double* si, const double* sj,
                const double* sk, const double* xdot, const double* ydot,
                const double* zdot, double* lambdas, int imax, int jmax,
                int kmax);
void or_bdeltat(double* dtvol, const double* lambdas, double* betaxi,
                double* betaeta, double* betazeta, int imax, int jmax,
                int kmax);
void or_muscl_b(int idir, int nl, const int32_t* bctopo, int nbcs,
                const double* q, double* qp, double* qm, double* dq,
                double* dqp, double* dqm, int imax, int jmax, int kmax);
void or_roflux_b(int idir, const double* qr, const double* ql, double* flux,
                 const double* sijk, int imax, int jmax, int kmax);
void or_bresid(int idir, const double* flux, double* res, int imax, int jmax,
               int kmax);
void or_src_bhb(const double* q, const double* vol, double* res, double* qhb,
                int imax, int jmax, int kmax);
void or_resg_brms(const double* res, const double* rhs, double* resrms_b,
                  int imax, int jmax, int kmax);
void or_resg_brms_hb(const double* work1, double* resrms_b, int imax,
                     int jmax, int kmax);
void or_cirs_b(double* res, const double* dtvol, const double* betaxi,
               const double* betaeta, const double* betazeta, int imax,
               int jmax, int kmax);
void or_tridi(const double* a, const double* b, const double* c, double* x,
              int dim);
void or_update_b(double* res, double* q, const double* qold,
                 const double* dtvol, double ar, int imax, int jmax,
                 int kmax);
void or_rest1_b(int imax1, int jmax1, int kmax1, int imax2, int jmax2,
                int kmax2, const double* q1, double* q2, const double* vol1,
                const double* vol2);
void or_rest2_b(int ho, int imax1, int jmax1, int kmax1, int imax2,
                int jmax2, int kmax2, double* rhs2, const double* res1);
void or_prol_b(int bln, int imax1, int jmax1, int kmax1, int imax2,
               int jmax2, int kmax2, double* q1, const double* q2,
               const double* vol1);
void or_q_bedges(double* q, int imax, int jmax, int kmax);
/* viscous / Wilcox k-omega */
void or_bbc_dist(double* q, const double* si, const double* sj, const double* sk, const double* dist,
                 const int32_t* bctopo, int nbcs, int imax, int jmax, int kmax);
void or_lambdab_v(const double* q, const double* si, const double* sj, const double* sk, const double* xdot,
                  const double* ydot, const double* zdot, const double* vol, const double* mut, double* lambdas,
                  double* lamv, int imax, int jmax, int kmax);
void or_bdeltat_v(double* dtvol, double* dtvolt, const double* lambdas, const double* lamv, double* betaxi,
                  double* betaeta, double* betazeta, double* tbetaxi, double* tbetaeta, double* tbetazeta, int imax,
                  int jmax, int kmax);
void or_cirs_kb(double* res, const double* dtvol, const double* betaxi, const double* betaeta,
                const double* betazeta, const double* dtvolt, const double* tbetaxi, const double* tbetaeta,
                const double* tbetazeta, int imax, int jmax, int kmax);
void or_q_face_b(int idir, const int32_t* bctopo, int nbcs, const double* q, double* qf, const double* mut,
                 double* mutf, int calmut, int imax, int jmax, int kmax);
void or_q_der_b(int idir, const int32_t* bctopo, int nbcs, const double* q, double* dqxi, double* dqeta,
                double* dqzeta, int imax, int jmax, int kmax);
void or_rtst_b(int idir, const double* qf, const double* dqxi, const double* dqeta, const double* dqzeta,
               const double* xider, const double* etader, const double* zetader, double* tau1, const double* mutf,
               int imax, int jmax, int kmax);
void or_vflux_b(int idir, const double* qf, const double* dqxi, const double* dqeta, const double* dqzeta,
                const double* tau1, double* flux, const double* fdist, const double* xider, const double* etader,
                const double* zetader, const double* sijk, int imax, int jmax, int kmax);
void or_q_der_c_b(int idir, const double* qf, double* dqc, int imax, int jmax, int kmax);
void or_turprod_b(const double* q, const double* dqxi, const double* dqeta, const double* dqzeta, double* prd,
                  double* divvel, double* delpls, const double* const* xd, int imax, int jmax, int kmax);
void or_srckw_b(const double* q, const double* mutg, const double* prd, const double* divvel, double* res,
                const double* vol, int nl, int imax, int jmax, int kmax);
void or_lim_prod_mg_b(const double* q, double* prd, const double* mutg, int imax, int jmax, int kmax);
void or_fixq_b(double* q, int imax, int jmax, int kmax);
void or_tur_vis_b(const double* q, double* mut, int imax, int jmax, int kmax);
void or_facedist_b(int idir, const double* dist, double* fd, int imax, int jmax, int kmax);
void or_update_ttrm_b(const double* q, const double* dqxi, const double* dqeta, const double* dqzeta,
                      const double* const* xd, double* ttrm, const double* dist, int imax, int jmax, int kmax);
void or_srckw_sst_b(const double* q, const double* prd, const double* divvel, double* res, const double* mutg,
                    const double* ttrm, const double* vol, int nl, int imax, int jmax, int kmax);
void or_tur_vis_sst_b(const double* q, double* mut, const double* dqxi, const double* dqeta, const double* dqzeta,
                      const double* const* xd, const double* dist, int imax, int jmax, int kmax);
void or_update_kb(double* res, double* q, const double* qold, const double* vol, const double* dtvol,
                  const double* dtvolt, const double* delpls, const double* prd, double ar, int imax, int jmax,
                  int kmax, const double* ttrm);
void or_lim_res_ome_b(double* res, const double* q, const double* vol, const double* dtvolt, const double* delpls,
                      const double* prd, double ar, int imax, int jmax, int kmax, const double* ttrm);
void or_ccirs_b(double* res, const double* ores, const double* betaxi, const double* betaeta, const double* betazeta,
                const double* tbetaxi, const double* tbetaeta, const double* tbetazeta, int imax, int jmax,
                int kmax);
void or_scale_b(double* res, const double* dtvol, const double* dtvolt, int back, int imax, int jmax, int kmax);
int or_lu_factor(double* a, int m, int* ipiv);
void or_lu_solve(const double* a, int m, const int* ipiv, double* b, int nrhs);
int or_lu_solve_dense(int m, int nrhs, double* a, double* b);
void or_update_bhb_rkim(double* res, double* q, const double* qold, const double* vol, const double* dtvol,
                        const double* dtvolt, double* qhb0, const double* delpls, const double* prd, double ar,
                        int imax, int jmax, int kmax, const double* ttrm);
void or_lim_res_ome_bhb(double* res, const double* q, const double* qold, const double* vol, const double* dtvolt,
                        const double* delpls, const double* prd, int imax, int jmax, int kmax, const double* ttrm);
void or_rms_hb_term(double* work1, const double* q, const double* vol, int imax, int jmax, int kmax);
/* force_b (src/output.f90:4388-4886) on one block: pressure and viscous
 * force / moment coefficients of surfaces s = 1..nsurface (BC patches of
 * type 1500 + s - 1; inv14 != 0 also counts type-14 walls as surface 1,
 * pressure part only).  q conservative, x/y/z node coordinates and
 * xdot/ydot/zdot grid velocities (0:imax+1,...,0:2N*hbmove), momr the
 * moment centre (xmomr, ymomr, zmomr) per snapshot, lngth(1:3).
 * coef(12, 0:2N, nsurface): cl_p cd_p cz_p cmx_p cmy_p cm_p cl_v cd_v cz_v
 * cmx_v cmy_v cm_v.                                                       */
void or_force_b(const double* q, const double* x, const double* y, const double* z, const double* xdot,
                const double* ydot, const double* zdot, const double* si, const double* sj, const double* sk,
                const double* xideri, const double* etaderj, const double* zetaderk, int nbcs,
                const int32_t* bctopo, int nsurface, int inv14, const double* lngth, const double* momr,
                double* coef, int imax, int jmax, int kmax);
void or_rest4_b(int imax1, int jmax1, int kmax1, int imax2, int jmax2, int kmax2, const double* prd1, double* prd2,
                const double* vol1, const double* vol2);
void or_q_bcorns(double* q, int imax, int jmax, int kmax);

/* whole-level routines (loop over blocks, as the reference wrappers) */
void or_cons2prim(or_case* c, int nl, double* q);
void or_prim2cons(or_case* c, int nl, double* q);
void or_bc(or_case* c, int nl, double* q);
void or_cutman_q(or_case* c, int nl, double* q);
void or_edges_corners(or_case* c, int nl, double* q);
int or_prtest(or_case* c, int nl, int32_t* where);
void or_lambda_deltat(or_case* c, int nl);
void or_zero_res(or_case* c, int nl);
void or_residual(or_case* c, int nl);
void or_src_hb(or_case* c, int nl);
void or_resg_rms(or_case* c, int nl, double* resrms);
void or_add_array(or_case* c, int nl, double* dst, const double* src,
                  double fact);
void or_copy_array(or_case* c, int nl, double* dst, const double* src);
/* dual time stepping: per block and the physical step around mg() */
void or_src_uns_b(const double* q, const double* vol, double* res, int imax, int jmax, int kmax);
void or_rhs_uns_b(double* rhs, const double* q1, const double* q2, const double* vol, int imax, int jmax,
                  int kmax);
void or_resg_brms_dual(const double* q, const double* res, const double* rhs, const double* vol, double* rb,
                       int imax, int jmax, int kmax);
void or_update_dual_b(double* res, double* q, const double* qold, const double* vol, const double* dtvol,
                      const double* dtvolt, const double* delpls, const double* prd, double ar, int imax, int jmax,
                      int kmax, const double* ttrm);
void or_lim_res_ome_dual_b(double* res, const double* q, const double* vol, const double* dtvolt,
                           const double* delpls, const double* prd, double ar, int imax, int jmax, int kmax,
                           const double* ttrm);
void or_dual_begin(or_case* c, int itime);
void or_dual_end(or_case* c);
void or_cirs(or_case* c, int nl);
void or_update(or_case* c, int nl, double ar);
void or_tur_vis(or_case* c, int nl);
void or_fixq(or_case* c, int nl);
void or_turb_limits(or_case* c, int nl, double ar);

/* drivers */
int or_smooth(or_case* c, int nl, int from_fine, int to_coarse, int niter,
              int mgit, int32_t* done, int32_t* where);
void or_restrict(or_case* c, int nl);
void or_prolong(or_case* c, int nl);
/* mg (src/cosa.f90:806-930) for ncycle cycles; hist receives resrms(1:8)
 * per cycle (rows of 8); returns cycles run, or -error.                  */
int or_mg(or_case* c, int nstart, int npre, int npost, int ncrs,
          double* hist, int32_t* where);

/* grid motion and metric set-up (grid_oracle.c; src/grid.f90, src/cosa.f90) */
int or_imetrix_b(const double* x, const double* y, const double* z, const double* xdot, const double* ydot,
                 const double* zdot, double* si, double* sj, double* sk, int imax, int jmax, int kmax, int lmet,
                 int nsm, int moving);
int or_bvols(const double* x, const double* y, const double* z, double* vol, int imax, int jmax, int kmax);
void or_volex_b(double* vol, int imax, int jmax, int kmax);
void or_vol_bedges(double* vol, int imax, int jmax, int kmax);
void or_cordex_b(double* x, double* y, double* z, int imax, int jmax, int kmax, int nsm);
void or_vmetrix_b(const double* x, const double* y, const double* z, double* const* d, const double* vol, int imax,
                  int jmax, int kmax, int nsm);
void or_coarsen_b(const double* x, const double* y, const double* z, int imax, int jmax, int kmax, double* x2,
                  double* y2, double* z2, int imax2, int jmax2, int kmax2);
void or_pitch(double cth, double sth, double ca, double dtheta0, double omegas, double xh, double yh,
              const double* x0, const double* y0, const double* z0, const double* dx, const double* dy, double* x,
              double* y, double* z, double* xdot, double* ydot, double* zdot, int imax, int jmax, int kmax);
void or_cut_layer(int kind, int nsm, int imax1, int jmax1, int kmax1, double* a1, int imax2, int jmax2, int kmax2,
                  const double* a2, const int32_t* cutopo);

#ifdef __cplusplus
}
#endif
#endif
