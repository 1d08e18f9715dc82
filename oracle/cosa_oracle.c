/*
 * cosa_oracle.c -- CPU restatement of COSA's pseudo-time residual/update
 * path.  TEST INFRASTRUCTURE ONLY (see cosa_oracle.h): it is the checker of
 * libcosahip and is never linked into the product.
 *
 * Citations are file:line into LANCASTER-CFD/COSA src/.  Index conventions
 * follow the Fortran explicit-shape dummies exactly:
 *   P3 arrays  a(-1:imax+1,-1:jmax+1,-1:kmax+1,ncomp,0:2N)
 *   P2 metrics s(lmet,0:imax+1,0:jmax+1,0:kmax+1,0:2N*hbmove)
 *   P1 arrays  a(0:imax,0:jmax,0:kmax[,0:2N])
 *   face flux  f(imax,jmax,kmax,npde,0:2N)
 * Expressions keep the reference's evaluation order (Fortran and C agree on
 * precedence and left-to-right association for + - * /).
 */
#include "cosa_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static cosahip_config P;
static int NPDE = 5, NH = 0, DIM5 = 1, DIM5H = 1, LMET = 4;
static double DHB[131 * 131];
static double RESRMS[9];
/* dual time stepping: dualt, dt, dual_prop_start, itime of common_variables
 * (src/common_variables.f90:28, 39, 59) */
static int DUALT = 0, DUAL_PROP = 0, ITIME = 1;
static double DT = 1.0;

typedef long long ll;

typedef struct {
  int imax, jmax, kmax;
  ll n3, n2, n1, n0;
} bg;

static bg mkbg(int imax, int jmax, int kmax) {
  bg g;
  g.imax = imax;
  g.jmax = jmax;
  g.kmax = kmax;
  g.n3 = (ll)(imax + 3) * (jmax + 3) * (kmax + 3);
  g.n2 = (ll)(imax + 2) * (jmax + 2) * (kmax + 2);
  g.n1 = (ll)(imax + 1) * (jmax + 1) * (kmax + 1);
  g.n0 = (ll)imax * jmax * kmax;
  return g;
}

#define I3(g, i, j, k) \
  ((ll)((i) + 1) + (ll)((g).imax + 3) * ((ll)((j) + 1) + (ll)((g).jmax + 3) * (ll)((k) + 1)))
#define Q(a, g, i, j, k, v, n) (a)[I3(g, i, j, k) + (g).n3 * ((ll)((v) - 1) + (ll)NPDE * (n))]
#define S1(a, g, i, j, k, n) (a)[I3(g, i, j, k) + (g).n3 * (ll)(n)]
#define M(a, g, m, i, j, k, nh)                                              \
  (a)[((m) - 1) + (ll)LMET * ((ll)(i) + (ll)((g).imax + 2) *                 \
                                            ((ll)(j) + (ll)((g).jmax + 2) *  \
                                                           ((ll)(k) + (ll)((g).kmax + 2) * (nh))))]
#define X2(a, g, i, j, k, nh) \
  (a)[(ll)(i) + (ll)((g).imax + 2) * ((ll)(j) + (ll)((g).jmax + 2) * ((ll)(k) + (ll)((g).kmax + 2) * (nh)))]
#define P1(a, g, i, j, k, n) \
  (a)[(ll)(i) + (ll)((g).imax + 1) * ((ll)(j) + (ll)((g).jmax + 1) * ((ll)(k) + (ll)((g).kmax + 1) * (n)))]
#define V1(a, g, i, j, k) (a)[(ll)(i) + (ll)((g).imax + 1) * ((ll)(j) + (ll)((g).jmax + 1) * (ll)(k))]
#define F0(a, g, i, j, k, v, n)                                                          \
  (a)[(ll)((i) - 1) + (ll)(g).imax * ((ll)((j) - 1) + (ll)(g).jmax *                      \
                                                      ((ll)((k) - 1) + (ll)(g).kmax *     \
                                                                          ((ll)((v) - 1) + (ll)NPDE * (n))))]

static int cyc[3][3] = {{1, 2, 3}, {2, 3, 1}, {3, 1, 2}}; /* cyc(idir,m) */
#define CYC(d, m) cyc[(d) - 1][(m) - 1]
#define KRD(a, b) ((a) == (b) ? 1 : 0)

static int is_kom(void) { return P.kom || P.kom_bsl || P.kom_sst; }
static int is_sst2(void) { return P.kom_bsl || P.kom_sst; }

/* BSL / SST constants, fixed by input (src/input.f90:2523-2541) */
static double SST_A1 = 0.31, SIGK1 = 0.5, SIGK2 = 1.0, SIGW1 = 0.5, SIGW2 = 0.856, B1 = 3.0 / 40.0, B2 = 0.0828,
              G1 = 0.0, G2 = 0.0;
static void set_sst_constants(void) {
  const double kappa = 0.41, bstr = 0.09;
  SIGK1 = P.kom_sst ? 0.85 : 0.5;
  G1 = B1 / bstr - SIGW1 * (kappa * kappa) / sqrt(bstr);
  G2 = B2 / bstr - SIGW2 * (kappa * kappa) / sqrt(bstr);
}
/* ttrm(-1:imax+1,-1:jmax+1,-1:kmax+1,3,0:2N): cross diffusion, gamma, beta */
#define T3(a, g, i, j, k, m, n) (a)[I3(g, i, j, k) + (g).n3 * ((ll)((m) - 1) + 3LL * (n))]

void or_set_config(const cosahip_config* cfg) {
  P = *cfg;
  NPDE = P.npde;
  NH = P.nharms;
  DIM5 = 2 * NH + 1;
  DIM5H = 2 * NH * P.hbmove + 1;
  LMET = P.lmet;
  set_sst_constants();
}

void or_set_dhb(int nsnap, const double* dhb) {
  memcpy(DHB, dhb, sizeof(double) * (size_t)nsnap * nsnap);
}
#define D(i, j) DHB[(i) + (ll)DIM5 * (j)] /* dhb(i,j), column-major, 0-based */

void or_set_dual(int dualt, double dt, int prop_start, int itime) {
  DUALT = dualt;
  DT = dt;
  DUAL_PROP = prop_start;
  ITIME = itime;
}
/* 1 on the first physical step of a backward-Euler start, else the BDF2
 * 3/2 (src/update.f90:650-653, 2753-2756) */
static double dual_fact(void) { return (ITIME == 1 && !DUAL_PROP) ? 1.0 : 1.5; }

double or_get_resrms(int i) { return RESRMS[i]; }
void or_set_resrms(int i, double v) { RESRMS[i] = v; }

/* ---------------------------------------------------------------------- */
/* cons2prim_b  (src/update.f90:2136-2188)                                 */
void or_cons2prim_b(double* q, int imax, int jmax, int kmax, int xtnd) {
  bg g = mkbg(imax, jmax, kmax);
  int istr = 1 - xtnd, jstr = 1 - xtnd, kstr = 1 - xtnd;
  int iend = imax - 1 + xtnd, jend = jmax - 1 + xtnd, kend = kmax - 1 + xtnd;
  for (int n = 0; n <= 2 * NH; n++) {
    for (int ipde = 2; ipde <= NPDE; ipde++) {
      if (ipde == 5) continue;
      for (int k = kstr; k <= kend; k++)
        for (int j = jstr; j <= jend; j++)
          for (int i = istr; i <= iend; i++)
            Q(q, g, i, j, k, ipde, n) = Q(q, g, i, j, k, ipde, n) / Q(q, g, i, j, k, 1, n);
    }
    for (int k = kstr; k <= kend; k++)
      for (int j = jstr; j <= jend; j++)
        for (int i = istr; i <= iend; i++) {
          double rho = Q(q, g, i, j, k, 1, n);
          double u = Q(q, g, i, j, k, 2, n);
          double v = Q(q, g, i, j, k, 3, n);
          double w = Q(q, g, i, j, k, 4, n);
          double rhoe = Q(q, g, i, j, k, 5, n);
          double tke = is_kom() ? Q(q, g, i, j, k, 6, n) : 0.0;
          Q(q, g, i, j, k, 5, n) = (P.gamma - 1) * (rhoe - rho * (u * u + v * v + w * w) / 2 - rho * tke);
        }
  }
}

/* prim2cons_b  (src/update.f90:1875-2084): interior + i-halo rows,
 * j-halo and k-halo strips (no edges/corners).                           */
static void p2c_range(double* q, bg g, int n, int i0, int i1, int j0, int j1, int k0, int k1) {
  /* tke saved before the momentum scaling (work(...,1) = q(...,6)) */
  for (int ipde = 2; ipde <= NPDE; ipde++) {
    if (ipde == 5) continue;
    for (int k = k0; k <= k1; k++)
      for (int j = j0; j <= j1; j++)
        for (int i = i0; i <= i1; i++) Q(q, g, i, j, k, ipde, n) = Q(q, g, i, j, k, ipde, n) * Q(q, g, i, j, k, 1, n);
  }
}
static void p2c_energy(double* q, const double* tke_save, bg g, int n, int i0, int i1, int j0, int j1, int k0,
                       int k1) {
  for (int k = k0; k <= k1; k++)
    for (int j = j0; j <= j1; j++)
      for (int i = i0; i <= i1; i++) {
        double rho = Q(q, g, i, j, k, 1, n);
        double rhou = Q(q, g, i, j, k, 2, n);
        double rhov = Q(q, g, i, j, k, 3, n);
        double rhow = Q(q, g, i, j, k, 4, n);
        double p = Q(q, g, i, j, k, 5, n);
        double tke = tke_save ? tke_save[I3(g, i, j, k)] : 0.0;
        Q(q, g, i, j, k, 5, n) = p / (P.gamma - 1) + (rhou * rhou + rhov * rhov + rhow * rhow) / 2 / rho + rho * tke;
      }
}
void or_prim2cons_b(double* q, int imax, int jmax, int kmax) {
  bg g = mkbg(imax, jmax, kmax);
  double* tk = NULL;
  if (is_kom()) tk = (double*)malloc(sizeof(double) * (size_t)g.n3);
  for (int n = 0; n <= 2 * NH; n++) {
    int R[5][6] = {{-1, imax + 1, 1, jmax - 1, 1, kmax - 1},
                   {1, imax - 1, -1, 0, 1, kmax - 1},
                   {1, imax - 1, jmax, jmax + 1, 1, kmax - 1},
                   {1, imax - 1, 1, jmax - 1, -1, 0},
                   {1, imax - 1, 1, jmax - 1, kmax, kmax + 1}};
    if (tk)
      for (int r = 0; r < 5; r++)
        for (int k = R[r][4]; k <= R[r][5]; k++)
          for (int j = R[r][2]; j <= R[r][3]; j++)
            for (int i = R[r][0]; i <= R[r][1]; i++) tk[I3(g, i, j, k)] = Q(q, g, i, j, k, 6, n);
    /* the reference scales each variable over all five ranges in turn; the
       ranges are disjoint, so range-by-range is the same arithmetic */
    for (int r = 0; r < 5; r++) p2c_range(q, g, n, R[r][0], R[r][1], R[r][2], R[r][3], R[r][4], R[r][5]);
    for (int r = 0; r < 5; r++) p2c_energy(q, tk, g, n, R[r][0], R[r][1], R[r][2], R[r][3], R[r][4], R[r][5]);
  }
  free(tk);
}

/* bprtest (src/cosa.f90:960-1120): interior + face halo strips; returns 0
 * or 1 (pressure) / 2 (density) with where = {i,j,k,n}.                   */
int or_bprtest(const double* q, int imax, int jmax, int kmax, int32_t* where) {
  bg g = mkbg(imax, jmax, kmax);
  int R[5][6] = {{-1, imax + 1, 1, jmax - 1, 1, kmax - 1},
                 {1, imax - 1, -1, 0, 1, kmax - 1},
                 {1, imax - 1, jmax, jmax + 1, 1, kmax - 1},
                 {1, imax - 1, 1, jmax - 1, -1, 0},
                 {1, imax - 1, 1, jmax - 1, kmax, kmax + 1}};
  for (int n = 0; n <= 2 * NH; n++)
    for (int r = 0; r < 5; r++)
      for (int k = R[r][4]; k <= R[r][5]; k++)
        for (int j = R[r][2]; j <= R[r][3]; j++)
          for (int i = R[r][0]; i <= R[r][1]; i++) {
            double rho = Q(q, g, i, j, k, 1, n), p = Q(q, g, i, j, k, 5, n);
            int bad = 0;
            if (p <= P.qmin[4]) bad = 1;
            else if (rho <= P.qmin[0]) bad = 2;
            if (bad) {
              if (where) { where[0] = i; where[1] = j; where[2] = k; where[3] = n; }
              return bad;
            }
          }
  return 0;
}

/* ---------------------------------------------------------------------- */
/* boundary conditions                                                     */

typedef struct {
  int idir, inrout, istrt[4], iend[4], ic1, ic2, ic3;
} bcgeo;

static bcgeo bc_decode(const int32_t* t) {
  bcgeo b;
  b.idir = t[1];
  b.inrout = t[2];
  b.istrt[1] = t[3]; b.iend[1] = t[4];
  b.istrt[2] = t[5]; b.iend[2] = t[6];
  b.istrt[3] = t[7]; b.iend[3] = t[8];
  b.ic1 = CYC(b.idir, 1);
  b.ic2 = CYC(b.idir, 2);
  b.ic3 = CYC(b.idir, 3);
  return b;
}
/* ijk index of (face coordinate a along idir, i2, i3 transverse) */
static void bc_ijk(const bcgeo* b, int a, int i2, int i3, int* i, int* j, int* k) {
  *i = a * KRD(b->ic1, 1) + i2 * KRD(b->ic2, 1) + i3 * KRD(b->ic3, 1);
  *j = a * KRD(b->ic1, 2) + i2 * KRD(b->ic2, 2) + i3 * KRD(b->ic3, 2);
  *k = a * KRD(b->ic1, 3) + i2 * KRD(b->ic2, 3) + i3 * KRD(b->ic3, 3);
}

/* bc_far0 (src/bc.f90:124-483): Riemann-invariant farfield (aircraft/vawt
 * freestream), high-speed branch.                                         */
static void or_bc_far0(double* q, const double* si, const double* sj, const double* sk, const int32_t* t, bg g) {
  double qfst[5];
  qfst[0] = P.rhoinf;
  qfst[1] = P.machfs * cos(P.alpha) * cos(P.beta);
  qfst[2] = P.machfs * sin(P.alpha) * cos(P.beta);
  qfst[3] = P.machfs * sin(P.beta);
  qfst[4] = P.pinf;
  double gamma = P.gamma;
  double sinf = P.pinf / pow(P.rhoinf, gamma);
  int ijkmax[4] = {0, g.imax, g.jmax, g.kmax};
  bcgeo b = bc_decode(t);
  int ibcpt, ibcpt2, ibcn, ibcm;
  double sgnm;
  if (b.inrout == 1) {
    ibcpt = 0; ibcpt2 = -1; ibcn = 1; ibcm = 1; sgnm = -1.0;
  } else {
    ibcpt = ijkmax[b.idir]; ibcpt2 = ijkmax[b.idir] + 1; ibcn = ijkmax[b.idir] - 1; ibcm = ijkmax[b.idir]; sgnm = 1.0;
  }
  const double* s = b.idir == 1 ? si : (b.idir == 2 ? sj : sk);
  int kom = is_kom();
  for (int n = 0; n <= 2 * NH; n++) {
    int nh = n * P.hbmove;
    for (int i3 = b.istrt[b.ic3]; i3 <= b.iend[b.ic3]; i3++)
      for (int i2 = b.istrt[b.ic2]; i2 <= b.iend[b.ic2]; i2++) {
        int ib, jb, kb, ib2, jb2, kb2, in, jn, kn, im, jm, km;
        bc_ijk(&b, ibcpt, i2, i3, &ib, &jb, &kb);
        bc_ijk(&b, ibcpt2, i2, i3, &ib2, &jb2, &kb2);
        bc_ijk(&b, ibcn, i2, i3, &in, &jn, &kn);
        bc_ijk(&b, ibcm, i2, i3, &im, &jm, &km);
        double sx = M(s, g, 1, im, jm, km, nh) * sgnm;
        double sy = M(s, g, 2, im, jm, km, nh) * sgnm;
        double sz = M(s, g, 3, im, jm, km, nh) * sgnm;
        double qext = Q(q, g, in, jn, kn, 2, n) * sx + Q(q, g, in, jn, kn, 3, n) * sy + Q(q, g, in, jn, kn, 4, n) * sz;
        double qinf = qfst[1] * sx + qfst[2] * sy + qfst[3] * sz;
        double rhoext = Q(q, g, in, jn, kn, 1, n);
        double pext = Q(q, g, in, jn, kn, 5, n);
        double aext = sqrt(gamma * pext / rhoext);
        double rext = qext + 2 * aext / (gamma - 1);
        double rinf = qinf - 2 * P.ainf / (gamma - 1);
        double qdotn = (rext + rinf) / 2;
        double a = (gamma - 1) * (rext - rinf) / 4;
        double mach = qdotn / a;
        double sext = pext / pow(rhoext, gamma);
        double hv[5];
        int set = 0;
        if (qdotn <= 0.0) {
          if (fabs(mach) >= 1.0) {
            for (int v = 0; v < 5; v++) hv[v] = qfst[v];
            set = 1;
          } else if (fabs(mach) < 1.0) {
            double rhobc = pow(a * a / (gamma * sinf), 1.0 / (gamma - 1));
            double p = rhobc * a * a / gamma;
            hv[0] = rhobc;
            hv[1] = qfst[1] + (qdotn - qinf) * sx;
            hv[2] = qfst[2] + (qdotn - qinf) * sy;
            hv[3] = qfst[3] + (qdotn - qinf) * sz;
            hv[4] = p;
            set = 1;
          }
        } else if (qdotn > 0.0) {
          if (fabs(mach) >= 1.0) {
            for (int v = 0; v < 5; v++) hv[v] = Q(q, g, in, jn, kn, v + 1, n);
            set = 1;
          } else if (fabs(mach) < 1.0) {
            double rhobc = pow(a * a / (gamma * sext), 1.0 / (gamma - 1));
            double p = rhobc * a * a / gamma;
            hv[0] = rhobc;
            hv[1] = Q(q, g, in, jn, kn, 2, n) + (qdotn - qext) * sx;
            hv[2] = Q(q, g, in, jn, kn, 3, n) + (qdotn - qext) * sy;
            hv[3] = Q(q, g, in, jn, kn, 4, n) + (qdotn - qext) * sz;
            hv[4] = p;
            set = 1;
          }
        }
        if (set)
          for (int v = 0; v < 5; v++) {
            Q(q, g, ib, jb, kb, v + 1, n) = hv[v];
            Q(q, g, ib2, jb2, kb2, v + 1, n) = hv[v];
          }
        if (kom) {
          /* src/bc.f90:385-479: k, omega fixed on inflow, extrapolated on outflow */
          double t6, t7;
          if (qdotn <= 0.0) {
            t6 = P.tkefar; t7 = P.omefar;
          } else {
            t6 = Q(q, g, in, jn, kn, 6, n); t7 = Q(q, g, in, jn, kn, 7, n);
          }
          if (qdotn <= 0.0 || qdotn > 0.0) {
            Q(q, g, ib, jb, kb, 6, n) = t6; Q(q, g, ib, jb, kb, 7, n) = t7;
            Q(q, g, ib2, jb2, kb2, 6, n) = t6; Q(q, g, ib2, jb2, kb2, 7, n) = t7;
          }
        }
      }
  }
}

/* bc_walinv_14 (src/bc.f90:1457-1675): second-order extrapolation minus
 * the wall-normal (relative) velocity.                                    */
static void or_bc_walinv14(double* q, const double* si, const double* sj, const double* sk, const int32_t* t, bg g) {
  int ijkmax[4] = {0, g.imax, g.jmax, g.kmax};
  bcgeo b = bc_decode(t);
  int ibcpt, ibcpt2, ibcn, ibcn2, ibcm;
  if (b.inrout == 1) {
    ibcpt = 0; ibcpt2 = -1; ibcn = 1; ibcn2 = 2; ibcm = 1;
  } else {
    ibcpt = ijkmax[b.idir]; ibcpt2 = ijkmax[b.idir] + 1; ibcn = ijkmax[b.idir] - 1; ibcn2 = ijkmax[b.idir] - 2;
    ibcm = ijkmax[b.idir];
  }
  const double* s = b.idir == 1 ? si : (b.idir == 2 ? sj : sk);
  for (int n = 0; n <= 2 * NH; n++) {
    int nh = n * P.hbmove;
    for (int i3 = b.istrt[b.ic3]; i3 <= b.iend[b.ic3]; i3++)
      for (int i2 = b.istrt[b.ic2]; i2 <= b.iend[b.ic2]; i2++) {
        int ib, jb, kb, ib2, jb2, kb2, in, jn, kn, in2, jn2, kn2, im, jm, km;
        bc_ijk(&b, ibcpt, i2, i3, &ib, &jb, &kb);
        bc_ijk(&b, ibcpt2, i2, i3, &ib2, &jb2, &kb2);
        bc_ijk(&b, ibcn, i2, i3, &in, &jn, &kn);
        bc_ijk(&b, ibcn2, i2, i3, &in2, &jn2, &kn2);
        bc_ijk(&b, ibcm, i2, i3, &im, &jm, &km);
        double nx = M(s, g, 1, im, jm, km, nh), ny = M(s, g, 2, im, jm, km, nh), nz = M(s, g, 3, im, jm, km, nh);
        double xwdot = 0, ywdot = 0, zwdot = 0;
        if (P.moving) {
          xwdot = M(s, g, 5, im, jm, km, nh);
          ywdot = M(s, g, 6, im, jm, km, nh);
          zwdot = M(s, g, 7, im, jm, km, nh);
        }
        double rho1 = Q(q, g, in, jn, kn, 1, n), u1 = Q(q, g, in, jn, kn, 2, n), v1 = Q(q, g, in, jn, kn, 3, n);
        double w1 = Q(q, g, in, jn, kn, 4, n), p1 = Q(q, g, in, jn, kn, 5, n);
        double rho2 = Q(q, g, in2, jn2, kn2, 1, n), u2 = Q(q, g, in2, jn2, kn2, 2, n), v2 = Q(q, g, in2, jn2, kn2, 3, n);
        double w2 = Q(q, g, in2, jn2, kn2, 4, n), p2 = Q(q, g, in2, jn2, kn2, 5, n);
        double rhow = (3 * rho1 - rho2) / 2;
        double uw = (3 * u1 - u2) / 2;
        double vw = (3 * v1 - v2) / 2;
        double ww = (3 * w1 - w2) / 2;
        double pw = (3 * p1 - p2) / 2;
        double un1 = u1 * nx + v1 * ny + w1 * nz;
        double un2 = u2 * nx + v2 * ny + w2 * nz;
        double unw = (3 * un1 - un2) / 2;
        if (P.moving) unw = unw - (xwdot * nx + ywdot * ny + zwdot * nz);
        double hv[5] = {rhow, uw - unw * nx, vw - unw * ny, ww - unw * nz, pw};
        for (int v = 0; v < 5; v++) {
          Q(q, g, ib, jb, kb, v + 1, n) = hv[v];
          Q(q, g, ib2, jb2, kb2, v + 1, n) = hv[v];
        }
      }
  }
  if (is_kom()) {
    for (int n = 0; n <= 2 * NH; n++)
      for (int i3 = b.istrt[b.ic3]; i3 <= b.iend[b.ic3]; i3++)
        for (int i2 = b.istrt[b.ic2]; i2 <= b.iend[b.ic2]; i2++) {
          int ib, jb, kb, ib2, jb2, kb2, in, jn, kn;
          bc_ijk(&b, ibcpt, i2, i3, &ib, &jb, &kb);
          bc_ijk(&b, ibcpt2, i2, i3, &ib2, &jb2, &kb2);
          bc_ijk(&b, ibcn, i2, i3, &in, &jn, &kn);
          for (int v = 6; v <= 7; v++) {
            Q(q, g, ib, jb, kb, v, n) = Q(q, g, in, jn, kn, v, n);
            Q(q, g, ib2, jb2, kb2, v, n) = Q(q, g, ib, jb, kb, v, n);
          }
        }
  }
}

/* bc_extrapol (src/bc.f90:2494-2592): zeroth-order extrapolation. */
static void or_bc_extrapol(double* q, const int32_t* t, bg g) {
  int ijkmax[4] = {0, g.imax, g.jmax, g.kmax};
  bcgeo b = bc_decode(t);
  int ibcpt, ibcpt2, ibcn;
  if (b.inrout == 1) {
    ibcpt = 0; ibcpt2 = -1; ibcn = 1;
  } else {
    ibcpt = ijkmax[b.idir]; ibcpt2 = ijkmax[b.idir] + 1; ibcn = ijkmax[b.idir] - 1;
  }
  for (int n = 0; n <= 2 * NH; n++)
    for (int ipde = 1; ipde <= NPDE; ipde++)
      for (int i3 = b.istrt[b.ic3]; i3 <= b.iend[b.ic3]; i3++)
        for (int i2 = b.istrt[b.ic2]; i2 <= b.iend[b.ic2]; i2++) {
          int ib, jb, kb, ib2, jb2, kb2, in, jn, kn;
          bc_ijk(&b, ibcpt, i2, i3, &ib, &jb, &kb);
          bc_ijk(&b, ibcpt2, i2, i3, &ib2, &jb2, &kb2);
          bc_ijk(&b, ibcn, i2, i3, &in, &jn, &kn);
          Q(q, g, ib, jb, kb, ipde, n) = Q(q, g, in, jn, kn, ipde, n);
          Q(q, g, ib2, jb2, kb2, ipde, n) = Q(q, g, in, jn, kn, ipde, n);
        }
}

/* bc_sym (src/bc.f90:2595-2710): mirror layers 1,2 with the velocity
 * component idir negated.                                                 */
static void or_bc_sym(double* q, const int32_t* t, bg g) {
  int ijkmax[4] = {0, g.imax, g.jmax, g.kmax};
  bcgeo b = bc_decode(t);
  int ibcpt, ibcpt2, ibcn, ibcn2;
  if (b.inrout == 1) {
    ibcpt = 0; ibcpt2 = -1; ibcn = 1; ibcn2 = 2;
  } else {
    ibcpt = ijkmax[b.idir]; ibcpt2 = ijkmax[b.idir] + 1; ibcn = ijkmax[b.idir] - 1; ibcn2 = ijkmax[b.idir] - 2;
  }
  int isym[8];
  for (int v = 1; v <= NPDE; v++) isym[v] = 1;
  isym[b.idir + 1] = -1;
  for (int n = 0; n <= 2 * NH; n++)
    for (int ipde = 1; ipde <= NPDE; ipde++)
      for (int i3 = b.istrt[b.ic3]; i3 <= b.iend[b.ic3]; i3++)
        for (int i2 = b.istrt[b.ic2]; i2 <= b.iend[b.ic2]; i2++) {
          int ib, jb, kb, ib2, jb2, kb2, in, jn, kn, in2, jn2, kn2;
          bc_ijk(&b, ibcpt, i2, i3, &ib, &jb, &kb);
          bc_ijk(&b, ibcpt2, i2, i3, &ib2, &jb2, &kb2);
          bc_ijk(&b, ibcn, i2, i3, &in, &jn, &kn);
          bc_ijk(&b, ibcn2, i2, i3, &in2, &jn2, &kn2);
          Q(q, g, ib, jb, kb, ipde, n) = isym[ipde] * Q(q, g, in, jn, kn, ipde, n);
          Q(q, g, ib2, jb2, kb2, ipde, n) = isym[ipde] * Q(q, g, in2, jn2, kn2, ipde, n);
        }
}

/* bbc (src/bc.f90:35-121): patches in input order. */
static void or_bc_walvis15(double* q, const double* si, const double* sj, const double* sk, const double* dist,
                           const int32_t* t, bg g);
void or_bbc(double* q, const double* si, const double* sj, const double* sk, const int32_t* bctopo, int nbcs,
            int imax, int jmax, int kmax) {
  or_bbc_dist(q, si, sj, sk, NULL, bctopo, nbcs, imax, jmax, kmax);
}
/* bbc with the cell-centre wall distance dist(0:imax,0:jmax,0:kmax) that
 * viscous walls (types 15xx) need.                                        */
void or_bbc_dist(double* q, const double* si, const double* sj, const double* sk, const double* dist,
                 const int32_t* bctopo, int nbcs, int imax, int jmax, int kmax) {
  bg g = mkbg(imax, jmax, kmax);
  for (int ibc = 0; ibc < nbcs; ibc++) {
    const int32_t* t = bctopo + 10 * ibc;
    if (t[0] == 80) or_bc_sym(q, t, g);
    else if (t[0] == 4) or_bc_extrapol(q, t, g);
    else if (t[0] == 30) or_bc_far0(q, si, sj, sk, t, g);
    else if (t[0] == 14) or_bc_walinv14(q, si, sj, sk, t, g);
    else if (t[0] / 100 == 15 && dist) or_bc_walvis15(q, si, sj, sk, dist, t, g);
    /* other types are not built here; the library rejects them at define time */
  }
}

/* cut_q, same-rank branch (src/bc.f90:1894-2141): two interior layers of
 * block 2 into the two halo layers of block 1.                            */
void or_cut_q(int imax1, int jmax1, int kmax1, double* q1, int imax2, int jmax2, int kmax2, const double* q2,
              const int32_t* cutopo) {
  bg g1 = mkbg(imax1, jmax1, kmax1), g2 = mkbg(imax2, jmax2, kmax2);
  int ijkmax1[4] = {0, imax1, jmax1, kmax1}, ijkmax2[4] = {0, imax2, jmax2, kmax2};
  int idir1 = cutopo[1], inout1 = cutopo[2];
  int istr1[4], iend1[4], istr2[4], iend2[4], iord[4], isgn1[4], isgn2[4], len[4];
  istr1[1] = cutopo[3]; iend1[1] = cutopo[4];
  istr1[2] = cutopo[5]; iend1[2] = cutopo[6];
  istr1[3] = cutopo[7]; iend1[3] = cutopo[8];
  int idir2 = cutopo[10], inout2 = cutopo[11];
  istr2[1] = cutopo[12]; iend2[1] = cutopo[13];
  istr2[2] = cutopo[14]; iend2[2] = cutopo[15];
  istr2[3] = cutopo[16]; iend2[3] = cutopo[17];
  iord[1] = cutopo[18]; iord[2] = cutopo[19]; iord[3] = cutopo[20];
  int ibcpt, ibcinc, inr, inrinc;
  if (inout1 == 1) { ibcpt = 0; ibcinc = -1; } else { ibcpt = ijkmax1[idir1]; ibcinc = 1; }
  if (inout2 == 1) { inr = 1; inrinc = 1; } else { inr = ijkmax2[idir2] - 1; inrinc = -1; }
  for (int l = 1; l <= 3; l++) {
    len[l] = abs(iend1[l] - istr1[l]);
    isgn1[l] = iend1[l] > istr1[l] ? 1 : -1;
    isgn2[l] = iend2[l] > istr2[l] ? 1 : -1;
  }
  int ic1 = CYC(idir1, 1), ic2 = CYC(idir1, 2), ic3 = CYC(idir1, 3);
  int jc1 = iord[ic1], jc2 = iord[ic2], jc3 = iord[ic3];
  for (int i3 = 0; i3 <= len[ic3]; i3++)
    for (int i2 = 0; i2 <= len[ic2]; i2++) {
      int a2 = istr1[ic2] + isgn1[ic2] * i2, a3 = istr1[ic3] + isgn1[ic3] * i3;
      int ii1 = ibcpt * KRD(ic1, 1) + a2 * KRD(ic2, 1) + a3 * KRD(ic3, 1);
      int jj1 = ibcpt * KRD(ic1, 2) + a2 * KRD(ic2, 2) + a3 * KRD(ic3, 2);
      int kk1 = ibcpt * KRD(ic1, 3) + a2 * KRD(ic2, 3) + a3 * KRD(ic3, 3);
      int ii2 = ii1 + ibcinc * KRD(ic1, 1), jj2 = jj1 + ibcinc * KRD(ic1, 2), kk2 = kk1 + ibcinc * KRD(ic1, 3);
      int b2 = istr2[jc2] + isgn2[jc2] * i2, b3 = istr2[jc3] + isgn2[jc3] * i3;
      int in1 = inr * KRD(jc1, 1) + b2 * KRD(jc2, 1) + b3 * KRD(jc3, 1);
      int jn1 = inr * KRD(jc1, 2) + b2 * KRD(jc2, 2) + b3 * KRD(jc3, 2);
      int kn1 = inr * KRD(jc1, 3) + b2 * KRD(jc2, 3) + b3 * KRD(jc3, 3);
      int in2 = in1 + inrinc * KRD(jc1, 1), jn2 = jn1 + inrinc * KRD(jc1, 2), kn2 = kn1 + inrinc * KRD(jc1, 3);
      for (int n = 0; n <= 2 * NH; n++)
        for (int ipde = 1; ipde <= NPDE; ipde++) {
          Q(q1, g1, ii1, jj1, kk1, ipde, n) = Q(q2, g2, in1, jn1, kn1, ipde, n);
          Q(q1, g1, ii2, jj2, kk2, ipde, n) = Q(q2, g2, in2, jn2, kn2, ipde, n);
        }
    }
}

/* ---------------------------------------------------------------------- */
/* spectral radii and local time step                                      */

/* eig_eul (src/cosa.f90:4183-4237) */
static void eig_eul(const double* qc, const double* sim, const double* sip, const double* sjm, const double* sjp,
                    const double* skm, const double* skp, double xdot, double ydot, double zdot, double* lami,
                    double* lamj, double* lamk) {
  double rho = qc[0], u = qc[1], v = qc[2], w = qc[3], p = qc[4];
  double a = sqrt(P.gamma * p / rho);
  double xix = 0.5 * (sim[0] * sim[3] + sip[0] * sip[3]);
  double xiy = 0.5 * (sim[1] * sim[3] + sip[1] * sip[3]);
  double xiz = 0.5 * (sim[2] * sim[3] + sip[2] * sip[3]);
  double etax = 0.5 * (sjm[0] * sjm[3] + sjp[0] * sjp[3]);
  double etay = 0.5 * (sjm[1] * sjm[3] + sjp[1] * sjp[3]);
  double etaz = 0.5 * (sjm[2] * sjm[3] + sjp[2] * sjp[3]);
  double zetax = 0.5 * (skm[0] * skm[3] + skp[0] * skp[3]);
  double zetay = 0.5 * (skm[1] * skm[3] + skp[1] * skp[3]);
  double zetaz = 0.5 * (skm[2] * skm[3] + skp[2] * skp[3]);
  double axi = (sim[3] + sip[3]) / 2;
  double aeta = (sjm[3] + sjp[3]) / 2;
  double azeta = (skm[3] + skp[3]) / 2;
  if (P.moving) {
    u = u - xdot;
    v = v - ydot;
    w = w - zdot;
  }
  *lami = fabs(u * xix + v * xiy + w * xiz) + a * axi;
  *lamj = fabs(u * etax + v * etay + w * etaz) + a * aeta;
  *lamk = fabs(u * zetax + v * zetay + w * zetaz) + a * azeta;
}

/* lambdab, inviscid branch (src/cosa.f90:1248-1346); lambdas is a P3
 * array with npde components of which 1..3 are written.                  */
void or_lambdab(const double* q, const double* si, const double* sj, const double* sk, const double* xdot,
                const double* ydot, const double* zdot, double* lambdas, int imax, int jmax, int kmax) {
  bg g = mkbg(imax, jmax, kmax);
  for (int n = 0; n <= 2 * NH; n++) {
    int nh = n * P.hbmove;
    for (int k = 1; k <= kmax - 1; k++)
      for (int j = 1; j <= jmax - 1; j++)
        for (int i = 1; i <= imax - 1; i++) {
          double qc[5], sim[4], sip[4], sjm[4], sjp[4], skm[4], skp[4];
          for (int v = 0; v < 5; v++) qc[v] = Q(q, g, i, j, k, v + 1, n);
          for (int m = 0; m < 4; m++) {
            sim[m] = M(si, g, m + 1, i, j, k, nh);
            sip[m] = M(si, g, m + 1, i + 1, j, k, nh);
            sjm[m] = M(sj, g, m + 1, i, j, k, nh);
            sjp[m] = M(sj, g, m + 1, i, j + 1, k, nh);
            skm[m] = M(sk, g, m + 1, i, j, k, nh);
            skp[m] = M(sk, g, m + 1, i, j, k + 1, nh);
          }
          double xd = 0, yd = 0, zd = 0;
          if (P.moving) {
            xd = X2(xdot, g, i, j, k, nh);
            yd = X2(ydot, g, i, j, k, nh);
            zd = X2(zdot, g, i, j, k, nh);
          }
          double li, lj, lk;
          eig_eul(qc, sim, sip, sjm, sjp, skm, skp, xd, yd, zd, &li, &lj, &lk);
          Q(lambdas, g, i, j, k, 1, n) = li;
          Q(lambdas, g, i, j, k, 2, n) = lj;
          Q(lambdas, g, i, j, k, 3, n) = lk;
        }
  }
}

/* lambdab, viscous branch (src/cosa.f90:1350-1430) with eig_ns
 * (src/cosa.f90:4240-4297): inviscid radii into lambdas(1:3), viscous
 * ones into lamv(1:3) (the reference's work = dq).  mut enters through
 * work(:,:,:,4) for k-omega and is zero otherwise.  As in the reference,
 * the temperature of the +1 neighbour divides by the density of the -1
 * neighbour.                                                              */
void or_lambdab_v(const double* q, const double* si, const double* sj, const double* sk, const double* xdot,
                  const double* ydot, const double* zdot, const double* vol, const double* mut, double* lambdas,
                  double* lamv, int imax, int jmax, int kmax) {
  bg g = mkbg(imax, jmax, kmax);
  double gamma = P.gamma;
  int kom = is_kom();
  for (int n = 0; n <= 2 * NH; n++) {
    int nh = n * P.hbmove;
    for (int k = 1; k <= kmax - 1; k++)
      for (int j = 1; j <= jmax - 1; j++)
        for (int i = 1; i <= imax - 1; i++) {
          int e[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
          double rho = Q(q, g, i, j, k, 1, n), p = Q(q, g, i, j, k, 5, n);
          double t = gamma * p / rho;
          double mu = (P.stemp + 1.0) / (t + P.stemp) * pow(t, 1.5);
          double mutc = kom ? S1(mut, g, i, j, k, n) : 0.0;
          double mud[3], mutd[3];
          for (int d = 0; d < 3; d++) {
            int a = e[d][0], b = e[d][1], c = e[d][2];
            double rhom = Q(q, g, i - a, j - b, k - c, 1, n), pm = Q(q, g, i - a, j - b, k - c, 5, n);
            double tm = gamma * pm / rhom;
            double mum = (P.stemp + 1.0) / (tm + P.stemp) * pow(tm, 1.5);
            double mutm = kom ? S1(mut, g, i - a, j - b, k - c, n) : 0.0;
            double pp = Q(q, g, i + a, j + b, k + c, 5, n);
            double tp = gamma * pp / rhom;
            double mup = (P.stemp + 1.0) / (tp + P.stemp) * pow(tp, 1.5);
            double mutp = kom ? S1(mut, g, i + a, j + b, k + c, n) : 0.0;
            mud[d] = (mum + 2 * mu + mup) / 4;
            mutd[d] = (mutm + 2 * mutc + mutp) / 4;
          }
          double sim[4], sip[4], sjm[4], sjp[4], skm[4], skp[4];
          for (int m = 0; m < 4; m++) {
            sim[m] = M(si, g, m + 1, i, j, k, nh);
            sip[m] = M(si, g, m + 1, i + 1, j, k, nh);
            sjm[m] = M(sj, g, m + 1, i, j, k, nh);
            sjp[m] = M(sj, g, m + 1, i, j + 1, k, nh);
            skm[m] = M(sk, g, m + 1, i, j, k, nh);
            skp[m] = M(sk, g, m + 1, i, j, k + 1, nh);
          }
          double u = Q(q, g, i, j, k, 2, n), v = Q(q, g, i, j, k, 3, n), w = Q(q, g, i, j, k, 4, n);
          double a = sqrt(t);
          double xix = 0.5 * (sim[0] * sim[3] + sip[0] * sip[3]);
          double xiy = 0.5 * (sim[1] * sim[3] + sip[1] * sip[3]);
          double xiz = 0.5 * (sim[2] * sim[3] + sip[2] * sip[3]);
          double etax = 0.5 * (sjm[0] * sjm[3] + sjp[0] * sjp[3]);
          double etay = 0.5 * (sjm[1] * sjm[3] + sjp[1] * sjp[3]);
          double etaz = 0.5 * (sjm[2] * sjm[3] + sjp[2] * sjp[3]);
          double zetax = 0.5 * (skm[0] * skm[3] + skp[0] * skp[3]);
          double zetay = 0.5 * (skm[1] * skm[3] + skp[1] * skp[3]);
          double zetaz = 0.5 * (skm[2] * skm[3] + skp[2] * skp[3]);
          double axi = (sim[3] + sip[3]) / 2, aeta = (sjm[3] + sjp[3]) / 2, azeta = (skm[3] + skp[3]) / 2;
          if (P.moving) {
            u = u - X2(xdot, g, i, j, k, nh);
            v = v - X2(ydot, g, i, j, k, nh);
            w = w - X2(zdot, g, i, j, k, nh);
          }
          double volc = V1(vol, g, i, j, k);
          Q(lambdas, g, i, j, k, 1, n) = fabs(u * xix + v * xiy + w * xiz) + a * axi;
          Q(lambdas, g, i, j, k, 2, n) = fabs(u * etax + v * etay + w * etaz) + a * aeta;
          Q(lambdas, g, i, j, k, 3, n) = fabs(u * zetax + v * zetay + w * zetaz) + a * azeta;
          double ar_[3] = {axi, aeta, azeta};
          for (int d = 0; d < 3; d++)
            Q(lamv, g, i, j, k, d + 1, n) = P.cdff * ar_[d] * ar_[d] * gamma * P.machfs / P.reyno *
                                            (mud[d] / P.pranl + mutd[d] / P.prant) / (rho * volc);
        }
  }
}

/* CFL ramping of k-omega runs (bdeltat, src/cosa.f90:4760-4820): cfl and
 * cflt (kept, as the reference's globals) from the cycle counter MGIT;
 * ramping1 refreshes at cycle 1 and every 100th, ramping2 at every call. */
static int MGIT = 1;
void or_set_mgit(int m) { MGIT = m; }
static void cfl_ramp(void) {
  if (!is_kom()) return;
  const int mgit = MGIT;
  const int* nc = P.ncycr; /* ncycr(0:3) */
  if (P.ramping == 1 && (mgit == 1 || mgit / 100 * 100 == mgit)) {
    const double sqrt3 = sqrt(3.0);
    const double s33 = sqrt3 * sqrt3 * sqrt3;
    const double dcfli1 = (P.vcfli[1] - P.vcfli[0]) * s33, dcfli2 = P.vcfli[2] - P.vcfli[1];
    const double dcflit1 = (P.vcflit[1] - P.vcflit[0]) * s33, dcflit2 = P.vcflit[2] - P.vcflit[1];
    if (mgit <= nc[1]) {
      P.cfl = P.vcfli[0];
      P.cflt = P.vcflit[0];
    } else if (mgit > nc[1] && mgit <= nc[2]) {
      const double x = (double)(mgit - nc[1]) / (nc[2] - nc[1]) / sqrt3;
      P.cfl = P.vcfli[0] + dcfli1 * (x * x * x);
      P.cflt = P.vcflit[0] + dcflit1 * (x * x * x);
    } else if (mgit > nc[2] && mgit <= nc[3]) {
      P.cfl = P.vcfli[1] + dcfli2 * ((double)(mgit - nc[2]) / (nc[3] - nc[2]));
      P.cflt = P.vcflit[1] + dcflit2 * ((double)(mgit - nc[2]) / (nc[3] - nc[2]));
    } else if (mgit > nc[3]) {
      P.cfl = P.vcfli[2];
      P.cflt = P.vcflit[2];
    }
  } else if (P.ramping == 2) {
    const double dcfli1 = P.vcfli[1] - P.vcfli[0], dcfli2 = P.vcfli[2] - P.vcfli[1];
    const double dcflit1 = P.vcflit[1] - P.vcflit[0], dcflit2 = P.vcflit[2] - P.vcflit[1];
    if (mgit <= nc[0]) {
      P.cfl = P.vcfli[0];
      P.cflt = P.vcflit[0];
    } else if (mgit > nc[0] && mgit <= nc[1]) {
      P.cfl = P.vcfli[0] + dcfli1 * ((double)(mgit - nc[0]) / (nc[1] - nc[0]));
      P.cflt = P.vcflit[0] + dcflit1 * ((double)(mgit - nc[0]) / (nc[1] - nc[0]));
    } else if (mgit > nc[1] && mgit <= nc[2]) {
      P.cfl = P.vcfli[1];
      P.cflt = P.vcflit[1];
    } else if (mgit > nc[2] && mgit <= nc[3]) {
      P.cfl = P.vcfli[1] + dcfli2 * ((double)(mgit - nc[2]) / (nc[3] - nc[2]));
      P.cflt = P.vcflit[1] + dcflit2 * ((double)(mgit - nc[2]) / (nc[3] - nc[2]));
    } else if (mgit > nc[3]) {
      P.cfl = P.vcfli[2];
      P.cflt = P.vcflit[2];
    }
  }
}
double or_get_cfl(int turb) { return turb ? P.cflt : P.cfl; }

/* bdeltat (src/cosa.f90:4733-5203) without CFL ramping: lamv (viscous
 * radii, NULL when inviscid) join the sum of dtvol; k-omega adds
 * dtvolt = cflt/cfl * dtvol and the turbulence CIRS coefficients tbeta*
 * (same forms with ratt).                                                 */
void or_bdeltat_v(double* dtvol, double* dtvolt, const double* lambdas, const double* lamv, double* betaxi,
                  double* betaeta, double* betazeta, double* tbetaxi, double* tbetaeta, double* tbetazeta, int imax,
                  int jmax, int kmax) {
  cfl_ramp();
  bg g = mkbg(imax, jmax, kmax);
  const double zeta = 0.4;
  int kom = is_kom();
  for (int n = 0; n <= 2 * NH; n++) {
    for (int k = 1; k <= kmax - 1; k++)
      for (int j = 1; j <= jmax - 1; j++)
        for (int i = 1; i <= imax - 1; i++) {
          double sum = Q(lambdas, g, i, j, k, 1, n) + Q(lambdas, g, i, j, k, 2, n) + Q(lambdas, g, i, j, k, 3, n);
          /* irsop 5/6: viscous radii in the smoothing only (src/cosa.f90:4831-4890) */
          if (lamv && P.irsop != 5 && P.irsop != 6)
            sum = sum + Q(lamv, g, i, j, k, 1, n) + Q(lamv, g, i, j, k, 2, n) + Q(lamv, g, i, j, k, 3, n);
          P1(dtvol, g, i, j, k, n) = P.cfl / sum;
        }
    if (kom)
      for (int k = 1; k <= kmax - 1; k++)
        for (int j = 1; j <= jmax - 1; j++)
          for (int i = 1; i <= imax - 1; i++) P1(dtvolt, g, i, j, k, n) = P.cflt / P.cfl * P1(dtvol, g, i, j, k, n);
  }
  for (int pass = 0; pass < (kom ? 2 : 1); pass++) {
    double rr = pass == 0 ? P.rat : P.ratt;
    double* bx_ = pass == 0 ? betaxi : tbetaxi;
    double* by_ = pass == 0 ? betaeta : tbetaeta;
    double* bz_ = pass == 0 ? betazeta : tbetazeta;
    if (P.irsop == 2 || P.irsop == 5) {
      for (int n = 0; n <= 2 * NH; n++)
        for (int k = 1; k <= kmax - 1; k++)
          for (int j = 1; j <= jmax - 1; j++)
            for (int i = 1; i <= imax - 1; i++) {
              double lami = Q(lambdas, g, i, j, k, 1, n), lamj = Q(lambdas, g, i, j, k, 2, n),
                     lamk = Q(lambdas, g, i, j, k, 3, n);
              double phixi = 1 + pow(lamj / lami, zeta) + pow(lamk / lami, zeta);
              double phieta = 1 + pow(lami / lamj, zeta) + pow(lamk / lamj, zeta);
              double phizeta = 1 + pow(lami / lamk, zeta) + pow(lamj / lamk, zeta);
              double t;
              t = rr * lami * phixi / (lami + lamj + lamk);
              P1(bx_, g, i, j, k, n) = fmax(0.0, (t * t - 1) / 4);
              t = rr * lamj * phieta / (lami + lamj + lamk);
              P1(by_, g, i, j, k, n) = fmax(0.0, (t * t - 1) / 4);
              t = rr * lamk * phizeta / (lami + lamj + lamk);
              P1(bz_, g, i, j, k, n) = fmax(0.0, (t * t - 1) / 4);
              if (P.irsop == 5) { /* cirs_v1v, src/cosa.f90:4981-5060 */
                const double c1 = 40.0;
                double lv1 = lamv ? Q(lamv, g, i, j, k, 1, n) : 0.0, lv2 = lamv ? Q(lamv, g, i, j, k, 2, n) : 0.0,
                       lv3 = lamv ? Q(lamv, g, i, j, k, 3, n) : 0.0;
                P1(bx_, g, i, j, k, n) = fmax(P1(bx_, g, i, j, k, n), c1 / 4 * lv1 / (lami + lamj + lamk));
                P1(by_, g, i, j, k, n) = fmax(P1(by_, g, i, j, k, n), c1 / 4 * lv2 / (lami + lamj + lamk));
                P1(bz_, g, i, j, k, n) = fmax(P1(bz_, g, i, j, k, n), c1 / 4 * lv3 / (lami + lamj + lamk));
              }
            }
    } else if (P.irsop == 3 || P.irsop == 6) {
      for (int n = 0; n <= 2 * NH; n++)
        for (int k = 1; k <= kmax - 1; k++)
          for (int j = 1; j <= jmax - 1; j++)
            for (int i = 1; i <= imax - 1; i++) {
              double lami = Q(lambdas, g, i, j, k, 1, n), lamj = Q(lambdas, g, i, j, k, 2, n),
                     lamk = Q(lambdas, g, i, j, k, 3, n);
              double rji = lamj / lami, rki = lamk / lami, rij = lami / lamj, rkj = lamk / lamj, rik = lami / lamk,
                     rjk = lamj / lamk;
              double t;
              t = rr / (1 + P.psirs * (rji + rki));
              P1(bx_, g, i, j, k, n) = fmax(0.0, 0.25 * (t * t - 1));
              t = rr / (1 + P.psirs * (rij + rkj));
              P1(by_, g, i, j, k, n) = fmax(0.0, 0.25 * (t * t - 1));
              t = rr / (1 + P.psirs * (rik + rjk));
              P1(bz_, g, i, j, k, n) = fmax(0.0, 0.25 * (t * t - 1));
              if (P.irsop == 6) { /* cirs_v2v, src/cosa.f90:4896-4975 */
                const double c1 = 40.0;
                double lv1 = lamv ? Q(lamv, g, i, j, k, 1, n) : 0.0, lv2 = lamv ? Q(lamv, g, i, j, k, 2, n) : 0.0,
                       lv3 = lamv ? Q(lamv, g, i, j, k, 3, n) : 0.0;
                P1(bx_, g, i, j, k, n) = fmax(P1(bx_, g, i, j, k, n), 0.25 * c1 * lv1 / (lami + lamj + lamk));
                P1(by_, g, i, j, k, n) = fmax(P1(by_, g, i, j, k, n), 0.25 * c1 * lv2 / (lami + lamj + lamk));
                P1(bz_, g, i, j, k, n) = fmax(P1(bz_, g, i, j, k, n), 0.25 * c1 * lv3 / (lami + lamj + lamk));
              }
            }
    }
  }
}

/* bdeltat (src/cosa.f90:4733-5203), inviscid, no k-omega ramping:
 * dtvol = cfl / sum(lambda); CIRS coefficients for irsop = 2 and 3.       */
void or_bdeltat(double* dtvol, const double* lambdas, double* betaxi, double* betaeta, double* betazeta, int imax,
                int jmax, int kmax) {
  bg g = mkbg(imax, jmax, kmax);
  const double zeta = 0.4;
  for (int n = 0; n <= 2 * NH; n++)
    for (int k = 1; k <= kmax - 1; k++)
      for (int j = 1; j <= jmax - 1; j++)
        for (int i = 1; i <= imax - 1; i++)
          P1(dtvol, g, i, j, k, n) = P.cfl / (Q(lambdas, g, i, j, k, 1, n) + Q(lambdas, g, i, j, k, 2, n) +
                                              Q(lambdas, g, i, j, k, 3, n));
  if (P.irsop == 2 || P.irsop == 5) {
    for (int n = 0; n <= 2 * NH; n++)
      for (int k = 1; k <= kmax - 1; k++)
        for (int j = 1; j <= jmax - 1; j++)
          for (int i = 1; i <= imax - 1; i++) {
            double lami = Q(lambdas, g, i, j, k, 1, n), lamj = Q(lambdas, g, i, j, k, 2, n),
                   lamk = Q(lambdas, g, i, j, k, 3, n);
            double phixi = 1 + pow(lamj / lami, zeta) + pow(lamk / lami, zeta);
            double phieta = 1 + pow(lami / lamj, zeta) + pow(lamk / lamj, zeta);
            double phizeta = 1 + pow(lami / lamk, zeta) + pow(lamj / lamk, zeta);
            double t, bx, by, bz;
            t = P.rat * lami * phixi / (lami + lamj + lamk);
            bx = (t * t - 1) / 4;
            P1(betaxi, g, i, j, k, n) = fmax(0.0, bx);
            t = P.rat * lamj * phieta / (lami + lamj + lamk);
            by = (t * t - 1) / 4;
            P1(betaeta, g, i, j, k, n) = fmax(0.0, by);
            t = P.rat * lamk * phizeta / (lami + lamj + lamk);
            bz = (t * t - 1) / 4;
            P1(betazeta, g, i, j, k, n) = fmax(0.0, bz);
          }
  } else if (P.irsop == 3 || P.irsop == 6) {
    for (int n = 0; n <= 2 * NH; n++)
      for (int k = 1; k <= kmax - 1; k++)
        for (int j = 1; j <= jmax - 1; j++)
          for (int i = 1; i <= imax - 1; i++) {
            double lami = Q(lambdas, g, i, j, k, 1, n), lamj = Q(lambdas, g, i, j, k, 2, n),
                   lamk = Q(lambdas, g, i, j, k, 3, n);
            double rji = lamj / lami, rki = lamk / lami, rij = lami / lamj, rkj = lamk / lamj, rik = lami / lamk,
                   rjk = lamj / lamk;
            double t, bx, by, bz;
            t = P.rat / (1 + P.psirs * (rji + rki));
            bx = 0.25 * (t * t - 1);
            P1(betaxi, g, i, j, k, n) = fmax(0.0, bx);
            t = P.rat / (1 + P.psirs * (rij + rkj));
            by = 0.25 * (t * t - 1);
            P1(betaeta, g, i, j, k, n) = fmax(0.0, by);
            t = P.rat / (1 + P.psirs * (rik + rjk));
            bz = 0.25 * (t * t - 1);
            P1(betazeta, g, i, j, k, n) = fmax(0.0, bz);
          }
  }
}

/* ---------------------------------------------------------------------- */
/* MUSCL reconstruction, limiters, Roe flux, flux difference               */

/* direction-generic indexing: position a along idir, (b,c) transverse in
 * the reference's loop order (idir 1: j,k; idir 2: i,k; idir 3: i,j).     */
static ll dI3(bg g, int idir, int a, int t1, int t2) {
  if (idir == 1) return I3(g, a, t1, t2);
  if (idir == 2) return I3(g, t1, a, t2);
  return I3(g, t1, t2, a);
}
#define QD(arr, v, n, idx) (arr)[(idx) + g.n3 * ((ll)((v) - 1) + (ll)NPDE * (n))]

static int strong_bc(int t) {
  return t == 13 || t == 14 || t / 100 == 15 || t == 22 || t == 23 || t == 92 || t == 93 || t == 5;
}

/* limit_i/j/k (src/flux.f90:3472-3888) along idir for variables vs..ve.  */
static void or_limit(int idir, const double* dq, double* dqp, double* dqm, double epslim, double rkap, int vs, int ve,
                     int ilim, bg g) {
  int L = idir == 1 ? g.imax : (idir == 2 ? g.jmax : g.kmax);
  int T1 = idir == 1 ? g.jmax : g.imax;
  int T2 = idir == 3 ? g.jmax : g.kmax;
  for (int n = 0; n <= 2 * NH; n++)
    for (int v = vs; v <= ve; v++)
      for (int t2 = 1; t2 <= T2 - 1; t2++)
        for (int t1 = 1; t1 <= T1 - 1; t1++) {
          if (ilim == 0) {
            double rkp = 1 + rkap, rkm = 1 - rkap;
            for (int a = 1; a <= L; a++) {
              double dm1 = QD(dq, v, n, dI3(g, idir, a - 1, t1, t2)), d0 = QD(dq, v, n, dI3(g, idir, a, t1, t2)),
                     dp1 = QD(dq, v, n, dI3(g, idir, a + 1, t1, t2));
              QD(dqp, v, n, dI3(g, idir, a, t1, t2)) = -(rkp * d0 + rkm * dp1) / 4;
              QD(dqm, v, n, dI3(g, idir, a, t1, t2)) = (rkm * dm1 + rkp * d0) / 4;
            }
          } else if (ilim == 1) {
            for (int a = 1; a <= L; a++) {
              double dm1 = QD(dq, v, n, dI3(g, idir, a - 1, t1, t2)), d0 = QD(dq, v, n, dI3(g, idir, a, t1, t2)),
                     dp1 = QD(dq, v, n, dI3(g, idir, a + 1, t1, t2));
              double s = (2 * dp1 * d0 + epslim) / (dp1 * dp1 + d0 * d0 + epslim);
              double rkp = 1 + rkap * s, rkm = 1 - rkap * s;
              QD(dqp, v, n, dI3(g, idir, a, t1, t2)) = -s * (rkp * d0 + rkm * dp1) / 4;
              s = (2 * d0 * dm1 + epslim) / (d0 * d0 + dm1 * dm1 + epslim);
              rkp = 1 + rkap * s;
              rkm = 1 - rkap * s;
              QD(dqm, v, n, dI3(g, idir, a, t1, t2)) = s * (rkm * dm1 + rkp * d0) / 4;
            }
          } else if (ilim == 2) {
            double rkp = 1 + rkap, rkm = 1 - rkap, b = (3.0 - rkap) / (1.0 - rkap);
            for (int a = 0; a <= L; a++) {
              double d0 = QD(dq, v, n, dI3(g, idir, a, t1, t2)), d1 = QD(dq, v, n, dI3(g, idir, a + 1, t1, t2));
              double dqmx = (copysign(1.0, d0) + copysign(1.0, d1)) / 2 * fmin(fabs(d0), fabs(d1) * b);
              double dqpx = (copysign(1.0, d1) + copysign(1.0, d0)) / 2 * fmin(fabs(d1), fabs(d0) * b);
              QD(dqp, v, n, dI3(g, idir, a, t1, t2)) = -(rkp * dqmx + rkm * dqpx) / 4;
              QD(dqm, v, n, dI3(g, idir, a + 1, t1, t2)) = (rkm * dqmx + rkp * dqpx) / 4;
            }
          } else if (ilim == 3) {
            double epsq = epslim * epslim;
            for (int a = 0; a <= L; a++) {
              double d1 = QD(dq, v, n, dI3(g, idir, a + 1, t1, t2)), d2 = QD(dq, v, n, dI3(g, idir, a, t1, t2));
              double dqmx = (d1 * (d2 * d2 + 2 * epsq) + d2 * (2 * d1 * d1 + epsq)) /
                            (2 * d1 * d1 - d1 * d2 + 2 * d2 * d2 + 3 * epsq);
              double dqpx = (d2 * (d1 * d1 + 2 * epsq) + d1 * (2 * d2 * d2 + epsq)) /
                            (2 * d2 * d2 - d2 * d1 + 2 * d1 * d1 + 3 * epsq);
              QD(dqp, v, n, dI3(g, idir, a, t1, t2)) = -dqpx / 2;
              QD(dqm, v, n, dI3(g, idir, a + 1, t1, t2)) = dqmx / 2;
            }
          } else if (ilim == 4) {
            double epsq = epslim * epslim;
            for (int a = 0; a <= L; a++) {
              double d1 = QD(dq, v, n, dI3(g, idir, a + 1, t1, t2)), d2 = QD(dq, v, n, dI3(g, idir, a, t1, t2));
              double dqmx = (d1 * (d2 * d2 + epsq) + d2 * (d1 * d1 + epsq)) / (d1 * d1 + d2 * d2 + 2 * epsq);
              double dqpx = dqmx;
              QD(dqp, v, n, dI3(g, idir, a, t1, t2)) = -dqpx / 2;
              QD(dqm, v, n, dI3(g, idir, a + 1, t1, t2)) = dqmx / 2;
            }
          }
        }
}

/* muscl_bi/bj/bk (src/flux.f90:1221-1812) */
void or_muscl_b(int idir, int nl, const int32_t* bctopo, int nbcs, const double* q, double* qp, double* qm, double* dq,
                double* dqp, double* dqm, int imax, int jmax, int kmax) {
  bg g = mkbg(imax, jmax, kmax);
  int L = idir == 1 ? imax : (idir == 2 ? jmax : kmax);
  int T1 = idir == 1 ? jmax : imax;
  int T2 = idir == 3 ? jmax : kmax;
  int fo_beg = 0, fo_end = -1, so_beg = 1, so_end = NPDE, dofirst = 0, doscnd = 1;
  if (nl >= P.nl_crs) {
    fo_beg = 1; fo_end = NPDE; so_beg = 0; so_end = -1; dofirst = 1; doscnd = 0;
  } else if (P.foturb) {
    fo_beg = 6; fo_end = NPDE; so_beg = 1; so_end = 5; dofirst = 1; doscnd = 1;
  }
  /* transverse index ranges of a patch: idir 1 -> (6,7),(8,9); idir 2 ->
     (4,5),(8,9); idir 3 -> (4,5),(6,7) */
  int r1 = idir == 1 ? 5 : 3, r2 = idir == 3 ? 5 : 7;
  if (dofirst)
    for (int n = 0; n <= 2 * NH; n++)
      for (int v = fo_beg; v <= fo_end; v++)
        for (int t2 = 1; t2 <= T2 - 1; t2++)
          for (int t1 = 1; t1 <= T1 - 1; t1++)
            for (int a = 1; a <= L; a++) {
              QD(qp, v, n, dI3(g, idir, a, t1, t2)) = QD(q, v, n, dI3(g, idir, a, t1, t2));
              QD(qm, v, n, dI3(g, idir, a, t1, t2)) = QD(q, v, n, dI3(g, idir, a - 1, t1, t2));
            }
  if (doscnd) {
    for (int n = 0; n <= 2 * NH; n++)
      for (int v = so_beg; v <= so_end; v++)
        for (int t2 = 1; t2 <= T2 - 1; t2++)
          for (int t1 = 1; t1 <= T1 - 1; t1++)
            for (int a = 0; a <= L + 1; a++)
              QD(dq, v, n, dI3(g, idir, a, t1, t2)) =
                  QD(q, v, n, dI3(g, idir, a, t1, t2)) - QD(q, v, n, dI3(g, idir, a - 1, t1, t2));
    for (int ibc = 0; ibc < nbcs; ibc++) {
      const int32_t* t = bctopo + 10 * ibc;
      if (t[1] != idir || !strong_bc(t[0])) continue;
      int abc = t[2] == 1 ? 1 : L;
      for (int n = 0; n <= 2 * NH; n++)
        for (int v = so_beg; v <= so_end; v++)
          for (int t2 = t[r2]; t2 <= t[r2 + 1]; t2++)
            for (int t1 = t[r1]; t1 <= t[r1 + 1]; t1++)
              QD(dq, v, n, dI3(g, idir, abc, t1, t2)) =
                  (QD(q, v, n, dI3(g, idir, abc, t1, t2)) - QD(q, v, n, dI3(g, idir, abc - 1, t1, t2))) * 2;
    }
    or_limit(idir, dq, dqp, dqm, P.epslimf, P.rkap, 1, 5, P.ilimf, g);
    or_limit(idir, dq, dqp, dqm, P.epslimt, P.rkap, 6, NPDE, P.ilimt, g);
    for (int n = 0; n <= 2 * NH; n++)
      for (int v = so_beg; v <= so_end; v++)
        for (int t2 = 1; t2 <= T2 - 1; t2++)
          for (int t1 = 1; t1 <= T1 - 1; t1++)
            for (int a = 1; a <= L; a++) {
              QD(qp, v, n, dI3(g, idir, a, t1, t2)) =
                  QD(q, v, n, dI3(g, idir, a, t1, t2)) + QD(dqp, v, n, dI3(g, idir, a, t1, t2));
              QD(qm, v, n, dI3(g, idir, a, t1, t2)) =
                  QD(q, v, n, dI3(g, idir, a - 1, t1, t2)) + QD(dqm, v, n, dI3(g, idir, a, t1, t2));
            }
  }
  for (int ibc = 0; ibc < nbcs; ibc++) {
    const int32_t* t = bctopo + 10 * ibc;
    if (t[1] != idir || !strong_bc(t[0])) continue;
    int aqpm = t[2] == 1 ? 1 : L;
    int abc = t[2] == 1 ? 0 : L;
    for (int n = 0; n <= 2 * NH; n++)
      for (int v = 1; v <= NPDE; v++)
        for (int t2 = t[r2]; t2 <= t[r2 + 1]; t2++)
          for (int t1 = t[r1]; t1 <= t[r1 + 1]; t1++) {
            double val = QD(q, v, n, dI3(g, idir, abc, t1, t2));
            QD(qp, v, n, dI3(g, idir, aqpm, t1, t2)) = val;
            QD(qm, v, n, dI3(g, idir, aqpm, t1, t2)) = val;
          }
  }
}

/* turbulence part of roflux_b (src/flux.f90:250-383) for snapshot n */
static void or_roflux_turb(const double* qr, const double* ql, double* flux, const double* sijk, bg g, int n, int ic,
                           int jc, int kc, double h0l, double h0r) {
  double gamma = P.gamma;
  int nh = n * P.hbmove;
  for (int k = 1; k <= kc; k++)
    for (int j = 1; j <= jc; j++)
      for (int i = 1; i <= ic; i++) {
        double ubn = 0.0;
        double kx = M(sijk, g, 1, i, j, k, nh), ky = M(sijk, g, 2, i, j, k, nh), kz = M(sijk, g, 3, i, j, k, nh),
               ds = M(sijk, g, 4, i, j, k, nh);
        if (P.moving)
          ubn = M(sijk, g, 5, i, j, k, nh) * kx + M(sijk, g, 6, i, j, k, nh) * ky + M(sijk, g, 7, i, j, k, nh) * kz;
        double rhor = Q(qr, g, i, j, k, 1, n), ur = Q(qr, g, i, j, k, 2, n), vr = Q(qr, g, i, j, k, 3, n),
               wr = Q(qr, g, i, j, k, 4, n), pr = Q(qr, g, i, j, k, 5, n), omer = Q(qr, g, i, j, k, 7, n);
        double unr = ur * kx + vr * ky + wr * kz - ubn;
        double tker = Q(qr, g, i, j, k, 6, n);
        double fr1 = rhor * unr * tker, fr2 = rhor * unr * omer;
        double rhol = Q(ql, g, i, j, k, 1, n), ul = Q(ql, g, i, j, k, 2, n), vl = Q(ql, g, i, j, k, 3, n),
               wl = Q(ql, g, i, j, k, 4, n), pl = Q(ql, g, i, j, k, 5, n), omel = Q(ql, g, i, j, k, 7, n);
        double unl = ul * kx + vl * ky + wl * kz - ubn;
        double tkel = Q(ql, g, i, j, k, 6, n);
        double fl1 = rhol * unl * tkel, fl2 = rhol * unl * omel;
        double drho = rhor - rhol, dp = pr - pl, dun = unr - unl, dtke = tker - tkel, dome = omer - omel;
        double rhols = sqrt(rhol), rhors = sqrt(rhor);
        double rho = rhols * rhors;
        double denom = rhols + rhors;
        double u = (rhols * ul + rhors * ur) / denom;
        double v = (rhols * vl + rhors * vr) / denom;
        double w = (rhols * wl + rhors * wr) / denom;
        double un = u * kx + v * ky + w * kz;
        double h0 = (rhols * h0l + rhors * h0r) / denom;
        double tke = (rhols * tkel + rhors * tker) / denom;
        double ome = (rhols * omel + rhors * omer) / denom;
        double q2 = u * u + v * v + w * w;
        double a2 = (gamma - 1) * (h0 - q2 / 2 - tke);
        double a = sqrt(a2);
        double rm123 = fabs(un - ubn), rm4 = fabs(un + a - ubn), rm5 = fabs(un - a - ubn);
        if (P.efx_setup_t) {
          double rdelta = P.cntrpy_t * (P.entfx_t_w[0] * P.entfxctf_t + P.entfx_t_w[1] * (sqrt(u * u + v * v + w * w) + a) +
                                        P.entfx_t_w[2] * 0.2 * (fabs(un) + a + fabs(ubn)));
          rdelta = fmax(rdelta, 1.e-12);
          if (rm123 <= rdelta && P.efx_all_t) rm123 = 0.5 * (rm123 * rm123 + rdelta * rdelta) / rdelta;
          if (rm4 <= rdelta) rm4 = 0.5 * (rm4 * rm4 + rdelta * rdelta) / rdelta;
          if (rm5 <= rdelta) rm5 = 0.5 * (rm5 * rm5 + rdelta * rdelta) / rdelta;
        }
        if (P.eig_limit_t) {
          if (rm123 <= P.eig_cutoff_t) rm123 = P.eig_cutoff_t;
          if (P.eig_lim_all_t && rm4 <= P.eig_cutoff_t) rm4 = P.eig_cutoff_t;
          if (P.eig_lim_all_t && rm5 <= P.eig_cutoff_t) rm5 = P.eig_cutoff_t;
        }
        double al1 = rm123 * (drho - dp / a2);
        double al2 = rm123 * rho;
        double al3 = rm4 * (dp / a2 + rho * dun / a) / 2;
        double al4 = rm5 * (dp / a2 - rho * dun / a) / 2;
        double df1 = (al1 + al3 + al4) * tke + al2 * dtke;
        double df2 = (al1 + al3 + al4) * ome + al2 * dome;
        F0(flux, g, i, j, k, 6, n) = (fr1 + fl1 - df1) / 2 * ds;
        F0(flux, g, i, j, k, 7, n) = (fr2 + fl2 - df2) / 2 * ds;
      }
}

/* roflux_b (src/flux.f90:45-387).  With k-omega the turbulent variables
 * are upwinded with the mean-flow waves (src/flux.f90:250-383); as in the
 * reference, that second loop reuses h0l/h0r of the last face of the
 * mean-flow loop of the same snapshot for its Roe-averaged h0.            */
void or_roflux_b(int idir, const double* qr, const double* ql, double* flux, const double* sijk, int imax, int jmax,
                 int kmax) {
  bg g = mkbg(imax, jmax, kmax);
  int ic = idir == 1 ? imax : imax - 1, jc = idir == 2 ? jmax : jmax - 1, kc = idir == 3 ? kmax : kmax - 1;
  double gamma = P.gamma;
  int kom = is_kom();
  double h0l_last = 0.0, h0r_last = 0.0;
  for (int n = 0; n <= 2 * NH; n++) {
    int nh = n * P.hbmove;
    for (int k = 1; k <= kc; k++)
      for (int j = 1; j <= jc; j++)
        for (int i = 1; i <= ic; i++) {
          double ubn = 0.0;
          double kx = M(sijk, g, 1, i, j, k, nh), ky = M(sijk, g, 2, i, j, k, nh), kz = M(sijk, g, 3, i, j, k, nh),
                 ds = M(sijk, g, 4, i, j, k, nh);
          if (P.moving)
            ubn = M(sijk, g, 5, i, j, k, nh) * kx + M(sijk, g, 6, i, j, k, nh) * ky + M(sijk, g, 7, i, j, k, nh) * kz;
          double fr[5], fl[5], df[5];
          double rhor = Q(qr, g, i, j, k, 1, n), ur = Q(qr, g, i, j, k, 2, n), vr = Q(qr, g, i, j, k, 3, n),
                 wr = Q(qr, g, i, j, k, 4, n), pr = Q(qr, g, i, j, k, 5, n);
          double unr = ur * kx + vr * ky + wr * kz - ubn;
          double q2r = ur * ur + vr * vr + wr * wr;
          double tker = kom ? Q(qr, g, i, j, k, 6, n) : 0.0;
          double h0r = gamma / (gamma - 1) * pr / rhor + q2r / 2 + tker;
          double rh0r = rhor * h0r;
          fr[0] = rhor * unr;
          fr[1] = (rhor * ur * unr + pr * kx);
          fr[2] = (rhor * vr * unr + pr * ky);
          fr[3] = (rhor * wr * unr + pr * kz);
          fr[4] = (rh0r * unr + pr * ubn);
          double rhol = Q(ql, g, i, j, k, 1, n), ul = Q(ql, g, i, j, k, 2, n), vl = Q(ql, g, i, j, k, 3, n),
                 wl = Q(ql, g, i, j, k, 4, n), pl = Q(ql, g, i, j, k, 5, n);
          double unl = ul * kx + vl * ky + wl * kz - ubn;
          double q2l = ul * ul + vl * vl + wl * wl;
          double tkel = kom ? Q(ql, g, i, j, k, 6, n) : 0.0;
          double h0l = gamma / (gamma - 1) * pl / rhol + q2l / 2 + tkel;
          double rh0l = rhol * h0l;
          fl[0] = rhol * unl;
          fl[1] = (rhol * ul * unl + pl * kx);
          fl[2] = (rhol * vl * unl + pl * ky);
          fl[3] = (rhol * wl * unl + pl * kz);
          fl[4] = (rh0l * unl + pl * ubn);
          double drho = rhor - rhol, dp = pr - pl, du = ur - ul, dv = vr - vl, dw = wr - wl, dun = unr - unl,
                 dtke = tker - tkel;
          double rhols = sqrt(rhol), rhors = sqrt(rhor);
          double rho = rhols * rhors;
          double denom = rhols + rhors;
          double u = (rhols * ul + rhors * ur) / denom;
          double v = (rhols * vl + rhors * vr) / denom;
          double w = (rhols * wl + rhors * wr) / denom;
          double un = u * kx + v * ky + w * kz;
          double h0 = (rhols * h0l + rhors * h0r) / denom;
          double tke = (rhols * tkel + rhors * tker) / denom;
          double q2 = u * u + v * v + w * w;
          double a2 = (gamma - 1) * (h0 - q2 / 2 - tke);
          double a = sqrt(a2);
          double rm123 = fabs(un - ubn), rm4 = fabs(un + a - ubn), rm5 = fabs(un - a - ubn);
          if (P.efx_setup_f) {
            double rdelta = P.cntrpy_f * (P.entfx_f_w[0] * P.entfxctf_f + P.entfx_f_w[1] * (sqrt(u * u + v * v + w * w) + a) +
                                          P.entfx_f_w[2] * 0.2 * (fabs(un) + a + fabs(ubn)));
            rdelta = fmax(rdelta, 1.e-12);
            if (rm123 <= rdelta && P.efx_all_f) rm123 = 0.5 * (rm123 * rm123 + rdelta * rdelta) / rdelta;
            if (rm4 <= rdelta) rm4 = 0.5 * (rm4 * rm4 + rdelta * rdelta) / rdelta;
            if (rm5 <= rdelta) rm5 = 0.5 * (rm5 * rm5 + rdelta * rdelta) / rdelta;
          }
          if (P.eig_limit_f) {
            if (rm123 <= P.eig_cutoff_f) rm123 = P.eig_cutoff_f;
            if (P.eig_lim_all_f && rm4 <= P.eig_cutoff_f) rm4 = P.eig_cutoff_f;
            if (P.eig_lim_all_f && rm5 <= P.eig_cutoff_f) rm5 = P.eig_cutoff_f;
          }
          double al1 = rm123 * (drho - dp / a2);
          double al2 = rm123 * rho;
          double al3 = rm4 * (dp / a2 + rho * dun / a) / 2;
          double al4 = rm5 * (dp / a2 - rho * dun / a) / 2;
          df[0] = al1 + al3 + al4;
          df[1] = al1 * u + al2 * (du - dun * kx) + al3 * (u + a * kx) + al4 * (u - a * kx);
          df[2] = al1 * v + al2 * (dv - dun * ky) + al3 * (v + a * ky) + al4 * (v - a * ky);
          df[3] = al1 * w + al2 * (dw - dun * kz) + al3 * (w + a * kz) + al4 * (w - a * kz);
          df[4] = al1 * (q2 / 2 + tke) + al2 * (u * du + v * dv + w * dw - un * dun + dtke) + al3 * (h0 + a * un) +
                  al4 * (h0 - a * un);
          for (int v5 = 0; v5 < 5; v5++) F0(flux, g, i, j, k, v5 + 1, n) = (fr[v5] + fl[v5] - df[v5]) / 2 * ds;
          h0l_last = h0l;
          h0r_last = h0r;
        }
    if (kom) or_roflux_turb(qr, ql, flux, sijk, g, n, ic, jc, kc, h0l_last, h0r_last);
  }
}

/* bresid (src/cosa.f90:1149-1211) */
void or_bresid(int idir, const double* flux, double* res, int imax, int jmax, int kmax) {
  bg g = mkbg(imax, jmax, kmax);
  int di = idir == 1, dj = idir == 2, dk = idir == 3;
  for (int n = 0; n <= 2 * NH; n++)
    for (int ipde = 1; ipde <= NPDE; ipde++)
      for (int k = 1; k <= kmax - 1; k++)
        for (int j = 1; j <= jmax - 1; j++)
          for (int i = 1; i <= imax - 1; i++)
            Q(res, g, i, j, k, ipde, n) =
                Q(res, g, i, j, k, ipde, n) + F0(flux, g, i + di, j + dj, k + dk, ipde, n) - F0(flux, g, i, j, k, ipde, n);
}

/* ---------------------------------------------------------------------- */
/* harmonic balance source  src_bhb (src/src.f90:556-613)                  */
void or_src_bhb(const double* q, const double* vol, double* res, double* qhb, int imax, int jmax, int kmax) {
  bg g = mkbg(imax, jmax, kmax);
  for (int ni = 0; ni <= 2 * NH; ni++)
    for (int ipde = 1; ipde <= NPDE; ipde++)
      for (int k = 1; k <= kmax - 1; k++)
        for (int j = 1; j <= jmax - 1; j++)
          for (int i = 1; i <= imax - 1; i++) Q(qhb, g, i, j, k, ipde, ni) = 0.0;
  for (int ni = 0; ni <= 2 * NH; ni++)
    for (int nj = 0; nj <= 2 * NH; nj++)
      for (int ipde = 1; ipde <= NPDE; ipde++)
        for (int k = 1; k <= kmax - 1; k++)
          for (int j = 1; j <= jmax - 1; j++)
            for (int i = 1; i <= imax - 1; i++)
              Q(qhb, g, i, j, k, ipde, ni) =
                  Q(qhb, g, i, j, k, ipde, ni) + P.omega * V1(vol, g, i, j, k) * D(ni, nj) * Q(q, g, i, j, k, ipde, nj);
  for (int ni = 0; ni <= 2 * NH; ni++)
    for (int ipde = 1; ipde <= NPDE; ipde++)
      for (int k = 1; k <= kmax - 1; k++)
        for (int j = 1; j <= jmax - 1; j++)
          for (int i = 1; i <= imax - 1; i++)
            Q(res, g, i, j, k, ipde, ni) = Q(res, g, i, j, k, ipde, ni) + Q(qhb, g, i, j, k, ipde, ni);
}

/* ---------------------------------------------------------------------- */
/* residual norms  resg_brms (src/cosa.f90:4503-4582), explicit branch     */
void or_resg_brms(const double* res, const double* rhs, double* rb, int imax, int jmax, int kmax) {
  bg g = mkbg(imax, jmax, kmax);
  for (int p = 0; p <= NPDE; p++) rb[p] = 0;
  for (int ipde = 1; ipde <= NPDE; ipde++)
    for (int k = 1; k <= kmax - 1; k++)
      for (int j = 1; j <= jmax - 1; j++)
        for (int i = 1; i <= imax - 1; i++) {
          double d = Q(res, g, i, j, k, ipde, 0) - Q(rhs, g, i, j, k, ipde, 0);
          rb[ipde - 1] = rb[ipde - 1] + d * d;
        }
  for (int ipde = 1; ipde <= NPDE; ipde++) rb[NPDE] = rb[NPDE] + rb[ipde - 1];
}

/* resg_brms_hb (src/cosa.f90:4629-4697), explicit branch */
void or_resg_brms_hb(const double* work1, double* rb, int imax, int jmax, int kmax) {
  bg g = mkbg(imax, jmax, kmax);
  for (int p = 0; p <= NPDE; p++) rb[p] = 0;
  for (int n = 0; n <= 2 * NH; n++)
    for (int ipde = 1; ipde <= NPDE; ipde++)
      for (int k = 1; k <= kmax - 1; k++)
        for (int j = 1; j <= jmax - 1; j++)
          for (int i = 1; i <= imax - 1; i++) {
            double d = Q(work1, g, i, j, k, ipde, n);
            rb[ipde - 1] = rb[ipde - 1] + d * d;
          }
  for (int ipde = 1; ipde <= NPDE; ipde++) rb[NPDE] = rb[NPDE] + rb[ipde - 1];
}

/* ---------------------------------------------------------------------- */
/* implicit residual smoothing  cirs_b + tridi (src/cosa.f90:5238-5623,
 * 2418-2443), Euler variables 1..5 (k-omega variables are not built).     */

/* tridi (src/cosa.f90:2418-2443).  For dim = 1 the reference reads r(0)
 * and t(0) out of bounds; here they are taken as 0 (x = x/b).             */
void or_tridi(const double* a, const double* b, const double* c, double* x, int dim) {
  double* r = (double*)malloc(sizeof(double) * (dim + 1));
  double* t = (double*)malloc(sizeof(double) * (dim + 1));
  r[0] = 0; t[0] = 0; /* 1-based below */
  r[1] = c[0] / b[0];
  t[1] = x[0] / b[0];
  for (int j = 2; j <= dim - 1; j++) {
    r[j] = c[j - 1] / (b[j - 1] - r[j - 1] * a[j - 1]);
    t[j] = (x[j - 1] - t[j - 1] * a[j - 1]) / (b[j - 1] - r[j - 1] * a[j - 1]);
  }
  if (dim >= 2)
    t[dim] = (x[dim - 1] - t[dim - 1] * a[dim - 1]) / (b[dim - 1] - r[dim - 1] * a[dim - 1]);
  x[dim - 1] = t[dim];
  for (int j = dim - 1; j >= 1; j--) x[j - 1] = t[j] - x[j] * r[j];
  free(r);
  free(t);
}

/* ccirs_b (src/cosa.f90:2067-2415): CIRS with cut boundaries (cutcirs = 1).
 * The xi sweep takes the scaled residual before smoothing, ores, as its
 * right-hand side, eta and zeta the current res; every line adds
 * eps * res(halo) at both ends (the cut-exchanged neighbour residual, zero
 * elsewhere).  irsop 1: one constant eps for all variables; irsop >= 2:
 * beta (mean flow) and tbeta (k-omega) per cell.  res is already scaled by
 * dtvol (scalef).                                                          */
void or_ccirs_b(double* res, const double* ores, const double* betaxi, const double* betaeta, const double* betazeta,
                const double* tbetaxi, const double* tbetaeta, const double* tbetazeta, int imax, int jmax,
                int kmax) {
  bg g = mkbg(imax, jmax, kmax);
  const int im = imax - 1, jm = jmax - 1, km = kmax - 1;
  int nmax = imax > jmax ? imax : jmax;
  if (kmax > nmax) nmax = kmax;
  double *wa = (double*)malloc(sizeof(double) * nmax), *wb = (double*)malloc(sizeof(double) * nmax),
         *wc = (double*)malloc(sizeof(double) * nmax), *wx = (double*)malloc(sizeof(double) * nmax);
  const int kom = is_kom();
  const double* bset[2][4] = {{NULL, betaxi, betaeta, betazeta}, {NULL, tbetaxi, tbetaeta, tbetazeta}};
  const int nsets = P.irsop == 1 ? 1 : (kom ? 2 : 1);
  for (int set = 0; set < nsets; set++)
    for (int idir = 1; idir <= 3; idir++) {
      const int v0 = set == 0 ? 1 : 6;
      const int v1 = set == 1 ? NPDE : (P.irsop == 1 ? NPDE : 5);
      const int L = idir == 1 ? im : (idir == 2 ? jm : km);
      const int T1 = idir == 1 ? jm : im;
      const int T2 = idir == 3 ? jm : km;
      const int hi = idir == 1 ? imax : (idir == 2 ? jmax : kmax);
      for (int n = 0; n <= 2 * NH; n++)
        for (int ipde = v0; ipde <= v1; ipde++)
          for (int t2 = 1; t2 <= T2; t2++)
            for (int t1 = 1; t1 <= T1; t1++) {
              for (int a = 1; a <= L; a++) {
                const int i = idir == 1 ? a : t1, j = idir == 2 ? a : (idir == 1 ? t1 : t2), k = idir == 3 ? a : t2;
                const double eps =
                    P.irsop == 1 ? 0.25 * (P.rat * P.rat - 1.0) : P1(bset[set][idir], g, i, j, k, n);
                wa[a - 1] = a == 1 ? 0.0 : -eps;
                wb[a - 1] = 1 + 2 * eps;
                wc[a - 1] = a == L ? 0.0 : -eps;
                const double rhs = idir == 1 ? Q(ores, g, i, j, k, ipde, n) : Q(res, g, i, j, k, ipde, n);
                /* the last cell's assignment comes after the first's in the
                 * reference, so a one-cell line takes the upper halo term */
                if (a == L) {
                  const int i0 = idir == 1 ? hi : i, j0 = idir == 2 ? hi : j, k0 = idir == 3 ? hi : k;
                  wx[a - 1] = rhs + eps * Q(res, g, i0, j0, k0, ipde, n);
                } else if (a == 1) {
                  const int i0 = idir == 1 ? 0 : i, j0 = idir == 2 ? 0 : j, k0 = idir == 3 ? 0 : k;
                  wx[a - 1] = rhs + eps * Q(res, g, i0, j0, k0, ipde, n);
                } else {
                  wx[a - 1] = rhs;
                }
              }
              or_tridi(wa, wb, wc, wx, L);
              for (int a = 1; a <= L; a++) {
                const int i = idir == 1 ? a : t1, j = idir == 2 ? a : (idir == 1 ? t1 : t2), k = idir == 3 ? a : t2;
                Q(res, g, i, j, k, ipde, n) = wx[a - 1];
              }
            }
    }
  free(wa);
  free(wb);
  free(wc);
  free(wx);
}

/* scalef_b / scaleb_b (src/cosa.f90:5652-5773): res times / over dtvol
 * (dtvolt for k and omega) on the interior.                              */
void or_scale_b(double* res, const double* dtvol, const double* dtvolt, int back, int imax, int jmax, int kmax) {
  bg g = mkbg(imax, jmax, kmax);
  for (int n = 0; n <= 2 * NH; n++)
    for (int ipde = 1; ipde <= NPDE; ipde++) {
      if (ipde > 5 && !is_kom()) break;
      const double* d = ipde > 5 ? dtvolt : dtvol;
      for (int k = 1; k <= kmax - 1; k++)
        for (int j = 1; j <= jmax - 1; j++)
          for (int i = 1; i <= imax - 1; i++)
            Q(res, g, i, j, k, ipde, n) =
                back ? Q(res, g, i, j, k, ipde, n) / P1(d, g, i, j, k, n) : Q(res, g, i, j, k, ipde, n) * P1(d, g, i, j, k, n);
    }
}

void or_cirs_b(double* res, const double* dtvol, const double* betaxi, const double* betaeta, const double* betazeta,
               int imax, int jmax, int kmax) {
  or_cirs_kb(res, dtvol, betaxi, betaeta, betazeta, NULL, NULL, NULL, NULL, imax, jmax, kmax);
}

/* cirs_b with the k-omega equations (src/cosa.f90:5238-5623): res(6:7) is
 * scaled by dtvolt; irsop 1 smooths every equation with the same epsilon,
 * irsop 2/3 smooth 6:7 with tbeta*.                                        */
void or_cirs_kb(double* res, const double* dtvol, const double* betaxi, const double* betaeta,
                const double* betazeta, const double* dtvolt, const double* tbetaxi, const double* tbetaeta,
                const double* tbetazeta, int imax, int jmax, int kmax) {
  bg g = mkbg(imax, jmax, kmax);
  int kom = is_kom() && dtvolt;
  int im = imax - 1, jm = jmax - 1, km = kmax - 1;
  int nmax = imax > jmax ? imax : jmax;
  if (kmax > nmax) nmax = kmax;
  double *wa = (double*)malloc(sizeof(double) * nmax), *wb = (double*)malloc(sizeof(double) * nmax),
         *wc = (double*)malloc(sizeof(double) * nmax), *wx = (double*)malloc(sizeof(double) * nmax);
  for (int n = 0; n <= 2 * NH; n++)
    for (int ipde = 1; ipde <= 5; ipde++)
      for (int k = 1; k <= km; k++)
        for (int j = 1; j <= jm; j++)
          for (int i = 1; i <= im; i++) Q(res, g, i, j, k, ipde, n) = Q(res, g, i, j, k, ipde, n) * P1(dtvol, g, i, j, k, n);
  if (kom)
    for (int n = 0; n <= 2 * NH; n++)
      for (int ipde = 6; ipde <= NPDE; ipde++)
        for (int k = 1; k <= km; k++)
          for (int j = 1; j <= jm; j++)
            for (int i = 1; i <= im; i++)
              Q(res, g, i, j, k, ipde, n) = Q(res, g, i, j, k, ipde, n) * P1(dtvolt, g, i, j, k, n);
  const double* betas[2][4] = {{NULL, betaxi, betaeta, betazeta}, {NULL, tbetaxi, tbetaeta, tbetazeta}};
  int nsets = P.irsop == 1 ? 1 : (kom ? 2 : 1);
  for (int set = 0; set < nsets; set++)
  for (int idir = 1; idir <= 3; idir++) {
    const double* const* beta = betas[set];
    int v0 = set == 0 ? 1 : 6;
    int v1 = set == 1 ? NPDE : (P.irsop == 1 && kom ? NPDE : 5);
    int L = idir == 1 ? im : (idir == 2 ? jm : km);
    int T1 = idir == 1 ? jm : im;
    int T2 = idir == 3 ? jm : km;
    for (int n = 0; n <= 2 * NH; n++)
      for (int ipde = v0; ipde <= v1; ipde++)
        for (int t2 = 1; t2 <= T2; t2++)
          for (int t1 = 1; t1 <= T1; t1++) {
            for (int a = 1; a <= L; a++) {
              int i = idir == 1 ? a : t1, j = idir == 2 ? a : (idir == 1 ? t1 : t2), k = idir == 3 ? a : t2;
              double eps = P.irsop == 1 ? 0.25 * (P.rat * P.rat - 1.0) : P1(beta[idir], g, i, j, k, n);
              wa[a - 1] = -eps;
              wb[a - 1] = 1 + 2 * eps;
              wc[a - 1] = -eps;
              wx[a - 1] = Q(res, g, i, j, k, ipde, n);
            }
            wa[0] = 0;
            if (L >= 1) {
              /* the reference writes work(1,1)=0 first and then the last row,
                 so for L = 1 the last-row values win */
              int i = idir == 1 ? L : t1, j = idir == 2 ? L : (idir == 1 ? t1 : t2), k = idir == 3 ? L : t2;
              double eps = P.irsop == 1 ? 0.25 * (P.rat * P.rat - 1.0) : P1(beta[idir], g, i, j, k, n);
              wa[L - 1] = -eps;
              wc[L - 1] = 0;
            }
            or_tridi(wa, wb, wc, wx, L);
            for (int a = 1; a <= L; a++) {
              int i = idir == 1 ? a : t1, j = idir == 2 ? a : (idir == 1 ? t1 : t2), k = idir == 3 ? a : t2;
              Q(res, g, i, j, k, ipde, n) = wx[a - 1];
            }
          }
  }
  for (int n = 0; n <= 2 * NH; n++)
    for (int ipde = 1; ipde <= 5; ipde++)
      for (int k = 1; k <= km; k++)
        for (int j = 1; j <= jm; j++)
          for (int i = 1; i <= im; i++) Q(res, g, i, j, k, ipde, n) = Q(res, g, i, j, k, ipde, n) / P1(dtvol, g, i, j, k, n);
  if (kom)
    for (int n = 0; n <= 2 * NH; n++)
      for (int ipde = 6; ipde <= NPDE; ipde++)
        for (int k = 1; k <= km; k++)
          for (int j = 1; j <= jm; j++)
            for (int i = 1; i <= im; i++)
              Q(res, g, i, j, k, ipde, n) = Q(res, g, i, j, k, ipde, n) / P1(dtvolt, g, i, j, k, n);
  free(wa); free(wb); free(wc); free(wx);
}

/* ---------------------------------------------------------------------- */
/* explicit RK update: update_b / update_bhb, .not.rkimp, Euler variables
 * (src/update.f90:568-624 and 1359-1408): q = qold - ar*res*dtvol.       */
void or_update_b(double* res, double* q, const double* qold, const double* dtvol, double ar, int imax, int jmax,
                 int kmax) {
  bg g = mkbg(imax, jmax, kmax);
  (void)res;
  for (int n = 0; n <= 2 * NH; n++)
    for (int ipde = 1; ipde <= 5; ipde++)
      for (int k = 1; k <= kmax - 1; k++)
        for (int j = 1; j <= jmax - 1; j++)
          for (int i = 1; i <= imax - 1; i++)
            Q(q, g, i, j, k, ipde, n) =
                Q(qold, g, i, j, k, ipde, n) - ar * Q(res, g, i, j, k, ipde, n) * P1(dtvol, g, i, j, k, n);
}

/* ---------------------------------------------------------------------- */
/* viscous fluxes and Wilcox k-omega: per-block restatements              */

/* metric derivative arrays xider(3,imax,jmax,kmax,0:2N*hbmove) */
#define MD(a, g, c, i, j, k, nh)                                                        \
  (a)[((c) - 1) + 3 * ((ll)((i) - 1) + (ll)(g).imax * ((ll)((j) - 1) + (ll)(g).jmax *    \
                                                        ((ll)((k) - 1) + (ll)(g).kmax * (nh))))]

/* ---------------------------------------------------------------------- */
/* force_b  (src/output.f90:4388-4886), aircraft conventions              */
void or_force_b(const double* q, const double* x, const double* y, const double* z, const double* xdot,
                const double* ydot, const double* zdot, const double* si, const double* sj, const double* sk,
                const double* xideri, const double* etaderj, const double* zetaderk, int nbcs,
                const int32_t* bctopo, int nsurface, int inv14, const double* lngth, const double* momr,
                double* coef, int imax, int jmax, int kmax) {
  bg g = mkbg(imax, jmax, kmax);
  const double gamma = P.gamma;
  const int ijkmax[3] = {imax, jmax, kmax};
  const double mach = P.machfs > 1e-14 ? P.machfs : 1.0;
  const double alpha = P.alpha, beta = P.beta;
  /* src/output.f90:4451-4462 */
  const double ssx = cos(alpha) * cos(beta), ssy = sin(alpha) * cos(beta), ssz = sin(beta);
  const double den = sqrt(sin(beta) * sin(beta) + cos(alpha) * cos(alpha) * (cos(beta) * cos(beta)));
  const double aax = -sin(alpha) * cos(alpha) * (cos(beta) * cos(beta)) / den;
  const double aay = den;
  const double aaz = sin(alpha) * sin(beta) * cos(beta) / den;
  const double bbx = -sin(beta) / den, bby = 0.0, bbz = cos(alpha) * cos(beta) / den;
  for (int n = 0; n < DIM5; n++) {
    const int nh = n * P.hbmove;
    const double xmomr = momr[3 * n], ymomr = momr[3 * n + 1], zmomr = momr[3 * n + 2];
    for (int isurface = 1; isurface <= nsurface; isurface++) {
      double fp[6] = {0, 0, 0, 0, 0, 0}, fv[6] = {0, 0, 0, 0, 0, 0};
      for (int ib = 0; ib < nbcs; ib++) {
        const int32_t* t = bctopo + 10 * ib;
        const int visc_wall = t[0] == 1500 + (isurface - 1);
        const int inv_wall = inv14 && isurface == 1 && t[0] == 14;
        if (!visc_wall && !inv_wall) continue;
        const int idir = t[1], inrout = t[2];
        const int istrt[3] = {t[3], t[5], t[7]}, iend[3] = {t[4], t[6], t[8]};
        int ibcpt, ibcw, ibcm;
        if (inrout == 1) {
          ibcpt = 0;
          ibcm = 1;
          ibcw = 0;
        } else {
          ibcpt = ijkmax[idir - 1];
          ibcm = ijkmax[idir - 1];
          ibcw = ijkmax[idir - 1];
        }
        const int sgn = inrout == 1 ? 1 : -1; /* (-1)**(1+1/inrout) */
        const int ioff1 = idir == 1 ? sgn : 0, joff1 = idir == 2 ? sgn : 0, koff1 = idir == 3 ? sgn : 0;
        const int ioff2 = 2 * ioff1, joff2 = 2 * joff1, koff2 = 2 * koff1;
        const int ic1 = CYC(idir, 1), ic2 = CYC(idir, 2), ic3 = CYC(idir, 3);
        const double* sm = idir == 1 ? si : (idir == 2 ? sj : sk);
        const double* ed = idir == 1 ? xideri : (idir == 2 ? etaderj : zetaderk);
        for (int pass = 0; pass < 2; pass++) {
          if (pass == 1 && !(P.viscous && visc_wall)) break;
          for (int i3 = istrt[ic3 - 1]; i3 <= iend[ic3 - 1]; i3++)
            for (int i2 = istrt[ic2 - 1]; i2 <= iend[ic2 - 1]; i2++) {
              const int ibc = ibcpt * KRD(ic1, 1) + i2 * KRD(ic2, 1) + i3 * KRD(ic3, 1);
              const int jbc = ibcpt * KRD(ic1, 2) + i2 * KRD(ic2, 2) + i3 * KRD(ic3, 2);
              const int kbc = ibcpt * KRD(ic1, 3) + i2 * KRD(ic2, 3) + i3 * KRD(ic3, 3);
              const int im = ibcm * KRD(ic1, 1) + i2 * KRD(ic2, 1) + i3 * KRD(ic3, 1);
              const int jm = ibcm * KRD(ic1, 2) + i2 * KRD(ic2, 2) + i3 * KRD(ic3, 2);
              const int km = ibcm * KRD(ic1, 3) + i2 * KRD(ic2, 3) + i3 * KRD(ic3, 3);
              const int iw = ibcw * KRD(ic1, 1) + i2 * KRD(ic2, 1) + i3 * KRD(ic3, 1);
              const int jw = ibcw * KRD(ic1, 2) + i2 * KRD(ic2, 2) + i3 * KRD(ic3, 2);
              const int kw = ibcw * KRD(ic1, 3) + i2 * KRD(ic2, 3) + i3 * KRD(ic3, 3);
              const double nx = M(sm, g, 1, im, jm, km, nh), ny = M(sm, g, 2, im, jm, km, nh);
              const double nz = M(sm, g, 3, im, jm, km, nh), ds = M(sm, g, 4, im, jm, km, nh);
              const double kx = sgn * nx, ky = sgn * ny, kz = sgn * nz;
              const double r1 = Q(q, g, ibc, jbc, kbc, 1, n), r2 = Q(q, g, ibc, jbc, kbc, 2, n);
              const double r3 = Q(q, g, ibc, jbc, kbc, 3, n), r4 = Q(q, g, ibc, jbc, kbc, 4, n);
              const double pw = (gamma - 1) * (Q(q, g, ibc, jbc, kbc, 5, n) - r2 * r2 / r1 / 2 - r3 * r3 / r1 / 2 -
                                               r4 * r4 / r1 / 2);
              const double xarm = X2(x, g, iw, jw, kw, nh) - xmomr;
              const double yarm = X2(y, g, iw, jw, kw, nh) - ymomr;
              const double zarm = X2(z, g, iw, jw, kw, nh) - zmomr;
              if (pass == 0) { /* src/output.f90:4672-4684 */
                fp[0] = fp[0] - pw * kx * ds;
                fp[1] = fp[1] - pw * ky * ds;
                fp[2] = fp[2] - pw * kz * ds;
                fp[3] = fp[3] + (ky * zarm - kz * yarm) * pw * ds;
                fp[4] = fp[4] + (kz * xarm - kx * zarm) * pw * ds;
                fp[5] = fp[5] + (kx * yarm - ky * xarm) * pw * ds;
                continue;
              }
              /* viscous force, src/output.f90:4727-4850 */
              const double etax = MD(ed, g, 1, im, jm, km, nh), etay = MD(ed, g, 2, im, jm, km, nh);
              const double etaz = MD(ed, g, 3, im, jm, km, nh);
              const double rhow = r1;
              const double tw = pw * gamma / rhow;
              const double muw = (P.stemp + 1) / (tw + P.stemp) * pow(tw, 1.5);
              const int a1 = ibc + ioff1, b1 = jbc + joff1, c1 = kbc + koff1;
              const int a2 = ibc + ioff2, b2 = jbc + joff2, c2 = kbc + koff2;
              double uw = Q(q, g, ibc, jbc, kbc, 2, n) / Q(q, g, ibc, jbc, kbc, 1, n);
              double u1 = Q(q, g, a1, b1, c1, 2, n) / Q(q, g, a1, b1, c1, 1, n);
              double u2 = Q(q, g, a2, b2, c2, 2, n) / Q(q, g, a2, b2, c2, 1, n);
              double vw = Q(q, g, ibc, jbc, kbc, 3, n) / Q(q, g, ibc, jbc, kbc, 1, n);
              double v1 = Q(q, g, a1, b1, c1, 3, n) / Q(q, g, a1, b1, c1, 1, n);
              double v2 = Q(q, g, a2, b2, c2, 3, n) / Q(q, g, a2, b2, c2, 1, n);
              double ww = Q(q, g, ibc, jbc, kbc, 4, n) / Q(q, g, ibc, jbc, kbc, 1, n);
              double w1 = Q(q, g, a1, b1, c1, 4, n) / Q(q, g, a1, b1, c1, 1, n);
              double w2 = Q(q, g, a2, b2, c2, 4, n) / Q(q, g, a2, b2, c2, 1, n);
              if (P.moving) {
                uw = uw - X2(xdot, g, ibc, jbc, kbc, nh);
                u1 = u1 - X2(xdot, g, a1, b1, c1, nh);
                u2 = u2 - X2(xdot, g, a2, b2, c2, nh);
                vw = vw - X2(ydot, g, ibc, jbc, kbc, nh);
                v1 = v1 - X2(ydot, g, a1, b1, c1, nh);
                v2 = v2 - X2(ydot, g, a2, b2, c2, nh);
                ww = ww - X2(zdot, g, ibc, jbc, kbc, nh);
                w1 = w1 - X2(zdot, g, a1, b1, c1, nh);
                w2 = w2 - X2(zdot, g, a2, b2, c2, nh);
              }
              const double ueta = sgn * (-8 * uw + 9 * u1 - u2) / 3;
              const double veta = sgn * (-8 * vw + 9 * v1 - v2) / 3;
              const double weta = sgn * (-8 * ww + 9 * w1 - w2) / 3;
              const double dudx = ueta * etax, dudy = ueta * etay, dudz = ueta * etaz;
              const double dvdx = veta * etax, dvdy = veta * etay, dvdz = veta * etaz;
              const double dwdx = weta * etax, dwdy = weta * etay, dwdz = weta * etaz;
              const double divv = dudx + dvdy + dwdz;
              const double txx = 2 * muw * (dudx - divv / 3);
              const double txy = muw * (dudy + dvdx);
              const double txz = muw * (dudz + dwdx);
              const double tyy = 2 * muw * (dvdy - divv / 3);
              const double tyz = muw * (dvdz + dwdy);
              const double tzz = 2 * muw * (dwdz - divv / 3);
              const double mr = P.machfs / P.reyno;
              fv[0] = fv[0] + mr * ds * (txx * kx + txy * ky + txz * kz);
              fv[1] = fv[1] + mr * ds * (txy * kx + tyy * ky + tyz * kz);
              fv[2] = fv[2] + mr * ds * (txz * kx + tyz * ky + tzz * kz);
              fv[3] = fv[3] + mr * ds * (-(txy * kx + tyy * ky + tyz * kz) * zarm + (txz * kx + tyz * ky + tzz * kz) * yarm);
              fv[4] = fv[4] + mr * ds * (-(txz * kx + tyz * ky + tzz * kz) * xarm + (txx * kx + txy * ky + txz * kz) * zarm);
              fv[5] = fv[5] + mr * ds * (-(txx * kx + txy * ky + txz * kz) * yarm + (txy * kx + tyy * ky + tyz * kz) * xarm);
            }
        }
      }
      /* src/output.f90:4857-4880 */
      double* c = coef + 12 * ((ll)n + (ll)DIM5 * (isurface - 1));
      c[0] = (fp[0] * aax + fp[1] * aay + fp[2] * aaz) * 2 / (lngth[0] * lngth[1] * (mach * mach));
      c[1] = (fp[0] * ssx + fp[1] * ssy + fp[2] * ssz) * 2 / (lngth[0] * lngth[1] * (mach * mach));
      c[2] = (fp[0] * bbx + fp[1] * bby + fp[2] * bbz) * 2 / (lngth[0] * lngth[1] * (mach * mach));
      c[3] = fp[3] * 2 / (lngth[0] * lngth[1] * lngth[2] * (mach * mach));
      c[4] = fp[4] * 2 / (lngth[0] * lngth[1] * lngth[2] * (mach * mach));
      c[5] = fp[5] * 2 / (lngth[0] * lngth[1] * lngth[2] * (mach * mach));
      c[6] = (fv[0] * aax + fv[1] * aay + fv[2] * aaz) * 2 / (lngth[0] * lngth[1] * (mach * mach));
      c[7] = (fv[0] * ssx + fv[1] * ssy + fv[2] * ssz) * 2 / (lngth[0] * lngth[1] * (mach * mach));
      c[8] = (fv[0] * bbx + fv[1] * bby + fv[2] * bbz) * 2 / (lngth[0] * lngth[1] * (mach * mach));
      c[9] = fv[3] * 2 / (lngth[0] * lngth[1] * lngth[2] * (mach * mach));
      c[10] = fv[4] * 2 / (lngth[0] * lngth[1] * lngth[2] * (mach * mach));
      c[11] = fv[5] * 2 / (lngth[0] * lngth[1] * lngth[2] * (mach * mach));
    }
  }
}

/* the strong (wall-like) BC types of muscl_b and bq_face/bq_der
 * (src/flux.f90:1320, 1898, 2471)                                         */
static int wall_like(int t) {
  return t == 13 || t == 14 || t / 100 == 15 || t == 22 || t == 23 || t == 92 || t == 93 || t == 5;
}

/* Sutherland viscosity (src/flux.f90:936-937) */
static double suth(double t) { return (P.stemp + 1) / (t + P.stemp) * pow(t, 1.5); }

/* face ranges of direction idir: idir index 1..ijkmax, others 1..n-1 */
static void face_ranges(int idir, bg g, int* ic, int* jc, int* kc) {
  *ic = idir == 1 ? g.imax : g.imax - 1;
  *jc = idir == 2 ? g.jmax : g.jmax - 1;
  *kc = idir == 3 ? g.kmax : g.kmax - 1;
}

/* bc_walvis_15 (src/bc.f90:2144-2492): no-slip wall (wall velocity from
 * the face metrics on moving grids), pressure extrapolated linearly and
 * density from isentropic scaling; k = 0 and omega from Wilcox's rough
 * wall (wallwilc) or Menter's formula on the wall halo layers.           */
static void or_bc_walvis15(double* q, const double* si, const double* sj, const double* sk, const double* dist,
                           const int32_t* t, bg g) {
  int ijkmax[4] = {0, g.imax, g.jmax, g.kmax};
  bcgeo b = bc_decode(t);
  int ibcpt, ibcpt2, ibcn, ibcn2, ibcm;
  const double small = 1.e-12, smasq = 1.e-24;
  double uw = 0.0, vw = 0.0, ww = 0.0;
  const double mor = P.machfs / P.reyno;
  if (b.inrout == 1) {
    ibcpt = 0; ibcpt2 = -1; ibcn = 1; ibcn2 = 2; ibcm = 1;
  } else {
    ibcpt = ijkmax[b.idir]; ibcpt2 = ijkmax[b.idir] + 1; ibcn = ijkmax[b.idir] - 1; ibcn2 = ijkmax[b.idir] - 2;
    ibcm = ijkmax[b.idir];
  }
  const double* s = b.idir == 1 ? si : (b.idir == 2 ? sj : sk);
  for (int n = 0; n <= 2 * NH; n++) {
    int nh = n * P.hbmove;
    for (int i3 = b.istrt[b.ic3]; i3 <= b.iend[b.ic3]; i3++)
      for (int i2 = b.istrt[b.ic2]; i2 <= b.iend[b.ic2]; i2++) {
        int ib, jb, kb, ib2, jb2, kb2, in, jn, kn, in2, jn2, kn2, im, jm, km;
        bc_ijk(&b, ibcpt, i2, i3, &ib, &jb, &kb);
        bc_ijk(&b, ibcpt2, i2, i3, &ib2, &jb2, &kb2);
        bc_ijk(&b, ibcn, i2, i3, &in, &jn, &kn);
        bc_ijk(&b, ibcn2, i2, i3, &in2, &jn2, &kn2);
        bc_ijk(&b, ibcm, i2, i3, &im, &jm, &km);
        if (P.moving) {
          uw = M(s, g, 5, im, jm, km, nh);
          vw = M(s, g, 6, im, jm, km, nh);
          ww = M(s, g, 7, im, jm, km, nh);
        }
        double p1 = Q(q, g, in, jn, kn, 5, n), p2 = Q(q, g, in2, jn2, kn2, 5, n);
        double pw = 0.5 * (3 * p1 - p2);
        double rhow = pow(pw / p1, 1 / P.gamma) * Q(q, g, in, jn, kn, 1, n);
        double hv[5] = {rhow, uw, vw, ww, pw};
        for (int v = 0; v < 5; v++) {
          Q(q, g, ib, jb, kb, v + 1, n) = hv[v];
          Q(q, g, ib2, jb2, kb2, v + 1, n) = hv[v];
        }
      }
  }
  if (!is_kom()) return;
  for (int n = 0; n <= 2 * NH; n++) {
    int nh = n * P.hbmove;
    for (int i3 = b.istrt[b.ic3]; i3 <= b.iend[b.ic3]; i3++)
      for (int i2 = b.istrt[b.ic2]; i2 <= b.iend[b.ic2]; i2++) {
        int ib, jb, kb, ib2, jb2, kb2, in, jn, kn, im, jm, km;
        bc_ijk(&b, ibcpt, i2, i3, &ib, &jb, &kb);
        bc_ijk(&b, ibcpt2, i2, i3, &ib2, &jb2, &kb2);
        bc_ijk(&b, ibcn, i2, i3, &in, &jn, &kn);
        bc_ijk(&b, ibcm, i2, i3, &im, &jm, &km);
        double rhow = Q(q, g, ib, jb, kb, 1, n), pw = Q(q, g, ib, jb, kb, 5, n);
        double tw = pw * P.gamma / rhow;
        double muw = suth(tw);
        double omew;
        if (P.wallwilc) {
          double nx = M(s, g, 1, im, jm, km, nh), ny = M(s, g, 2, im, jm, km, nh), nz = M(s, g, 3, im, jm, km, nh);
          if (P.moving) {
            uw = M(s, g, 5, im, jm, km, nh);
            vw = M(s, g, 6, im, jm, km, nh);
            ww = M(s, g, 7, im, jm, km, nh);
          }
          double uc = Q(q, g, in, jn, kn, 2, n) - uw, vc = Q(q, g, in, jn, kn, 3, n) - vw,
                 wc = Q(q, g, in, jn, kn, 4, n) - ww;
          double uorth = nx * uc + ny * vc + nz * wc;
          double a1 = uc - uorth * nx, a2 = vc - uorth * ny, a3 = wc - uorth * nz;
          double upar = sqrt(a1 * a1 + a2 * a2 + a3 * a3);
          double dupardn = upar / (V1(dist, g, in, jn, kn) + small);
          double tauw = muw * fabs(dupardn);
          double rkrpls;
          if (P.roughk > 0.0) rkrpls = fmax(1.0, sqrt(rhow * tauw / mor) * P.roughk / muw);
          else rkrpls = 1.0;
          double sr = rkrpls >= 25.0 ? 100.0 / rkrpls : 2500.0 / (rkrpls * rkrpls);
          omew = fabs(dupardn) * sr;
        } else {
          double d = V1(dist, g, in, jn, kn);
          omew = 60.0 * muw / (P.bkw * rhow * (d * d) + smasq) * P.machfs / P.reyno;
        }
        Q(q, g, ib, jb, kb, 6, n) = 0.0;
        Q(q, g, ib, jb, kb, 7, n) = omew;
        Q(q, g, ib2, jb2, kb2, 6, n) = 0.0;
        Q(q, g, ib2, jb2, kb2, 7, n) = omew;
      }
  }
}

/* bq_face_{i,j,k} (src/flux.f90:2432-2801): face averages of q into qf
 * (and of mut into mutf when calmut); wall-like patches of direction idir
 * take the halo value at the wall face.                                   */
void or_q_face_b(int idir, const int32_t* bctopo, int nbcs, const double* q, double* qf, const double* mut,
                 double* mutf, int calmut, int imax, int jmax, int kmax) {
  bg g = mkbg(imax, jmax, kmax);
  int ic, jc, kc;
  face_ranges(idir, g, &ic, &jc, &kc);
  int di = idir == 1, dj = idir == 2, dk = idir == 3;
  int ijkmax[4] = {0, imax, jmax, kmax};
  for (int n = 0; n <= 2 * NH; n++)
    for (int ipde = 1; ipde <= NPDE; ipde++)
      for (int k = 1; k <= kc; k++)
        for (int j = 1; j <= jc; j++)
          for (int i = 1; i <= ic; i++)
            Q(qf, g, i, j, k, ipde, n) = (Q(q, g, i - di, j - dj, k - dk, ipde, n) + Q(q, g, i, j, k, ipde, n)) / 2;
  for (int pass = 0; pass < (calmut ? 2 : 1); pass++) {
    if (pass == 1)
      for (int n = 0; n <= 2 * NH; n++)
        for (int k = 1; k <= kc; k++)
          for (int j = 1; j <= jc; j++)
            for (int i = 1; i <= ic; i++)
              F0(mutf, g, i, j, k, 1, n) = (S1(mut, g, i - di, j - dj, k - dk, n) + S1(mut, g, i, j, k, n)) / 2;
    for (int ibc = 0; ibc < nbcs; ibc++) {
      const int32_t* t = bctopo + 10 * ibc;
      if (t[1] != idir || !wall_like(t[0])) continue;
      bcgeo b = bc_decode(t);
      int bcpt = b.inrout == 1 ? 0 : ijkmax[idir];
      int face = b.inrout == 1 ? 1 : ijkmax[idir];
      for (int n = 0; n <= 2 * NH; n++)
        for (int i3 = b.istrt[b.ic3]; i3 <= b.iend[b.ic3]; i3++)
          for (int i2 = b.istrt[b.ic2]; i2 <= b.iend[b.ic2]; i2++) {
            int ia, ja, ka, iff, jf, kf;
            bc_ijk(&b, bcpt, i2, i3, &ia, &ja, &ka);
            bc_ijk(&b, face, i2, i3, &iff, &jf, &kf);
            if (pass == 0)
              for (int ipde = 1; ipde <= NPDE; ipde++) Q(qf, g, iff, jf, kf, ipde, n) = Q(q, g, ia, ja, ka, ipde, n);
            else
              F0(mutf, g, iff, jf, kf, 1, n) = S1(mut, g, ia, ja, ka, n);
          }
    }
  }
}

/* bq_der_{i,j,k} (src/flux.f90:1848-2390): xi, eta, zeta differences of
 * q at the faces of direction idir.  Across the face: q(p) - q(m); along
 * a transverse direction t: (q(m)+q(p)+q(p+t)+q(m+t))/4 minus
 * (q(m-t)+q(p-t)+q(p)+q(m))/4.  At wall-like patches the normal difference
 * is the one-sided (-8 q_w + 9 q_1 - q_2)/3 and the transverse ones use the
 * halo layer with one-sided ends.                                         */
void or_q_der_b(int idir, const int32_t* bctopo, int nbcs, const double* q, double* dqxi, double* dqeta,
                double* dqzeta, int imax, int jmax, int kmax) {
  bg g = mkbg(imax, jmax, kmax);
  int ic, jc, kc;
  face_ranges(idir, g, &ic, &jc, &kc);
  int ijkmax[4] = {0, imax, jmax, kmax};
  double* outs[4] = {NULL, dqxi, dqeta, dqzeta};
  int e[4][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  for (int n = 0; n <= 2 * NH; n++)
    for (int ipde = 1; ipde <= NPDE; ipde++)
      for (int k = 1; k <= kc; k++)
        for (int j = 1; j <= jc; j++)
          for (int i = 1; i <= ic; i++) {
            int im = i - e[idir][0], jm = j - e[idir][1], km = k - e[idir][2];
            for (int d = 1; d <= 3; d++) {
              double val;
              if (d == idir) {
                val = Q(q, g, i, j, k, ipde, n) - Q(q, g, im, jm, km, ipde, n);
              } else {
                int ti = e[d][0], tj = e[d][1], tk = e[d][2];
                double q2 = (Q(q, g, im, jm, km, ipde, n) + Q(q, g, i, j, k, ipde, n) +
                             Q(q, g, i + ti, j + tj, k + tk, ipde, n) + Q(q, g, im + ti, jm + tj, km + tk, ipde, n)) /
                            4;
                double q1 = (Q(q, g, im - ti, jm - tj, km - tk, ipde, n) + Q(q, g, i - ti, j - tj, k - tk, ipde, n) +
                             Q(q, g, i, j, k, ipde, n) + Q(q, g, im, jm, km, ipde, n)) /
                            4;
                val = q2 - q1;
              }
              Q(outs[d], g, i, j, k, ipde, n) = val;
            }
          }
  for (int ibc = 0; ibc < nbcs; ibc++) {
    const int32_t* t = bctopo + 10 * ibc;
    if (t[1] != idir || !wall_like(t[0])) continue;
    bcgeo b = bc_decode(t);
    int aux, n1, n2, face;
    double sgnm;
    if (b.inrout == 1) {
      aux = 0; n1 = 1; n2 = 2; face = 1; sgnm = 1.0;
    } else {
      aux = ijkmax[idir]; n1 = ijkmax[idir] - 1; n2 = ijkmax[idir] - 2; face = ijkmax[idir]; sgnm = -1.0;
    }
    for (int n = 0; n <= 2 * NH; n++)
      for (int ipde = 1; ipde <= NPDE; ipde++)
        for (int i3 = b.istrt[b.ic3]; i3 <= b.iend[b.ic3]; i3++)
          for (int i2 = b.istrt[b.ic2]; i2 <= b.iend[b.ic2]; i2++) {
            int ia, ja, ka, i1_, j1_, k1_, i2_, j2_, k2_, iff, jf, kf;
            bc_ijk(&b, aux, i2, i3, &ia, &ja, &ka);
            bc_ijk(&b, n1, i2, i3, &i1_, &j1_, &k1_);
            bc_ijk(&b, n2, i2, i3, &i2_, &j2_, &k2_);
            bc_ijk(&b, face, i2, i3, &iff, &jf, &kf);
            Q(outs[idir], g, iff, jf, kf, ipde, n) =
                sgnm / 3 * (-Q(q, g, ia, ja, ka, ipde, n) * 8 + Q(q, g, i1_, j1_, k1_, ipde, n) * 9 -
                            Q(q, g, i2_, j2_, k2_, ipde, n));
            for (int d = 1; d <= 3; d++) {
              if (d == idir) continue;
              int ti = e[d][0], tj = e[d][1], tk = e[d][2];
              int c = d == 1 ? ia : (d == 2 ? ja : ka);
              double qc = Q(q, g, ia, ja, ka, ipde, n);
              double qp1 = Q(q, g, ia + ti, ja + tj, ka + tk, ipde, n);
              double qm1 = Q(q, g, ia - ti, ja - tj, ka - tk, ipde, n);
              double q2, q1;
              if (c == b.istrt[d]) {
                q2 = (qc + qp1) / 2;
                q1 = (qc * 3 - qp1) / 2;
              } else if (c == b.iend[d]) {
                q2 = (qc * 3 - qm1) / 2;
                q1 = (qc + qm1) / 2;
              } else {
                q2 = (qc + qp1) / 2;
                q1 = (qm1 + qc) / 2;
              }
              Q(outs[d], g, iff, jf, kf, ipde, n) = q2 - q1;
            }
          }
  }
}

/* b_rtst (src/flux.f90:2846-2947): Reynolds stresses at the faces of
 * direction idir into tau1(:,:,:,1:6) (txx, txy, txz, tyy, tyz, tzz).     */
void or_rtst_b(int idir, const double* qf, const double* dqxi, const double* dqeta, const double* dqzeta,
               const double* xider, const double* etader, const double* zetader, double* tau1, const double* mutf,
               int imax, int jmax, int kmax) {
  bg g = mkbg(imax, jmax, kmax);
  int ic, jc, kc;
  face_ranges(idir, g, &ic, &jc, &kc);
  const double fact = 2.0 / 3.0;
  for (int n = 0; n <= 2 * NH; n++) {
    int nh = n * P.hbmove;
    for (int k = 1; k <= kc; k++)
      for (int j = 1; j <= jc; j++)
        for (int i = 1; i <= ic; i++) {
          double rho = Q(qf, g, i, j, k, 1, n), tke = Q(qf, g, i, j, k, 6, n);
          double mut = F0(mutf, g, i, j, k, 1, n);
          double xix = MD(xider, g, 1, i, j, k, nh), xiy = MD(xider, g, 2, i, j, k, nh), xiz = MD(xider, g, 3, i, j, k, nh);
          double etax = MD(etader, g, 1, i, j, k, nh), etay = MD(etader, g, 2, i, j, k, nh),
                 etaz = MD(etader, g, 3, i, j, k, nh);
          double zetax = MD(zetader, g, 1, i, j, k, nh), zetay = MD(zetader, g, 2, i, j, k, nh),
                 zetaz = MD(zetader, g, 3, i, j, k, nh);
          double uxi = Q(dqxi, g, i, j, k, 2, n), ueta = Q(dqeta, g, i, j, k, 2, n), uzeta = Q(dqzeta, g, i, j, k, 2, n);
          double vxi = Q(dqxi, g, i, j, k, 3, n), veta = Q(dqeta, g, i, j, k, 3, n), vzeta = Q(dqzeta, g, i, j, k, 3, n);
          double wxi = Q(dqxi, g, i, j, k, 4, n), weta = Q(dqeta, g, i, j, k, 4, n), wzeta = Q(dqzeta, g, i, j, k, 4, n);
          double dudx = uxi * xix + ueta * etax + uzeta * zetax;
          double dudy = uxi * xiy + ueta * etay + uzeta * zetay;
          double dudz = uxi * xiz + ueta * etaz + uzeta * zetaz;
          double dvdx = vxi * xix + veta * etax + vzeta * zetax;
          double dvdy = vxi * xiy + veta * etay + vzeta * zetay;
          double dvdz = vxi * xiz + veta * etaz + vzeta * zetaz;
          double dwdx = wxi * xix + weta * etax + wzeta * zetax;
          double dwdy = wxi * xiy + weta * etay + wzeta * zetay;
          double dwdz = wxi * xiz + weta * etaz + wzeta * zetaz;
          double divv = (dudx + dvdy + dwdz) * (double)P.turcmp;
          double txx = 2 * mut * (dudx - divv / 3);
          double txy = mut * (dudy + dvdx);
          double txz = mut * (dudz + dwdx);
          double tyy = 2 * mut * (dvdy - divv / 3);
          double tyz = mut * (dvdz + dwdy);
          double tzz = 2 * mut * (dwdz - divv / 3);
          double mr = P.machfs / P.reyno;
          Q(tau1, g, i, j, k, 1, n) = mr * txx - fact * rho * tke;
          Q(tau1, g, i, j, k, 2, n) = mr * txy;
          Q(tau1, g, i, j, k, 3, n) = mr * txz;
          Q(tau1, g, i, j, k, 4, n) = mr * tyy - fact * rho * tke;
          Q(tau1, g, i, j, k, 5, n) = mr * tyz;
          Q(tau1, g, i, j, k, 6, n) = mr * tzz - fact * rho * tke;
        }
  }
}

/* vflux_b (src/flux.f90:878-1182), Wilcox k-omega and laminar branches:
 * viscous and turbulent diffusion fluxes at the faces of direction idir.
 * On entry flux(:,:,:,1) holds mut at the faces (q_face with calmut).
 * Kept as in the reference: the omega diffusion uses dk/dz in place of
 * domega/dz, and the energy flux's z term repeats the y-row stresses.    */
void or_vflux_b(int idir, const double* qf, const double* dqxi, const double* dqeta, const double* dqzeta,
                const double* tau1, double* flux, const double* fdist, const double* xider, const double* etader,
                const double* zetader, const double* sijk, int imax, int jmax, int kmax) {
  bg g = mkbg(imax, jmax, kmax);
  int ic, jc, kc;
  face_ranges(idir, g, &ic, &jc, &kc);
  double gamma = P.gamma;
  if (is_kom()) {
    for (int n = 0; n <= 2 * NH; n++) {
      int nh = n * P.hbmove;
      for (int k = 1; k <= kc; k++)
        for (int j = 1; j <= jc; j++)
          for (int i = 1; i <= ic; i++) {
            double kx = M(sijk, g, 1, i, j, k, nh), ky = M(sijk, g, 2, i, j, k, nh), kz = M(sijk, g, 3, i, j, k, nh),
                   ds = M(sijk, g, 4, i, j, k, nh);
            double rho = Q(qf, g, i, j, k, 1, n), p = Q(qf, g, i, j, k, 5, n);
            double t = gamma * p / rho;
            double mu = suth(t);
            double mut = F0(flux, g, i, j, k, 1, n);
            double xix = MD(xider, g, 1, i, j, k, nh), xiy = MD(xider, g, 2, i, j, k, nh),
                   xiz = MD(xider, g, 3, i, j, k, nh);
            double etax = MD(etader, g, 1, i, j, k, nh), etay = MD(etader, g, 2, i, j, k, nh),
                   etaz = MD(etader, g, 3, i, j, k, nh);
            double zetax = MD(zetader, g, 1, i, j, k, nh), zetay = MD(zetader, g, 2, i, j, k, nh),
                   zetaz = MD(zetader, g, 3, i, j, k, nh);
            double tkexi = Q(dqxi, g, i, j, k, 6, n), tkeeta = Q(dqeta, g, i, j, k, 6, n),
                   tkezeta = Q(dqzeta, g, i, j, k, 6, n);
            double omexi = Q(dqxi, g, i, j, k, 7, n), omeeta = Q(dqeta, g, i, j, k, 7, n),
                   omezeta = Q(dqzeta, g, i, j, k, 7, n);
            double dtkedx = tkexi * xix + tkeeta * etax + tkezeta * zetax;
            double dtkedy = tkexi * xiy + tkeeta * etay + tkezeta * zetay;
            double dtkedz = tkexi * xiz + tkeeta * etaz + tkezeta * zetaz;
            double domedx = omexi * xix + omeeta * etax + omezeta * zetax;
            double domedy = omexi * xiy + omeeta * etay + omezeta * zetay;
            double txx = Q(tau1, g, i, j, k, 1, n), txy = Q(tau1, g, i, j, k, 2, n), txz = Q(tau1, g, i, j, k, 3, n);
            double tyy = Q(tau1, g, i, j, k, 4, n), tyz = Q(tau1, g, i, j, k, 5, n), tzz = Q(tau1, g, i, j, k, 6, n);
            F0(flux, g, i, j, k, 2, n) = -((txx * kx + txy * ky + txz * kz) * ds);
            F0(flux, g, i, j, k, 3, n) = -((txy * kx + tyy * ky + tyz * kz) * ds);
            F0(flux, g, i, j, k, 4, n) = -((txz * kx + tyz * ky + tzz * kz) * ds);
            double sk = P.sigk, sw = P.sigw;
            if (is_sst2()) {
              /* BSL/SST: sigma_k, sigma_omega blended with F1 of the face
               * (src/flux.f90:1035-1046), fdist the face wall distance */
              const double fd = fdist[(ll)(i - 1) + (ll)imax * ((ll)(j - 1) + (ll)jmax * (ll)(k - 1))];
              const double tke = Q(qf, g, i, j, k, 6, n), ome = Q(qf, g, i, j, k, 7, n);
              double arg1 = fmax(sqrt(tke) / 0.09 / ome / (fd + P.epst),
                                 500 * mu / rho / ome / ((fd + P.epst) * (fd + P.epst)) * P.machfs / P.reyno);
              const double dmz = omexi * xiz + omeeta * etaz + omezeta * zetaz;
              const double cdt = 2 * rho * SIGW2 / ome * (dtkedx * domedx + dtkedy * domedy + dtkedz * dmz);
              const double acdt = fmax(cdt, 1e-20);
              arg1 = fmin(arg1, 4 * rho * SIGW2 * tke / acdt / ((fd + P.epst) * (fd + P.epst)));
              /* arg1**4 as the reference build evaluates it: ((a a) a) a */
              arg1 = fmin(arg1 * arg1 * arg1 * arg1, 1e2);
              const double f1 = (exp(2 * arg1) - 1) / (exp(2 * arg1) + 1);
              sk = f1 * SIGK1 + (1 - f1) * SIGK2;
              sw = f1 * SIGW1 + (1 - f1) * SIGW2;
            }
            double ctke = (mu + sk * mut) * P.machfs / P.reyno;
            double come = (mu + sw * mut) * P.machfs / P.reyno;
            F0(flux, g, i, j, k, 5, n) = -(ctke * (dtkedx * kx + dtkedy * ky + dtkedz * kz) * ds);
            F0(flux, g, i, j, k, 6, n) = -(ctke * (dtkedx * kx + dtkedy * ky + dtkedz * kz) * ds);
            F0(flux, g, i, j, k, 7, n) = -(come * (domedx * kx + domedy * ky + dtkedz * kz) * ds);
          }
    }
  } else {
    for (int n = 0; n <= 2 * NH; n++)
      for (int k = 1; k <= kc; k++)
        for (int j = 1; j <= jc; j++)
          for (int i = 1; i <= ic; i++)
            for (int v = 1; v <= 5; v++) F0(flux, g, i, j, k, v, n) = 0;
  }
  for (int n = 0; n <= 2 * NH; n++) {
    int nh = n * P.hbmove;
    for (int k = 1; k <= kc; k++)
      for (int j = 1; j <= jc; j++)
        for (int i = 1; i <= ic; i++) {
          double kx = M(sijk, g, 1, i, j, k, nh), ky = M(sijk, g, 2, i, j, k, nh), kz = M(sijk, g, 3, i, j, k, nh),
                 ds = M(sijk, g, 4, i, j, k, nh);
          double rho = Q(qf, g, i, j, k, 1, n), u = Q(qf, g, i, j, k, 2, n), v = Q(qf, g, i, j, k, 3, n),
                 w = Q(qf, g, i, j, k, 4, n), p = Q(qf, g, i, j, k, 5, n);
          double t = gamma * p / rho;
          double mu = suth(t);
          double mut = F0(flux, g, i, j, k, 1, n);
          double xix = MD(xider, g, 1, i, j, k, nh), xiy = MD(xider, g, 2, i, j, k, nh), xiz = MD(xider, g, 3, i, j, k, nh);
          double etax = MD(etader, g, 1, i, j, k, nh), etay = MD(etader, g, 2, i, j, k, nh),
                 etaz = MD(etader, g, 3, i, j, k, nh);
          double zetax = MD(zetader, g, 1, i, j, k, nh), zetay = MD(zetader, g, 2, i, j, k, nh),
                 zetaz = MD(zetader, g, 3, i, j, k, nh);
          double uxi = Q(dqxi, g, i, j, k, 2, n), ueta = Q(dqeta, g, i, j, k, 2, n), uzeta = Q(dqzeta, g, i, j, k, 2, n);
          double vxi = Q(dqxi, g, i, j, k, 3, n), veta = Q(dqeta, g, i, j, k, 3, n), vzeta = Q(dqzeta, g, i, j, k, 3, n);
          double wxi = Q(dqxi, g, i, j, k, 4, n), weta = Q(dqeta, g, i, j, k, 4, n), wzeta = Q(dqzeta, g, i, j, k, 4, n);
          double txi = t * (Q(dqxi, g, i, j, k, 5, n) / p - Q(dqxi, g, i, j, k, 1, n) / rho);
          double teta = t * (Q(dqeta, g, i, j, k, 5, n) / p - Q(dqeta, g, i, j, k, 1, n) / rho);
          double tzeta = t * (Q(dqzeta, g, i, j, k, 5, n) / p - Q(dqzeta, g, i, j, k, 1, n) / rho);
          double dudx = uxi * xix + ueta * etax + uzeta * zetax;
          double dudy = uxi * xiy + ueta * etay + uzeta * zetay;
          double dudz = uxi * xiz + ueta * etaz + uzeta * zetaz;
          double dvdx = vxi * xix + veta * etax + vzeta * zetax;
          double dvdy = vxi * xiy + veta * etay + vzeta * zetay;
          double dvdz = vxi * xiz + veta * etaz + vzeta * zetaz;
          double dwdx = wxi * xix + weta * etax + wzeta * zetax;
          double dwdy = wxi * xiy + weta * etay + wzeta * zetay;
          double dwdz = wxi * xiz + weta * etaz + wzeta * zetaz;
          double dtdx = txi * xix + teta * etax + tzeta * zetax;
          double dtdy = txi * xiy + teta * etay + tzeta * zetay;
          double dtdz = txi * xiz + teta * etaz + tzeta * zetaz;
          double divv = dudx + dvdy + dwdz;
          double mr = P.machfs / P.reyno;
          double txx = mr * 2 * mu * (dudx - divv / 3);
          double txy = mr * mu * (dudy + dvdx);
          double txz = mr * mu * (dudz + dwdx);
          double tyy = mr * 2 * mu * (dvdy - divv / 3);
          double tyz = mr * mu * (dvdz + dwdy);
          double tzz = mr * 2 * mu * (dwdz - divv / 3);
          double cq = mr * (mu / P.pranl + mut / P.prant) / (gamma - 1);
          double qx = -(cq * dtdx), qy = -(cq * dtdy), qz = -(cq * dtdz);
          F0(flux, g, i, j, k, 1, n) = 0;
          F0(flux, g, i, j, k, 2, n) = F0(flux, g, i, j, k, 2, n) - (txx * kx + txy * ky + txz * kz) * ds;
          F0(flux, g, i, j, k, 3, n) = F0(flux, g, i, j, k, 3, n) - (txy * kx + tyy * ky + tyz * kz) * ds;
          F0(flux, g, i, j, k, 4, n) = F0(flux, g, i, j, k, 4, n) - (txz * kx + tyz * ky + tzz * kz) * ds;
          F0(flux, g, i, j, k, 5, n) = F0(flux, g, i, j, k, 5, n) -
                                       ((u * txx + v * txy + w * txz) * kx + (u * txy + v * tyy + w * tyz) * ky +
                                        (u * txy + v * tyy + w * tyz) * kz) * ds +
                                       (qx * kx + qy * ky + qz * kz) * ds;
        }
  }
}

/* bq_der_c{i,j,k} (src/flux.f90:2979-3131): cell-centre differences of the
 * face averages along idir, velocity components (kom) only.              */
void or_q_der_c_b(int idir, const double* qf, double* dqc, int imax, int jmax, int kmax) {
  bg g = mkbg(imax, jmax, kmax);
  int di = idir == 1, dj = idir == 2, dk = idir == 3;
  for (int n = 0; n <= 2 * NH; n++)
    for (int ipde = 2; ipde <= 7; ipde++) {
      if (ipde == 5 || (ipde > 5 && !is_sst2())) continue;   /* k, omega for BSL/SST */
      for (int k = 1; k <= kmax - 1; k++)
        for (int j = 1; j <= jmax - 1; j++)
          for (int i = 1; i <= imax - 1; i++)
            Q(dqc, g, i, j, k, ipde, n) = Q(qf, g, i + di, j + dj, k + dk, ipde, n) - Q(qf, g, i, j, k, ipde, n);
    }
}

/* turprod_b (src/src.f90:40-175): P_d (second, low-round-off form),
 * div(u) and delpls = 2/3 max(div u, 0) at cell centres.                   */
void or_turprod_b(const double* q, const double* dqxi, const double* dqeta, const double* dqzeta, double* prd,
                  double* divvel, double* delpls, const double* const* xd, int imax, int jmax, int kmax) {
  /* xd: xideri, etaderi, zetaderi, xiderj, etaderj, zetaderj, xiderk, etaderk, zetaderk */
  bg g = mkbg(imax, jmax, kmax);
  const double fact = 2.0 / 3.0;
  for (int n = 0; n <= 2 * NH; n++) {
    int nh = n * P.hbmove;
    for (int k = 1; k <= kmax - 1; k++)
      for (int j = 1; j <= jmax - 1; j++)
        for (int i = 1; i <= imax - 1; i++) {
          double xix = (MD(xd[0], g, 1, i + 1, j, k, nh) + MD(xd[0], g, 1, i, j, k, nh)) / 2;
          double xiy = (MD(xd[3], g, 2, i, j + 1, k, nh) + MD(xd[3], g, 2, i, j, k, nh)) / 2;
          double xiz = (MD(xd[6], g, 3, i, j, k + 1, nh) + MD(xd[6], g, 3, i, j, k, nh)) / 2;
          double etax = (MD(xd[1], g, 1, i + 1, j, k, nh) + MD(xd[1], g, 1, i, j, k, nh)) / 2;
          double etay = (MD(xd[4], g, 2, i, j + 1, k, nh) + MD(xd[4], g, 2, i, j, k, nh)) / 2;
          double etaz = (MD(xd[7], g, 3, i, j, k + 1, nh) + MD(xd[7], g, 3, i, j, k, nh)) / 2;
          double zetax = (MD(xd[2], g, 1, i + 1, j, k, nh) + MD(xd[2], g, 1, i, j, k, nh)) / 2;
          double zetay = (MD(xd[5], g, 2, i, j + 1, k, nh) + MD(xd[5], g, 2, i, j, k, nh)) / 2;
          double zetaz = (MD(xd[8], g, 3, i, j, k + 1, nh) + MD(xd[8], g, 3, i, j, k, nh)) / 2;
          double uxi = Q(dqxi, g, i, j, k, 2, n), ueta = Q(dqeta, g, i, j, k, 2, n), uzeta = Q(dqzeta, g, i, j, k, 2, n);
          double vxi = Q(dqxi, g, i, j, k, 3, n), veta = Q(dqeta, g, i, j, k, 3, n), vzeta = Q(dqzeta, g, i, j, k, 3, n);
          double wxi = Q(dqxi, g, i, j, k, 4, n), weta = Q(dqeta, g, i, j, k, 4, n), wzeta = Q(dqzeta, g, i, j, k, 4, n);
          double dudx = uxi * xix + ueta * etax + uzeta * zetax;
          double dudy = uxi * xiy + ueta * etay + uzeta * zetay;
          double dudz = uxi * xiz + ueta * etaz + uzeta * zetaz;
          double dvdx = vxi * xix + veta * etax + vzeta * zetax;
          double dvdy = vxi * xiy + veta * etay + vzeta * zetay;
          double dvdz = vxi * xiz + veta * etaz + vzeta * zetaz;
          double dwdx = wxi * xix + weta * etax + wzeta * zetax;
          double dwdy = wxi * xiy + weta * etay + wzeta * zetay;
          double dwdz = wxi * xiz + weta * etaz + wzeta * zetaz;
          double divv = (dudx + dvdy + dwdz) * (double)P.turcmp;
          double txx = dudx, txy = (dudy + dvdx) / 2, txz = (dudz + dwdx) / 2, tyy = dvdy, tyz = (dvdz + dwdy) / 2,
                 tzz = dwdz;
          double prod = 2 * (txx * txx + tyy * tyy + tzz * tzz + 2 * (txy * txy + txz * txz + tyz * tyz) -
                             divv * divv / 3);
          (void)q;
          S1(prd, g, i, j, k, n) = prod;
          S1(divvel, g, i, j, k, n) = divv;
          S1(delpls, g, i, j, k, n) = fact * fmax(divv, 0.0);
        }
  }
}

/* srckw_b (src/src.f90:209-272): Wilcox k-omega sources into res(6:7);
 * q primitive.                                                           */
void or_srckw_b(const double* q, const double* mutg, const double* prd, const double* divvel, double* res,
                const double* vol, int nl, int imax, int jmax, int kmax) {
  bg g = mkbg(imax, jmax, kmax);
  const double fact = 2.0 / 3.0, mor = P.machfs / P.reyno;
  for (int n = 0; n <= 2 * NH; n++)
    for (int k = 1; k <= kmax - 1; k++)
      for (int j = 1; j <= jmax - 1; j++)
        for (int i = 1; i <= imax - 1; i++) {
          double rho = Q(q, g, i, j, k, 1, n), tke = Q(q, g, i, j, k, 6, n), ome = Q(q, g, i, j, k, 7, n);
          double mut = S1(mutg, g, i, j, k, n);
          double divv = S1(divvel, g, i, j, k, n);
          double prod = S1(prd, g, i, j, k, n), prodw = S1(prd, g, i, j, k, n);
          if (P.lim_prodtke && (nl == 1 || P.prd_all_lvls)) {
            double tstdis = P.prdlim * P.bstrkw * rho * tke * ome;
            if (mor * mut * prod > tstdis) {
              prod = tstdis / mor / mut;
              if (P.lim_prodomega) prodw = prod;
            }
          }
          double v = V1(vol, g, i, j, k);
          Q(res, g, i, j, k, 6, n) =
              Q(res, g, i, j, k, 6, n) - v * (mor * mut * prod - fact * divv * rho * tke - P.bstrkw * rho * tke * ome);
          Q(res, g, i, j, k, 7, n) = Q(res, g, i, j, k, 7, n) - v * (P.gkw * P.cmu * rho * prodw -
                                                                     P.gkw * fact * divv * rho * ome -
                                                                     P.bkw * rho * (ome * ome));
        }
}

/* lim_prod_mg_b (src/src.f90:708-757); q conservative as in the caller. */
void or_lim_prod_mg_b(const double* q, double* prd, const double* mutg, int imax, int jmax, int kmax) {
  bg g = mkbg(imax, jmax, kmax);
  const double mor = P.machfs / P.reyno;
  for (int n = 0; n <= 2 * NH; n++)
    for (int k = 1; k <= kmax - 1; k++)
      for (int j = 1; j <= jmax - 1; j++)
        for (int i = 1; i <= imax - 1; i++) {
          double rho = Q(q, g, i, j, k, 1, n), tke = Q(q, g, i, j, k, 6, n), ome = Q(q, g, i, j, k, 7, n);
          double mut = S1(mutg, g, i, j, k, n), prod = S1(prd, g, i, j, k, n);
          double tstdis = P.prdlim * P.bstrkw * rho * tke * ome;
          if (mor * mut * prod > tstdis) S1(prd, g, i, j, k, n) = tstdis / mor / mut;
        }
}

/* fixq_b (src/update.f90:2216-2269): floor k and omega at qmin(6:7). */
void or_fixq_b(double* q, int imax, int jmax, int kmax) {
  bg g = mkbg(imax, jmax, kmax);
  for (int n = 0; n <= 2 * NH; n++)
    for (int k = 1; k <= kmax - 1; k++)
      for (int j = 1; j <= jmax - 1; j++)
        for (int i = 1; i <= imax - 1; i++) {
          double rho = Q(q, g, i, j, k, 1, n);
          double tke = Q(q, g, i, j, k, 6, n) / rho;
          double ome = Q(q, g, i, j, k, 7, n) / rho;
          if (tke < P.qmin[5]) {
            tke = P.qmin[5];
            Q(q, g, i, j, k, 6, n) = rho * tke;
          }
          if (ome < P.qmin[6]) {
            ome = P.qmin[6];
            Q(q, g, i, j, k, 7, n) = rho * ome;
          }
        }
}

/* tur_vis_b (src/update.f90:2297-2337): mut over 0..imax, 0..jmax, 0..kmax. */
void or_tur_vis_b(const double* q, double* mut, int imax, int jmax, int kmax) {
  bg g = mkbg(imax, jmax, kmax);
  for (int n = 0; n <= 2 * NH; n++)
    for (int k = 0; k <= kmax; k++)
      for (int j = 0; j <= jmax; j++)
        for (int i = 0; i <= imax; i++) {
          double rho = Q(q, g, i, j, k, 1, n), tke = Q(q, g, i, j, k, 6, n), ome = Q(q, g, i, j, k, 7, n);
          S1(mut, g, i, j, k, n) = P.cmu * rho * tke / (ome + P.epst) * P.reyno / P.machfs;
        }
}

/* the cell-centre metric derivatives of turprod_b / update_ttrm_b /
 * tur_vis_sst_b: averages of the two face values along each direction   */
static void cell_metrics(const double* const* xd, bg g, int i, int j, int k, int nh, double m[9]) {
  m[0] = (MD(xd[0], g, 1, i + 1, j, k, nh) + MD(xd[0], g, 1, i, j, k, nh)) / 2; /* xix   */
  m[1] = (MD(xd[3], g, 2, i, j + 1, k, nh) + MD(xd[3], g, 2, i, j, k, nh)) / 2; /* xiy   */
  m[2] = (MD(xd[6], g, 3, i, j, k + 1, nh) + MD(xd[6], g, 3, i, j, k, nh)) / 2; /* xiz   */
  m[3] = (MD(xd[1], g, 1, i + 1, j, k, nh) + MD(xd[1], g, 1, i, j, k, nh)) / 2; /* etax  */
  m[4] = (MD(xd[4], g, 2, i, j + 1, k, nh) + MD(xd[4], g, 2, i, j, k, nh)) / 2; /* etay  */
  m[5] = (MD(xd[7], g, 3, i, j, k + 1, nh) + MD(xd[7], g, 3, i, j, k, nh)) / 2; /* etaz  */
  m[6] = (MD(xd[2], g, 1, i + 1, j, k, nh) + MD(xd[2], g, 1, i, j, k, nh)) / 2; /* zetax */
  m[7] = (MD(xd[5], g, 2, i, j + 1, k, nh) + MD(xd[5], g, 2, i, j, k, nh)) / 2; /* zetay */
  m[8] = (MD(xd[8], g, 3, i, j, k + 1, nh) + MD(xd[8], g, 3, i, j, k, nh)) / 2; /* zetaz */
}

/* facedist_b (src/grid.f90:2834-2884): wall distance of the idir faces,
 * face array (imax, jmax, kmax), from the cell-centre distance.           */
void or_facedist_b(int idir, const double* dist, double* fd, int imax, int jmax, int kmax) {
  bg g = mkbg(imax, jmax, kmax);
#define FD(i, j, k) fd[(ll)((i) - 1) + (ll)imax * ((ll)((j) - 1) + (ll)jmax * (ll)((k) - 1))]
  if (idir == 1) {
    for (int k = 1; k <= kmax - 1; k++)
      for (int j = 1; j <= jmax - 1; j++) {
        FD(1, j, k) = V1(dist, g, 0, j, k);
        for (int i = 2; i <= imax - 1; i++) FD(i, j, k) = (V1(dist, g, i - 1, j, k) + V1(dist, g, i, j, k)) / 2;
        FD(imax, j, k) = V1(dist, g, imax, j, k);
      }
  } else if (idir == 2) {
    for (int k = 1; k <= kmax - 1; k++)
      for (int i = 1; i <= imax - 1; i++) {
        FD(i, 1, k) = V1(dist, g, i, 0, k);
        for (int j = 2; j <= jmax - 1; j++) FD(i, j, k) = (V1(dist, g, i, j - 1, k) + V1(dist, g, i, j, k)) / 2;
        FD(i, jmax, k) = V1(dist, g, i, jmax, k);
      }
  } else {
    for (int j = 1; j <= jmax - 1; j++)
      for (int i = 1; i <= imax - 1; i++) {
        FD(i, j, 1) = V1(dist, g, i, j, 0);
        for (int k = 2; k <= kmax - 1; k++) FD(i, j, k) = (V1(dist, g, i, j, k - 1) + V1(dist, g, i, j, k)) / 2;
        FD(i, j, kmax) = V1(dist, g, i, j, kmax);
      }
  }
#undef FD
}

/* update_ttrm_b (src/update.f90:3036-3161): Wilcox -- ttrm = (0, gkw, bkw);
 * BSL/SST -- F1 from the wall distance, k and omega gradients (cell-centre
 * differences dq of the face averages), ttrm = (cross diffusion (1-F1),
 * blended gamma, blended beta).  q primitive.                           */
void or_update_ttrm_b(const double* q, const double* dqxi, const double* dqeta, const double* dqzeta,
                      const double* const* xd, double* ttrm, const double* dist, int imax, int jmax, int kmax) {
  bg g = mkbg(imax, jmax, kmax);
  for (int n = 0; n <= 2 * NH; n++) {
    int nh = n * P.hbmove;
    for (int k = 1; k <= kmax - 1; k++)
      for (int j = 1; j <= jmax - 1; j++)
        for (int i = 1; i <= imax - 1; i++) {
          if (P.kom) {
            T3(ttrm, g, i, j, k, 1, n) = 0.0;
            T3(ttrm, g, i, j, k, 2, n) = P.gkw;
            T3(ttrm, g, i, j, k, 3, n) = P.bkw;
            continue;
          }
          double rho = Q(q, g, i, j, k, 1, n), tke = Q(q, g, i, j, k, 6, n), ome = Q(q, g, i, j, k, 7, n);
          double p = Q(q, g, i, j, k, 5, n);
          double t = P.gamma * p / rho;
          double mu = suth(t);
          double m[9];
          cell_metrics(xd, g, i, j, k, nh, m);
          double tkexi = Q(dqxi, g, i, j, k, 6, n), tkeeta = Q(dqeta, g, i, j, k, 6, n),
                 tkezeta = Q(dqzeta, g, i, j, k, 6, n);
          double omexi = Q(dqxi, g, i, j, k, 7, n), omeeta = Q(dqeta, g, i, j, k, 7, n),
                 omezeta = Q(dqzeta, g, i, j, k, 7, n);
          double dtkedx = tkexi * m[0] + tkeeta * m[3] + tkezeta * m[6];
          double dtkedy = tkexi * m[1] + tkeeta * m[4] + tkezeta * m[7];
          double dtkedz = tkexi * m[2] + tkeeta * m[5] + tkezeta * m[8];
          double domedx = omexi * m[0] + omeeta * m[3] + omezeta * m[6];
          double domedy = omexi * m[1] + omeeta * m[4] + omezeta * m[7];
          double domedz = omexi * m[2] + omeeta * m[5] + omezeta * m[8];
          double d = V1(dist, g, i, j, k) + P.epst;
          double arg1 = fmax(sqrt(tke) / 0.09 / ome / d, 500 * mu / rho / ome / (d * d) * P.machfs / P.reyno);
          double cdt = 2 * rho * SIGW2 / ome * (dtkedx * domedx + dtkedy * domedy + dtkedz * domedz);
          double acdt = fmax(cdt, 1e-20);
          arg1 = fmin(arg1, 4 * rho * SIGW2 * tke / acdt / (d * d));
          arg1 = arg1 * arg1 * arg1 * arg1;   /* arg1**4: ((a a) a) a */
          arg1 = fmin(arg1, 1e2);
          double f1 = (exp(2 * arg1) - 1) / (exp(2 * arg1) + 1);
          T3(ttrm, g, i, j, k, 1, n) = cdt * (1 - f1);
          T3(ttrm, g, i, j, k, 2, n) = f1 * G1 + (1 - f1) * G2;
          T3(ttrm, g, i, j, k, 3, n) = f1 * B1 + (1 - f1) * B2;
        }
  }
}

/* srckw_sst_b (src/src.f90:306-378): BSL/SST sources with the blended
 * coefficients and cross diffusion of ttrm; q primitive.                 */
void or_srckw_sst_b(const double* q, const double* prd, const double* divvel, double* res, const double* mutg,
                    const double* ttrm, const double* vol, int nl, int imax, int jmax, int kmax) {
  bg g = mkbg(imax, jmax, kmax);
  const double fact = 2.0 / 3.0, mor = P.machfs / P.reyno;
  for (int n = 0; n <= 2 * NH; n++)
    for (int k = 1; k <= kmax - 1; k++)
      for (int j = 1; j <= jmax - 1; j++)
        for (int i = 1; i <= imax - 1; i++) {
          double rho = Q(q, g, i, j, k, 1, n), tke = Q(q, g, i, j, k, 6, n), ome = Q(q, g, i, j, k, 7, n);
          double mut = S1(mutg, g, i, j, k, n);
          double divv = S1(divvel, g, i, j, k, n);
          double prod = S1(prd, g, i, j, k, n), prodw = S1(prd, g, i, j, k, n);
          double cdif = T3(ttrm, g, i, j, k, 1, n), gav = T3(ttrm, g, i, j, k, 2, n), bav = T3(ttrm, g, i, j, k, 3, n);
          if (P.lim_prodtke && (nl == 1 || P.prd_all_lvls)) {
            double tstdis = P.prdlim * P.bstrkw * rho * tke * ome;
            if (mor * mut * prod > tstdis) {
              prod = tstdis / mor / mut;
              if (P.lim_prodomega) prodw = prod;
            }
          }
          double v = V1(vol, g, i, j, k);
          Q(res, g, i, j, k, 6, n) =
              Q(res, g, i, j, k, 6, n) - v * (mor * mut * prod - fact * divv * rho * tke - P.bstrkw * rho * tke * ome);
          Q(res, g, i, j, k, 7, n) = Q(res, g, i, j, k, 7, n) - v * (gav * P.cmu * rho * prodw -
                                                                     gav * fact * divv * rho * ome -
                                                                     bav * rho * (ome * ome) + cdif);
        }
}

/* tur_vis_sst_b (src/update.f90:2375-2524): SST eddy viscosity with the F2
 * limiter on the strain rate; q conservative; velocity differences dq of
 * the cell-centre face averages; halo faces copied from the first interior
 * layer (k, then j, then i faces; no edges).                             */
void or_tur_vis_sst_b(const double* q, double* mut, const double* dqxi, const double* dqeta, const double* dqzeta,
                      const double* const* xd, const double* dist, int imax, int jmax, int kmax) {
  bg g = mkbg(imax, jmax, kmax);
  for (int n = 0; n <= 2 * NH; n++) {
    int nh = n * P.hbmove;
    for (int k = 1; k <= kmax - 1; k++)
      for (int j = 1; j <= jmax - 1; j++)
        for (int i = 1; i <= imax - 1; i++) {
          double q1 = Q(q, g, i, j, k, 1, n), q2 = Q(q, g, i, j, k, 2, n), q3 = Q(q, g, i, j, k, 3, n),
                 q4 = Q(q, g, i, j, k, 4, n);
          double p = (P.gamma - 1) * (Q(q, g, i, j, k, 5, n) - (q2 * q2 + q3 * q3 + q4 * q4) / q1 / 2 -
                                      Q(q, g, i, j, k, 6, n));
          double rho = q1;
          double tke = Q(q, g, i, j, k, 6, n) / rho;
          double ome = Q(q, g, i, j, k, 7, n) / rho;
          double t = P.gamma * p / rho;
          double mu = suth(t);
          double m[9];
          cell_metrics(xd, g, i, j, k, nh, m);
          double uxi = Q(dqxi, g, i, j, k, 2, n), ueta = Q(dqeta, g, i, j, k, 2, n), uzeta = Q(dqzeta, g, i, j, k, 2, n);
          double vxi = Q(dqxi, g, i, j, k, 3, n), veta = Q(dqeta, g, i, j, k, 3, n), vzeta = Q(dqzeta, g, i, j, k, 3, n);
          double wxi = Q(dqxi, g, i, j, k, 4, n), weta = Q(dqeta, g, i, j, k, 4, n), wzeta = Q(dqzeta, g, i, j, k, 4, n);
          double dudx = uxi * m[0] + ueta * m[3] + uzeta * m[6];
          double dudy = uxi * m[1] + ueta * m[4] + uzeta * m[7];
          double dudz = uxi * m[2] + ueta * m[5] + uzeta * m[8];
          double dvdx = vxi * m[0] + veta * m[3] + vzeta * m[6];
          double dvdy = vxi * m[1] + veta * m[4] + vzeta * m[7];
          double dvdz = vxi * m[2] + veta * m[5] + vzeta * m[8];
          double dwdx = wxi * m[0] + weta * m[3] + wzeta * m[6];
          double dwdy = wxi * m[1] + weta * m[4] + wzeta * m[7];
          double dwdz = wxi * m[2] + weta * m[5] + wzeta * m[8];
          double sxx = dudx, sxy = (dudy + dvdx) / 2, sxz = (dudz + dwdx) / 2, syy = dvdy, syz = (dvdz + dwdy) / 2,
                 szz = dwdz;
          double sv = sxx * sxx + syy * syy + szz * szz + 2 * (sxy * sxy + sxz * sxz + syz * syz);
          double d = V1(dist, g, i, j, k) + P.epst;
          double arg2 = fmax(2 * sqrt(tke) / 0.09 / ome / d, 500 * mu / rho / ome / (d * d) * P.machfs / P.reyno);
          arg2 = arg2 * arg2;
          arg2 = fmin(arg2, 1e2);
          double f2 = (exp(2 * arg2) - 1) / (exp(2 * arg2) + 1);
          double fact = sqrt(2 * sv);
          S1(mut, g, i, j, k, n) = SST_A1 * rho * tke / fmax(SST_A1 * ome, fact * f2) * P.reyno / P.machfs;
        }
    for (int k = 0; k <= kmax; k += kmax)
      for (int j = 1; j <= jmax - 1; j++)
        for (int i = 1; i <= imax - 1; i++) S1(mut, g, i, j, k, n) = S1(mut, g, i, j, k == 0 ? 1 : kmax - 1, n);
    for (int k = 1; k <= kmax - 1; k++)
      for (int j = 0; j <= jmax; j += jmax)
        for (int i = 1; i <= imax - 1; i++) S1(mut, g, i, j, k, n) = S1(mut, g, i, j == 0 ? 1 : jmax - 1, k, n);
    for (int k = 1; k <= kmax - 1; k++)
      for (int j = 1; j <= jmax - 1; j++)
        for (int i = 0; i <= imax; i += imax) S1(mut, g, i, j, k, n) = S1(mut, g, i == 0 ? 1 : imax - 1, j, k, n);
  }
}

/* update_b / update_bhb explicit branch with the Wilcox k-omega point-
 * implicit rho*omega, rho*k update (src/update.f90:596-634, 1327-1366);
 * ttrm(:,:,:,2:3) of kom are the constants gkw, bkw (update_ttrm_b,
 * src/update.f90:3069-3081).                                              */
void or_update_kb(double* res, double* q, const double* qold, const double* vol, const double* dtvol,
                  const double* dtvolt, const double* delpls, const double* prd, double ar, int imax, int jmax,
                  int kmax, const double* ttrm) {
  bg g = mkbg(imax, jmax, kmax);
  for (int n = 0; n <= 2 * NH; n++) {
    if (is_kom())
      for (int k = 1; k <= kmax - 1; k++)
        for (int j = 1; j <= jmax - 1; j++)
          for (int i = 1; i <= imax - 1; i++) {
            double rho = Q(q, g, i, j, k, 1, n);
            double ome = Q(q, g, i, j, k, 7, n) / Q(q, g, i, j, k, 1, n);
            double tke = Q(q, g, i, j, k, 6, n) / Q(q, g, i, j, k, 1, n);
            double dtt = P1(dtvolt, g, i, j, k, n);
            double dtau = dtt * V1(vol, g, i, j, k);
            double gav = ttrm ? T3(ttrm, g, i, j, k, 2, n) : P.gkw, bav = ttrm ? T3(ttrm, g, i, j, k, 3, n) : P.bkw;
            double dl = S1(delpls, g, i, j, k, n);
            double rhomega = (Q(qold, g, i, j, k, 7, n) - ar * Q(res, g, i, j, k, 7, n) * dtt +
                              ar * dtau * rho * ome * (gav * dl + 2 * bav * ome)) /
                             (1 + ar * dtau * (gav * dl + 2 * bav * ome));
            double rhomegamin = gav * P.cmu * rho * sqrt(S1(prd, g, i, j, k, n));
            if (rhomega < rhomegamin) rhomega = rhomegamin;
            Q(q, g, i, j, k, 7, n) = rhomega;
            double drhomega = Q(q, g, i, j, k, 7, n) - rho * ome;
            Q(q, g, i, j, k, 6, n) = (Q(qold, g, i, j, k, 6, n) - ar * Q(res, g, i, j, k, 6, n) * dtt +
                                      ar * dtau * rho * tke * (dl + P.bstrkw * ome) -
                                      ar * dtau * P.bstrkw * tke * drhomega) /
                                     (1 + ar * dtau * (dl + P.bstrkw * ome));
          }
    for (int ipde = 1; ipde <= 5; ipde++)
      for (int k = 1; k <= kmax - 1; k++)
        for (int j = 1; j <= jmax - 1; j++)
          for (int i = 1; i <= imax - 1; i++)
            Q(q, g, i, j, k, ipde, n) =
                Q(qold, g, i, j, k, ipde, n) - ar * Q(res, g, i, j, k, ipde, n) * P1(dtvol, g, i, j, k, n);
  }
}

/* lim_res_ome_b (src/update.f90:2585-2835), explicit branch without
 * low-speed preconditioning; q conservative.                              */
void or_lim_res_ome_b(double* res, const double* q, const double* vol, const double* dtvolt, const double* delpls,
                      const double* prd, double ar, int imax, int jmax, int kmax, const double* ttrm) {
  bg g = mkbg(imax, jmax, kmax);
  for (int k = 1; k <= kmax - 1; k++)
    for (int j = 1; j <= jmax - 1; j++)
      for (int i = 1; i <= imax - 1; i++) {
        double gav = ttrm ? T3(ttrm, g, i, j, k, 2, 0) : P.gkw, bav = ttrm ? T3(ttrm, g, i, j, k, 3, 0) : P.bkw;
        double dtt = P1(dtvolt, g, i, j, k, 0);
        double dtau = dtt * V1(vol, g, i, j, k);
        double rho = Q(q, g, i, j, k, 1, 0);
        double ome = Q(q, g, i, j, k, 7, 0) / Q(q, g, i, j, k, 1, 0);
        double den1 = 1 + ar * dtau * (gav * S1(delpls, g, i, j, k, 0) + 2 * bav * ome);
        double rhomega = Q(q, g, i, j, k, 7, 0) - ar * Q(res, g, i, j, k, 7, 0) * dtt / den1;
        double rhomegamin = gav * P.cmu * rho * sqrt(S1(prd, g, i, j, k, 0));
        if (rhomega < rhomegamin) Q(res, g, i, j, k, 7, 0) = den1 * (Q(q, g, i, j, k, 7, 0) - rhomegamin) / dtt / ar;
      }
}

/* ---------------------------------------------------------------------- */
/* dual time stepping, per block (one snapshot: dualt excludes harbal)     */

/* src_uns_b (src/src.f90:406-450): the physical-time term of the current
 * state added to the residual, backward Euler on the first step of a
 * first-order start, else BDF2's 3/2.                                     */
void or_src_uns_b(const double* q, const double* vol, double* res, int imax, int jmax, int kmax) {
  bg g = mkbg(imax, jmax, kmax);
  const int first = ITIME == 1 && !DUAL_PROP;
  for (int ipde = 1; ipde <= NPDE; ipde++)
    for (int k = 1; k <= kmax - 1; k++)
      for (int j = 1; j <= jmax - 1; j++)
        for (int i = 1; i <= imax - 1; i++) {
          const double qv = Q(q, g, i, j, k, ipde, 0), v = V1(vol, g, i, j, k);
          Q(res, g, i, j, k, ipde, 0) = first ? Q(res, g, i, j, k, ipde, 0) + qv * v / DT
                                              : Q(res, g, i, j, k, ipde, 0) + 3 * qv * v / 2 / DT;
        }
}

/* rhs_uns_b (src/src.f90:480-525): the known physical-time levels,
 * q1 = q^n, q2 = q^(n-1).                                                 */
void or_rhs_uns_b(double* rhs, const double* q1, const double* q2, const double* vol, int imax, int jmax,
                  int kmax) {
  bg g = mkbg(imax, jmax, kmax);
  const int first = ITIME == 1 && !DUAL_PROP;
  for (int ipde = 1; ipde <= NPDE; ipde++)
    for (int k = 1; k <= kmax - 1; k++)
      for (int j = 1; j <= jmax - 1; j++)
        for (int i = 1; i <= imax - 1; i++) {
          const double a = Q(q1, g, i, j, k, ipde, 0), v = V1(vol, g, i, j, k);
          Q(rhs, g, i, j, k, ipde, 0) = first ? a * v / DT : (2 * a - Q(q2, g, i, j, k, ipde, 0) / 2) * v / DT;
        }
}

/* resg_brms, rkimp without IRS on one level (src/cosa.f90:4534-4549): the
 * physical-time term the implicit update carries, added back.            */
void or_resg_brms_dual(const double* q, const double* res, const double* rhs, const double* vol, double* rb,
                       int imax, int jmax, int kmax) {
  bg g = mkbg(imax, jmax, kmax);
  const double fact = dual_fact();
  for (int p = 0; p <= NPDE; p++) rb[p] = 0;
  for (int ipde = 1; ipde <= NPDE; ipde++)
    for (int k = 1; k <= kmax - 1; k++)
      for (int j = 1; j <= jmax - 1; j++)
        for (int i = 1; i <= imax - 1; i++) {
          double d = fact * Q(q, g, i, j, k, ipde, 0) * V1(vol, g, i, j, k) / DT + Q(res, g, i, j, k, ipde, 0) -
                     Q(rhs, g, i, j, k, ipde, 0);
          rb[ipde - 1] = rb[ipde - 1] + d * d;
        }
  for (int ipde = 1; ipde <= NPDE; ipde++) rb[NPDE] = rb[NPDE] + rb[ipde - 1];
}

/* update_b, rkimp branches (src/update.f90:649-762): the stage with the
 * physical-time term implicit.  Without IRS on one level (plain) the
 * residual carries no src_uns and the term enters the denominator only;
 * otherwise the residual holds it and the update adds it back explicitly
 * before dividing.  k-omega: the point-implicit rho*omega, rho*k of the
 * steady update with the same physical-time term.                         */
void or_update_dual_b(double* res, double* q, const double* qold, const double* vol, const double* dtvol,
                      const double* dtvolt, const double* delpls, const double* prd, double ar, int imax, int jmax,
                      int kmax, const double* ttrm) {
  bg g = mkbg(imax, jmax, kmax);
  const double fact = dual_fact(), dt = DT;
  const int plain = P.irsop == 0 && P.nlevel == 1;
  if (is_kom())
    for (int k = 1; k <= kmax - 1; k++)
      for (int j = 1; j <= jmax - 1; j++)
        for (int i = 1; i <= imax - 1; i++) {
          const double rho = Q(q, g, i, j, k, 1, 0);
          const double ome = Q(q, g, i, j, k, 7, 0) / Q(q, g, i, j, k, 1, 0);
          const double dtt = P1(dtvolt, g, i, j, k, 0);
          const double dtau = dtt * V1(vol, g, i, j, k);
          const double gav = ttrm ? T3(ttrm, g, i, j, k, 2, 0) : P.gkw, bav = ttrm ? T3(ttrm, g, i, j, k, 3, 0) : P.bkw;
          const double dl = S1(delpls, g, i, j, k, 0);
          double rhomega;
          if (plain)
            rhomega = (Q(qold, g, i, j, k, 7, 0) - ar * Q(res, g, i, j, k, 7, 0) * dtt +
                       ar * dtau * Q(q, g, i, j, k, 7, 0) * (gav * dl + 2 * bav * ome)) /
                      (1 + ar * dtau * (gav * dl + 2 * bav * ome) + fact * ar * dtau / dt);
          else
            rhomega = (Q(qold, g, i, j, k, 7, 0) + ar * fact * dtau / dt * Q(q, g, i, j, k, 7, 0) -
                       ar * Q(res, g, i, j, k, 7, 0) * dtt +
                       ar * dtau * Q(q, g, i, j, k, 7, 0) * (gav * dl + 2 * bav * ome)) /
                      (1 + ar * dtau * (gav * dl + 2 * bav * ome) + ar * fact * dtau / dt);
          const double rhomegamin = gav * P.cmu * rho * sqrt(S1(prd, g, i, j, k, 0));
          if (rhomega < rhomegamin) rhomega = rhomegamin;
          Q(q, g, i, j, k, 7, 0) = rhomega;
          const double drhomega = Q(q, g, i, j, k, 7, 0) - rho * ome;
          if (plain)
            Q(q, g, i, j, k, 6, 0) =
                (Q(qold, g, i, j, k, 6, 0) - ar * Q(res, g, i, j, k, 6, 0) * dtt +
                 ar * dtau * Q(q, g, i, j, k, 6, 0) * (dl + P.bstrkw * ome - P.bstrkw * drhomega / rho)) /
                (1 + ar * dtau * (dl + P.bstrkw * ome) + ar * fact * dtau / dt);
          else
            Q(q, g, i, j, k, 6, 0) =
                (Q(qold, g, i, j, k, 6, 0) + ar * fact * dtau / dt * Q(q, g, i, j, k, 6, 0) -
                 ar * Q(res, g, i, j, k, 6, 0) * dtt +
                 ar * dtau * Q(q, g, i, j, k, 6, 0) * (dl + P.bstrkw * ome - P.bstrkw * drhomega / rho)) /
                (1 + ar * dtau * (dl + P.bstrkw * ome) + ar * fact * dtau / dt);
        }
  for (int ipde = 1; ipde <= 5; ipde++)
    for (int k = 1; k <= kmax - 1; k++)
      for (int j = 1; j <= jmax - 1; j++)
        for (int i = 1; i <= imax - 1; i++) {
          const double dv = P1(dtvol, g, i, j, k, 0), v = V1(vol, g, i, j, k);
          double* qq = &Q(q, g, i, j, k, ipde, 0);
          if (plain)
            *qq = (Q(qold, g, i, j, k, ipde, 0) - ar * Q(res, g, i, j, k, ipde, 0) * dv) / (1 + fact * ar * dv * v / dt);
          else
            *qq = (Q(qold, g, i, j, k, ipde, 0) - ar * dv * (-(fact * v / dt * *qq) + Q(res, g, i, j, k, ipde, 0))) /
                  (1 + fact * ar * dv * v / dt);
        }
}

/* lim_res_ome_b, rkimp branch without low-speed preconditioning
 * (src/update.f90:2753-2794): the multigrid transfer's omega residual
 * limit with the physical-time term in the denominator.  den1 carries no
 * RK coefficient here, as in the reference.                               */
void or_lim_res_ome_dual_b(double* res, const double* q, const double* vol, const double* dtvolt,
                           const double* delpls, const double* prd, double ar, int imax, int jmax, int kmax,
                           const double* ttrm) {
  bg g = mkbg(imax, jmax, kmax);
  if (!(P.irsop > 0 || (P.irsop == 0 && P.nlevel > 1))) return;
  const double fact = dual_fact();
  for (int k = 1; k <= kmax - 1; k++)
    for (int j = 1; j <= jmax - 1; j++)
      for (int i = 1; i <= imax - 1; i++) {
        const double gav = ttrm ? T3(ttrm, g, i, j, k, 2, 0) : P.gkw, bav = ttrm ? T3(ttrm, g, i, j, k, 3, 0) : P.bkw;
        const double dtt = P1(dtvolt, g, i, j, k, 0);
        const double dtau = dtt * V1(vol, g, i, j, k);
        const double rho = Q(q, g, i, j, k, 1, 0);
        const double ome = Q(q, g, i, j, k, 7, 0) / Q(q, g, i, j, k, 1, 0);
        const double den1 = 1 + dtau * (gav * S1(delpls, g, i, j, k, 0) + 2 * bav * ome);
        const double den2 = ar * fact * dtau / DT;
        const double rhomega = Q(q, g, i, j, k, 7, 0) - ar * Q(res, g, i, j, k, 7, 0) * dtt / (den1 + den2);
        const double rhomegamin = gav * P.cmu * rho * sqrt(S1(prd, g, i, j, k, 0));
        if (rhomega < rhomegamin)
          Q(res, g, i, j, k, 7, 0) = ((den1 + den2) * (Q(q, g, i, j, k, 7, 0) - rhomegamin)) / dtt / ar;
      }
}

/* ---------------------------------------------------------------------- */
/* Dense LU with partial pivoting and its solve, as the reference's calls
 * DGETRF / DGETRS('N') compute them (src/update.f90:1460-1478): LAPACK's
 * right-looking elimination -- pivot = first row of largest |a| in the
 * column, rows swapped whole, multipliers a(i,j) * (1/a(j,j)) when
 * |a(j,j)| >= the smallest normal number, rank-1 update a(i,l) - a(i,j) a(j,l)
 * in column order -- then row interchanges on b, forward substitution with
 * the unit lower factor and back substitution with the upper one, both by
 * columns.  a is m x m column-major, b is m x nrhs column-major.  LAPACK is
 * not part of the reference tree (an external library of the build); this is
 * its published algorithm.  Returns info as DGETRF: 0, or j+1 for the first
 * zero pivot.                                                              */
int or_lu_factor(double* a, int m, int* ipiv) {
  int info = 0;
  for (int j = 0; j < m; j++) {
    int p = j;
    double amax = fabs(a[j + (ll)m * j]);
    for (int i = j + 1; i < m; i++)
      if (fabs(a[i + (ll)m * j]) > amax) {
        amax = fabs(a[i + (ll)m * j]);
        p = i;
      }
    ipiv[j] = p;
    if (a[p + (ll)m * j] != 0.0) {
      if (p != j)
        for (int l = 0; l < m; l++) {
          double t = a[j + (ll)m * l];
          a[j + (ll)m * l] = a[p + (ll)m * l];
          a[p + (ll)m * l] = t;
        }
      const double piv = a[j + (ll)m * j];
      if (fabs(piv) >= 2.2250738585072014e-308) {
        const double r = 1.0 / piv;
        for (int i = j + 1; i < m; i++) a[i + (ll)m * j] = a[i + (ll)m * j] * r;
      } else {
        for (int i = j + 1; i < m; i++) a[i + (ll)m * j] = a[i + (ll)m * j] / piv;
      }
    } else if (info == 0) {
      info = j + 1;
    }
    for (int l = j + 1; l < m; l++) {
      const double u = a[j + (ll)m * l];
      if (u != 0.0)
        for (int i = j + 1; i < m; i++) a[i + (ll)m * l] = a[i + (ll)m * l] - a[i + (ll)m * j] * u;
    }
  }
  return info;
}
void or_lu_solve(const double* a, int m, const int* ipiv, double* b, int nrhs) {
  for (int c = 0; c < nrhs; c++) {
    double* x = b + (ll)m * c;
    for (int j = 0; j < m; j++)
      if (ipiv[j] != j) {
        double t = x[j];
        x[j] = x[ipiv[j]];
        x[ipiv[j]] = t;
      }
    for (int k = 0; k < m; k++)
      if (x[k] != 0.0)
        for (int i = k + 1; i < m; i++) x[i] = x[i] - x[k] * a[i + (ll)m * k];
    for (int k = m - 1; k >= 0; k--)
      if (x[k] != 0.0) {
        x[k] = x[k] / a[k + (ll)m * k];
        for (int i = 0; i < k; i++) x[i] = x[i] - x[k] * a[i + (ll)m * k];
      }
  }
}
/* m x m system with nrhs right-hand sides, solved in place in b */
int or_lu_solve_dense(int m, int nrhs, double* a, double* b) {
  int ipiv[131];
  if (m < 1 || m > 131) return -1;
  int info = or_lu_factor(a, m, ipiv);
  or_lu_solve(a, m, ipiv, b, nrhs);
  return info;
}

/* update_bhb, implicit branch rkim (src/update.f90:1410-1841): per cell the
 * 2N+1 snapshots are coupled through the spectral operator and solved with
 * an LU, first rho*omega (limited at rhomegamin), then rho*k, then the mean
 * flow (5 right-hand sides).  Without IRS and with one grid level the
 * right-hand side is the explicit stage (src_hb was not added to res,
 * src/cosa.f90:1916-1921); otherwise the HB term qhb0 (kept by src_hb) is
 * taken back out of the smoothed residual.  gav, bav = ttrm(2:3) = gkw, bkw
 * for Wilcox k-omega.  qhb0 is overwritten, as in the reference.          */
void or_update_bhb_rkim(double* res, double* q, const double* qold, const double* vol, const double* dtvol,
                        const double* dtvolt, double* qhb0, const double* delpls, const double* prd, double ar,
                        int imax, int jmax, int kmax, const double* ttrm) {
  bg g = mkbg(imax, jmax, kmax);
  const int m = DIM5;
  const int plain = P.irsop == 0 && P.nlevel == 1;
  double* a = (double*)malloc(sizeof(double) * (size_t)m * m);
  double* w = (double*)malloc(sizeof(double) * (size_t)m * 5);
  double* old7 = (double*)malloc(sizeof(double) * (size_t)m);
  double* drho = (double*)malloc(sizeof(double) * (size_t)m);
  int ipiv[131];
  for (int k = 1; k <= kmax - 1; k++)
    for (int j = 1; j <= jmax - 1; j++)
      for (int i = 1; i <= imax - 1; i++) {
        const double vv = V1(vol, g, i, j, k);
        if (is_kom()) {
          /* rho*omega */
          for (int ni = 0; ni < m; ni++) {
            const double gav = ttrm ? T3(ttrm, g, i, j, k, 2, ni) : P.gkw;
            const double bav = ttrm ? T3(ttrm, g, i, j, k, 3, ni) : P.bkw;
            double ome = Q(q, g, i, j, k, 7, ni) / Q(q, g, i, j, k, 1, ni);
            double dtt = P1(dtvolt, g, i, j, k, ni);
            double dtau = dtt * vv;
            double dps = S1(delpls, g, i, j, k, ni);
            if (plain)
              Q(qhb0, g, i, j, k, 7, ni) = Q(qold, g, i, j, k, 7, ni) - ar * Q(res, g, i, j, k, 7, ni) * dtt +
                                           ar * dtau * Q(q, g, i, j, k, 7, ni) * (gav * dps + 2 * bav * ome);
            else
              Q(qhb0, g, i, j, k, 7, ni) = Q(qold, g, i, j, k, 7, ni) +
                                           ar * dtt * (Q(qhb0, g, i, j, k, 7, ni) - Q(res, g, i, j, k, 7, ni)) +
                                           ar * dtau * Q(q, g, i, j, k, 7, ni) * (gav * dps + 2 * bav * ome);
            old7[ni] = Q(q, g, i, j, k, 7, ni);
            for (int nj = 0; nj < m; nj++)
              a[ni + (ll)m * nj] = (double)KRD(ni, nj) * (1 + ar * dtau * (gav * dps + 2 * bav * ome)) +
                                   ar * P.omega * dtau * D(ni, nj);
          }
          or_lu_factor(a, m, ipiv);
          for (int ni = 0; ni < m; ni++) w[ni] = Q(qhb0, g, i, j, k, 7, ni);
          or_lu_solve(a, m, ipiv, w, 1);
          for (int ni = 0; ni < m; ni++) {
            const double gav = ttrm ? T3(ttrm, g, i, j, k, 2, ni) : P.gkw;
            double rho = Q(q, g, i, j, k, 1, ni);
            double rhomegamin = gav * P.cmu * rho * sqrt(S1(prd, g, i, j, k, ni));
            if (w[ni] < rhomegamin) w[ni] = rhomegamin;
            Q(q, g, i, j, k, 7, ni) = w[ni];
            drho[ni] = Q(q, g, i, j, k, 7, ni) - old7[ni];
          }
          /* rho*k */
          for (int ni = 0; ni < m; ni++) {
            double rho = Q(q, g, i, j, k, 1, ni);
            double ome = old7[ni] / Q(q, g, i, j, k, 1, ni);
            double dtt = P1(dtvolt, g, i, j, k, ni);
            double dtau = dtt * vv;
            double dps = S1(delpls, g, i, j, k, ni);
            if (plain)
              Q(qhb0, g, i, j, k, 6, ni) =
                  (Q(qold, g, i, j, k, 6, ni) - ar * Q(res, g, i, j, k, 6, ni) * dtt +
                   ar * dtau * Q(q, g, i, j, k, 6, ni) * (dps + P.bstrkw * ome - P.bstrkw * drho[ni] / rho));
            else
              Q(qhb0, g, i, j, k, 6, ni) =
                  (Q(qold, g, i, j, k, 6, ni) + ar * dtt * (Q(qhb0, g, i, j, k, 6, ni) - Q(res, g, i, j, k, 6, ni)) +
                   ar * dtau * Q(q, g, i, j, k, 6, ni) * (dps + P.bstrkw * ome - P.bstrkw * drho[ni] / rho));
            for (int nj = 0; nj < m; nj++)
              a[ni + (ll)m * nj] =
                  (double)KRD(ni, nj) * (1 + ar * dtau * (dps + P.bstrkw * ome)) + ar * P.omega * dtau * D(ni, nj);
          }
          or_lu_factor(a, m, ipiv);
          for (int ni = 0; ni < m; ni++) w[ni] = Q(qhb0, g, i, j, k, 6, ni);
          or_lu_solve(a, m, ipiv, w, 1);
          for (int ni = 0; ni < m; ni++) Q(q, g, i, j, k, 6, ni) = w[ni];
        }
        /* mean flow */
        for (int ni = 0; ni < m; ni++) {
          double dt = P1(dtvol, g, i, j, k, ni);
          for (int v = 1; v <= 5; v++) {
            if (plain)
              Q(qhb0, g, i, j, k, v, ni) = Q(qold, g, i, j, k, v, ni) - ar * Q(res, g, i, j, k, v, ni) * dt;
            else
              Q(qhb0, g, i, j, k, v, ni) =
                  Q(qold, g, i, j, k, v, ni) + (Q(qhb0, g, i, j, k, v, ni) - Q(res, g, i, j, k, v, ni)) * ar * dt;
          }
          double dtau = dt * vv;
          for (int nj = 0; nj < m; nj++)
            a[ni + (ll)m * nj] = (double)KRD(ni, nj) + ar * P.omega * dtau * D(ni, nj);
        }
        or_lu_factor(a, m, ipiv);
        for (int ni = 0; ni < m; ni++)
          for (int v = 1; v <= 5; v++) w[ni + (ll)m * (v - 1)] = Q(qhb0, g, i, j, k, v, ni);
        or_lu_solve(a, m, ipiv, w, 5);
        for (int ni = 0; ni < m; ni++)
          for (int v = 1; v <= 5; v++) Q(q, g, i, j, k, v, ni) = w[ni + (ll)m * (v - 1)];
      }
  free(a);
  free(w);
  free(old7);
  free(drho);
}

/* lim_res_ome_bhb (src/update.f90:2836-2998) without low-speed
 * preconditioning: the omega residual limited so that the stage would not
 * take rho*omega below rhomegamin; explicit (per snapshot, no RK
 * coefficient, from qold) or rkim (per cell, the snapshots coupled as in
 * the update, right-hand side res*dtvolt).                                */
void or_lim_res_ome_bhb(double* res, const double* q, const double* qold, const double* vol, const double* dtvolt,
                        const double* delpls, const double* prd, int imax, int jmax, int kmax, const double* ttrm) {
  bg g = mkbg(imax, jmax, kmax);
  const int m = DIM5;
#define GAV(n) (ttrm ? T3(ttrm, g, i, j, k, 2, n) : P.gkw)
#define BAV(n) (ttrm ? T3(ttrm, g, i, j, k, 3, n) : P.bkw)
  if (!P.rkimp) {
    for (int n = 0; n < m; n++)
      for (int k = 1; k <= kmax - 1; k++)
        for (int j = 1; j <= jmax - 1; j++)
          for (int i = 1; i <= imax - 1; i++) {
            double rho = Q(q, g, i, j, k, 1, n);
            double ome = Q(q, g, i, j, k, 7, n) / Q(q, g, i, j, k, 1, n);
            double dtt = P1(dtvolt, g, i, j, k, n);
            double dtau = dtt * V1(vol, g, i, j, k);
            double dl = S1(delpls, g, i, j, k, n);
            const double gav = GAV(n), bav = BAV(n);
            double rhomega = Q(qold, g, i, j, k, 7, n) - Q(res, g, i, j, k, 7, n) * dtt / (1 + dtau * (gav * dl + 2 * bav * ome));
            double rhomegamin = gav * P.cmu * rho * sqrt(S1(prd, g, i, j, k, n));
            double res_ome = Q(res, g, i, j, k, 7, n) -
                             (1 + dtau * (gav * dl + 2 * bav * ome)) * fmax(0.0, (rhomegamin - rhomega) / dtt);
            Q(res, g, i, j, k, 7, n) = res_ome;
          }
    return;
  }
  double* a = (double*)malloc(sizeof(double) * (size_t)m * m);
  double* w = (double*)malloc(sizeof(double) * (size_t)m);
  int ipiv[131];
  for (int k = 1; k <= kmax - 1; k++)
    for (int j = 1; j <= jmax - 1; j++)
      for (int i = 1; i <= imax - 1; i++) {
        for (int ni = 0; ni < m; ni++) {
          double dtau = P1(dtvolt, g, i, j, k, ni) * V1(vol, g, i, j, k);
          double ome = Q(q, g, i, j, k, 7, ni) / Q(q, g, i, j, k, 1, ni);
          double dps = S1(delpls, g, i, j, k, ni);
          const double gav = GAV(ni), bav = BAV(ni);
          for (int nj = 0; nj < m; nj++)
            a[ni + (ll)m * nj] =
                (double)KRD(ni, nj) * (1 + dtau * (gav * dps + 2 * bav * ome)) + P.omega * dtau * D(ni, nj);
        }
        or_lu_factor(a, m, ipiv);
        for (int n = 0; n < m; n++) w[n] = Q(res, g, i, j, k, 7, n) * P1(dtvolt, g, i, j, k, n);
        or_lu_solve(a, m, ipiv, w, 1);
        for (int n = 0; n < m; n++) {
          double rho = Q(q, g, i, j, k, 1, n);
          double rhomegamin = GAV(n) * P.cmu * rho * sqrt(S1(prd, g, i, j, k, n));
          double rhomeganew = Q(qold, g, i, j, k, 7, n) - w[n];
          double res_ome = Q(res, g, i, j, k, 7, n) - fmax(0.0, (rhomegamin - rhomeganew)) / P1(dtvolt, g, i, j, k, n);
          Q(res, g, i, j, k, 7, n) = res_ome;
        }
      }
  free(a);
  free(w);
}
#undef GAV
#undef BAV

/* resg_brms_hb with rkim, no IRS and one grid level (src/cosa.f90:4657-
 * 4674): the HB term, which src_hb did not add to res, is added to the copy
 * work1 before the sums: work1 += omega dhb(ni,nj) q(nj) vol in nj order. */
void or_rms_hb_term(double* work1, const double* q, const double* vol, int imax, int jmax, int kmax) {
  bg g = mkbg(imax, jmax, kmax);
  for (int ni = 0; ni < DIM5; ni++)
    for (int nj = 0; nj < DIM5; nj++)
      for (int v = 1; v <= NPDE; v++)
        for (int k = 1; k <= kmax - 1; k++)
          for (int j = 1; j <= jmax - 1; j++)
            for (int i = 1; i <= imax - 1; i++)
              Q(work1, g, i, j, k, v, ni) =
                  Q(work1, g, i, j, k, v, ni) + P.omega * D(ni, nj) * Q(q, g, i, j, k, v, nj) * V1(vol, g, i, j, k);
}

/* rest4_b (src/cosa.f90:3122-3164): volume-weighted restriction of a
 * scalar P3 field (production, divergence, delpls).                       */
void or_rest4_b(int imax1, int jmax1, int kmax1, int imax2, int jmax2, int kmax2, const double* prd1, double* prd2,
                const double* vol1, const double* vol2) {
  bg g1 = mkbg(imax1, jmax1, kmax1), g2 = mkbg(imax2, jmax2, kmax2);
  for (int n = 0; n <= 2 * NH; n++)
    for (int k1 = 1; k1 <= kmax1 - 2; k1 += 2)
      for (int j1 = 1; j1 <= jmax1 - 2; j1 += 2)
        for (int i1 = 1; i1 <= imax1 - 2; i1 += 2) {
          int i2 = 1 + i1 / 2, j2 = 1 + j1 / 2, k2 = 1 + k1 / 2;
          S1(prd2, g2, i2, j2, k2, n) =
              (V1(vol1, g1, i1, j1, k1) * S1(prd1, g1, i1, j1, k1, n) +
               V1(vol1, g1, i1, j1 + 1, k1) * S1(prd1, g1, i1, j1 + 1, k1, n) +
               V1(vol1, g1, i1 + 1, j1, k1) * S1(prd1, g1, i1 + 1, j1, k1, n) +
               V1(vol1, g1, i1 + 1, j1 + 1, k1) * S1(prd1, g1, i1 + 1, j1 + 1, k1, n) +
               V1(vol1, g1, i1, j1, k1 + 1) * S1(prd1, g1, i1, j1, k1 + 1, n) +
               V1(vol1, g1, i1, j1 + 1, k1 + 1) * S1(prd1, g1, i1, j1 + 1, k1 + 1, n) +
               V1(vol1, g1, i1 + 1, j1, k1 + 1) * S1(prd1, g1, i1 + 1, j1, k1 + 1, n) +
               V1(vol1, g1, i1 + 1, j1 + 1, k1 + 1) * S1(prd1, g1, i1 + 1, j1 + 1, k1 + 1, n)) /
              V1(vol2, g2, i2, j2, k2);
        }
}

/* ---------------------------------------------------------------------- */
/* multigrid transfer                                                      */

/* rest1_b (src/cosa.f90:2788-2832) */
void or_rest1_b(int imax1, int jmax1, int kmax1, int imax2, int jmax2, int kmax2, const double* q1, double* q2,
                const double* vol1, const double* vol2) {
  bg g1 = mkbg(imax1, jmax1, kmax1), g2 = mkbg(imax2, jmax2, kmax2);
  for (int n = 0; n <= 2 * NH; n++)
    for (int k1 = 1; k1 <= kmax1 - 2; k1 += 2)
      for (int j1 = 1; j1 <= jmax1 - 2; j1 += 2)
        for (int i1 = 1; i1 <= imax1 - 2; i1 += 2) {
          int i2 = 1 + i1 / 2, j2 = 1 + j1 / 2, k2 = 1 + k1 / 2;
          for (int ipde = 1; ipde <= NPDE; ipde++) {
#define R1(di, dj, dk) (V1(vol1, g1, i1 + di, j1 + dj, k1 + dk) * Q(q1, g1, i1 + di, j1 + dj, k1 + dk, ipde, n))
            Q(q2, g2, i2, j2, k2, ipde, n) = (R1(0, 0, 0) + R1(0, 1, 0) + R1(1, 0, 0) + R1(1, 1, 0) + R1(0, 0, 1) +
                                              R1(0, 1, 1) + R1(1, 0, 1) + R1(1, 1, 1)) /
                                             V1(vol2, g2, i2, j2, k2);
#undef R1
          }
        }
}

/* rest2_lo_b / rest2_ho_b (src/cosa.f90:2871-3011) */
void or_rest2_b(int ho, int imax1, int jmax1, int kmax1, int imax2, int jmax2, int kmax2, double* rhs2,
                const double* res1) {
  bg g1 = mkbg(imax1, jmax1, kmax1), g2 = mkbg(imax2, jmax2, kmax2);
  for (int n = 0; n <= 2 * NH; n++)
    for (int ipde = 1; ipde <= NPDE; ipde++)
      for (int k1 = 1; k1 <= kmax1 - 2; k1 += 2)
        for (int j1 = 1; j1 <= jmax1 - 2; j1 += 2)
          for (int i1 = 1; i1 <= imax1 - 2; i1 += 2) {
            int i2 = 1 + i1 / 2, j2 = 1 + j1 / 2, k2 = 1 + k1 / 2;
#define RR(di, dj, dk) Q(res1, g1, i1 + (di), j1 + (dj), k1 + (dk), ipde, n)
            if (!ho) {
              Q(rhs2, g2, i2, j2, k2, ipde, n) =
                  -(RR(0, 0, 0) + RR(0, 1, 0) + RR(1, 0, 0) + RR(1, 1, 0) + RR(0, 0, 1) + RR(0, 1, 1) + RR(1, 0, 1) +
                    RR(1, 1, 1));
            } else {
              double s27 = RR(0, 0, 0) + RR(0, 1, 0) + RR(1, 0, 0) + RR(1, 1, 0) + RR(0, 0, 1) + RR(0, 1, 1) +
                           RR(1, 0, 1) + RR(1, 1, 1);
              double s9 = RR(-1, 0, 0) + RR(-1, 1, 0) + RR(2, 0, 0) + RR(2, 1, 0) + RR(-1, 0, 1) + RR(-1, 1, 1) +
                          RR(2, 0, 1) + RR(2, 1, 1) + RR(0, -1, 0) + RR(1, -1, 0) + RR(0, 2, 0) + RR(1, 2, 0) +
                          RR(0, -1, 1) + RR(1, -1, 1) + RR(0, 2, 1) + RR(1, 2, 1) + RR(0, 0, -1) + RR(0, 1, -1) +
                          RR(0, 0, 2) + RR(0, 1, 2) + RR(1, 0, -1) + RR(1, 1, -1) + RR(1, 0, 2) + RR(1, 1, 2);
              double s3 = RR(0, -1, -1) + RR(1, -1, -1) + RR(0, 2, -1) + RR(1, 2, -1) + RR(-1, 0, -1) + RR(-1, 1, -1) +
                          RR(2, 0, -1) + RR(2, 1, -1) + RR(0, -1, 2) + RR(1, -1, 2) + RR(0, 2, 2) + RR(1, 2, 2) +
                          RR(-1, 0, 2) + RR(-1, 1, 2) + RR(2, 0, 2) + RR(2, 1, 2) + RR(-1, -1, 0) + RR(-1, -1, 1) +
                          RR(-1, 2, 0) + RR(-1, 2, 1) + RR(2, -1, 0) + RR(2, -1, 1) + RR(2, 2, 0) + RR(2, 2, 1);
              double s1 = RR(-1, -1, -1) + RR(2, -1, -1) + RR(-1, 2, -1) + RR(2, 2, -1) + RR(-1, -1, 2) +
                          RR(2, -1, 2) + RR(-1, 2, 2) + RR(2, 2, 2);
              Q(rhs2, g2, i2, j2, k2, ipde, n) = -(27 * s27 + 9 * s9 + 3 * s3 + s1) / 64;
            }
#undef RR
          }
}

/* prol_lin_b / prol_bln_b (src/cosa.f90:3315-3887), Euler variables. */
void or_prol_b(int bln, int imax1, int jmax1, int kmax1, int imax2, int jmax2, int kmax2, double* q1, const double* q2,
               const double* vol1) {
  bg g1 = mkbg(imax1, jmax1, kmax1), g2 = mkbg(imax2, jmax2, kmax2);
  int nv = bln ? 5 : NPDE;
  for (int n = 0; n <= 2 * NH; n++)
    for (int ipde = 1; ipde <= nv; ipde++)
      for (int k1 = 0; k1 <= kmax1 - 1; k1 += 2)
        for (int j1 = 0; j1 <= jmax1 - 1; j1 += 2)
          for (int i1 = 0; i1 <= imax1 - 1; i1 += 2) {
            int i2 = i1 / 2, j2 = j1 / 2, k2 = k1 / 2;
#define C2(di, dj, dk) Q(q2, g2, i2 + (di), j2 + (dj), k2 + (dk), ipde, n)
#define QF(di, dj, dk) Q(q1, g1, i1 + (di), j1 + (dj), k1 + (dk), ipde, n)
#define VV(di, dj, dk) V1(vol1, g1, i1 + (di), j1 + (dj), k1 + (dk))
            if (bln) {
              double cnec_l = C2(1, 1, 0), cnwc_l = C2(0, 1, 0), csec_l = C2(1, 0, 0), cswc_l = C2(0, 0, 0);
              double cnec_u = C2(1, 1, 1), cnwc_u = C2(0, 1, 1), csec_u = C2(1, 0, 1), cswc_u = C2(0, 0, 1);
              double cswf_l = (27 * cswc_l + 9 * csec_l + 9 * cnwc_l + 3 * cnec_l + 9 * cswc_u + 3 * csec_u +
                               3 * cnwc_u + cnec_u) / 64;
              double csef_l = (27 * csec_l + 9 * cswc_l + 9 * cnec_l + 3 * cnwc_l + 9 * csec_u + 3 * cswc_u +
                               3 * cnec_u + cnwc_u) / 64;
              double cnef_l = (27 * cnec_l + 9 * cnwc_l + 9 * csec_l + 3 * cswc_l + 9 * cnec_u + 3 * cnwc_u +
                               3 * csec_u + cswc_u) / 64;
              double cnwf_l = (27 * cnwc_l + 9 * cnec_l + 9 * cswc_l + 3 * csec_l + 9 * cnwc_u + 3 * cnec_u +
                               3 * cswc_u + csec_u) / 64;
              double cswf_u = (27 * cswc_u + 9 * csec_u + 9 * cnwc_u + 3 * cnec_u + 9 * cswc_l + 3 * csec_l +
                               3 * cnwc_l + cnec_l) / 64;
              double csef_u = (27 * csec_u + 9 * cswc_u + 9 * cnec_u + 3 * cnwc_u + 9 * csec_l + 3 * cswc_l +
                               3 * cnec_l + cnwc_l) / 64;
              double cnef_u = (27 * cnec_u + 9 * cnwc_u + 9 * csec_u + 3 * cswc_u + 9 * cnec_l + 3 * cnwc_l +
                               3 * csec_l + cswc_l) / 64;
              double cnwf_u = (27 * cnwc_u + 9 * cnec_u + 9 * cswc_u + 3 * csec_u + 9 * cnwc_l + 3 * cnec_l +
                               3 * cswc_l + csec_l) / 64;
              QF(0, 0, 0) = QF(0, 0, 0) + cswf_l;
              QF(1, 0, 0) = QF(1, 0, 0) + csef_l;
              QF(1, 1, 0) = QF(1, 1, 0) + cnef_l;
              QF(0, 1, 0) = QF(0, 1, 0) + cnwf_l;
              QF(0, 0, 1) = QF(0, 0, 1) + cswf_u;
              QF(1, 0, 1) = QF(1, 0, 1) + csef_u;
              QF(1, 1, 1) = QF(1, 1, 1) + cnef_u;
              QF(0, 1, 1) = QF(0, 1, 1) + cnwf_u;
            } else {
              double f[2][4]; /* [layer][sw,se,ne,nw] */
              for (int l = 0; l < 2; l++) {
                double cnec = C2(1, 1, l), cnwc = C2(0, 1, l), csec = C2(1, 0, l), cswc = C2(0, 0, l);
                double dcnc = cnec - cnwc, dcsc = csec - cswc, dcec = cnec - csec, dcwc = cnwc - cswc;
                f[l][0] = cswc + dcsc / 2 * VV(0, 0, l) / (VV(0, 0, l) + VV(1, 0, l)) +
                          dcwc / 2 * VV(0, 0, l) / (VV(0, 0, l) + VV(0, 1, l));
                f[l][1] = csec - dcsc / 2 * VV(1, 0, l) / (VV(0, 0, l) + VV(1, 0, l)) +
                          dcec / 2 * VV(1, 0, l) / (VV(1, 0, l) + VV(1, 1, l));
                f[l][3] = cnwc + dcnc / 2 * VV(0, 1, l) / (VV(0, 1, l) + VV(1, 1, l)) -
                          dcwc / 2 * VV(0, 1, l) / (VV(0, 0, l) + VV(0, 1, l));
                f[l][2] = cnec - dcnc / 2 * VV(1, 1, l) / (VV(1, 1, l) + VV(0, 1, l)) -
                          dcec / 2 * VV(1, 1, l) / (VV(1, 1, l) + VV(1, 0, l));
              }
              int di[4] = {0, 1, 1, 0}, dj[4] = {0, 0, 1, 1};
              for (int c = 0; c < 4; c++) {
                double fact = VV(di[c], dj[c], 0) / 2 / (VV(di[c], dj[c], 0) + VV(di[c], dj[c], 1));
                double lo = f[0][c], up = f[1][c];
                QF(di[c], dj[c], 0) = QF(di[c], dj[c], 0) + lo + fact / 2 * (up - lo);
                QF(di[c], dj[c], 1) = QF(di[c], dj[c], 1) + up - (1 - fact) / 2 * (up - lo);
              }
            }
#undef C2
#undef QF
#undef VV
          }
}

/* q_bedges (src/flux.f90:3914-4263, calmut = .false.): first-layer edge
 * cells by plane extrapolation from three neighbours.                     */
static const int EDG[12][4][2] = {
    /* ic:  p0, p1, p2, p3 as (first, second) index pairs; symbols resolved
       at run time: 100 = imax, 99 = imax-1, 200 = jmax, 199 = jmax-1,
       300 = kmax, 299 = kmax-1 */
    {{0, 0}, {0, 1}, {1, 0}, {1, 1}},
    {{100, 0}, {100, 1}, {99, 0}, {99, 1}},
    {{100, 200}, {100, 199}, {99, 200}, {99, 199}},
    {{0, 200}, {0, 199}, {1, 200}, {1, 199}},
    {{0, 0}, {0, 1}, {1, 0}, {1, 1}},
    {{100, 0}, {100, 1}, {99, 0}, {99, 1}},
    {{100, 300}, {100, 299}, {99, 300}, {99, 299}},
    {{0, 300}, {0, 299}, {1, 300}, {1, 299}},
    {{0, 0}, {0, 1}, {1, 0}, {1, 1}},
    {{200, 0}, {200, 1}, {199, 0}, {199, 1}},
    {{200, 300}, {200, 299}, {199, 300}, {199, 299}},
    {{0, 300}, {0, 299}, {1, 300}, {1, 299}}};
static int sym_res(int s, int imax, int jmax, int kmax) {
  switch (s) {
    case 100: return imax;
    case 99: return imax - 1;
    case 200: return jmax;
    case 199: return jmax - 1;
    case 300: return kmax;
    case 299: return kmax - 1;
    default: return s;
  }
}
void or_q_bedges(double* q, int imax, int jmax, int kmax) {
  bg g = mkbg(imax, jmax, kmax);
  int c[12][4][2];
  for (int e = 0; e < 12; e++)
    for (int p = 0; p < 4; p++)
      for (int d = 0; d < 2; d++) c[e][p][d] = sym_res(EDG[e][p][d], imax, jmax, kmax);
  for (int grp = 0; grp < 3; grp++) {
    int len = grp == 0 ? kmax - 1 : (grp == 1 ? jmax - 1 : imax - 1);
    for (int n = 0; n <= 2 * NH; n++)
      for (int ipde = 1; ipde <= NPDE; ipde++)
        for (int s = 1; s <= len; s++)
          for (int e = grp * 4; e < grp * 4 + 4; e++) {
            int x[4], y[4];
            for (int p = 0; p < 4; p++) { x[p] = c[e][p][0]; y[p] = c[e][p][1]; }
#define QE(p)                                                        \
  (grp == 0 ? Q(q, g, x[p], y[p], s, ipde, n)                        \
            : (grp == 1 ? Q(q, g, x[p], s, y[p], ipde, n) : Q(q, g, s, x[p], y[p], ipde, n)))
            double b = ((QE(3) - QE(1)) * (x[2] - x[1]) - (QE(2) - QE(1)) * (x[3] - x[1])) /
                       ((y[3] - y[1]) * (x[2] - x[1]) - (y[2] - y[1]) * (x[3] - x[1]));
            double a = (QE(2) - QE(1) - b * (y[2] - y[1])) / (x[2] - x[1]);
            double cc = QE(1) - a * x[1] - b * y[1];
            double val = a * x[0] + b * y[0] + cc;
            if (grp == 0) Q(q, g, x[0], y[0], s, ipde, n) = val;
            else if (grp == 1) Q(q, g, x[0], s, y[0], ipde, n) = val;
            else Q(q, g, s, x[0], y[0], ipde, n) = val;
#undef QE
          }
  }
}

/* q_bcorns (src/flux.f90:3156-3471, calmut = .false.): first-layer corner
 * cells by linear extrapolation from four neighbours.  The integer
 * coefficient expressions are evaluated with Fortran integer semantics (all
 * operands are unit offsets, so every integer division is exact).          */
void or_q_bcorns(double* q, int imax, int jmax, int kmax) {
  bg g = mkbg(imax, jmax, kmax);
  int I = imax, J = jmax, K = kmax;
  int pc[8][5][3] = {
      {{0, 0, 0}, {0, 1, 1}, {1, 1, 0}, {1, 0, 1}, {1, 1, 1}},
      {{I, 0, 0}, {I, 1, 1}, {I - 1, 1, 0}, {I - 1, 0, 1}, {I - 1, 1, 1}},
      {{I, J, 0}, {I, J - 1, 1}, {I - 1, J - 1, 0}, {I - 1, J, 1}, {I - 1, J - 1, 1}},
      {{0, J, 0}, {0, J - 1, 1}, {1, J - 1, 0}, {1, J, 1}, {1, J - 1, 1}},
      {{0, 0, K}, {0, 1, K - 1}, {1, 1, K}, {1, 0, K - 1}, {1, 1, K - 1}},
      {{I, 0, K}, {I, 1, K - 1}, {I - 1, 1, K}, {I - 1, 0, K - 1}, {I - 1, 1, K - 1}},
      {{I, J, K}, {I, J - 1, K - 1}, {I - 1, J - 1, K}, {I - 1, J, K - 1}, {I - 1, J - 1, K - 1}},
      {{0, J, K}, {0, J - 1, K - 1}, {1, J - 1, K}, {1, J, K - 1}, {1, J - 1, K - 1}}};
  for (int n = 0; n <= 2 * NH; n++)
    for (int ipde = 1; ipde <= NPDE; ipde++)
      for (int ic = 0; ic < 8; ic++) {
        int(*p)[3] = pc[ic];
#define X(m, d) p[m][(d) - 1]
#define QC(m) Q(q, g, p[m][0], p[m][1], p[m][2], ipde, n)
        int c1 = ((X(3, 3) - X(1, 3)) * (X(4, 2) - X(1, 2)) - (X(4, 3) - X(1, 3)) * (X(3, 2) - X(1, 2))) /
                 (X(2, 3) - X(1, 3));
        int c2 = ((X(2, 3) - X(1, 3)) * (X(1, 2) - X(4, 2)) + (X(4, 3) - X(1, 3)) * (X(2, 2) - X(1, 2))) /
                 (X(2, 3) - X(1, 3));
        int c3 = ((X(2, 3) - X(1, 3)) * (X(3, 2) - X(1, 2)) + (X(3, 3) - X(1, 3)) * (X(1, 2) - X(2, 2))) /
                 (X(2, 3) - X(1, 3));
        int dn = ((X(4, 1) - X(1, 1)) * (X(3, 2) - X(1, 2)) + (X(1, 1) - X(3, 1)) * (X(4, 2) - X(1, 2))) +
                 ((X(4, 1) - X(1, 1)) * (X(1, 2) - X(2, 2)) + (X(2, 1) - X(1, 1)) * (X(4, 2) - X(1, 2))) *
                     ((X(3, 3) - X(1, 3))) / ((X(2, 3) - X(1, 3))) +
                 ((X(1, 1) - X(2, 1)) * (X(3, 2) - X(1, 2)) + (X(1, 1) - X(3, 1)) * (X(1, 2) - X(2, 2))) *
                     ((X(4, 3) - X(1, 3))) / ((X(2, 3) - X(1, 3)));
        double a = (c1 * (QC(2) - QC(1)) + c2 * (QC(3) - QC(1)) + c3 * (QC(4) - QC(1))) / dn;
        int e1 = (X(1, 1) - X(3, 1)) + (X(2, 1) - X(1, 1)) * (X(3, 3) - X(1, 3)) / (X(2, 3) - X(1, 3));
        int e2 = (X(3, 3) - X(1, 3)) / (X(2, 3) - X(1, 3));
        int e3 = (X(3, 2) - X(1, 2)) + (X(1, 2) - X(2, 2)) * (X(3, 3) - X(1, 3)) / (X(2, 3) - X(1, 3));
        double b = (e1 * a + e2 * (QC(1) - QC(2)) + (QC(3) - QC(1))) / e3;
        double c = ((X(1, 1) - X(2, 1)) * a + (X(1, 2) - X(2, 2)) * b + (QC(2) - QC(1))) / (X(2, 3) - X(1, 3));
        double d = QC(1) - a * X(1, 1) - b * X(1, 2) - c * X(1, 3);
        QC(0) = a * X(0, 1) + b * X(0, 2) + c * X(0, 3) + d;
#undef X
#undef QC
      }
}

/* ---------------------------------------------------------------------- */
/* whole-level wrappers (the reference's do iblock loops)                   */

static ll off_of(const or_level* L, int b, int fam) {
  ll o = 0;
  for (int bb = 0; bb < b; bb++) {
    ll im = L->imax[bb], jm = L->jmax[bb], km = L->kmax[bb];
    if (fam == 3) o += (im + 3) * (jm + 3) * (km + 3);
    else if (fam == 2) o += (im + 2) * (jm + 2) * (km + 2);
    else if (fam == 1) o += (im + 1) * (jm + 1) * (km + 1);
    else o += im * jm * km;
  }
  return o;
}
static ll off_bct(const or_level* L, int b) {
  ll o = 0;
  for (int bb = 0; bb < b; bb++) o += L->nbcs[bb];
  return o * 10;
}
#define QB(arr, L, b) ((arr) + off_of(L, b, 3) * NPDE * DIM5)
#define MB(arr, L, b) ((arr) + off_of(L, b, 2) * LMET * DIM5H)
#define XB(arr, L, b) ((arr) + off_of(L, b, 2) * DIM5H)
#define PB(arr, L, b) ((arr) + off_of(L, b, 1) * DIM5)
#define VB(arr, L, b) ((arr) + off_of(L, b, 1))
#define FB(arr, L, b) ((arr) + off_of(L, b, 0) * NPDE * DIM5)
#define SB(arr, L, b) ((arr) + off_of(L, b, 3) * DIM5)
#define DB(arr, L, b) ((arr) + off_of(L, b, 0) * 3 * DIM5H)
/* ttrm of block b, or NULL: Wilcox k-omega uses the constants gkw, bkw */
#define TTB(L, b) (is_sst2() && (L)->ttrm ? (L)->ttrm + off_of(L, b, 3) * 3 * DIM5 : (const double*)NULL)

void or_cons2prim(or_case* c, int nl, double* q) {
  or_level* L = &c->lev[nl - 1];
  for (int b = 0; b < L->nblocks; b++) or_cons2prim_b(QB(q, L, b), L->imax[b], L->jmax[b], L->kmax[b], 0);
}
void or_prim2cons(or_case* c, int nl, double* q) {
  or_level* L = &c->lev[nl - 1];
  for (int b = 0; b < L->nblocks; b++) or_prim2cons_b(QB(q, L, b), L->imax[b], L->jmax[b], L->kmax[b]);
}
void or_bc(or_case* c, int nl, double* q) {
  or_level* L = &c->lev[nl - 1];
  for (int b = 0; b < L->nblocks; b++)
    or_bbc_dist(QB(q, L, b), MB(L->si, L, b), MB(L->sj, L, b), MB(L->sk, L, b), L->dist ? VB(L->dist, L, b) : NULL,
                L->bctopo + off_bct(L, b), L->nbcs[b], L->imax[b], L->jmax[b], L->kmax[b]);
}
void or_cutman_q(or_case* c, int nl, double* q) {
  or_level* L = &c->lev[nl - 1];
  for (int icut = 0; icut < L->ncuts; icut++) {
    const int32_t* t = L->cutopo + 21 * icut;
    int b1 = t[0] - 1, b2 = t[9] - 1;
    or_cut_q(L->imax[b1], L->jmax[b1], L->kmax[b1], QB(q, L, b1), L->imax[b2], L->jmax[b2], L->kmax[b2],
             QB(q, L, b2), t);
  }
}
void or_edges_corners(or_case* c, int nl, double* q) {
  or_level* L = &c->lev[nl - 1];
  for (int b = 0; b < L->nblocks; b++) or_q_bedges(QB(q, L, b), L->imax[b], L->jmax[b], L->kmax[b]);
  for (int b = 0; b < L->nblocks; b++) or_q_bcorns(QB(q, L, b), L->imax[b], L->jmax[b], L->kmax[b]);
}
int or_prtest(or_case* c, int nl, int32_t* where) {
  or_level* L = &c->lev[nl - 1];
  for (int b = 0; b < L->nblocks; b++) {
    int r = or_bprtest(QB(L->q, L, b), L->imax[b], L->jmax[b], L->kmax[b], where ? where + 2 : NULL);
    if (r) {
      if (where) { where[0] = nl; where[1] = b + 1; where[6] = r; }
      return r;
    }
  }
  return 0;
}
void or_lambda_deltat(or_case* c, int nl) {
  or_level* L = &c->lev[nl - 1];
  if (P.viscous) {
    for (int b = 0; b < L->nblocks; b++)
      or_lambdab_v(QB(L->q, L, b), MB(L->si, L, b), MB(L->sj, L, b), MB(L->sk, L, b),
                   L->xdot ? XB(L->xdot, L, b) : NULL, L->ydot ? XB(L->ydot, L, b) : NULL,
                   L->zdot ? XB(L->zdot, L, b) : NULL, VB(L->vol, L, b), L->mut ? SB(L->mut, L, b) : NULL,
                   QB(L->qp, L, b), QB(L->dq, L, b), L->imax[b], L->jmax[b], L->kmax[b]);
    for (int b = 0; b < L->nblocks; b++)
      or_bdeltat_v(PB(L->dtvol, L, b), L->dtvolt ? PB(L->dtvolt, L, b) : NULL, QB(L->qp, L, b), QB(L->dq, L, b),
                   PB(L->betaxi, L, b), PB(L->betaeta, L, b), PB(L->betazeta, L, b),
                   L->tbetaxi ? PB(L->tbetaxi, L, b) : NULL, L->tbetaeta ? PB(L->tbetaeta, L, b) : NULL,
                   L->tbetazeta ? PB(L->tbetazeta, L, b) : NULL, L->imax[b], L->jmax[b], L->kmax[b]);
    return;
  }
  for (int b = 0; b < L->nblocks; b++)
    or_lambdab(QB(L->q, L, b), MB(L->si, L, b), MB(L->sj, L, b), MB(L->sk, L, b),
               L->xdot ? XB(L->xdot, L, b) : NULL, L->ydot ? XB(L->ydot, L, b) : NULL,
               L->zdot ? XB(L->zdot, L, b) : NULL, QB(L->qp, L, b), L->imax[b], L->jmax[b], L->kmax[b]);
  for (int b = 0; b < L->nblocks; b++)
    or_bdeltat(PB(L->dtvol, L, b), QB(L->qp, L, b), PB(L->betaxi, L, b), PB(L->betaeta, L, b),
               PB(L->betazeta, L, b), L->imax[b], L->jmax[b], L->kmax[b]);
}
/* zero(...,off_l=1): res(0:imax,0:jmax,0:kmax,:,:) = 0 (src/util.f90:238-369) */
void or_zero_res(or_case* c, int nl) {
  or_level* L = &c->lev[nl - 1];
  for (int b = 0; b < L->nblocks; b++) {
    bg g = mkbg(L->imax[b], L->jmax[b], L->kmax[b]);
    double* r = QB(L->res, L, b);
    for (int n = 0; n <= 2 * NH; n++)
      for (int v = 1; v <= NPDE; v++)
        for (int k = 0; k <= g.kmax; k++)
          for (int j = 0; j <= g.jmax; j++)
            for (int i = 0; i <= g.imax; i++) Q(r, g, i, j, k, v, n) = 0.0;
  }
}
/* viscous fluxes of direction idir (src/cosa.f90:1806-1821): q_face,
 * q_der, rtst (k-omega), vflux, resid; scratch as in the reference: qf in
 * dq, mutf in flux, derivatives in qp/qm/dqm, Reynolds stresses in dqp.    */
static void or_viscous_dir(or_case* c, int nl, int idir) {
  or_level* L = &c->lev[nl - 1];
  int kom = is_kom();
  const double* s = idir == 1 ? L->si : (idir == 2 ? L->sj : L->sk);
  for (int b = 0; b < L->nblocks; b++) {
    int im = L->imax[b], jm = L->jmax[b], km = L->kmax[b];
    const int32_t* bt = L->bctopo + off_bct(L, b);
    double* xd0 = L->xd[3 * (idir - 1)] ? DB(L->xd[3 * (idir - 1)], L, b) : NULL;
    double* xd1 = L->xd[3 * (idir - 1) + 1] ? DB(L->xd[3 * (idir - 1) + 1], L, b) : NULL;
    double* xd2 = L->xd[3 * (idir - 1) + 2] ? DB(L->xd[3 * (idir - 1) + 2], L, b) : NULL;
    or_q_face_b(idir, bt, L->nbcs[b], QB(L->q, L, b), QB(L->dq, L, b), kom ? SB(L->mut, L, b) : NULL,
                FB(L->flux, L, b), kom, im, jm, km);
    or_q_der_b(idir, bt, L->nbcs[b], QB(L->q, L, b), QB(L->qp, L, b), QB(L->qm, L, b), QB(L->dqm, L, b), im, jm, km);
    if (kom)
      or_rtst_b(idir, QB(L->dq, L, b), QB(L->qp, L, b), QB(L->qm, L, b), QB(L->dqm, L, b), xd0, xd1, xd2,
                QB(L->dqp, L, b), FB(L->flux, L, b), im, jm, km);
    double* fd = NULL;
    if (is_sst2()) {
      fd = (double*)malloc(sizeof(double) * (size_t)im * jm * km);
      or_facedist_b(idir, VB(L->dist, L, b), fd, im, jm, km);
    }
    or_vflux_b(idir, QB(L->dq, L, b), QB(L->qp, L, b), QB(L->qm, L, b), QB(L->dqm, L, b), QB(L->dqp, L, b),
               FB(L->flux, L, b), fd, xd0, xd1, xd2, MB(s, L, b), im, jm, km);
    free(fd);
    or_bresid(idir, FB(L->flux, L, b), QB(L->res, L, b), im, jm, km);
  }
}

/* k-omega sources (src/cosa.f90:1869-1902): cell-centre velocity
 * differences from the face averages, turprod (finest level), srckw.     */
static void or_komega_sources(or_case* c, int nl) {
  or_level* L = &c->lev[nl - 1];
  double* outs[3] = {L->qp, L->qm, L->dqm};
  for (int idir = 1; idir <= 3; idir++)
    for (int b = 0; b < L->nblocks; b++) {
      or_q_face_b(idir, L->bctopo + off_bct(L, b), L->nbcs[b], QB(L->q, L, b), QB(L->dq, L, b), NULL, NULL, 0,
                  L->imax[b], L->jmax[b], L->kmax[b]);
      or_q_der_c_b(idir, QB(L->dq, L, b), QB(outs[idir - 1], L, b), L->imax[b], L->jmax[b], L->kmax[b]);
    }
  if (nl == 1 || P.prd_all_lvls)
    for (int b = 0; b < L->nblocks; b++) {
      const double* xd[9];
      for (int m = 0; m < 9; m++) xd[m] = DB(L->xd[m], L, b);
      or_turprod_b(QB(L->q, L, b), QB(L->qp, L, b), QB(L->qm, L, b), QB(L->dqm, L, b), SB(L->prd, L, b),
                   SB(L->divvel, L, b), SB(L->delpls, L, b), xd, L->imax[b], L->jmax[b], L->kmax[b]);
    }
  if (P.kom)
    for (int b = 0; b < L->nblocks; b++)
      or_srckw_b(QB(L->q, L, b), SB(L->mut, L, b), SB(L->prd, L, b), SB(L->divvel, L, b), QB(L->res, L, b),
                 VB(L->vol, L, b), nl, L->imax[b], L->jmax[b], L->kmax[b]);
  else
    for (int b = 0; b < L->nblocks; b++) {
      /* update_ttrm then srckw_sst (src/cosa.f90:1896-1900) */
      const double* xd[9];
      for (int m = 0; m < 9; m++) xd[m] = DB(L->xd[m], L, b);
      double* tt = L->ttrm + off_of(L, b, 3) * 3 * DIM5;
      or_update_ttrm_b(QB(L->q, L, b), QB(L->qp, L, b), QB(L->qm, L, b), QB(L->dqm, L, b), xd, tt,
                       VB(L->dist, L, b), L->imax[b], L->jmax[b], L->kmax[b]);
      or_srckw_sst_b(QB(L->q, L, b), SB(L->prd, L, b), SB(L->divvel, L, b), QB(L->res, L, b), SB(L->mut, L, b), tt,
                     VB(L->vol, L, b), nl, L->imax[b], L->jmax[b], L->kmax[b]);
    }
}

void or_residual(or_case* c, int nl) {
  or_level* L = &c->lev[nl - 1];
  or_zero_res(c, nl);
  if (P.viscous)
    for (int b = 0; b < L->nblocks; b++) or_q_bedges(QB(L->q, L, b), L->imax[b], L->jmax[b], L->kmax[b]);
  for (int idir = 1; idir <= 3; idir++) {
    for (int b = 0; b < L->nblocks; b++)
      or_muscl_b(idir, nl, L->bctopo + off_bct(L, b), L->nbcs[b], QB(L->q, L, b), QB(L->qp, L, b), QB(L->qm, L, b),
                 QB(L->dq, L, b), QB(L->dqp, L, b), QB(L->dqm, L, b), L->imax[b], L->jmax[b], L->kmax[b]);
    const double* s = idir == 1 ? L->si : (idir == 2 ? L->sj : L->sk);
    for (int b = 0; b < L->nblocks; b++)
      or_roflux_b(idir, QB(L->qp, L, b), QB(L->qm, L, b), FB(L->flux, L, b), MB(s, L, b), L->imax[b], L->jmax[b],
                  L->kmax[b]);
    for (int b = 0; b < L->nblocks; b++)
      or_bresid(idir, FB(L->flux, L, b), QB(L->res, L, b), L->imax[b], L->jmax[b], L->kmax[b]);
    if (P.viscous) or_viscous_dir(c, nl, idir);
  }
  if (is_kom()) or_komega_sources(c, nl);
}
void or_src_hb(or_case* c, int nl) {
  or_level* L = &c->lev[nl - 1];
  for (int b = 0; b < L->nblocks; b++)
    or_src_bhb(QB(L->q, L, b), VB(L->vol, L, b), QB(L->res, L, b), QB(L->dq, L, b), L->imax[b], L->jmax[b],
               L->kmax[b]);
}
/* resg_rms / resg_rms_hb (src/cosa.f90:4458-4500, 4585-4626) */
void or_resg_rms(or_case* c, int nl, double* resrms) {
  or_level* L = &c->lev[nl - 1];
  double rb[8];
  ll ncell = 0;
  for (int p = 0; p <= NPDE; p++) resrms[p] = 0;
  for (int b = 0; b < L->nblocks; b++) {
    if (P.harbal) {
      /* work1 = copy of res on the interior (src/cosa.f90:1930-1931) */
      or_resg_brms_hb(QB(L->work1, L, b), rb, L->imax[b], L->jmax[b], L->kmax[b]);
    } else if (DUALT && P.rkimp && P.irsop == 0 && P.nlevel == 1) {
      or_resg_brms_dual(QB(L->q, L, b), QB(L->res, L, b), QB(L->rhs, L, b), VB(L->vol, L, b), rb, L->imax[b],
                        L->jmax[b], L->kmax[b]);
    } else {
      or_resg_brms(QB(L->res, L, b), QB(L->rhs, L, b), rb, L->imax[b], L->jmax[b], L->kmax[b]);
    }
    for (int p = 0; p <= NPDE; p++) resrms[p] = resrms[p] + rb[p];
    ncell = ncell + (ll)(L->imax[b] - 1) * (L->jmax[b] - 1) * (L->kmax[b] - 1);
  }
  if (P.harbal) {
    for (int p = 0; p < NPDE; p++) resrms[p] = sqrt(resrms[p] / ncell / DIM5);
    resrms[NPDE] = sqrt(resrms[NPDE] / ncell / DIM5 / NPDE);
  } else {
    for (int p = 0; p < NPDE; p++) resrms[p] = sqrt(resrms[p] / ncell);
    resrms[NPDE] = sqrt(resrms[NPDE] / ncell / NPDE);
  }
}
/* add_array / copy_array with off_l = -1 (interior) */
void or_add_array(or_case* c, int nl, double* dst, const double* src, double fact) {
  or_level* L = &c->lev[nl - 1];
  for (int b = 0; b < L->nblocks; b++) {
    bg g = mkbg(L->imax[b], L->jmax[b], L->kmax[b]);
    double* d = QB(dst, L, b);
    const double* s = QB(src, L, b);
    for (int n = 0; n <= 2 * NH; n++)
      for (int v = 1; v <= NPDE; v++)
        for (int k = 1; k <= g.kmax - 1; k++)
          for (int j = 1; j <= g.jmax - 1; j++)
            for (int i = 1; i <= g.imax - 1; i++) Q(d, g, i, j, k, v, n) = Q(d, g, i, j, k, v, n) + fact * Q(s, g, i, j, k, v, n);
  }
}
void or_copy_array(or_case* c, int nl, double* dst, const double* src) {
  or_level* L = &c->lev[nl - 1];
  for (int b = 0; b < L->nblocks; b++) {
    bg g = mkbg(L->imax[b], L->jmax[b], L->kmax[b]);
    double* d = QB(dst, L, b);
    const double* s = QB(src, L, b);
    for (int n = 0; n <= 2 * NH; n++)
      for (int v = 1; v <= NPDE; v++)
        for (int k = 1; k <= g.kmax - 1; k++)
          for (int j = 1; j <= g.jmax - 1; j++)
            for (int i = 1; i <= g.imax - 1; i++) Q(d, g, i, j, k, v, n) = Q(s, g, i, j, k, v, n);
  }
}
void or_cirs(or_case* c, int nl) {
  or_level* L = &c->lev[nl - 1];
  int kom = is_kom();
  if (P.cutcirs == 1) {
    /* src/cosa.f90:1984-1992: scalef, ores = res, two passes of cutman_q on
     * res and ccirs, scaleb */
    for (int b = 0; b < L->nblocks; b++)
      or_scale_b(QB(L->res, L, b), PB(L->dtvol, L, b), kom ? PB(L->dtvolt, L, b) : NULL, 0, L->imax[b], L->jmax[b],
                 L->kmax[b]);
    for (int b = 0; b < L->nblocks; b++) {
      bg g = mkbg(L->imax[b], L->jmax[b], L->kmax[b]);
      memcpy(QB(L->work1, L, b), QB(L->res, L, b), sizeof(double) * (size_t)g.n3 * NPDE * DIM5);
    }
    for (int it = 0; it < 2; it++) {
      or_cutman_q(c, nl, L->res);
      for (int b = 0; b < L->nblocks; b++)
        or_ccirs_b(QB(L->res, L, b), QB(L->work1, L, b), PB(L->betaxi, L, b), PB(L->betaeta, L, b),
                   PB(L->betazeta, L, b), kom ? PB(L->tbetaxi, L, b) : NULL, kom ? PB(L->tbetaeta, L, b) : NULL,
                   kom ? PB(L->tbetazeta, L, b) : NULL, L->imax[b], L->jmax[b], L->kmax[b]);
    }
    for (int b = 0; b < L->nblocks; b++)
      or_scale_b(QB(L->res, L, b), PB(L->dtvol, L, b), kom ? PB(L->dtvolt, L, b) : NULL, 1, L->imax[b], L->jmax[b],
                 L->kmax[b]);
    return;
  }
  for (int b = 0; b < L->nblocks; b++)
    or_cirs_kb(QB(L->res, L, b), PB(L->dtvol, L, b), PB(L->betaxi, L, b), PB(L->betaeta, L, b),
               PB(L->betazeta, L, b), kom ? PB(L->dtvolt, L, b) : NULL, kom ? PB(L->tbetaxi, L, b) : NULL,
               kom ? PB(L->tbetaeta, L, b) : NULL, kom ? PB(L->tbetazeta, L, b) : NULL, L->imax[b], L->jmax[b],
               L->kmax[b]);
}
void or_update(or_case* c, int nl, double ar) {
  or_level* L = &c->lev[nl - 1];
  for (int b = 0; b < L->nblocks; b++) {
    if (DUALT && P.rkimp)
      or_update_dual_b(QB(L->res, L, b), QB(L->q, L, b), QB(L->qold, L, b), VB(L->vol, L, b), PB(L->dtvol, L, b),
                       is_kom() ? PB(L->dtvolt, L, b) : NULL, is_kom() ? SB(L->delpls, L, b) : NULL,
                       is_kom() ? SB(L->prd, L, b) : NULL, ar, L->imax[b], L->jmax[b], L->kmax[b], TTB(L, b));
    else if (P.harbal && P.rkimp)
      or_update_bhb_rkim(QB(L->res, L, b), QB(L->q, L, b), QB(L->qold, L, b), VB(L->vol, L, b), PB(L->dtvol, L, b),
                         PB(L->dtvolt, L, b), QB(L->dq, L, b), SB(L->delpls, L, b), SB(L->prd, L, b), ar,
                         L->imax[b], L->jmax[b], L->kmax[b], TTB(L, b));
    else if (is_kom())
      or_update_kb(QB(L->res, L, b), QB(L->q, L, b), QB(L->qold, L, b), VB(L->vol, L, b), PB(L->dtvol, L, b),
                   PB(L->dtvolt, L, b), SB(L->delpls, L, b), SB(L->prd, L, b), ar, L->imax[b], L->jmax[b],
                   L->kmax[b], TTB(L, b));
    else
      or_update_b(QB(L->res, L, b), QB(L->q, L, b), QB(L->qold, L, b), PB(L->dtvol, L, b), ar, L->imax[b],
                  L->jmax[b], L->kmax[b]);
  }
}
/* tur_vis + cutman_mut (src/update.f90:2272-2295; src/bc.f90:2713-2800) */
/* cutman_mut (src/bc.f90:2713-2800): cut_q on the one-variable mut */
static void or_cutman_mut(or_case* c, int nl) {
  or_level* L = &c->lev[nl - 1];
  int save = NPDE;
  NPDE = 1;   /* cut_q with one variable per snapshot */
  for (int icut = 0; icut < L->ncuts; icut++) {
    const int32_t* t = L->cutopo + 21 * icut;
    int b1 = t[0] - 1, b2 = t[9] - 1;
    or_cut_q(L->imax[b1], L->jmax[b1], L->kmax[b1], SB(L->mut, L, b1), L->imax[b2], L->jmax[b2], L->kmax[b2],
             SB(L->mut, L, b2), t);
  }
  NPDE = save;
}

/* SST mut on the first visit of a coarse level (src/cosa.f90:1739-1757):
 * the cell-centre differences of a primitive copy of q (in res), that copy
 * made conservative, tur_vis_sst, res cleared over the whole box.        */
static void or_sst_first_mut(or_case* c, int nl) {
  or_level* L = &c->lev[nl - 1];
  double* outs[3] = {L->qp, L->qm, L->dqm};
  for (int b = 0; b < L->nblocks; b++) {
    bg g = mkbg(L->imax[b], L->jmax[b], L->kmax[b]);
    memcpy(QB(L->res, L, b), QB(L->q, L, b), sizeof(double) * (size_t)g.n3 * NPDE * DIM5);
  }
  for (int idir = 1; idir <= 3; idir++)
    for (int b = 0; b < L->nblocks; b++) {
      or_q_face_b(idir, L->bctopo + off_bct(L, b), L->nbcs[b], QB(L->res, L, b), QB(L->dq, L, b), NULL, NULL, 0,
                  L->imax[b], L->jmax[b], L->kmax[b]);
      or_q_der_c_b(idir, QB(L->dq, L, b), QB(outs[idir - 1], L, b), L->imax[b], L->jmax[b], L->kmax[b]);
    }
  or_prim2cons(c, nl, L->res);
  for (int b = 0; b < L->nblocks; b++) {
    const double* xd[9];
    for (int m = 0; m < 9; m++) xd[m] = DB(L->xd[m], L, b);
    or_tur_vis_sst_b(QB(L->res, L, b), SB(L->mut, L, b), QB(L->qp, L, b), QB(L->qm, L, b), QB(L->dqm, L, b), xd,
                     VB(L->dist, L, b), L->imax[b], L->jmax[b], L->kmax[b]);
  }
  for (int b = 0; b < L->nblocks; b++) {
    bg g = mkbg(L->imax[b], L->jmax[b], L->kmax[b]);
    memset(QB(L->res, L, b), 0, sizeof(double) * (size_t)g.n3 * NPDE * DIM5);
  }
  or_cutman_mut(c, nl);
}

void or_tur_vis(or_case* c, int nl) {
  or_level* L = &c->lev[nl - 1];
  for (int b = 0; b < L->nblocks; b++) {
    if (P.kom_sst) {
      /* tur_vis_sst with the differences q_der_c left in qp, qm, dqm by the
       * stage's residual (src/cosa.f90:2008-2011) */
      const double* xd[9];
      for (int m = 0; m < 9; m++) xd[m] = DB(L->xd[m], L, b);
      or_tur_vis_sst_b(QB(L->q, L, b), SB(L->mut, L, b), QB(L->qp, L, b), QB(L->qm, L, b), QB(L->dqm, L, b), xd,
                       VB(L->dist, L, b), L->imax[b], L->jmax[b], L->kmax[b]);
    } else {
      or_tur_vis_b(QB(L->q, L, b), SB(L->mut, L, b), L->imax[b], L->jmax[b], L->kmax[b]);
    }
  }
  or_cutman_mut(c, nl);
}
void or_fixq(or_case* c, int nl) {
  or_level* L = &c->lev[nl - 1];
  for (int b = 0; b < L->nblocks; b++) or_fixq_b(QB(L->q, L, b), L->imax[b], L->jmax[b], L->kmax[b]);
}
/* restriction stage limits of k-omega (src/cosa.f90:1965-1973) */
void or_turb_limits(or_case* c, int nl, double ar) {
  or_level* L = &c->lev[nl - 1];
  if (P.nlevel > 1 && nl != P.nlevel && !P.prd_all_lvls && P.lim_prodtke)
    for (int b = 0; b < L->nblocks; b++)
      or_lim_prod_mg_b(QB(L->q, L, b), SB(L->prd, L, b), SB(L->mut, L, b), L->imax[b], L->jmax[b], L->kmax[b]);
  for (int b = 0; b < L->nblocks; b++) {
    if (P.harbal)
      or_lim_res_ome_bhb(QB(L->res, L, b), QB(L->q, L, b), QB(L->qold, L, b), VB(L->vol, L, b), PB(L->dtvolt, L, b),
                         SB(L->delpls, L, b), SB(L->prd, L, b), L->imax[b], L->jmax[b], L->kmax[b], TTB(L, b));
    else if (DUALT && P.rkimp)
      or_lim_res_ome_dual_b(QB(L->res, L, b), QB(L->q, L, b), VB(L->vol, L, b), PB(L->dtvolt, L, b),
                            SB(L->delpls, L, b), SB(L->prd, L, b), ar, L->imax[b], L->jmax[b], L->kmax[b], TTB(L, b));
    else
      or_lim_res_ome_b(QB(L->res, L, b), QB(L->q, L, b), VB(L->vol, L, b), PB(L->dtvolt, L, b), SB(L->delpls, L, b),
                       SB(L->prd, L, b), ar, L->imax[b], L->jmax[b], L->kmax[b], TTB(L, b));
  }
}

/* ---------------------------------------------------------------------- */
/* smooth (src/cosa.f90:1594-2032), Euler / HB explicit path               */
int or_smooth(or_case* c, int nl, int from_fine, int to_coarse, int niter, int mgit, int32_t* done, int32_t* where) {
  or_level* L = &c->lev[nl - 1];
  MGIT = mgit;
  *done = 0;
  int niterp = to_coarse ? niter + 1 : niter;
  for (int iter = 1; iter <= niterp; iter++) {
    if (iter <= niter) or_copy_array(c, nl, L->qold, L->q);
    for (int irk = 1; irk <= P.nstage; irk++) {
      or_cons2prim(c, nl, L->q);
      or_bc(c, nl, L->q);
      or_cutman_q(c, nl, L->q);
      if (is_kom() && mgit == 1 && iter == 1 && irk == 1 && from_fine && !P.prd_all_lvls) {
        if (P.kom_sst) or_sst_first_mut(c, nl);
        else or_tur_vis(c, nl);
      }
      int bad = or_prtest(c, nl, where);
      if (bad) {
        if (where) where[5] = iter;
        return COSAHIP_ERR_POSITIVITY;
      }
      if (iter % P.iupdt == 0 && irk == 1) or_lambda_deltat(c, nl);
      or_residual(c, nl);
      or_prim2cons(c, nl, L->q);
      const int rkim_plain = P.rkimp && P.irsop == 0 && P.nlevel == 1;
      if (P.harbal && !rkim_plain) or_src_hb(c, nl);
      /* src/cosa.f90:1910-1915 */
      if (DUALT && (!P.rkimp || P.irsop > 0 || P.nlevel > 1))
        for (int b = 0; b < L->nblocks; b++)
          or_src_uns_b(QB(L->q, L, b), VB(L->vol, L, b), QB(L->res, L, b), L->imax[b], L->jmax[b], L->kmax[b]);
      if (nl == 1 && irk == 1 && iter != niter + 1 &&
          (mgit == 1 || mgit % P.iprint == 0 || mgit == P.ncycle)) {
        if (P.harbal) or_copy_array(c, nl, L->work1, L->res);
        if (P.harbal && rkim_plain)
          for (int b = 0; b < L->nblocks; b++)
            or_rms_hb_term(QB(L->work1, L, b), QB(L->q, L, b), VB(L->vol, L, b), L->imax[b], L->jmax[b], L->kmax[b]);
        double rr[8];
        or_resg_rms(c, nl, rr);
        for (int p = 0; p <= NPDE; p++) RESRMS[p + 1] = rr[p];
      }
      /* the overall residual of the finest grid (src/cosa.f90:1938-1941) */
      if (nl == 1 && DUALT) or_add_array(c, nl, L->res, L->rhs, -1.0);
      if (from_fine && iter == 1 && irk == 1) or_add_array(c, nl, L->rhs, L->res, 1.0);
      if (nl > 1) or_add_array(c, nl, L->res, L->rhs, -1.0);
      if (iter == niter + 1) {
        if (is_kom()) or_turb_limits(c, nl, P.alphas[irk - 1]);
        return 0;
      }
      if ((P.irsop > 0 && P.irsop <= 3) || P.irsop == 5 || P.irsop == 6) or_cirs(c, nl);
      or_update(c, nl, P.alphas[irk - 1]);
      if (is_kom()) {
        if (P.ifixq == 1) or_fixq(c, nl);
        if (irk == P.nstage && (nl == 1 || !P.prd_all_lvls)) or_tur_vis(c, nl);
      }
    }
    double errmax = RESRMS[1];
    for (int p = 2; p <= 5; p++) errmax = fmax(errmax, RESRMS[p]);
    if (errmax < P.toler) {
      *done = 1;
      return 0;
    }
  }
  return 0;
}

/* restrict (src/cosa.f90:2672-2754), Euler: rest1 + rest2 */
void or_restrict(or_case* c, int nl) {
  or_level *L1 = &c->lev[nl - 1], *L2 = &c->lev[nl];
  for (int b = 0; b < L1->nblocks; b++)
    or_rest1_b(L1->imax[b], L1->jmax[b], L1->kmax[b], L2->imax[b], L2->jmax[b], L2->kmax[b], QB(L1->q, L1, b),
               QB(L2->q, L2, b), VB(L1->vol, L1, b), VB(L2->vol, L2, b));
  if (is_kom() && !P.prd_all_lvls) {
    double* f1[3] = {L1->prd, L1->delpls, L1->divvel};
    double* f2[3] = {L2->prd, L2->delpls, L2->divvel};
    for (int m = 0; m < 3; m++)
      for (int b = 0; b < L1->nblocks; b++)
        or_rest4_b(L1->imax[b], L1->jmax[b], L1->kmax[b], L2->imax[b], L2->jmax[b], L2->kmax[b], SB(f1[m], L1, b),
                   SB(f2[m], L2, b), VB(L1->vol, L1, b), VB(L2->vol, L2, b));
  }
  for (int b = 0; b < L1->nblocks; b++)
    or_rest2_b(P.ho_rest, L1->imax[b], L1->jmax[b], L1->kmax[b], L2->imax[b], L2->jmax[b], L2->kmax[b],
               QB(L2->rhs, L2, b), QB(L1->res, L1, b));
}

/* prolong (src/cosa.f90:3167-3249) */
void or_prolong(or_case* c, int nl) {
  or_level *L1 = &c->lev[nl - 1], *L2 = &c->lev[nl];
  for (int b = 0; b < L1->nblocks; b++)
    or_rest1_b(L1->imax[b], L1->jmax[b], L1->kmax[b], L2->imax[b], L2->jmax[b], L2->kmax[b], QB(L1->q, L1, b),
               QB(L2->qold, L2, b), VB(L1->vol, L1, b), VB(L2->vol, L2, b));
  double* arrs[2] = {L2->qold, L2->q};
  for (int a = 0; a < 2; a++) {
    or_cons2prim(c, nl + 1, arrs[a]);
    or_bc(c, nl + 1, arrs[a]);
    or_cutman_q(c, nl + 1, arrs[a]);
    or_prim2cons(c, nl + 1, arrs[a]);
    or_edges_corners(c, nl + 1, arrs[a]);
  }
  /* add_array(..., q2, qold2, off_l = 1, -1): 0..imax incl. first halo */
  for (int b = 0; b < L2->nblocks; b++) {
    bg g = mkbg(L2->imax[b], L2->jmax[b], L2->kmax[b]);
    double* d = QB(L2->q, L2, b);
    const double* s = QB(L2->qold, L2, b);
    for (int n = 0; n <= 2 * NH; n++)
      for (int v = 1; v <= NPDE; v++)
        for (int k = 0; k <= g.kmax; k++)
          for (int j = 0; j <= g.jmax; j++)
            for (int i = 0; i <= g.imax; i++) Q(d, g, i, j, k, v, n) = Q(d, g, i, j, k, v, n) + -1.0 * Q(s, g, i, j, k, v, n);
  }
  for (int b = 0; b < L1->nblocks; b++)
    or_prol_b(P.bln_prol, L1->imax[b], L1->jmax[b], L1->kmax[b], L2->imax[b], L2->jmax[b], L2->kmax[b],
              QB(L1->q, L1, b), QB(L2->q, L2, b), VB(L1->vol, L1, b));
}

/* mg (src/cosa.f90:806-930) */
int or_mg(or_case* c, int nstart, int npre, int npost, int ncrs, double* hist, int32_t* where) {
  int nlevel = P.nlevel, rc;
  int32_t done = 0;
  int nl = 1;
  for (int mgit = 1; mgit <= P.ncycle; mgit++) {
    int niter = nlevel == 1 ? 1 : (mgit == 1 ? nstart : npost);
    nl = 1;
    rc = or_smooth(c, nl, 0, nlevel > 1, niter, mgit, &done, where);
    if (rc) return -rc;
    for (int p = 0; p < 8; p++) hist[(ll)(mgit - 1) * 8 + p] = RESRMS[p + 1];
    if (done) return mgit;
    if (nlevel > 1) {
      or_restrict(c, nl);
      nl = 2;
      niter = nlevel == 2 ? ncrs : npre;
      rc = or_smooth(c, nl, 1, nlevel > 2, niter, mgit, &done, where);
      if (rc) return -rc;
      if (nlevel == 3) {
        or_restrict(c, nl);
        nl = 3;
        rc = or_smooth(c, nl, 1, 0, ncrs, mgit, &done, where);
        if (rc) return -rc;
        nl = 2;
        or_prolong(c, nl);
        rc = or_smooth(c, nl, 0, 0, npost, mgit, &done, where);
        if (rc) return -rc;
      }
      nl = 1;
      or_prolong(c, nl);
    }
  }
  if (nlevel > 1) {
    rc = or_smooth(c, nl, 0, 0, npost, P.ncycle, &done, where);
    if (rc) return -rc;
  }
  return P.ncycle;
}

/* ---------------------------------------------------------------------- */
/* one physical step of dual time stepping around mg() (src/cosa.f90:640-
 * 651): q1 = q, rhs = rhs_uns(q1, q2) on the finest level, then after the
 * multigrid cycles q2 = q1.  copy_array with off_l = -1: interior cells.  */
void or_dual_begin(or_case* c, int itime) {
  or_level* L = &c->lev[0];
  ITIME = itime;
  or_copy_array(c, 1, L->q1, L->q);
  for (int b = 0; b < L->nblocks; b++)
    or_rhs_uns_b(QB(L->rhs, L, b), QB(L->q1, L, b), QB(L->q2, L, b), VB(L->vol, L, b), L->imax[b], L->jmax[b],
                 L->kmax[b]);
}
void or_dual_end(or_case* c) {
  or_level* L = &c->lev[0];
  or_copy_array(c, 1, L->q2, L->q1);
}
