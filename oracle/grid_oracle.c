/*
 * grid_oracle.c -- CPU restatement of COSA's grid-motion and metric set-up
 * routines, the producers of the hot path's geometric inputs.  TEST
 * INFRASTRUCTURE ONLY: the checker of libcosahip's device grid set-up
 * (cosahip_grid_*), never linked into the product.
 *
 * Citations are file:line into LANCASTER-CFD/COSA src/.  Index conventions
 * follow the Fortran explicit-shape dummies:
 *   node arrays x(0:imax+1,0:jmax+1,0:kmax+1,0:2N*hbmove)  (P2 scalar)
 *   face metrics s(lmet,0:imax+1,0:jmax+1,0:kmax+1,0:2N*hbmove)
 *   volumes vol(0:imax,0:jmax,0:kmax)                        (P1)
 *   metric derivatives d(3,imax,jmax,kmax,0:2N*hbmove)        (off_0)
 * Expressions keep the reference's evaluation order; the degenerate-cell
 * option (dgn_cell) is off, so a non-positive face area is an error.
 */
#include <math.h>
#include <stdlib.h>

#include "cosa_oracle.h"

typedef long long ll;

#define XN(a, im, jm, km, i, j, k, n) \
  (a)[(ll)(i) + (ll)((im) + 2) * ((ll)(j) + (ll)((jm) + 2) * ((ll)(k) + (ll)((km) + 2) * (ll)(n)))]
#define SM(a, lmet, im, jm, km, m, i, j, k, n) \
  (a)[((m) - 1) + (ll)(lmet) * ((ll)(i) + (ll)((im) + 2) * ((ll)(j) + (ll)((jm) + 2) * ((ll)(k) + (ll)((km) + 2) * (ll)(n))))]
#define VL(a, im, jm, km, i, j, k) (a)[(ll)(i) + (ll)((im) + 1) * ((ll)(j) + (ll)((jm) + 1) * (ll)(k))]
#define DR(a, im, jm, km, c, i, j, k, n) \
  (a)[((c) - 1) + 3 * ((ll)((i) - 1) + (ll)(im) * ((ll)((j) - 1) + (ll)(jm) * ((ll)((k) - 1) + (ll)(km) * (ll)(n))))]

/* ---- imetrix_b (src/grid.f90:848-1253), without the rotating-frame
 * moments (hawt .and. rotating).  Face normals from the cross product of
 * the two face diagonals, halved; area = |normal|; unit normal; with
 * moving grids (lmet = 7) the face-midpoint velocity as the mean of the
 * face's four node velocities.  Returns 0, or -1 at a non-positive area. */
int or_imetrix_b(const double* x, const double* y, const double* z, const double* xdot, const double* ydot,
                 const double* zdot, double* si, double* sj, double* sk, int imax, int jmax, int kmax, int lmet,
                 int nsm, int moving) {
  int bad = 0;
  const int im = imax, jm = jmax, km = kmax;
#define X(i, j, k) XN(x, im, jm, km, i, j, k, n)
#define Y(i, j, k) XN(y, im, jm, km, i, j, k, n)
#define Z(i, j, k) XN(z, im, jm, km, i, j, k, n)
  for (int n = 0; n < nsm; n++) {
    /* xi faces (src/grid.f90:898-933) */
    for (int k = 1; k <= kmax - 1; k++)
      for (int j = 1; j <= jmax - 1; j++)
        for (int i = 1; i <= imax; i++) {
          double s1 = ((Y(i, j + 1, k + 1) - Y(i, j, k)) * (Z(i, j, k + 1) - Z(i, j + 1, k)) -
                       (Z(i, j + 1, k + 1) - Z(i, j, k)) * (Y(i, j, k + 1) - Y(i, j + 1, k))) / 2;
          double s2 = ((Z(i, j + 1, k + 1) - Z(i, j, k)) * (X(i, j, k + 1) - X(i, j + 1, k)) -
                       (X(i, j + 1, k + 1) - X(i, j, k)) * (Z(i, j, k + 1) - Z(i, j + 1, k))) / 2;
          double s3 = ((X(i, j + 1, k + 1) - X(i, j, k)) * (Y(i, j, k + 1) - Y(i, j + 1, k)) -
                       (Y(i, j + 1, k + 1) - Y(i, j, k)) * (X(i, j, k + 1) - X(i, j + 1, k))) / 2;
          double a = sqrt(s1 * s1 + s2 * s2 + s3 * s3);
          if (a <= 0) bad = 1;
          SM(si, lmet, im, jm, km, 1, i, j, k, n) = s1 / a;
          SM(si, lmet, im, jm, km, 2, i, j, k, n) = s2 / a;
          SM(si, lmet, im, jm, km, 3, i, j, k, n) = s3 / a;
          SM(si, lmet, im, jm, km, 4, i, j, k, n) = a;
        }
    /* eta faces (src/grid.f90:948-983) */
    for (int k = 1; k <= kmax - 1; k++)
      for (int i = 1; i <= imax - 1; i++)
        for (int j = 1; j <= jmax; j++) {
          double s1 = ((Z(i + 1, j, k + 1) - Z(i, j, k)) * (Y(i, j, k + 1) - Y(i + 1, j, k)) -
                       (Y(i + 1, j, k + 1) - Y(i, j, k)) * (Z(i, j, k + 1) - Z(i + 1, j, k))) / 2;
          double s2 = ((X(i + 1, j, k + 1) - X(i, j, k)) * (Z(i, j, k + 1) - Z(i + 1, j, k)) -
                       (Z(i + 1, j, k + 1) - Z(i, j, k)) * (X(i, j, k + 1) - X(i + 1, j, k))) / 2;
          double s3 = ((Y(i + 1, j, k + 1) - Y(i, j, k)) * (X(i, j, k + 1) - X(i + 1, j, k)) -
                       (X(i + 1, j, k + 1) - X(i, j, k)) * (Y(i, j, k + 1) - Y(i + 1, j, k))) / 2;
          double a = sqrt(s1 * s1 + s2 * s2 + s3 * s3);
          if (a <= 0) bad = 1;
          SM(sj, lmet, im, jm, km, 1, i, j, k, n) = s1 / a;
          SM(sj, lmet, im, jm, km, 2, i, j, k, n) = s2 / a;
          SM(sj, lmet, im, jm, km, 3, i, j, k, n) = s3 / a;
          SM(sj, lmet, im, jm, km, 4, i, j, k, n) = a;
        }
    /* zeta faces (src/grid.f90:985-1021) */
    for (int j = 1; j <= jmax - 1; j++)
      for (int i = 1; i <= imax - 1; i++)
        for (int k = 1; k <= kmax; k++) {
          double s1 = ((Y(i + 1, j + 1, k) - Y(i, j, k)) * (Z(i, j + 1, k) - Z(i + 1, j, k)) -
                       (Z(i + 1, j + 1, k) - Z(i, j, k)) * (Y(i, j + 1, k) - Y(i + 1, j, k))) / 2;
          double s2 = ((Z(i + 1, j + 1, k) - Z(i, j, k)) * (X(i, j + 1, k) - X(i + 1, j, k)) -
                       (X(i + 1, j + 1, k) - X(i, j, k)) * (Z(i, j + 1, k) - Z(i + 1, j, k))) / 2;
          double s3 = ((X(i + 1, j + 1, k) - X(i, j, k)) * (Y(i, j + 1, k) - Y(i + 1, j, k)) -
                       (Y(i + 1, j + 1, k) - Y(i, j, k)) * (X(i, j + 1, k) - X(i + 1, j, k))) / 2;
          double a = sqrt(s1 * s1 + s2 * s2 + s3 * s3);
          if (a <= 0) bad = 1;
          SM(sk, lmet, im, jm, km, 1, i, j, k, n) = s1 / a;
          SM(sk, lmet, im, jm, km, 2, i, j, k, n) = s2 / a;
          SM(sk, lmet, im, jm, km, 3, i, j, k, n) = s3 / a;
          SM(sk, lmet, im, jm, km, 4, i, j, k, n) = a;
        }
    if (moving) {
#define XD(a, i, j, k) XN(a, im, jm, km, i, j, k, n)
      /* face-midpoint velocities (src/grid.f90:1025-1064) */
      for (int k = 1; k <= kmax - 1; k++)
        for (int j = 1; j <= jmax - 1; j++)
          for (int i = 1; i <= imax; i++) {
            SM(si, lmet, im, jm, km, 5, i, j, k, n) =
                (XD(xdot, i, j + 1, k + 1) + XD(xdot, i, j + 1, k) + XD(xdot, i, j, k + 1) + XD(xdot, i, j, k)) / 4;
            SM(si, lmet, im, jm, km, 6, i, j, k, n) =
                (XD(ydot, i, j + 1, k + 1) + XD(ydot, i, j + 1, k) + XD(ydot, i, j, k + 1) + XD(ydot, i, j, k)) / 4;
            SM(si, lmet, im, jm, km, 7, i, j, k, n) =
                (XD(zdot, i, j + 1, k + 1) + XD(zdot, i, j + 1, k) + XD(zdot, i, j, k + 1) + XD(zdot, i, j, k)) / 4;
          }
      for (int k = 1; k <= kmax - 1; k++)
        for (int i = 1; i <= imax - 1; i++)
          for (int j = 1; j <= jmax; j++) {
            SM(sj, lmet, im, jm, km, 5, i, j, k, n) =
                (XD(xdot, i + 1, j, k + 1) + XD(xdot, i + 1, j, k) + XD(xdot, i, j, k + 1) + XD(xdot, i, j, k)) / 4;
            SM(sj, lmet, im, jm, km, 6, i, j, k, n) =
                (XD(ydot, i + 1, j, k + 1) + XD(ydot, i + 1, j, k) + XD(ydot, i, j, k + 1) + XD(ydot, i, j, k)) / 4;
            SM(sj, lmet, im, jm, km, 7, i, j, k, n) =
                (XD(zdot, i + 1, j, k + 1) + XD(zdot, i + 1, j, k) + XD(zdot, i, j, k + 1) + XD(zdot, i, j, k)) / 4;
          }
      for (int j = 1; j <= jmax - 1; j++)
        for (int i = 1; i <= imax - 1; i++)
          for (int k = 1; k <= kmax; k++) {
            SM(sk, lmet, im, jm, km, 5, i, j, k, n) =
                (XD(xdot, i + 1, j + 1, k) + XD(xdot, i + 1, j, k) + XD(xdot, i, j + 1, k) + XD(xdot, i, j, k)) / 4;
            SM(sk, lmet, im, jm, km, 6, i, j, k, n) =
                (XD(ydot, i + 1, j + 1, k) + XD(ydot, i + 1, j, k) + XD(ydot, i, j + 1, k) + XD(ydot, i, j, k)) / 4;
            SM(sk, lmet, im, jm, km, 7, i, j, k, n) =
                (XD(zdot, i + 1, j + 1, k) + XD(zdot, i + 1, j, k) + XD(zdot, i, j + 1, k) + XD(zdot, i, j, k)) / 4;
          }
#undef XD
    }
  }
#undef X
#undef Y
#undef Z
  return bad ? -1 : 0;
}

/* ---- bvols (src/grid.f90:2162-2258): cell volumes of snapshot 0 by
 * Kordulla and Vinokur's formula, interior cells only.  Returns 0, or -1
 * at a non-positive volume (the reference only reports it). */
int or_bvols(const double* x, const double* y, const double* z, double* vol, int imax, int jmax, int kmax) {
  int bad = 0;
  const int im = imax, jm = jmax, km = kmax, n = 0;
#define X(i, j, k) XN(x, im, jm, km, i, j, k, n)
#define Y(i, j, k) XN(y, im, jm, km, i, j, k, n)
#define Z(i, j, k) XN(z, im, jm, km, i, j, k, n)
  for (int k = 1; k <= kmax - 1; k++)
    for (int j = 1; j <= jmax - 1; j++)
      for (int i = 1; i <= imax - 1; i++) {
        double si1 = ((Y(i, j + 1, k + 1) - Y(i, j, k)) * (Z(i, j, k + 1) - Z(i, j + 1, k)) -
                      (Z(i, j + 1, k + 1) - Z(i, j, k)) * (Y(i, j, k + 1) - Y(i, j + 1, k))) / 2;
        double sj1 = ((Z(i + 1, j, k + 1) - Z(i, j, k)) * (Y(i, j, k + 1) - Y(i + 1, j, k)) -
                      (Y(i + 1, j, k + 1) - Y(i, j, k)) * (Z(i, j, k + 1) - Z(i + 1, j, k))) / 2;
        double sk1 = ((Y(i + 1, j + 1, k) - Y(i, j, k)) * (Z(i, j + 1, k) - Z(i + 1, j, k)) -
                      (Z(i + 1, j + 1, k) - Z(i, j, k)) * (Y(i, j + 1, k) - Y(i + 1, j, k))) / 2;
        double dx = (si1 + sj1 + sk1) * (X(i + 1, j + 1, k + 1) - X(i, j, k));
        double si2 = ((Z(i, j + 1, k + 1) - Z(i, j, k)) * (X(i, j, k + 1) - X(i, j + 1, k)) -
                      (X(i, j + 1, k + 1) - X(i, j, k)) * (Z(i, j, k + 1) - Z(i, j + 1, k))) / 2;
        double sj2 = ((X(i + 1, j, k + 1) - X(i, j, k)) * (Z(i, j, k + 1) - Z(i + 1, j, k)) -
                      (Z(i + 1, j, k + 1) - Z(i, j, k)) * (X(i, j, k + 1) - X(i + 1, j, k))) / 2;
        double sk2 = ((Z(i + 1, j + 1, k) - Z(i, j, k)) * (X(i, j + 1, k) - X(i + 1, j, k)) -
                      (X(i + 1, j + 1, k) - X(i, j, k)) * (Z(i, j + 1, k) - Z(i + 1, j, k))) / 2;
        double dy = (si2 + sj2 + sk2) * (Y(i + 1, j + 1, k + 1) - Y(i, j, k));
        double si3 = ((X(i, j + 1, k + 1) - X(i, j, k)) * (Y(i, j, k + 1) - Y(i, j + 1, k)) -
                      (Y(i, j + 1, k + 1) - Y(i, j, k)) * (X(i, j, k + 1) - X(i, j + 1, k))) / 2;
        double sj3 = ((Y(i + 1, j, k + 1) - Y(i, j, k)) * (X(i, j, k + 1) - X(i + 1, j, k)) -
                      (X(i + 1, j, k + 1) - X(i, j, k)) * (Y(i, j, k + 1) - Y(i + 1, j, k))) / 2;
        double sk3 = ((X(i + 1, j + 1, k) - X(i, j, k)) * (Y(i, j + 1, k) - Y(i + 1, j, k)) -
                      (Y(i + 1, j + 1, k) - Y(i, j, k)) * (X(i, j + 1, k) - X(i + 1, j, k))) / 2;
        double dz = (si3 + sj3 + sk3) * (Z(i + 1, j + 1, k + 1) - Z(i, j, k));
        double v = (dx + dy + dz) / 3;
        if (v <= 0) bad = 1;
        VL(vol, im, jm, km, i, j, k) = v;
      }
#undef X
#undef Y
#undef Z
  return bad ? -1 : 0;
}

/* ---- volex_b (src/grid.f90:110-141): the face-halo cells take the
 * volume of the adjacent interior cell. */
void or_volex_b(double* vol, int imax, int jmax, int kmax) {
  const int im = imax, jm = jmax, km = kmax;
  for (int k = 1; k <= kmax - 1; k++) {
    for (int i = 1; i <= imax - 1; i++) {
      VL(vol, im, jm, km, i, 0, k) = VL(vol, im, jm, km, i, 1, k);
      VL(vol, im, jm, km, i, jmax, k) = VL(vol, im, jm, km, i, jmax - 1, k);
    }
    for (int j = 1; j <= jmax - 1; j++) {
      VL(vol, im, jm, km, 0, j, k) = VL(vol, im, jm, km, 1, j, k);
      VL(vol, im, jm, km, imax, j, k) = VL(vol, im, jm, km, imax - 1, j, k);
    }
  }
  for (int j = 1; j <= jmax - 1; j++)
    for (int i = 1; i <= imax - 1; i++) {
      VL(vol, im, jm, km, i, j, 0) = VL(vol, im, jm, km, i, j, 1);
      VL(vol, im, jm, km, i, j, kmax) = VL(vol, im, jm, km, i, j, kmax - 1);
    }
}

/* ---- vol_bedges (src/grid.f90:3607-3826): the edge-halo cells of the 12
 * block edges from the plane through the three neighbouring values of
 * each edge cross-section; integer offsets enter as integers and are
 * converted where they meet a real, as in the reference. */
static void edge_plane(const int c[2][4], double v1, double v2, double v3, double* out) {
  /* c[0][m], c[1][m]: the two in-plane indices of points m = 0..3 */
  const double b = ((v3 - v1) * (double)(c[0][2] - c[0][1]) - (v2 - v1) * (double)(c[0][3] - c[0][1])) /
                   (double)((c[1][3] - c[1][1]) * (c[0][2] - c[0][1]) - (c[1][2] - c[1][1]) * (c[0][3] - c[0][1]));
  const double a = (v2 - v1 - b * (double)(c[1][2] - c[1][1])) / (double)(c[0][2] - c[0][1]);
  const double cc = v1 - a * (double)c[0][1] - b * (double)c[1][1];
  *out = a * (double)c[0][0] + b * (double)c[1][0] + cc;
}

static void edge_table(int imax, int jmax, int kmax, int t[13][2][4]) {
  /* region r: (first index, second index) of points 0 (target) .. 3 */
  const int A[13][2][4] = {
      {{0}},
      {{0, 0, 1, 1}, {0, 1, 0, 1}},
      {{imax, imax, imax - 1, imax - 1}, {0, 1, 0, 1}},
      {{imax, imax, imax - 1, imax - 1}, {jmax, jmax - 1, jmax, jmax - 1}},
      {{0, 0, 1, 1}, {jmax, jmax - 1, jmax, jmax - 1}},
      {{0, 0, 1, 1}, {0, 1, 0, 1}},
      {{imax, imax, imax - 1, imax - 1}, {0, 1, 0, 1}},
      {{imax, imax, imax - 1, imax - 1}, {kmax, kmax - 1, kmax, kmax - 1}},
      {{0, 0, 1, 1}, {kmax, kmax - 1, kmax, kmax - 1}},
      {{0, 0, 1, 1}, {0, 1, 0, 1}},
      {{jmax, jmax, jmax - 1, jmax - 1}, {0, 1, 0, 1}},
      {{jmax, jmax, jmax - 1, jmax - 1}, {kmax, kmax - 1, kmax, kmax - 1}},
      {{0, 0, 1, 1}, {kmax, kmax - 1, kmax, kmax - 1}},
  };
  for (int r = 0; r < 13; r++)
    for (int a = 0; a < 2; a++)
      for (int m = 0; m < 4; m++) t[r][a][m] = A[r][a][m];
}

void or_vol_bedges(double* vol, int imax, int jmax, int kmax) {
  const int im = imax, jm = jmax, km = kmax;
  int t[13][2][4];
  edge_table(imax, jmax, kmax, t);
  for (int k = 1; k <= kmax - 1; k++)
    for (int r = 1; r <= 4; r++) {
      const int(*c)[4] = t[r];
      edge_plane(c, VL(vol, im, jm, km, c[0][1], c[1][1], k), VL(vol, im, jm, km, c[0][2], c[1][2], k),
                 VL(vol, im, jm, km, c[0][3], c[1][3], k), &VL(vol, im, jm, km, c[0][0], c[1][0], k));
    }
  for (int j = 1; j <= jmax - 1; j++)
    for (int r = 5; r <= 8; r++) {
      const int(*c)[4] = t[r];
      edge_plane(c, VL(vol, im, jm, km, c[0][1], j, c[1][1]), VL(vol, im, jm, km, c[0][2], j, c[1][2]),
                 VL(vol, im, jm, km, c[0][3], j, c[1][3]), &VL(vol, im, jm, km, c[0][0], j, c[1][0]));
    }
  for (int i = 1; i <= imax - 1; i++)
    for (int r = 9; r <= 12; r++) {
      const int(*c)[4] = t[r];
      edge_plane(c, VL(vol, im, jm, km, i, c[0][1], c[1][1]), VL(vol, im, jm, km, i, c[0][2], c[1][2]),
                 VL(vol, im, jm, km, i, c[0][3], c[1][3]), &VL(vol, im, jm, km, i, c[0][0], c[1][0]));
    }
}

/* ---- cordex_b (src/grid.f90:28-78): halo nodes copy the adjacent
 * boundary node, j and i halos of every k-plane first, then the k halos. */
void or_cordex_b(double* x, double* y, double* z, int imax, int jmax, int kmax, int nsm) {
  const int im = imax, jm = jmax, km = kmax;
  double* a[3] = {x, y, z};
  for (int n = 0; n < nsm; n++) {
    for (int k = 1; k <= kmax; k++) {
      for (int i = 1; i <= imax; i++)
        for (int c = 0; c < 3; c++) {
          XN(a[c], im, jm, km, i, 0, k, n) = XN(a[c], im, jm, km, i, 1, k, n);
          XN(a[c], im, jm, km, i, jmax + 1, k, n) = XN(a[c], im, jm, km, i, jmax, k, n);
        }
      for (int j = 1; j <= jmax; j++)
        for (int c = 0; c < 3; c++) {
          XN(a[c], im, jm, km, 0, j, k, n) = XN(a[c], im, jm, km, 1, j, k, n);
          XN(a[c], im, jm, km, imax + 1, j, k, n) = XN(a[c], im, jm, km, imax, j, k, n);
        }
    }
    for (int j = 1; j <= jmax; j++)
      for (int i = 1; i <= imax; i++)
        for (int c = 0; c < 3; c++) {
          XN(a[c], im, jm, km, i, j, 0, n) = XN(a[c], im, jm, km, i, j, 1, n);
          XN(a[c], im, jm, km, i, j, kmax + 1, n) = XN(a[c], im, jm, km, i, j, kmax, n);
        }
  }
}

/* ---- vmetrix_b (src/grid.f90:3150-3583): metric derivatives on the
 * faces as the inverse of the face Jacobian.  The derivative normal to
 * the face is the mean of central differences at the face's four nodes
 * (doubled where the halo node coincides with the boundary node, i.e. a
 * one-sided difference), the two tangential ones the means of the face's
 * edge differences; the Jacobian determinant is the mean volume of the two
 * cells beside the face.  d[0..8] = xideri, etaderi, zetaderi, xiderj,
 * etaderj, zetaderj, xiderk, etaderk, zetaderk. */
static double nd(double a1, double a2, double a3, double b1, double b2, double b3) {
  const double d1 = a1 - b1, d2 = a2 - b2, d3 = a3 - b3;
  return sqrt(d1 * d1 + d2 * d2 + d3 * d3);
}

static void jac_store(double xx, double yx, double zx, double xe, double ye, double ze, double xz, double yz,
                      double zz, double volvisc, double* dxi, double* deta, double* dzeta) {
  dxi[0] = (zz * ye - ze * yz) / volvisc;
  dxi[1] = (xz * ze - xe * zz) / volvisc;
  dxi[2] = (yz * xe - ye * xz) / volvisc;
  deta[0] = (zx * yz - zz * yx) / volvisc;
  deta[1] = (xx * zz - xz * zx) / volvisc;
  deta[2] = (yx * xz - yz * xx) / volvisc;
  dzeta[0] = (ze * yx - zx * ye) / volvisc;
  dzeta[1] = (xe * zx - xx * ze) / volvisc;
  dzeta[2] = (ye * xx - yx * xe) / volvisc;
}

void or_vmetrix_b(const double* x, const double* y, const double* z, double* const* d, const double* vol, int imax,
                  int jmax, int kmax, int nsm) {
  const int im = imax, jm = jmax, km = kmax;
  for (int n = 0; n < nsm; n++) {
#define X(i, j, k) XN(x, im, jm, km, i, j, k, n)
#define Y(i, j, k) XN(y, im, jm, km, i, j, k, n)
#define Z(i, j, k) XN(z, im, jm, km, i, j, k, n)
    /* xi faces (src/grid.f90:3204-3326) */
    for (int k = 1; k <= kmax - 1; k++)
      for (int j = 1; j <= jmax - 1; j++)
        for (int i = 1; i <= imax; i++) {
          const int jj[4] = {j, j + 1, j + 1, j}, kk[4] = {k, k, k + 1, k + 1};
          double xx[4], yx[4], zx[4];
          for (int m = 0; m < 4; m++) {
            xx[m] = (X(i + 1, jj[m], kk[m]) - X(i - 1, jj[m], kk[m])) / 2;
            yx[m] = (Y(i + 1, jj[m], kk[m]) - Y(i - 1, jj[m], kk[m])) / 2;
            zx[m] = (Z(i + 1, jj[m], kk[m]) - Z(i - 1, jj[m], kk[m])) / 2;
          }
          const double xe1 = X(i, j + 1, k) - X(i, j, k), ye1 = Y(i, j + 1, k) - Y(i, j, k),
                       ze1 = Z(i, j + 1, k) - Z(i, j, k);
          const double xe2 = X(i, j + 1, k + 1) - X(i, j, k + 1), ye2 = Y(i, j + 1, k + 1) - Y(i, j, k + 1),
                       ze2 = Z(i, j + 1, k + 1) - Z(i, j, k + 1);
          const double xz1 = X(i, j, k + 1) - X(i, j, k), yz1 = Y(i, j, k + 1) - Y(i, j, k),
                       zz1 = Z(i, j, k + 1) - Z(i, j, k);
          const double xz2 = X(i, j + 1, k + 1) - X(i, j + 1, k), yz2 = Y(i, j + 1, k + 1) - Y(i, j + 1, k),
                       zz2 = Z(i, j + 1, k + 1) - Z(i, j + 1, k);
          const double volvisc = (VL(vol, im, jm, km, i - 1, j, k) + VL(vol, im, jm, km, i, j, k)) / 2;
          if (i == 1 || i == imax) {
            const int h = i == 1 ? i - 1 : i + 1;
            for (int m = 0; m < 4; m++) {
              if (nd(X(h, jj[m], kk[m]), Y(h, jj[m], kk[m]), Z(h, jj[m], kk[m]), X(i, jj[m], kk[m]),
                     Y(i, jj[m], kk[m]), Z(i, jj[m], kk[m])) < 1e-14) {
                xx[m] = 2 * xx[m];
                yx[m] = 2 * yx[m];
                zx[m] = 2 * zx[m];
              }
            }
          }
          const double x0 = (xx[0] + xx[1] + xx[2] + xx[3]) / 4, y0 = (yx[0] + yx[1] + yx[2] + yx[3]) / 4,
                       z0 = (zx[0] + zx[1] + zx[2] + zx[3]) / 4;
          jac_store(x0, y0, z0, (xe1 + xe2) / 2, (ye1 + ye2) / 2, (ze1 + ze2) / 2, (xz1 + xz2) / 2, (yz1 + yz2) / 2,
                    (zz1 + zz2) / 2, volvisc, &DR(d[0], im, jm, km, 1, i, j, k, n),
                    &DR(d[1], im, jm, km, 1, i, j, k, n), &DR(d[2], im, jm, km, 1, i, j, k, n));
        }
    /* eta faces (src/grid.f90:3330-3451) */
    for (int k = 1; k <= kmax - 1; k++)
      for (int j = 1; j <= jmax; j++)
        for (int i = 1; i <= imax - 1; i++) {
          const int ii[4] = {i, i, i + 1, i + 1}, kk[4] = {k, k + 1, k + 1, k};
          double xe[4], ye[4], ze[4];
          for (int m = 0; m < 4; m++) {
            xe[m] = (X(ii[m], j + 1, kk[m]) - X(ii[m], j - 1, kk[m])) / 2;
            ye[m] = (Y(ii[m], j + 1, kk[m]) - Y(ii[m], j - 1, kk[m])) / 2;
            ze[m] = (Z(ii[m], j + 1, kk[m]) - Z(ii[m], j - 1, kk[m])) / 2;
          }
          const double xx1 = X(i + 1, j, k) - X(i, j, k), yx1 = Y(i + 1, j, k) - Y(i, j, k),
                       zx1 = Z(i + 1, j, k) - Z(i, j, k);
          const double xx2 = X(i + 1, j, k + 1) - X(i, j, k + 1), yx2 = Y(i + 1, j, k + 1) - Y(i, j, k + 1),
                       zx2 = Z(i + 1, j, k + 1) - Z(i, j, k + 1);
          const double xz1 = X(i, j, k + 1) - X(i, j, k), yz1 = Y(i, j, k + 1) - Y(i, j, k),
                       zz1 = Z(i, j, k + 1) - Z(i, j, k);
          const double xz2 = X(i + 1, j, k + 1) - X(i + 1, j, k), yz2 = Y(i + 1, j, k + 1) - Y(i + 1, j, k),
                       zz2 = Z(i + 1, j, k + 1) - Z(i + 1, j, k);
          const double volvisc = (VL(vol, im, jm, km, i, j, k) + VL(vol, im, jm, km, i, j - 1, k)) / 2;
          if (j == 1 || j == jmax) {
            const int h = j == 1 ? j - 1 : j + 1;
            for (int m = 0; m < 4; m++) {
              if (nd(X(ii[m], h, kk[m]), Y(ii[m], h, kk[m]), Z(ii[m], h, kk[m]), X(ii[m], j, kk[m]),
                     Y(ii[m], j, kk[m]), Z(ii[m], j, kk[m])) < 1e-14) {
                xe[m] = 2 * xe[m];
                ye[m] = 2 * ye[m];
                ze[m] = 2 * ze[m];
              }
            }
          }
          const double x0 = (xe[0] + xe[1] + xe[2] + xe[3]) / 4, y0 = (ye[0] + ye[1] + ye[2] + ye[3]) / 4,
                       z0 = (ze[0] + ze[1] + ze[2] + ze[3]) / 4;
          jac_store((xx1 + xx2) / 2, (yx1 + yx2) / 2, (zx1 + zx2) / 2, x0, y0, z0, (xz1 + xz2) / 2, (yz1 + yz2) / 2,
                    (zz1 + zz2) / 2, volvisc, &DR(d[3], im, jm, km, 1, i, j, k, n),
                    &DR(d[4], im, jm, km, 1, i, j, k, n), &DR(d[5], im, jm, km, 1, i, j, k, n));
        }
    /* zeta faces (src/grid.f90:3455-3578) */
    for (int k = 1; k <= kmax; k++)
      for (int j = 1; j <= jmax - 1; j++)
        for (int i = 1; i <= imax - 1; i++) {
          const int ii[4] = {i, i + 1, i + 1, i}, jj[4] = {j, j, j + 1, j + 1};
          double xz[4], yz[4], zz[4];
          for (int m = 0; m < 4; m++) {
            xz[m] = (X(ii[m], jj[m], k + 1) - X(ii[m], jj[m], k - 1)) / 2;
            yz[m] = (Y(ii[m], jj[m], k + 1) - Y(ii[m], jj[m], k - 1)) / 2;
            zz[m] = (Z(ii[m], jj[m], k + 1) - Z(ii[m], jj[m], k - 1)) / 2;
          }
          const double xx1 = X(i + 1, j, k) - X(i, j, k), yx1 = Y(i + 1, j, k) - Y(i, j, k),
                       zx1 = Z(i + 1, j, k) - Z(i, j, k);
          const double xx2 = X(i + 1, j + 1, k) - X(i, j + 1, k), yx2 = Y(i + 1, j + 1, k) - Y(i, j + 1, k),
                       zx2 = Z(i + 1, j + 1, k) - Z(i, j + 1, k);
          const double xe1 = X(i, j + 1, k) - X(i, j, k), ye1 = Y(i, j + 1, k) - Y(i, j, k),
                       ze1 = Z(i, j + 1, k) - Z(i, j, k);
          const double xe2 = X(i + 1, j + 1, k) - X(i + 1, j, k), ye2 = Y(i + 1, j + 1, k) - Y(i + 1, j, k),
                       ze2 = Z(i + 1, j + 1, k) - Z(i + 1, j, k);
          const double volvisc = (VL(vol, im, jm, km, i, j, k) + VL(vol, im, jm, km, i, j, k - 1)) / 2;
          if (k == 1 || k == kmax) {
            const int h = k == 1 ? k - 1 : k + 1;
            for (int m = 0; m < 4; m++) {
              if (nd(X(ii[m], jj[m], h), Y(ii[m], jj[m], h), Z(ii[m], jj[m], h), X(ii[m], jj[m], k),
                     Y(ii[m], jj[m], k), Z(ii[m], jj[m], k)) < 1e-14) {
                xz[m] = 2 * xz[m];
                yz[m] = 2 * yz[m];
                zz[m] = 2 * zz[m];
              }
            }
          }
          const double x0 = (xz[0] + xz[1] + xz[2] + xz[3]) / 4, y0 = (yz[0] + yz[1] + yz[2] + yz[3]) / 4,
                       z0 = (zz[0] + zz[1] + zz[2] + zz[3]) / 4;
          jac_store((xx1 + xx2) / 2, (yx1 + yx2) / 2, (zx1 + zx2) / 2, (xe1 + xe2) / 2, (ye1 + ye2) / 2,
                    (ze1 + ze2) / 2, x0, y0, z0, volvisc, &DR(d[6], im, jm, km, 1, i, j, k, n),
                    &DR(d[7], im, jm, km, 1, i, j, k, n), &DR(d[8], im, jm, km, 1, i, j, k, n));
        }
#undef X
#undef Y
#undef Z
  }
}

/* ---- coarsen_b (src/cosa.f90:2609-2669): every second node of the fine
 * level (snapshot 0; the node arrays carry no snapshot index there). */
void or_coarsen_b(const double* x, const double* y, const double* z, int imax, int jmax, int kmax, double* x2,
                  double* y2, double* z2, int imax2, int jmax2, int kmax2) {
  for (int k = 1; k <= kmax; k += 2) {
    const int k2 = 1 + (k - 1) / 2;
    for (int j = 1; j <= jmax; j += 2) {
      const int j2 = 1 + (j - 1) / 2;
      for (int i = 1; i <= imax; i += 2) {
        const int i2 = 1 + (i - 1) / 2;
        XN(x2, imax2, jmax2, kmax2, i2, j2, k2, 0) = XN(x, imax, jmax, kmax, i, j, k, 0);
        XN(y2, imax2, jmax2, kmax2, i2, j2, k2, 0) = XN(y, imax, jmax, kmax, i, j, k, 0);
        XN(z2, imax2, jmax2, kmax2, i2, j2, k2, 0) = XN(z, imax, jmax, kmax, i, j, k, 0);
      }
    }
  }
}

/* ---- pitch (src/grid.f90:1872-1964): rigid pitching about (xh, yh) of
 * one snapshot: node positions and velocities from the base grid x0 and
 * the displacement scales dx, dy (P1 extents, read at node indices).
 * cth = dcos(dtheta), sth = dsin(dtheta), ca = dcos(omegas*zeit), all as
 * the host's libm gives them (the reference evaluates the same calls). */
void or_pitch(double cth, double sth, double ca, double dtheta0, double omegas, double xh, double yh,
              const double* x0, const double* y0, const double* z0, const double* dx, const double* dy, double* x,
              double* y, double* z, double* xdot, double* ydot, double* zdot, int imax, int jmax, int kmax) {
  const int im = imax, jm = jmax, km = kmax;
  for (int k = 1; k <= kmax; k++)
    for (int i = 1; i <= imax; i++)
      for (int j = 1; j <= jmax; j++) {
        const double X0 = XN(x0, im, jm, km, i, j, k, 0), Y0 = XN(y0, im, jm, km, i, j, k, 0);
        XN(z, im, jm, km, i, j, k, 0) = XN(z0, im, jm, km, i, j, k, 0);
        XN(zdot, im, jm, km, i, j, k, 0) = 0.0;
        const double DX = VL(dx, im, jm, km, i, j, k), DY = VL(dy, im, jm, km, i, j, k);
        if (fabs(DX) > 1e-4 && fabs(DY) > 1e-4) {
          const double d1 = ((X0 - xh) * (cth - 1) - (Y0 - yh) * sth) / DX;
          const double d2 = ((X0 - xh) * sth + (Y0 - yh) * (cth - 1)) / DY;
          XN(x, im, jm, km, i, j, k, 0) = X0 + d1 * DX;
          XN(y, im, jm, km, i, j, k, 0) = Y0 + d2 * DY;
          const double v1 = (-(X0 - xh) * sth - (Y0 - yh) * cth) * dtheta0 * ca / DX;
          const double v2 = ((X0 - xh) * cth - (Y0 - yh) * sth) * dtheta0 * ca / DY;
          XN(xdot, im, jm, km, i, j, k, 0) = omegas * v1 * DX;
          XN(ydot, im, jm, km, i, j, k, 0) = omegas * v2 * DY;
        } else {
          const double d1 = ((X0 - xh) * (cth - 1.0) - Y0 * sth);
          const double d2 = ((X0 - xh) * sth + Y0 * (cth - 1.0));
          XN(x, im, jm, km, i, j, k, 0) = X0 + d1;
          XN(y, im, jm, km, i, j, k, 0) = Y0 + d2;
          const double v1 = dtheta0 * ca * (-(X0 - xh) * sth - Y0 * cth);
          const double v2 = dtheta0 * ca * ((X0 - xh) * cth - Y0 * sth);
          XN(xdot, im, jm, km, i, j, k, 0) = omegas * v1;
          XN(ydot, im, jm, km, i, j, k, 0) = omegas * v2;
        }
      }
}

/* ---- cut_v / cut_x (src/grid.f90:612-838, 242-524), same-process branch:
 * one halo layer of block 1 from the first interior layer of block 2
 * across a 1:1 cut.  kind 0: volumes (P1), halo cell 0 or ijkmax1(idir1)
 * from cell 1 or ijkmax2(idir2) - 1; kind 1: nodes (P2 scalar, snapshot
 * n), halo node 0 or ijkmax1(idir1) + 1 from node 2 or ijkmax2(idir2) - 1,
 * with the transverse cell ranges widened by one node, every snapshot
 * 0..nsm-1 (cut_x is applied to x, y and z alike). */
static const int CYC[3][3] = {{1, 2, 3}, {2, 3, 1}, {3, 1, 2}};

void or_cut_layer(int kind, int nsm, int imax1, int jmax1, int kmax1, double* a1, int imax2, int jmax2, int kmax2,
                  const double* a2, const int32_t* cutopo) {
  const int ijk1[3] = {imax1, jmax1, kmax1}, ijk2[3] = {imax2, jmax2, kmax2};
  const int idir1 = cutopo[1], inout1 = cutopo[2], idir2 = cutopo[10], inout2 = cutopo[11];
  int istr1[3] = {cutopo[3], cutopo[5], cutopo[7]}, iend1[3] = {cutopo[4], cutopo[6], cutopo[8]};
  int istr2[3] = {cutopo[12], cutopo[14], cutopo[16]}, iend2[3] = {cutopo[13], cutopo[15], cutopo[17]};
  const int iord[3] = {cutopo[18], cutopo[19], cutopo[20]};
  int ibcpt, inr;
  if (kind == 0) {
    ibcpt = inout1 == 1 ? 0 : ijk1[idir1 - 1];
    inr = inout2 == 1 ? 1 : ijk2[idir2 - 1] - 1;
  } else {
    ibcpt = inout1 == 1 ? 0 : ijk1[idir1 - 1] + 1;
    inr = inout2 == 1 ? 2 : ijk2[idir2 - 1] - 1;
    for (int l = 1; l <= 3; l++)
      if (l != idir1) {
        if (iend1[l - 1] > istr1[l - 1]) iend1[l - 1] += 1;
        else istr1[l - 1] += 1;
      }
    for (int l = 1; l <= 3; l++)
      if (l != idir2) {
        if (iend2[l - 1] > istr2[l - 1]) iend2[l - 1] += 1;
        else istr2[l - 1] += 1;
      }
  }
  int len[3], sg1[3], sg2[3];
  for (int l = 0; l < 3; l++) {
    len[l] = abs(iend1[l] - istr1[l]);
    sg1[l] = iend1[l] > istr1[l] ? 1 : -1;
    sg2[l] = iend2[l] > istr2[l] ? 1 : -1;
  }
  const int ic1 = CYC[idir1 - 1][0], ic2 = CYC[idir1 - 1][1], ic3 = CYC[idir1 - 1][2];
  const int jc1 = iord[ic1 - 1], jc2 = iord[ic2 - 1], jc3 = iord[ic3 - 1];
  for (int i3 = 0; i3 <= len[ic3 - 1]; i3++)
    for (int i2 = 0; i2 <= len[ic2 - 1]; i2++) {
      int p1[3] = {0, 0, 0}, p2[3] = {0, 0, 0};
      p1[ic1 - 1] = ibcpt;
      p1[ic2 - 1] = istr1[ic2 - 1] + sg1[ic2 - 1] * i2;
      p1[ic3 - 1] = istr1[ic3 - 1] + sg1[ic3 - 1] * i3;
      p2[jc1 - 1] = inr;
      p2[jc2 - 1] = istr2[jc2 - 1] + sg2[jc2 - 1] * i2;
      p2[jc3 - 1] = istr2[jc3 - 1] + sg2[jc3 - 1] * i3;
      if (kind == 0)
        VL(a1, imax1, jmax1, kmax1, p1[0], p1[1], p1[2]) = VL(a2, imax2, jmax2, kmax2, p2[0], p2[1], p2[2]);
      else
        for (int n = 0; n < nsm; n++)
          XN(a1, imax1, jmax1, kmax1, p1[0], p1[1], p1[2], n) = XN(a2, imax2, jmax2, kmax2, p2[0], p2[1], p2[2], n);
    }
}
