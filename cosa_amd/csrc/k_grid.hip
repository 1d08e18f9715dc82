// k_grid.hip -- grid motion and metric set-up on the device: the producers
// of the hot path's geometric inputs (face metrics, volumes, metric
// derivatives, node positions and velocities of moving grids), so a solve
// can be admitted from node coordinates alone.
//
// Each kernel keeps the reference routine's arithmetic term by term, so the
// results are the same bits as COSA's CPU set-up:
//   k_imetrix<DIR>  imetrix_b   (src/grid.f90:848-1253)   one thread per face
//   k_bvols         bvols       (src/grid.f90:2162-2258)  one thread per cell
//   k_volex         volex_b     (src/grid.f90:110-141)    one thread per halo cell
//   k_vol_bedges    vol_bedges  (src/grid.f90:3607-3826)  one thread per edge cell
//   k_cordex        cordex_b    (src/grid.f90:28-78)      one thread per halo node
//   k_vmetrix<DIR>  vmetrix_b   (src/grid.f90:3150-3583)  one thread per face
//   k_coarsen       coarsen_b   (src/cosa.f90:2609-2669)  one thread per coarse node
//   k_pitch         pitch       (src/grid.f90:1872-1964)  one thread per node
//   k_cut_layer     cut_v / cut_x (src/grid.f90:612-838, 242-524) one thread per cut point
// Work items come from per-level tile tables (block, first item), one launch
// per kind over every block of the level.  Set-up work: memory-bound and
// run once per grid (or per admitted solve), not per stage.
#include "kernels.h"

namespace cosahip {


// node (i, j, k) of snapshot n; the level array holds the snapshots of a
// block contiguously: block base o2 * nsm, then snapshot n * n2
__device__ __forceinline__ ll nidx(const Blk& b, int nsm, int i, int j, int k, int n) {
  return b.o2 * nsm + n2of(b) * (ll)n + i2(b, i, j, k);
}

// unnormalised normal of face (i, j, k) of direction DIR as imetrix_b and
// bvols form it: half the cross product of the two face diagonals, in the
// reference's component order.
template <int DIR>
__device__ __forceinline__ void face_normal(const Blk& b, int nsm, const double* x, const double* y, const double* z,
                                            int i, int j, int k, int n, double& s1, double& s2, double& s3) {
#define X(a, bb, c) x[nidx(b, nsm, a, bb, c, n)]
#define Y(a, bb, c) y[nidx(b, nsm, a, bb, c, n)]
#define Z(a, bb, c) z[nidx(b, nsm, a, bb, c, n)]
  if (DIR == 1) {
    s1 = ((Y(i, j + 1, k + 1) - Y(i, j, k)) * (Z(i, j, k + 1) - Z(i, j + 1, k)) -
          (Z(i, j + 1, k + 1) - Z(i, j, k)) * (Y(i, j, k + 1) - Y(i, j + 1, k))) / 2;
    s2 = ((Z(i, j + 1, k + 1) - Z(i, j, k)) * (X(i, j, k + 1) - X(i, j + 1, k)) -
          (X(i, j + 1, k + 1) - X(i, j, k)) * (Z(i, j, k + 1) - Z(i, j + 1, k))) / 2;
    s3 = ((X(i, j + 1, k + 1) - X(i, j, k)) * (Y(i, j, k + 1) - Y(i, j + 1, k)) -
          (Y(i, j + 1, k + 1) - Y(i, j, k)) * (X(i, j, k + 1) - X(i, j + 1, k))) / 2;
  } else if (DIR == 2) {
    s1 = ((Z(i + 1, j, k + 1) - Z(i, j, k)) * (Y(i, j, k + 1) - Y(i + 1, j, k)) -
          (Y(i + 1, j, k + 1) - Y(i, j, k)) * (Z(i, j, k + 1) - Z(i + 1, j, k))) / 2;
    s2 = ((X(i + 1, j, k + 1) - X(i, j, k)) * (Z(i, j, k + 1) - Z(i + 1, j, k)) -
          (Z(i + 1, j, k + 1) - Z(i, j, k)) * (X(i, j, k + 1) - X(i + 1, j, k))) / 2;
    s3 = ((Y(i + 1, j, k + 1) - Y(i, j, k)) * (X(i, j, k + 1) - X(i + 1, j, k)) -
          (X(i + 1, j, k + 1) - X(i, j, k)) * (Y(i, j, k + 1) - Y(i + 1, j, k))) / 2;
  } else {
    s1 = ((Y(i + 1, j + 1, k) - Y(i, j, k)) * (Z(i, j + 1, k) - Z(i + 1, j, k)) -
          (Z(i + 1, j + 1, k) - Z(i, j, k)) * (Y(i, j + 1, k) - Y(i + 1, j, k))) / 2;
    s2 = ((Z(i + 1, j + 1, k) - Z(i, j, k)) * (X(i, j + 1, k) - X(i + 1, j, k)) -
          (X(i + 1, j + 1, k) - X(i, j, k)) * (Z(i, j + 1, k) - Z(i + 1, j, k))) / 2;
    s3 = ((X(i + 1, j + 1, k) - X(i, j, k)) * (Y(i, j + 1, k) - Y(i + 1, j, k)) -
          (Y(i + 1, j + 1, k) - Y(i, j, k)) * (X(i, j + 1, k) - X(i + 1, j, k))) / 2;
  }
#undef X
#undef Y
#undef Z
}

// item -> (snapshot, face) for the face family of direction DIR:
// i faces imax x (jmax-1) x (kmax-1), j faces (imax-1) x jmax x (kmax-1),
// k faces (imax-1) x (jmax-1) x kmax, i fastest
template <int DIR>
__device__ __forceinline__ bool face_item(const Blk& b, int nsm, ll item, int& n, int& i, int& j, int& k) {
  const int ni = DIR == 1 ? b.imax : b.imax - 1;
  const int nj = DIR == 2 ? b.jmax : b.jmax - 1;
  const int nk = DIR == 3 ? b.kmax : b.kmax - 1;
  const ll per = (ll)ni * nj * nk;
  if (item >= per * nsm) return false;
  n = (int)(item / per);
  ll r = item - per * n;
  i = 1 + (int)(r % ni);
  r /= ni;
  j = 1 + (int)(r % nj);
  k = 1 + (int)(r / nj);
  return true;
}

template <int DIR>
__global__ void __launch_bounds__(256) k_imetrix(const Blk* __restrict__ blks, const Tile* __restrict__ tiles,
                                                 int nsm, int lmet, int moving, GridPtrs g, double* __restrict__ s,
                                                 int* __restrict__ bad) {
  const Tile t = tiles[blockIdx.x];
  const Blk b = blks[t.block];
  int n, i, j, k;
  if (!face_item<DIR>(b, nsm, (ll)t.base + threadIdx.x, n, i, j, k)) return;
  double s1, s2, s3;
  face_normal<DIR>(b, nsm, g.x, g.y, g.z, i, j, k, n, s1, s2, s3);
  const double a = sqrt(s1 * s1 + s2 * s2 + s3 * s3);
  if (!(a > 0)) atomicOr(bad, 1);
  double* o = s + (b.o2 * nsm + n2of(b) * (ll)n + i2(b, i, j, k)) * lmet;
  o[0] = s1 / a;
  o[1] = s2 / a;
  o[2] = s3 / a;
  o[3] = a;
  if (moving) {
    // the face's four node velocities, in the reference's order
    // (src/grid.f90:1025-1064)
    const double* v[3] = {g.xd, g.yd, g.zd};
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const double* w = v[c];
      double r;
      if (DIR == 1)
        r = (w[nidx(b, nsm, i, j + 1, k + 1, n)] + w[nidx(b, nsm, i, j + 1, k, n)] + w[nidx(b, nsm, i, j, k + 1, n)] +
             w[nidx(b, nsm, i, j, k, n)]) / 4;
      else if (DIR == 2)
        r = (w[nidx(b, nsm, i + 1, j, k + 1, n)] + w[nidx(b, nsm, i + 1, j, k, n)] + w[nidx(b, nsm, i, j, k + 1, n)] +
             w[nidx(b, nsm, i, j, k, n)]) / 4;
      else
        r = (w[nidx(b, nsm, i + 1, j + 1, k, n)] + w[nidx(b, nsm, i + 1, j, k, n)] + w[nidx(b, nsm, i, j + 1, k, n)] +
             w[nidx(b, nsm, i, j, k, n)]) / 4;
      o[4 + c] = r;
    }
  }
}

// interior cells of snapshot 0 (Kordulla and Vinokur's formula, as bvols)
__global__ void __launch_bounds__(256) k_bvols(const Blk* __restrict__ blks, const Tile* __restrict__ tiles, int nsm,
                                               GridPtrs g, double* __restrict__ vol) {
  const Tile t = tiles[blockIdx.x];
  const Blk b = blks[t.block];
  const ll item = (ll)t.base + threadIdx.x;
  const int ni = b.imax - 1, nj = b.jmax - 1, nk = b.kmax - 1;
  if (item >= (ll)ni * nj * nk) return;
  const int i = 1 + (int)(item % ni);
  const ll r = item / ni;
  const int j = 1 + (int)(r % nj), k = 1 + (int)(r / nj);
  const int n = 0;
  double si1, si2, si3, sj1, sj2, sj3, sk1, sk2, sk3;
  face_normal<1>(b, nsm, g.x, g.y, g.z, i, j, k, n, si1, si2, si3);
  face_normal<2>(b, nsm, g.x, g.y, g.z, i, j, k, n, sj1, sj2, sj3);
  face_normal<3>(b, nsm, g.x, g.y, g.z, i, j, k, n, sk1, sk2, sk3);
  const ll c0 = nidx(b, nsm, i, j, k, n), c1 = nidx(b, nsm, i + 1, j + 1, k + 1, n);
  const double dx = (si1 + sj1 + sk1) * (g.x[c1] - g.x[c0]);
  const double dy = (si2 + sj2 + sk2) * (g.y[c1] - g.y[c0]);
  const double dz = (si3 + sj3 + sk3) * (g.z[c1] - g.z[c0]);
  // a non-positive volume is reported by the reference but not fatal there
  vol[b.o1 + i1(b, i, j, k)] = (dx + dy + dz) / 3;
}

// face-halo cells of vol from the adjacent interior cell (volex_b): items
// j halos 2 (imax-1)(kmax-1), then i halos 2 (jmax-1)(kmax-1), then k halos
// 2 (imax-1)(jmax-1)
__global__ void __launch_bounds__(256) k_volex(const Blk* __restrict__ blks, const Tile* __restrict__ tiles,
                                               double* __restrict__ vol) {
  const Tile t = tiles[blockIdx.x];
  const Blk b = blks[t.block];
  ll item = (ll)t.base + threadIdx.x;
  const int ni = b.imax - 1, nj = b.jmax - 1, nk = b.kmax - 1;
  double* v = vol + b.o1;
  const ll nJ = 2ll * ni * nk, nI = 2ll * nj * nk, nK = 2ll * ni * nj;
  if (item < nJ) {
    const int side = (int)(item % 2);
    const ll r = item / 2;
    const int i = 1 + (int)(r % ni), k = 1 + (int)(r / ni);
    v[i1(b, i, side ? b.jmax : 0, k)] = v[i1(b, i, side ? b.jmax - 1 : 1, k)];
    return;
  }
  item -= nJ;
  if (item < nI) {
    const int side = (int)(item % 2);
    const ll r = item / 2;
    const int j = 1 + (int)(r % nj), k = 1 + (int)(r / nj);
    v[i1(b, side ? b.imax : 0, j, k)] = v[i1(b, side ? b.imax - 1 : 1, j, k)];
    return;
  }
  item -= nI;
  if (item < nK) {
    const int side = (int)(item % 2);
    const ll r = item / 2;
    const int i = 1 + (int)(r % ni), j = 1 + (int)(r / ni);
    v[i1(b, i, j, side ? b.kmax : 0)] = v[i1(b, i, j, side ? b.kmax - 1 : 1)];
  }
}

// the plane through the three neighbours of an edge cross-section
// (vol_bedges, src/grid.f90:3745-3824); integer offsets as integers
__device__ __forceinline__ double edge_plane(const int c0[4], const int c1[4], double v1, double v2, double v3) {
  const double b = ((v3 - v1) * (double)(c0[2] - c0[1]) - (v2 - v1) * (double)(c0[3] - c0[1])) /
                   (double)((c1[3] - c1[1]) * (c0[2] - c0[1]) - (c1[2] - c1[1]) * (c0[3] - c0[1]));
  const double a = (v2 - v1 - b * (double)(c1[2] - c1[1])) / (double)(c0[2] - c0[1]);
  const double cc = v1 - a * (double)c0[1] - b * (double)c1[1];
  return a * (double)c0[0] + b * (double)c1[0] + cc;
}

// items: regions 1-4 x (kmax-1), then regions 5-8 x (jmax-1), then 9-12 x (imax-1)
__global__ void __launch_bounds__(256) k_vol_bedges(const Blk* __restrict__ blks, const Tile* __restrict__ tiles,
                                                    double* __restrict__ vol) {
  const Tile t = tiles[blockIdx.x];
  const Blk b = blks[t.block];
  ll item = (ll)t.base + threadIdx.x;
  const int im = b.imax, jm = b.jmax, km = b.kmax;
  double* v = vol + b.o1;
  int fam, r, line;
  if (item < 4ll * (km - 1)) {
    fam = 0; r = (int)(item % 4); line = 1 + (int)(item / 4);
  } else if ((item -= 4ll * (km - 1)) < 4ll * (jm - 1)) {
    fam = 1; r = (int)(item % 4); line = 1 + (int)(item / 4);
  } else if ((item -= 4ll * (jm - 1)) < 4ll * (im - 1)) {
    fam = 2; r = (int)(item % 4); line = 1 + (int)(item / 4);
  } else {
    return;
  }
  // the two in-plane index limits of the family: (i, j), (i, k), (j, k)
  const int A = fam == 2 ? jm : im, B = fam == 0 ? jm : km;
  // regions in the reference's order: (0,0), (A,0), (A,B), (0,B); points
  // 0 (target), 1 (the target's neighbour along the second index), 2 (along
  // the first), 3 (diagonal)
  const int fa = r == 1 || r == 2, fb = r == 2 || r == 3;
  const int c0[4] = {fa ? A : 0, fa ? A : 0, fa ? A - 1 : 1, fa ? A - 1 : 1};
  const int c1[4] = {fb ? B : 0, fb ? B - 1 : 1, fb ? B : 0, fb ? B - 1 : 1};
  ll idx[4];
#pragma unroll
  for (int m = 0; m < 4; m++)
    idx[m] = fam == 0 ? i1(b, c0[m], c1[m], line) : fam == 1 ? i1(b, c0[m], line, c1[m]) : i1(b, line, c0[m], c1[m]);
  v[idx[0]] = edge_plane(c0, c1, v[idx[1]], v[idx[2]], v[idx[3]]);
}

// halo nodes copy the adjacent boundary node (cordex_b): per snapshot, j
// and i halos of k = 1..kmax, then k halos of (1..imax, 1..jmax)
__global__ void __launch_bounds__(256) k_cordex(const Blk* __restrict__ blks, const Tile* __restrict__ tiles,
                                                int nsm, double* __restrict__ x, double* __restrict__ y,
                                                double* __restrict__ z) {
  const Tile t = tiles[blockIdx.x];
  const Blk b = blks[t.block];
  const ll nJ = 2ll * b.imax * b.kmax, nI = 2ll * b.jmax * b.kmax, nK = 2ll * b.imax * b.jmax;
  const ll per = nJ + nI + nK;
  ll item = (ll)t.base + threadIdx.x;
  if (item >= per * nsm) return;
  const int n = (int)(item / per);
  item -= per * n;
  ll dst, src;
  if (item < nJ) {
    const int side = (int)(item % 2);
    const ll r = item / 2;
    const int i = 1 + (int)(r % b.imax), k = 1 + (int)(r / b.imax);
    dst = nidx(b, nsm, i, side ? b.jmax + 1 : 0, k, n);
    src = nidx(b, nsm, i, side ? b.jmax : 1, k, n);
  } else if ((item -= nJ) < nI) {
    const int side = (int)(item % 2);
    const ll r = item / 2;
    const int j = 1 + (int)(r % b.jmax), k = 1 + (int)(r / b.jmax);
    dst = nidx(b, nsm, side ? b.imax + 1 : 0, j, k, n);
    src = nidx(b, nsm, side ? b.imax : 1, j, k, n);
  } else {
    item -= nI;
    const int side = (int)(item % 2);
    const ll r = item / 2;
    const int i = 1 + (int)(r % b.imax), j = 1 + (int)(r / b.imax);
    dst = nidx(b, nsm, i, j, side ? b.kmax + 1 : 0, n);
    src = nidx(b, nsm, i, j, side ? b.kmax : 1, n);
  }
  x[dst] = x[src];
  y[dst] = y[src];
  z[dst] = z[src];
}

// one halo layer across a 1:1 cut (cut_v: kind 0, volumes; cut_x: kind 1,
// nodes of every snapshot), one thread per (record, point); the record's
// ranges are pre-adjusted on the host (GeomCut)
__global__ void __launch_bounds__(256) k_cut_layer(const Blk* __restrict__ blks, const GeomCut* __restrict__ cuts,
                                                   int ncut, int maxcount, int kind, int nsm, double* __restrict__ a,
                                                   double* __restrict__ b2y, double* __restrict__ b2z) {
  const ll gid = (ll)blockIdx.x * blockDim.x + threadIdx.x;
  const int rec = (int)(gid / maxcount), pt = (int)(gid - (ll)rec * maxcount);
  if (rec >= ncut) return;
  const GeomCut c = cuts[rec];
  if (pt >= c.count) return;
  const int i2 = pt % c.n2, i3 = pt / c.n2;
  int p1[3], p2[3];
  p1[c.ic1] = c.ibcpt;
  p1[c.ic2] = c.str1[c.ic2] + c.sg1[c.ic2] * i2;
  p1[c.ic3] = c.str1[c.ic3] + c.sg1[c.ic3] * i3;
  p2[c.jc1] = c.inr;
  p2[c.jc2] = c.str2[c.jc2] + c.sg2[c.jc2] * i2;
  p2[c.jc3] = c.str2[c.jc3] + c.sg2[c.jc3] * i3;
  const Blk B1 = blks[c.b1], B2 = blks[c.b2];
  if (kind == 0) {
    a[B1.o1 + i1(B1, p1[0], p1[1], p1[2])] = a[B2.o1 + i1(B2, p2[0], p2[1], p2[2])];
  } else {
    double* arr[3] = {a, b2y, b2z};
    for (int n = 0; n < nsm; n++)
#pragma unroll
      for (int m = 0; m < 3; m++)
        arr[m][nidx(B1, nsm, p1[0], p1[1], p1[2], n)] = arr[m][nidx(B2, nsm, p2[0], p2[1], p2[2], n)];
  }
}

// metric derivatives as the inverse of the face Jacobian (vmetrix_b)
__device__ __forceinline__ double ndist(const GridPtrs& g, ll a, ll c) {
  const double d1 = g.x[a] - g.x[c], d2 = g.y[a] - g.y[c], d3 = g.z[a] - g.z[c];
  return sqrt(d1 * d1 + d2 * d2 + d3 * d3);
}

template <int DIR>
__global__ void __launch_bounds__(256) k_vmetrix(const Blk* __restrict__ blks, const Tile* __restrict__ tiles, int nsm,
                                                 GridPtrs g, const double* __restrict__ vol,
                                                 double* const* __restrict__ d) {
  const Tile t = tiles[blockIdx.x];
  const Blk b = blks[t.block];
  int n, i, j, k;
  if (!face_item<DIR>(b, nsm, (ll)t.base + threadIdx.x, n, i, j, k)) return;
  // the four nodes of the face in the reference's order, the node offset
  // along DIR, and the two edge pairs of each tangential direction
  int P[4][3];
  if (DIR == 1) {
    const int jj[4] = {j, j + 1, j + 1, j}, kk[4] = {k, k, k + 1, k + 1};
    for (int m = 0; m < 4; m++) { P[m][0] = i; P[m][1] = jj[m]; P[m][2] = kk[m]; }
  } else if (DIR == 2) {
    const int ii[4] = {i, i, i + 1, i + 1}, kk[4] = {k, k + 1, k + 1, k};
    for (int m = 0; m < 4; m++) { P[m][0] = ii[m]; P[m][1] = j; P[m][2] = kk[m]; }
  } else {
    const int ii[4] = {i, i + 1, i + 1, i}, jj[4] = {j, j, j + 1, j + 1};
    for (int m = 0; m < 4; m++) { P[m][0] = ii[m]; P[m][1] = jj[m]; P[m][2] = k; }
  }
  const int e[3] = {DIR == 1, DIR == 2, DIR == 3};
  const int a = DIR == 1 ? i : DIR == 2 ? j : k;
  const int L = DIR == 1 ? b.imax : DIR == 2 ? b.jmax : b.kmax;
  double cx[4], cy[4], cz[4];
#pragma unroll
  for (int m = 0; m < 4; m++) {
    const ll pp = nidx(b, nsm, P[m][0] + e[0], P[m][1] + e[1], P[m][2] + e[2], n);
    const ll pm = nidx(b, nsm, P[m][0] - e[0], P[m][1] - e[1], P[m][2] - e[2], n);
    cx[m] = (g.x[pp] - g.x[pm]) / 2;
    cy[m] = (g.y[pp] - g.y[pm]) / 2;
    cz[m] = (g.z[pp] - g.z[pm]) / 2;
    if (a == 1 || a == L) {
      const ll pc = nidx(b, nsm, P[m][0], P[m][1], P[m][2], n);
      if (ndist(g, a == 1 ? pm : pp, pc) < 1e-14) {
        cx[m] = 2 * cx[m];
        cy[m] = 2 * cy[m];
        cz[m] = 2 * cz[m];
      }
    }
  }
  const double nx = (cx[0] + cx[1] + cx[2] + cx[3]) / 4, ny = (cy[0] + cy[1] + cy[2] + cy[3]) / 4,
               nz = (cz[0] + cz[1] + cz[2] + cz[3]) / 4;
  // tangential derivatives: the two edge differences of each transverse
  // direction, first the edge through (i, j, k), then the parallel one
  // (src/grid.f90:3223-3236, 3349-3362, 3475-3488)
  double tx[2], ty[2], tz[2];
  int tcount = 0;
#pragma unroll
  for (int d2 = 1; d2 <= 3; d2++) {
    if (d2 == DIR) continue;
    const int f[3] = {d2 == 1, d2 == 2, d2 == 3};
    // the other transverse direction
    const int o = 6 - DIR - d2;
    const int h[3] = {o == 1, o == 2, o == 3};
    const ll a0 = nidx(b, nsm, i, j, k, n), a1 = nidx(b, nsm, i + f[0], j + f[1], k + f[2], n);
    const ll b0 = nidx(b, nsm, i + h[0], j + h[1], k + h[2], n),
             b1 = nidx(b, nsm, i + h[0] + f[0], j + h[1] + f[1], k + h[2] + f[2], n);
    tx[tcount] = ((g.x[a1] - g.x[a0]) + (g.x[b1] - g.x[b0])) / 2;
    ty[tcount] = ((g.y[a1] - g.y[a0]) + (g.y[b1] - g.y[b0])) / 2;
    tz[tcount] = ((g.z[a1] - g.z[a0]) + (g.z[b1] - g.z[b0])) / 2;
    tcount++;
  }
  double xx, yx, zx, xe, ye, ze, xz, yz, zz;
  if (DIR == 1) {
    xx = nx; yx = ny; zx = nz; xe = tx[0]; ye = ty[0]; ze = tz[0]; xz = tx[1]; yz = ty[1]; zz = tz[1];
  } else if (DIR == 2) {
    xx = tx[0]; yx = ty[0]; zx = tz[0]; xe = nx; ye = ny; ze = nz; xz = tx[1]; yz = ty[1]; zz = tz[1];
  } else {
    xx = tx[0]; yx = ty[0]; zx = tz[0]; xe = tx[1]; ye = ty[1]; ze = tz[1]; xz = nx; yz = ny; zz = nz;
  }
  const double* vb = vol + b.o1;
  const double volvisc = DIR == 1 ? (vb[i1(b, i - 1, j, k)] + vb[i1(b, i, j, k)]) / 2
                         : DIR == 2 ? (vb[i1(b, i, j, k)] + vb[i1(b, i, j - 1, k)]) / 2
                                    : (vb[i1(b, i, j, k)] + vb[i1(b, i, j, k - 1)]) / 2;
  const ll o = (b.o0 * nsm + (ll)n * b.imax * b.jmax * b.kmax +
                ((ll)(i - 1) + (ll)b.imax * ((ll)(j - 1) + (ll)b.jmax * (ll)(k - 1)))) * 3;
  double* dxi = d[3 * (DIR - 1)] + o;
  double* deta = d[3 * (DIR - 1) + 1] + o;
  double* dzeta = d[3 * (DIR - 1) + 2] + o;
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

// every second node of the fine level, snapshot 0 (coarsen_b)
__global__ void __launch_bounds__(256) k_coarsen(const Blk* __restrict__ fb, const Blk* __restrict__ cb,
                                                 const Tile* __restrict__ tiles, int nsm_f, int nsm_c,
                                                 GridPtrs f, double* __restrict__ x2, double* __restrict__ y2,
                                                 double* __restrict__ z2) {
  const Tile t = tiles[blockIdx.x];
  const Blk C = cb[t.block], F = fb[t.block];
  const ll item = (ll)t.base + threadIdx.x;
  if (item >= (ll)C.imax * C.jmax * C.kmax) return;
  const int i2c = 1 + (int)(item % C.imax);
  const ll r = item / C.imax;
  const int j2c = 1 + (int)(r % C.jmax), k2c = 1 + (int)(r / C.jmax);
  const ll s = nidx(F, nsm_f, 2 * i2c - 1, 2 * j2c - 1, 2 * k2c - 1, 0), d = nidx(C, nsm_c, i2c, j2c, k2c, 0);
  x2[d] = f.x[s];
  y2[d] = f.y[s];
  z2[d] = f.z[s];
}

// rigid pitching of every snapshot about (xh, yh) from the base pose, with
// unit displacement scales dx = dy = 1 (pitch's first branch; multiplying
// and dividing by 1 is exact, so those steps are left out).  pose[3n ..]:
// dcos(dtheta), dsin(dtheta), dcos(omegas*zeit) of snapshot n.
__global__ void __launch_bounds__(256) k_pitch(const Blk* __restrict__ blks, const Tile* __restrict__ tiles, int nsm,
                                               const double* __restrict__ pose, double dtheta0, double omegas,
                                               double xh, double yh, const double* __restrict__ x0,
                                               const double* __restrict__ y0, const double* __restrict__ z0,
                                               double* __restrict__ x, double* __restrict__ y, double* __restrict__ z,
                                               double* __restrict__ xd, double* __restrict__ yd,
                                               double* __restrict__ zd) {
  const Tile t = tiles[blockIdx.x];
  const Blk b = blks[t.block];
  const ll per = (ll)b.imax * b.jmax * b.kmax;
  ll item = (ll)t.base + threadIdx.x;
  if (item >= per * nsm) return;
  const int n = (int)(item / per);
  item -= per * n;
  const int i = 1 + (int)(item % b.imax);
  const ll r = item / b.imax;
  const int j = 1 + (int)(r % b.jmax), k = 1 + (int)(r / b.jmax);
  const ll s0 = b.o2 + i2(b, i, j, k);   // base pose: one snapshot per block
  const ll d = nidx(b, nsm, i, j, k, n);
  const double cth = pose[3 * n], sth = pose[3 * n + 1], ca = pose[3 * n + 2];
  const double X0 = x0[s0], Y0 = y0[s0];
  z[d] = z0[s0];
  zd[d] = 0.0;
  const double d1 = (X0 - xh) * (cth - 1) - (Y0 - yh) * sth;
  const double d2 = (X0 - xh) * sth + (Y0 - yh) * (cth - 1);
  x[d] = X0 + d1;
  y[d] = Y0 + d2;
  const double v1 = (-(X0 - xh) * sth - (Y0 - yh) * cth) * dtheta0 * ca;
  const double v2 = ((X0 - xh) * cth - (Y0 - yh) * sth) * dtheta0 * ca;
  xd[d] = omegas * v1;
  yd[d] = omegas * v2;
}

// ---------------------------------------------------------------------------
// launchers

static dim3 g1(int n) { return dim3((unsigned)std::max(n, 1)); }

hipError_t launch_grid_metrics(Solve& s, Level& L, const GridTiles& gt, int* dbad, int do_volex, int do_bedges,
                               int do_vmetrix) {
  const Par& p = s.par;
  const int nsm = p.nsnap_met;
  GridPtrs g{L.f[COSAHIP_X], L.f[COSAHIP_Y], L.f[COSAHIP_Z], L.f[COSAHIP_XDOT], L.f[COSAHIP_YDOT],
             L.f[COSAHIP_ZDOT]};
  if (gt.n[GT_CELLS])
    hipLaunchKernelGGL(k_bvols, g1(gt.n[GT_CELLS]), dim3(256), 0, s.stream, L.dblk, gt.t[GT_CELLS], nsm, g,
                       L.f[COSAHIP_VOL]);
  if (gt.n[GT_FI])
    hipLaunchKernelGGL(k_imetrix<1>, g1(gt.n[GT_FI]), dim3(256), 0, s.stream, L.dblk, gt.t[GT_FI], nsm, p.lmet,
                       p.moving, g, L.f[COSAHIP_SI], dbad);
  if (gt.n[GT_FJ])
    hipLaunchKernelGGL(k_imetrix<2>, g1(gt.n[GT_FJ]), dim3(256), 0, s.stream, L.dblk, gt.t[GT_FJ], nsm, p.lmet,
                       p.moving, g, L.f[COSAHIP_SJ], dbad);
  if (gt.n[GT_FK])
    hipLaunchKernelGGL(k_imetrix<3>, g1(gt.n[GT_FK]), dim3(256), 0, s.stream, L.dblk, gt.t[GT_FK], nsm, p.lmet,
                       p.moving, g, L.f[COSAHIP_SK], dbad);
  if (do_volex) {
    if (gt.n[GT_VOLEX])
      hipLaunchKernelGGL(k_volex, g1(gt.n[GT_VOLEX]), dim3(256), 0, s.stream, L.dblk, gt.t[GT_VOLEX],
                         L.f[COSAHIP_VOL]);
    if (gt.ncut)
      hipLaunchKernelGGL(k_cut_layer, g1((int)(((ll)gt.ncut * gt.maxcut_v + 255) / 256)), dim3(256), 0, s.stream,
                         L.dblk, gt.cut_v, gt.ncut, gt.maxcut_v, 0, 1, L.f[COSAHIP_VOL], nullptr, nullptr);
  }
  if (do_bedges && gt.n[GT_EDGES])
    hipLaunchKernelGGL(k_vol_bedges, g1(gt.n[GT_EDGES]), dim3(256), 0, s.stream, L.dblk, gt.t[GT_EDGES],
                       L.f[COSAHIP_VOL]);
  if (do_vmetrix) {
    if (gt.n[GT_NODEHALO])
      hipLaunchKernelGGL(k_cordex, g1(gt.n[GT_NODEHALO]), dim3(256), 0, s.stream, L.dblk, gt.t[GT_NODEHALO], nsm,
                         L.f[COSAHIP_X], L.f[COSAHIP_Y], L.f[COSAHIP_Z]);
    if (gt.ncut)
      hipLaunchKernelGGL(k_cut_layer, g1((int)(((ll)gt.ncut * gt.maxcut_x + 255) / 256)), dim3(256), 0, s.stream,
                         L.dblk, gt.cut_x, gt.ncut, gt.maxcut_x, 1, nsm, L.f[COSAHIP_X], L.f[COSAHIP_Y],
                         L.f[COSAHIP_Z]);
    if (gt.n[GT_FI])
      hipLaunchKernelGGL(k_vmetrix<1>, g1(gt.n[GT_FI]), dim3(256), 0, s.stream, L.dblk, gt.t[GT_FI], nsm, g,
                         L.f[COSAHIP_VOL], L.dxd);
    if (gt.n[GT_FJ])
      hipLaunchKernelGGL(k_vmetrix<2>, g1(gt.n[GT_FJ]), dim3(256), 0, s.stream, L.dblk, gt.t[GT_FJ], nsm, g,
                         L.f[COSAHIP_VOL], L.dxd);
    if (gt.n[GT_FK])
      hipLaunchKernelGGL(k_vmetrix<3>, g1(gt.n[GT_FK]), dim3(256), 0, s.stream, L.dblk, gt.t[GT_FK], nsm, g,
                         L.f[COSAHIP_VOL], L.dxd);
  }
  return hipGetLastError();
}

hipError_t launch_grid_coarsen(Solve& s, Level& F, Level& C, const GridTiles& gtc) {
  const int nsm = s.par.nsnap_met;
  GridPtrs f{F.f[COSAHIP_X], F.f[COSAHIP_Y], F.f[COSAHIP_Z], nullptr, nullptr, nullptr};
  if (gtc.n[GT_NODES0])
    hipLaunchKernelGGL(k_coarsen, g1(gtc.n[GT_NODES0]), dim3(256), 0, s.stream, F.dblk, C.dblk, gtc.t[GT_NODES0],
                       nsm, nsm, f, C.f[COSAHIP_X], C.f[COSAHIP_Y], C.f[COSAHIP_Z]);
  return hipGetLastError();
}

hipError_t launch_grid_pitch(Solve& s, Level& L, const GridTiles& gt, const double* pose_dev, double dtheta0,
                             double omegas, double xh, double yh, const double* base) {
  const int nsm = s.par.nsnap_met;
  const ll n2 = L.n2tot;
  if (gt.n[GT_NODES])
    hipLaunchKernelGGL(k_pitch, g1(gt.n[GT_NODES]), dim3(256), 0, s.stream, L.dblk, gt.t[GT_NODES], nsm, pose_dev,
                       dtheta0, omegas, xh, yh, base, base + n2, base + 2 * n2, L.f[COSAHIP_X], L.f[COSAHIP_Y],
                       L.f[COSAHIP_Z], L.f[COSAHIP_XDOT], L.f[COSAHIP_YDOT], L.f[COSAHIP_ZDOT]);
  return hipGetLastError();
}

}  // namespace cosahip
