// k_core.hip -- state conversions, boundary/halo fills, the RK update,
// residual norms, the harmonic-balance source and multigrid transfer of
// COSA's pseudo-time loop, for gfx950.  Each kernel restates one reference
// routine (cited); arithmetic keeps the reference's operation order.
#include "kernels.h"

namespace cosahip {

static inline dim3 grid1(int n) { return dim3(n > 0 ? n : 1); }

// decode an interior-cell work item (i fastest) of TK_CELLS
struct CellIdx {
  int n, i, j, k;
};
__device__ __forceinline__ bool cell_of(const Par& p, const Blk& b, int item, CellIdx& c) {
  const int im = b.imax - 1, jm = b.jmax - 1, km = b.kmax - 1;
  const int nc = im * jm * km;
  if (item >= nc * p.nsnap) return false;
  c.n = item / nc;
  int rem = item - c.n * nc;
  c.i = rem % im + 1;
  rem /= im;
  c.j = rem % jm + 1;
  c.k = rem / jm + 1;
  return true;
}

// TK_HALO items: interior + face halo strips, in bprtest's loop order
//   region 0: k=1..km, j=1..jm, i=-1..imax+1
//   region 1: j=-1..0     region 2: j=jmax..jmax+1   (i=1..im, k=1..km)
//   region 3: k=-1..0     region 4: k=kmax..kmax+1   (i=1..im, j=1..jm)
__device__ __forceinline__ bool halo_of(const Par& p, const Blk& b, int item, int& n, int& reg, int& loc, int& i,
                                        int& j, int& k) {
  const int im = b.imax - 1, jm = b.jmax - 1, km = b.kmax - 1;
  const int r0 = (b.imax + 3) * jm * km, r1 = 2 * im * km, r3 = 2 * im * jm;
  const int per = r0 + 2 * r1 + 2 * r3;
  if (item >= per * p.nsnap) return false;
  n = item / per;
  int x = item - n * per;
  if (x < r0) {
    reg = 0;
    loc = x;
    i = x % (b.imax + 3) - 1;
    x /= (b.imax + 3);
    j = x % jm + 1;
    k = x / jm + 1;
    return true;
  }
  x -= r0;
  if (x < 2 * r1) {
    reg = 1 + x / r1;
    x -= (reg - 1) * r1;
    loc = x;
    i = x % im + 1;
    x /= im;
    int jj = x % 2;
    k = x / 2 + 1;
    j = reg == 1 ? jj - 1 : b.jmax + jj;
    return true;
  }
  x -= 2 * r1;
  reg = 3 + x / r3;
  x -= (reg - 3) * r3;
  loc = x;
  i = x % im + 1;
  x /= im;
  j = x % jm + 1;
  int kk = x / jm;
  k = reg == 3 ? kk - 1 : b.kmax + kk;
  return true;
}

__device__ __forceinline__ double* qblk(const Par& p, const Blk& b, double* q, int n) {
  return q + b.o3 * p.npde * p.nsnap + n3of(b) * (ll)p.npde * n;
}
__device__ __forceinline__ const double* qblk(const Par& p, const Blk& b, const double* q, int n) {
  return q + b.o3 * p.npde * p.nsnap + n3of(b) * (ll)p.npde * n;
}

// ---- cons2prim_b (src/update.f90:2136-2188), interior ----
__global__ void __launch_bounds__(256) k_cons2prim(Par p, const Blk* __restrict__ blks, const Tile* __restrict__ tiles,
                                                   double* __restrict__ q, float* __restrict__ q32,
                                                   double* __restrict__ qold) {
  const Tile t = tiles[blockIdx.x];
  const Blk b = blks[t.block];
  CellIdx c;
  if (!cell_of(p, b, t.base + threadIdx.x, c)) return;
  double* qb = qblk(p, b, q, c.n);
  const ll n3 = n3of(b), x = i3(b, c.i, c.j, c.k);
  if (qold) {   // copy_array of the conservative state first (the RK stage's qold)
    double* ob = qblk(p, b, qold, c.n);
    for (int m = 0; m < p.npde; m++) ob[x + m * n3] = qb[x + m * n3];
  }
  double rho = qb[x];
  double u = qb[x + n3] / rho, v = qb[x + 2 * n3] / rho, w = qb[x + 3 * n3] / rho;
  for (int m = 5; m < p.npde; m++) qb[x + m * n3] = qb[x + m * n3] / rho;
  double rhoe = qb[x + 4 * n3];
  double tke = p.kom ? qb[x + 5 * n3] : 0.0;
  qb[x + n3] = u;
  qb[x + 2 * n3] = v;
  qb[x + 3 * n3] = w;
  const double pp = (p.gamma - 1) * (rhoe - rho * (u * u + v * v + w * w) / 2 - rho * tke);
  qb[x + 4 * n3] = pp;
  if (q32) {   // FP32 shadow of the primitive state for the FP32-residual mode
    float* sb = q32 + (qb - q);
    sb[x] = (float)rho;
    sb[x + n3] = (float)u;
    sb[x + 2 * n3] = (float)v;
    sb[x + 3 * n3] = (float)w;
    sb[x + 4 * n3] = (float)pp;
    for (int m = 5; m < p.npde; m++) sb[x + m * n3] = (float)qb[x + m * n3];
  }
}

// FP32 shadow of the halo strips of q (the points of the prim2cons region
// outside the interior), after the BC and cut fills.
__global__ void __launch_bounds__(256) k_halo_f32(Par p, const Blk* __restrict__ blks, const Tile* __restrict__ tiles,
                                                  const double* __restrict__ q, float* __restrict__ q32) {
  const Tile t = tiles[blockIdx.x];
  const Blk b = blks[t.block];
  int n, reg, loc, i, j, k;
  if (!halo_of(p, b, t.base + threadIdx.x, n, reg, loc, i, j, k)) return;
  if (reg == 0 && i >= 1 && i <= b.imax - 1) return;
  const ll o = b.o3 * p.npde * p.nsnap + n3of(b) * (ll)p.npde * n;
  const ll n3 = n3of(b), x = i3(b, i, j, k);
  for (int m = 0; m < p.npde; m++) q32[o + x + m * n3] = (float)q[o + x + m * n3];
}

// dst = (float)src, n elements
__global__ void __launch_bounds__(256) k_to_f32(const double* __restrict__ src, float* __restrict__ dst, ll n) {
  for (ll x = (ll)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (ll)gridDim.x * blockDim.x) dst[x] = (float)src[x];
}

// ---- prim2cons_b (src/update.f90:1875-2084), interior + face halo strips ----
__global__ void __launch_bounds__(256) k_prim2cons(Par p, const Blk* __restrict__ blks, const Tile* __restrict__ tiles,
                                                   double* __restrict__ q, int halo_only) {
  const Tile t = tiles[blockIdx.x];
  const Blk b = blks[t.block];
  int n, reg, loc, i, j, k;
  if (!halo_of(p, b, t.base + threadIdx.x, n, reg, loc, i, j, k)) return;
  if (halo_only && reg == 0 && i >= 1 && i <= b.imax - 1) return;   // interior cell
  double* qb = qblk(p, b, q, n);
  const ll n3 = n3of(b), x = i3(b, i, j, k);
  double rho = qb[x];
  double rhou = qb[x + n3] * rho, rhov = qb[x + 2 * n3] * rho, rhow = qb[x + 3 * n3] * rho;
  double tke = p.kom ? qb[x + 5 * n3] : 0.0;   // k saved before the scaling (work(:,1) of prim2cons_b)
  for (int m = 5; m < p.npde; m++) qb[x + m * n3] = qb[x + m * n3] * rho;
  double pp = qb[x + 4 * n3];
  qb[x + n3] = rhou;
  qb[x + 2 * n3] = rhov;
  qb[x + 3 * n3] = rhow;
  qb[x + 4 * n3] = pp / (p.gamma - 1) + (rhou * rhou + rhov * rhov + rhow * rhow) / 2 / rho + rho * tke;
}

// ---- copy_array / add_array, off_l = -1 (src/util.f90:2-236) ----
__global__ void __launch_bounds__(256) k_copy_int(Par p, const Blk* __restrict__ blks, const Tile* __restrict__ tiles,
                                                  double* __restrict__ dst, const double* __restrict__ src) {
  const Tile t = tiles[blockIdx.x];
  const Blk b = blks[t.block];
  CellIdx c;
  if (!cell_of(p, b, t.base + threadIdx.x, c)) return;
  const ll n3 = n3of(b), x = i3(b, c.i, c.j, c.k);
  double* d = qblk(p, b, dst, c.n);
  const double* s = qblk(p, b, src, c.n);
  for (int v = 0; v < p.npde; v++) d[x + v * n3] = s[x + v * n3];
}
__global__ void __launch_bounds__(256) k_add_int(Par p, const Blk* __restrict__ blks, const Tile* __restrict__ tiles,
                                                 double* __restrict__ dst, const double* __restrict__ src,
                                                 double fact) {
  const Tile t = tiles[blockIdx.x];
  const Blk b = blks[t.block];
  CellIdx c;
  if (!cell_of(p, b, t.base + threadIdx.x, c)) return;
  const ll n3 = n3of(b), x = i3(b, c.i, c.j, c.k);
  double* d = qblk(p, b, dst, c.n);
  const double* s = qblk(p, b, src, c.n);
  for (int v = 0; v < p.npde; v++) d[x + v * n3] = d[x + v * n3] + s[x + v * n3] * fact;
}

// ---- zero (src/util.f90:238-369) on 0..imax x 0..jmax x 0..kmax ----
__global__ void __launch_bounds__(256) k_zero_box(Par p, const Blk* __restrict__ blks, const Tile* __restrict__ tiles,
                                                  double* __restrict__ res) {
  const Tile t = tiles[blockIdx.x];
  const Blk b = blks[t.block];
  const int item = t.base + threadIdx.x;
  const int per = (b.imax + 1) * (b.jmax + 1) * (b.kmax + 1);
  if (item >= per * p.nsnap) return;
  const int n = item / per;
  int x = item - n * per;
  const int i = x % (b.imax + 1);
  x /= (b.imax + 1);
  const int j = x % (b.jmax + 1), k = x / (b.jmax + 1);
  double* rb = qblk(p, b, res, n);
  const ll n3 = n3of(b), y = i3(b, i, j, k);
  for (int v = 0; v < p.npde; v++) rb[y + v * n3] = 0.0;
}

// res = 0 on the halo shell of 0..imax x 0..jmax x 0..kmax (the part of
// zero(res) that the residual kernels, which write every interior cell, do
// not cover): the faces i = 0, imax (all j, k), then j = 0, jmax (interior
// i, all k), then k = 0, kmax (interior i, j).
__global__ void __launch_bounds__(256) k_zero_shell(Par p, const Blk* __restrict__ blks,
                                                    const Tile* __restrict__ tiles, double* __restrict__ res) {
  const Tile t = tiles[blockIdx.x];
  const Blk b = blks[t.block];
  const int item = t.base + threadIdx.x;
  const int im = b.imax - 1, jm = b.jmax - 1;
  const int ra = 2 * (b.jmax + 1) * (b.kmax + 1), rb = 2 * im * (b.kmax + 1), rc = 2 * im * jm;
  const int per = ra + rb + rc;
  if (item >= per * p.nsnap) return;
  const int n = item / per;
  int x = item - n * per, i, j, k;
  if (x < ra) {
    const int side = x / ((b.jmax + 1) * (b.kmax + 1));
    x -= side * (b.jmax + 1) * (b.kmax + 1);
    i = side ? b.imax : 0;
    j = x % (b.jmax + 1);
    k = x / (b.jmax + 1);
  } else if (x < ra + rb) {
    x -= ra;
    const int side = x / (im * (b.kmax + 1));
    x -= side * im * (b.kmax + 1);
    j = side ? b.jmax : 0;
    i = x % im + 1;
    k = x / im;
  } else {
    x -= ra + rb;
    const int side = x / (im * jm);
    x -= side * im * jm;
    k = side ? b.kmax : 0;
    i = x % im + 1;
    j = x / im + 1;
  }
  double* rbk = qblk(p, b, res, n);
  const ll n3 = n3of(b), y = i3(b, i, j, k);
  for (int v = 0; v < p.npde; v++) rbk[y + v * n3] = 0.0;
}

// q2 = q2 + qold2 * (-1) on 0..imax (add_array off_l = 1, prolong)
__global__ void __launch_bounds__(256) k_sub_box(Par p, const Blk* __restrict__ blks, const Tile* __restrict__ tiles,
                                                 double* __restrict__ dst, const double* __restrict__ src) {
  const Tile t = tiles[blockIdx.x];
  const Blk b = blks[t.block];
  const int item = t.base + threadIdx.x;
  const int per = (b.imax + 1) * (b.jmax + 1) * (b.kmax + 1);
  if (item >= per * p.nsnap) return;
  const int n = item / per;
  int x = item - n * per;
  const int i = x % (b.imax + 1);
  x /= (b.imax + 1);
  const int j = x % (b.jmax + 1), k = x / (b.jmax + 1);
  double* d = qblk(p, b, dst, n);
  const double* s = qblk(p, b, src, n);
  const ll n3 = n3of(b), y = i3(b, i, j, k);
  for (int v = 0; v < p.npde; v++) d[y + v * n3] = d[y + v * n3] + s[y + v * n3] * -1.0;
}

// ---- update_b / update_bhb explicit (src/update.f90:568-624, 1359-1408) ----
// With Wilcox k-omega: the point-implicit rho*omega, rho*k update first
// (src/update.f90:596-634), reading rho, k, omega of the current state;
// ttrm(:,2:3) of kom are the constants gkw, bkw.
__global__ void __launch_bounds__(256) k_update(Par p, const Blk* __restrict__ blks, const Tile* __restrict__ tiles,
                                                const double* __restrict__ res, double* __restrict__ q,
                                                const double* __restrict__ qold, const double* __restrict__ dtvol,
                                                double ar, const double* __restrict__ vol,
                                                const double* __restrict__ dtvolt, const double* __restrict__ delpls,
                                                const double* __restrict__ prd, const double* __restrict__ ttrm) {
  const Tile t = tiles[blockIdx.x];
  const Blk b = blks[t.block];
  CellIdx c;
  if (!cell_of(p, b, t.base + threadIdx.x, c)) return;
  const ll n3 = n3of(b), x = i3(b, c.i, c.j, c.k);
  const ll po = b.o1 * p.nsnap + n1of(b) * (ll)c.n + i1(b, c.i, c.j, c.k);
  const double dt = dtvol[po];
  double* qb = qblk(p, b, q, c.n);
  const double* ob = qblk(p, b, qold, c.n);
  const double* rb = qblk(p, b, res, c.n);
  if (p.kom) {
    const double rho = qb[x];
    const double ome = qb[x + 6 * n3] / qb[x];
    const double tke = qb[x + 5 * n3] / qb[x];
    const double dtt = dtvolt[po];
    const double dtau = dtt * vol[b.o1 + i1(b, c.i, c.j, c.k)];
    const ll so = b.o3 * p.nsnap + n3 * (ll)c.n + x;
    // ttrm(2:3): the constants gkw, bkw of Wilcox k-omega, blended for BSL/SST
    const double* tb = p.sst ? ttrm + b.o3 * 3 * p.nsnap + n3 * 3 * (ll)c.n + x : nullptr;
    const double gav = tb ? tb[n3] : p.gkw, bav = tb ? tb[2 * n3] : p.bkw;
    const double dl = delpls[so];
    double rhomega = (ob[x + 6 * n3] - ar * rb[x + 6 * n3] * dtt + ar * dtau * rho * ome * (gav * dl + 2 * bav * ome)) /
                     (1 + ar * dtau * (gav * dl + 2 * bav * ome));
    const double rhomegamin = gav * p.cmu * rho * sqrt(prd[so]);
    if (rhomega < rhomegamin) rhomega = rhomegamin;
    qb[x + 6 * n3] = rhomega;
    const double drhomega = rhomega - rho * ome;
    qb[x + 5 * n3] = (ob[x + 5 * n3] - ar * rb[x + 5 * n3] * dtt + ar * dtau * rho * tke * (dl + p.bstrkw * ome) -
                      ar * dtau * p.bstrkw * tke * drhomega) /
                     (1 + ar * dtau * (dl + p.bstrkw * ome));
  }
  for (int v = 0; v < 5; v++) qb[x + v * n3] = ob[x + v * n3] - ar * rb[x + v * n3] * dt;
}

// ---- dual time stepping (one snapshot), interior cells ----------------------
// src_uns_b (src/src.f90:406-450): res += q vol / dt (first step of a
// backward-Euler start, dual_fact 1) or 3 q vol / 2 / dt; then, on the
// finest level, the overall residual res - rhs of src/cosa.f90:1938-1941.
__global__ void __launch_bounds__(256) k_src_uns(Par p, const Blk* __restrict__ blks, const Tile* __restrict__ tiles,
                                                 const double* __restrict__ q, const double* __restrict__ vol,
                                                 double* __restrict__ res, const double* __restrict__ rhs,
                                                 int sub_rhs) {
  const Tile t = tiles[blockIdx.x];
  const Blk b = blks[t.block];
  CellIdx c;
  if (!cell_of(p, b, t.base + threadIdx.x, c)) return;
  const ll n3 = n3of(b), x = i3(b, c.i, c.j, c.k);
  const double* qb = qblk(p, b, q, 0);
  double* rb = qblk(p, b, res, 0);
  const double* hb = qblk(p, b, rhs, 0);
  const double v = vol[b.o1 + i1(b, c.i, c.j, c.k)], dt = p.dual_dt;
  const bool first = p.dual_fact == 1.0;
  for (int m = 0; m < p.npde; m++) {
    const ll y = x + m * n3;
    double r = first ? rb[y] + qb[y] * v / dt : rb[y] + 3 * qb[y] * v / 2 / dt;
    if (sub_rhs) r = r + hb[y] * -1.0;
    rb[y] = r;
  }
}

// q1 = q (copy_array, interior) and rhs_uns_b (src/src.f90:480-525) of the
// finest level: q1 / dt vol, or (2 q1 - q2 / 2) vol / dt.
__global__ void __launch_bounds__(256) k_rhs_uns(Par p, const Blk* __restrict__ blks, const Tile* __restrict__ tiles,
                                                 const double* __restrict__ q, double* __restrict__ q1,
                                                 const double* __restrict__ q2, const double* __restrict__ vol,
                                                 double* __restrict__ rhs) {
  const Tile t = tiles[blockIdx.x];
  const Blk b = blks[t.block];
  CellIdx c;
  if (!cell_of(p, b, t.base + threadIdx.x, c)) return;
  const ll n3 = n3of(b), x = i3(b, c.i, c.j, c.k);
  const double* qb = qblk(p, b, q, 0);
  double* ab = qblk(p, b, q1, 0);
  const double* pb = qblk(p, b, q2, 0);
  double* hb = qblk(p, b, rhs, 0);
  const double v = vol[b.o1 + i1(b, c.i, c.j, c.k)], dt = p.dual_dt;
  const bool first = p.dual_fact == 1.0;
  for (int m = 0; m < p.npde; m++) {
    const ll y = x + m * n3;
    const double a = qb[y];
    ab[y] = a;
    hb[y] = first ? a * v / dt : (2 * a - pb[y] / 2) * v / dt;
  }
}

// resg_brms, rkimp without IRS on one level (src/cosa.f90:4534-4549): the
// summand fact q vol / dt + res - rhs, into work1 (interior)
__global__ void __launch_bounds__(256) k_rms_dual_term(Par p, const Blk* __restrict__ blks,
                                                       const Tile* __restrict__ tiles, const double* __restrict__ q,
                                                       const double* __restrict__ res, const double* __restrict__ rhs,
                                                       const double* __restrict__ vol, double* __restrict__ work1) {
  const Tile t = tiles[blockIdx.x];
  const Blk b = blks[t.block];
  CellIdx c;
  if (!cell_of(p, b, t.base + threadIdx.x, c)) return;
  const ll n3 = n3of(b), x = i3(b, c.i, c.j, c.k);
  const double* qb = qblk(p, b, q, 0);
  const double* rb = qblk(p, b, res, 0);
  const double* hb = qblk(p, b, rhs, 0);
  double* wb = qblk(p, b, work1, 0);
  const double v = vol[b.o1 + i1(b, c.i, c.j, c.k)];
  for (int m = 0; m < p.npde; m++) {
    const ll y = x + m * n3;
    wb[y] = p.dual_fact * qb[y] * v / p.dual_dt + rb[y] - hb[y];
  }
}

// update_b, rkimp branches (src/update.f90:649-762): the physical-time
// term implicit.  dual_plain (no IRS, one level): the residual carries no
// src_uns and the term enters the denominator; otherwise the residual holds
// it and the update adds it back before dividing.  k-omega first, with the
// point-implicit rho*omega, rho*k of the steady update.
__global__ void __launch_bounds__(256) k_update_dual(Par p, const Blk* __restrict__ blks,
                                                     const Tile* __restrict__ tiles, const double* __restrict__ res,
                                                     double* __restrict__ q, const double* __restrict__ qold,
                                                     const double* __restrict__ dtvol, double ar,
                                                     const double* __restrict__ vol,
                                                     const double* __restrict__ dtvolt,
                                                     const double* __restrict__ delpls,
                                                     const double* __restrict__ prd, const double* __restrict__ ttrm) {
  const Tile t = tiles[blockIdx.x];
  const Blk b = blks[t.block];
  CellIdx c;
  if (!cell_of(p, b, t.base + threadIdx.x, c)) return;
  const ll n3 = n3of(b), x = i3(b, c.i, c.j, c.k);
  const ll po = b.o1 + i1(b, c.i, c.j, c.k);
  const double dv = dtvol[po], v = vol[po], fact = p.dual_fact, dt = p.dual_dt;
  double* qb = qblk(p, b, q, 0);
  const double* ob = qblk(p, b, qold, 0);
  const double* rb = qblk(p, b, res, 0);
  if (p.kom) {
    const double rho = qb[x];
    const double ome = qb[x + 6 * n3] / qb[x];
    const double dtt = dtvolt[po];
    const double dtau = dtt * v;
    const ll so = b.o3 + x;
    const double* tb = p.sst ? ttrm + b.o3 * 3 + x : nullptr;
    const double gav = tb ? tb[n3] : p.gkw, bav = tb ? tb[2 * n3] : p.bkw;
    const double dl = delpls[so];
    const double q7 = qb[x + 6 * n3];
    double rhomega;
    if (p.dual_plain)
      rhomega = (ob[x + 6 * n3] - ar * rb[x + 6 * n3] * dtt + ar * dtau * q7 * (gav * dl + 2 * bav * ome)) /
                (1 + ar * dtau * (gav * dl + 2 * bav * ome) + fact * ar * dtau / dt);
    else
      rhomega = (ob[x + 6 * n3] + ar * fact * dtau / dt * q7 - ar * rb[x + 6 * n3] * dtt +
                 ar * dtau * q7 * (gav * dl + 2 * bav * ome)) /
                (1 + ar * dtau * (gav * dl + 2 * bav * ome) + ar * fact * dtau / dt);
    const double rhomegamin = gav * p.cmu * rho * sqrt(prd[so]);
    if (rhomega < rhomegamin) rhomega = rhomegamin;
    qb[x + 6 * n3] = rhomega;
    const double drhomega = rhomega - rho * ome;
    const double q6 = qb[x + 5 * n3];
    if (p.dual_plain)
      qb[x + 5 * n3] = (ob[x + 5 * n3] - ar * rb[x + 5 * n3] * dtt +
                        ar * dtau * q6 * (dl + p.bstrkw * ome - p.bstrkw * drhomega / rho)) /
                       (1 + ar * dtau * (dl + p.bstrkw * ome) + ar * fact * dtau / dt);
    else
      qb[x + 5 * n3] = (ob[x + 5 * n3] + ar * fact * dtau / dt * q6 - ar * rb[x + 5 * n3] * dtt +
                        ar * dtau * q6 * (dl + p.bstrkw * ome - p.bstrkw * drhomega / rho)) /
                       (1 + ar * dtau * (dl + p.bstrkw * ome) + ar * fact * dtau / dt);
  }
  const double den = 1 + fact * ar * dv * v / dt;
  for (int m = 0; m < 5; m++) {
    const ll y = x + m * n3;
    if (p.dual_plain) qb[y] = (ob[y] - ar * rb[y] * dv) / den;
    else qb[y] = (ob[y] - ar * dv * (-(fact * v / dt * qb[y]) + rb[y])) / den;
  }
}

// ---- src_bhb (src/src.f90:556-613): qhb = omega vol D q, res += qhb ----
// One thread per (cell, snapshot ni): the 2N+1 snapshots of a cell are
// read by 2N+1 threads (L2 serves the repeats), so a launch has as many
// threads as the level has cell-snapshots, not cells.  The per-term order
// ((omega * vol) * dhb(ni, nj)) * q(nj), summed over nj from zero, is the
// reference's.  The contraction is a (2N+1)^2 by (2N+1) x 5 product per
// cell: 0.6 flop per byte moved, HBM-bound, so it stays on the vector ALUs.
__global__ void __launch_bounds__(256) k_src_hb(Par p, const Blk* __restrict__ blks, const Tile* __restrict__ tiles,
                                                const double* __restrict__ q, const double* __restrict__ vol,
                                                double* __restrict__ res, double* __restrict__ qhb,
                                                const double* __restrict__ dhb) {
  const Tile t = tiles[blockIdx.x];
  const Blk b = blks[t.block];
  CellIdx c;
  if (!cell_of(p, b, t.base + threadIdx.x, c)) return;
  const ll n3 = n3of(b), x = i3(b, c.i, c.j, c.k);
  const ll stride_n = n3 * p.npde;
  const double* qb = q + b.o3 * p.npde * p.nsnap + x;
  double* rb = res + b.o3 * p.npde * p.nsnap + x;
  double* hb = qhb + b.o3 * p.npde * p.nsnap + x;
  const double vv = vol[b.o1 + i1(b, c.i, c.j, c.k)];
  const int ns = p.nsnap, ni = c.n;
  for (int v = 0; v < p.npde; v++) {
    double acc = 0.0;
    for (int nj = 0; nj < ns; nj++) acc = acc + p.omega * vv * dhb[ni + ns * nj] * qb[stride_n * nj + n3 * v];
    hb[stride_n * ni + n3 * v] = acc;
    rb[stride_n * ni + n3 * v] = rb[stride_n * ni + n3 * v] + acc;
  }
}

// ---- harmonic-balance implicit stage, rkim ----------------------------------
// Per interior cell the 2N+1 snapshots are coupled by the spectral operator:
// (I d + ar omega dtau D) q = rhs, solved with an LU per cell.  The system
// of each thread lives in LDS, thread-interleaved (element e of thread t at
// [e * cb + t]), so one work-group of cb threads holds cb systems.

// DGETRF's elimination (partial pivoting on the first largest |a|, whole
// rows swapped, multipliers times the reciprocal pivot, rank-1 updates in
// column order) with the swaps applied to the right-hand sides as they
// happen, then DGETRS's forward and back substitution by columns
// (src/update.f90:1460-1478 call LAPACK for this).  a: m x m, b: m x nrhs,
// column-major with stride cb, both already offset by the thread.
__device__ void lds_lu_solve(double* a, double* bb, int m, int nrhs, int cb) {
#define A_(r, c) a[((r) + m * (c)) * cb]
#define B_(r, c) bb[((r) + m * (c)) * cb]
  for (int j = 0; j < m; j++) {
    int pr = j;
    double amax = fabs(A_(j, j));
    for (int i = j + 1; i < m; i++) {
      const double v = fabs(A_(i, j));
      if (v > amax) {
        amax = v;
        pr = i;
      }
    }
    if (A_(pr, j) != 0.0) {
      if (pr != j) {
        for (int l = 0; l < m; l++) {
          const double t = A_(j, l);
          A_(j, l) = A_(pr, l);
          A_(pr, l) = t;
        }
        for (int l = 0; l < nrhs; l++) {
          const double t = B_(j, l);
          B_(j, l) = B_(pr, l);
          B_(pr, l) = t;
        }
      }
      const double piv = A_(j, j);
      if (fabs(piv) >= 2.2250738585072014e-308) {
        const double r = 1.0 / piv;
        for (int i = j + 1; i < m; i++) A_(i, j) = A_(i, j) * r;
      } else {
        for (int i = j + 1; i < m; i++) A_(i, j) = A_(i, j) / piv;
      }
    }
    for (int l = j + 1; l < m; l++) {
      const double u = A_(j, l);
      if (u != 0.0)
        for (int i = j + 1; i < m; i++) A_(i, l) = A_(i, l) - A_(i, j) * u;
    }
  }
  for (int c = 0; c < nrhs; c++) {
    for (int k = 0; k < m; k++) {
      const double x = B_(k, c);
      if (x != 0.0)
        for (int i = k + 1; i < m; i++) B_(i, c) = B_(i, c) - x * A_(i, k);
    }
    for (int k = m - 1; k >= 0; k--)
      if (B_(k, c) != 0.0) {
        const double x = B_(k, c) / A_(k, k);
        B_(k, c) = x;
        for (int i = 0; i < k; i++) B_(i, c) = B_(i, c) - x * A_(i, k);
      }
  }
}

// cell (i, j, k) of work item `item` of the single-snapshot items of a block
__device__ __forceinline__ bool cell_item(const Blk& b, int item, int& i, int& j, int& k) {
  const int im = b.imax - 1, jm = b.jmax - 1, km = b.kmax - 1;
  if (item >= im * jm * km) return false;
  i = item % im + 1;
  const int rem = item / im;
  j = rem % jm + 1;
  k = rem / jm + 1;
  return true;
}

// update_bhb, rkim branch (src/update.f90:1410-1841): rho*omega (floored at
// rhomegamin), rho*k, then the mean flow with 5 right-hand sides.  plain:
// no IRS and one grid level, the right-hand side is the explicit stage;
// otherwise the HB term qhb0 kept by src_hb is taken out of the smoothed
// residual.  qhb0 is overwritten with the right-hand sides, as in the
// reference.  Work-group: cb threads, tiles of TS single-snapshot items split
// in `sub` = TS / cb parts.
__global__ void __launch_bounds__(64) k_update_rkim(Par p, const Blk* __restrict__ blks, const Tile* __restrict__ tiles,
                                                    int sub, int cb, const double* __restrict__ res,
                                                    double* __restrict__ q, const double* __restrict__ qold,
                                                    const double* __restrict__ dtvol, double ar,
                                                    const double* __restrict__ vol, const double* __restrict__ dtvolt,
                                                    const double* __restrict__ delpls, const double* __restrict__ prd,
                                                    double* __restrict__ qhb0, const double* __restrict__ dhb,
                                                    int plain, const double* __restrict__ ttrm) {
  extern __shared__ double lds[];
  const Tile t = tiles[blockIdx.x / sub];
  const Blk b = blks[t.block];
  int i, j, k;
  if (!cell_item(b, t.base + (blockIdx.x % sub) * cb + threadIdx.x, i, j, k)) return;
  const int m = p.nsnap;
  double* a = lds + threadIdx.x;
  double* bb = lds + (ll)m * m * cb + threadIdx.x;
  const ll n3 = n3of(b), n1 = n1of(b), x = i3(b, i, j, k), x1 = i1(b, i, j, k);
  const ll sq = n3 * p.npde;
  double* qb = q + b.o3 * p.npde * p.nsnap + x;
  const double* ob = qold + b.o3 * p.npde * p.nsnap + x;
  const double* rb = res + b.o3 * p.npde * p.nsnap + x;
  double* hb = qhb0 + b.o3 * p.npde * p.nsnap + x;
  const double* dtb = dtvol + b.o1 * p.nsnap + x1;
  const double vv = vol[b.o1 + x1];
  if (p.kom) {
    const double* dttb = dtvolt + b.o1 * p.nsnap + x1;
    const double* dlb = delpls + b.o3 * p.nsnap + x;
    const double* pb = prd + b.o3 * p.nsnap + x;
    // ttrm(2:3) per snapshot: gkw, bkw (Wilcox) or blended (BSL/SST)
    const double* tb = p.sst ? ttrm + b.o3 * 3 * p.nsnap + x : nullptr;
#define GAV(n) (tb ? tb[n3 * (3 * (ll)(n) + 1)] : p.gkw)
#define BAV(n) (tb ? tb[n3 * (3 * (ll)(n) + 2)] : p.bkw)
    // rho*omega: rhs in column 0, the old rho*omega in column 1
    for (int ni = 0; ni < m; ni++) {
      const double gav = GAV(ni), bav = BAV(ni);
      const double ome = qb[sq * ni + 6 * n3] / qb[sq * ni];
      const double dtt = dttb[n1 * ni];
      const double dtau = dtt * vv;
      const double dps = dlb[n3 * ni];
      const double q7 = qb[sq * ni + 6 * n3];
      const double rhs = plain ? ob[sq * ni + 6 * n3] - ar * rb[sq * ni + 6 * n3] * dtt +
                                     ar * dtau * q7 * (gav * dps + 2 * bav * ome)
                               : ob[sq * ni + 6 * n3] + ar * dtt * (hb[sq * ni + 6 * n3] - rb[sq * ni + 6 * n3]) +
                                     ar * dtau * q7 * (gav * dps + 2 * bav * ome);
      hb[sq * ni + 6 * n3] = rhs;
      B_(ni, 0) = rhs;
      B_(ni, 1) = q7;
      const double dg = 1 + ar * dtau * (gav * dps + 2 * bav * ome);
      for (int nj = 0; nj < m; nj++) A_(ni, nj) = (ni == nj ? 1.0 : 0.0) * dg + ar * p.omega * dtau * dhb[ni + m * nj];
    }
    lds_lu_solve(a, bb, m, 1, cb);
    for (int ni = 0; ni < m; ni++) {
      const double rho = qb[sq * ni];
      const double rhomegamin = GAV(ni) * p.cmu * rho * sqrt(pb[n3 * ni]);
      double w = B_(ni, 0);
      if (w < rhomegamin) w = rhomegamin;
      qb[sq * ni + 6 * n3] = w;
      B_(ni, 2) = w - B_(ni, 1);   // d(rho*omega)
    }
    // rho*k
    for (int ni = 0; ni < m; ni++) {
      const double rho = qb[sq * ni];
      const double ome = B_(ni, 1) / qb[sq * ni];
      const double dtt = dttb[n1 * ni];
      const double dtau = dtt * vv;
      const double dps = dlb[n3 * ni];
      const double drho = B_(ni, 2);
      const double q6 = qb[sq * ni + 5 * n3];
      const double rhs =
          plain ? (ob[sq * ni + 5 * n3] - ar * rb[sq * ni + 5 * n3] * dtt +
                   ar * dtau * q6 * (dps + p.bstrkw * ome - p.bstrkw * drho / rho))
                : (ob[sq * ni + 5 * n3] + ar * dtt * (hb[sq * ni + 5 * n3] - rb[sq * ni + 5 * n3]) +
                   ar * dtau * q6 * (dps + p.bstrkw * ome - p.bstrkw * drho / rho));
      hb[sq * ni + 5 * n3] = rhs;
      B_(ni, 0) = rhs;
      const double dg = 1 + ar * dtau * (dps + p.bstrkw * ome);
      for (int nj = 0; nj < m; nj++) A_(ni, nj) = (ni == nj ? 1.0 : 0.0) * dg + ar * p.omega * dtau * dhb[ni + m * nj];
    }
    lds_lu_solve(a, bb, m, 1, cb);
    for (int ni = 0; ni < m; ni++) qb[sq * ni + 5 * n3] = B_(ni, 0);
  }
  // mean flow
  for (int ni = 0; ni < m; ni++) {
    const double dt = dtb[n1 * ni];
    for (int v = 0; v < 5; v++) {
      const ll o = sq * ni + n3 * v;
      const double rhs = plain ? ob[o] - ar * rb[o] * dt : ob[o] + (hb[o] - rb[o]) * ar * dt;
      hb[o] = rhs;
      B_(ni, v) = rhs;
    }
    const double dtau = dt * vv;
    for (int nj = 0; nj < m; nj++) A_(ni, nj) = (ni == nj ? 1.0 : 0.0) + ar * p.omega * dtau * dhb[ni + m * nj];
  }
  lds_lu_solve(a, bb, m, 5, cb);
  for (int ni = 0; ni < m; ni++)
    for (int v = 0; v < 5; v++) qb[sq * ni + n3 * v] = B_(ni, v);
}

// lim_res_ome_bhb, rkim branch (src/update.f90:2925-2990): the omega
// residual limited against the coupled solve of res7 * dtvolt.
__global__ void __launch_bounds__(64) k_lim_res_ome_rkim(Par p, const Blk* __restrict__ blks,
                                                         const Tile* __restrict__ tiles, int sub, int cb,
                                                         double* __restrict__ res, const double* __restrict__ q,
                                                         const double* __restrict__ qold,
                                                         const double* __restrict__ vol,
                                                         const double* __restrict__ dtvolt,
                                                         const double* __restrict__ delpls,
                                                         const double* __restrict__ prd,
                                                         const double* __restrict__ dhb,
                                                         const double* __restrict__ ttrm) {
  extern __shared__ double lds[];
  const Tile t = tiles[blockIdx.x / sub];
  const Blk b = blks[t.block];
  int i, j, k;
  if (!cell_item(b, t.base + (blockIdx.x % sub) * cb + threadIdx.x, i, j, k)) return;
  const int m = p.nsnap;
  double* a = lds + threadIdx.x;
  double* bb = lds + (ll)m * m * cb + threadIdx.x;
  const ll n3 = n3of(b), n1 = n1of(b), x = i3(b, i, j, k), x1 = i1(b, i, j, k);
  const ll sq = n3 * p.npde;
  const double* qb = q + b.o3 * p.npde * p.nsnap + x;
  const double* ob = qold + b.o3 * p.npde * p.nsnap + x;
  double* rb = res + b.o3 * p.npde * p.nsnap + x;
  const double* dttb = dtvolt + b.o1 * p.nsnap + x1;
  const double* dlb = delpls + b.o3 * p.nsnap + x;
  const double* pb = prd + b.o3 * p.nsnap + x;
  const double vv = vol[b.o1 + x1];
  const double* tb = p.sst ? ttrm + b.o3 * 3 * p.nsnap + x : nullptr;
  for (int ni = 0; ni < m; ni++) {
    const double gav = GAV(ni), bav = BAV(ni);
    const double dtau = dttb[n1 * ni] * vv;
    const double ome = qb[sq * ni + 6 * n3] / qb[sq * ni];
    const double dg = 1 + dtau * (gav * dlb[n3 * ni] + 2 * bav * ome);
    for (int nj = 0; nj < m; nj++) A_(ni, nj) = (ni == nj ? 1.0 : 0.0) * dg + p.omega * dtau * dhb[ni + m * nj];
    B_(ni, 0) = rb[sq * ni + 6 * n3] * dttb[n1 * ni];
  }
  lds_lu_solve(a, bb, m, 1, cb);
  for (int n = 0; n < m; n++) {
    const double rho = qb[sq * n];
    const double rhomegamin = GAV(n) * p.cmu * rho * sqrt(pb[n3 * n]);
    const double rhomeganew = ob[sq * n + 6 * n3] - B_(n, 0);
    rb[sq * n + 6 * n3] = rb[sq * n + 6 * n3] - fmax(0.0, (rhomegamin - rhomeganew)) / dttb[n1 * n];
  }
#undef A_
#undef B_
#undef GAV
#undef BAV
}

// resg_brms_hb with rkim, no IRS, one level (src/cosa.f90:4657-4674): the
// HB term added to the copy of the residual, snapshot by snapshot in order.
__global__ void __launch_bounds__(256) k_rms_hb_term(Par p, const Blk* __restrict__ blks,
                                                     const Tile* __restrict__ tiles, double* __restrict__ work1,
                                                     const double* __restrict__ q, const double* __restrict__ vol,
                                                     const double* __restrict__ dhb) {
  const Tile t = tiles[blockIdx.x];
  const Blk b = blks[t.block];
  int i, j, k;
  if (!cell_item(b, t.base + threadIdx.x, i, j, k)) return;
  const int m = p.nsnap;
  const ll n3 = n3of(b), x = i3(b, i, j, k), sq = n3 * p.npde;
  const double* qb = q + b.o3 * p.npde * p.nsnap + x;
  double* wb = work1 + b.o3 * p.npde * p.nsnap + x;
  const double vv = vol[b.o1 + i1(b, i, j, k)];
  for (int ni = 0; ni < m; ni++)
    for (int v = 0; v < p.npde; v++) {
      double acc = wb[sq * ni + n3 * v];
      for (int nj = 0; nj < m; nj++) acc = acc + p.omega * dhb[ni + m * nj] * qb[sq * nj + n3 * v] * vv;
      wb[sq * ni + n3 * v] = acc;
    }
}

// ---- resg_brms (src/cosa.f90:4503-4582): per-tile partial sums ----
__global__ void __launch_bounds__(256) k_rms_part(Par p, const Blk* __restrict__ blks, const Tile* __restrict__ tiles,
                                                  const double* __restrict__ res, const double* __restrict__ rhs,
                                                  int use_rhs, double* __restrict__ part) {
  __shared__ double sh[7][256];
  const Tile t = tiles[blockIdx.x];
  const Blk b = blks[t.block];
  CellIdx c;
  double s[7] = {0, 0, 0, 0, 0, 0, 0};
  const int nv = p.npde;
  if (cell_of(p, b, t.base + threadIdx.x, c)) {
    const ll n3 = n3of(b), x = i3(b, c.i, c.j, c.k);
    const double* rb = qblk(p, b, res, c.n);
    const double* hb = use_rhs ? qblk(p, b, rhs, c.n) : nullptr;
    for (int v = 0; v < nv; v++) {
      double d = use_rhs ? rb[x + v * n3] - hb[x + v * n3] : rb[x + v * n3];
      s[v] = d * d;
    }
  }
  for (int v = 0; v < nv; v++) sh[v][threadIdx.x] = s[v];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int v = 0; v < nv; v++) sh[v][threadIdx.x] += sh[v][threadIdx.x + w];
    __syncthreads();
  }
  if ((int)threadIdx.x < nv) part[(ll)blockIdx.x * 8 + threadIdx.x] = sh[threadIdx.x][0];
}
// fixed-order final reduction: one workgroup, strided partial sums then tree
__global__ void __launch_bounds__(256) k_rms_final(const double* __restrict__ part, int nt, int nv,
                                                   double* __restrict__ out) {
  __shared__ double sh[7][256];
  double s[7] = {0, 0, 0, 0, 0, 0, 0};
  // each thread sums partials t, t + 256, ... in that order; the loads of
  // eight of them are issued before their additions
  int t = threadIdx.x;
  for (; t + 7 * 256 < nt; t += 8 * 256) {
    double pv[8][7];
#pragma unroll
    for (int u = 0; u < 8; u++)
#pragma unroll
      for (int v = 0; v < 7; v++) pv[u][v] = v < nv ? part[(ll)(t + u * 256) * 8 + v] : 0.0;
#pragma unroll
    for (int u = 0; u < 8; u++)
#pragma unroll
      for (int v = 0; v < 7; v++)
        if (v < nv) s[v] += pv[u][v];
  }
  for (; t < nt; t += 256)
    for (int v = 0; v < nv; v++) s[v] += part[(ll)t * 8 + v];
  for (int v = 0; v < nv; v++) sh[v][threadIdx.x] = s[v];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int v = 0; v < nv; v++) sh[v][threadIdx.x] += sh[v][threadIdx.x + w];
    __syncthreads();
  }
  if ((int)threadIdx.x < nv) out[threadIdx.x] = sh[threadIdx.x][0];
}

// ---- bprtest (src/cosa.f90:960-1120): first offender in loop order ----
__global__ void __launch_bounds__(256) k_prtest(Par p, const Blk* __restrict__ blks, const Tile* __restrict__ tiles,
                                                const double* __restrict__ q, int stage_seq,
                                                unsigned long long* __restrict__ flag) {
  const Tile t = tiles[blockIdx.x];
  const Blk b = blks[t.block];
  int n, reg, loc, i, j, k;
  if (!halo_of(p, b, t.base + threadIdx.x, n, reg, loc, i, j, k)) return;
  const double* qb = qblk(p, b, q, n);
  const ll n3 = n3of(b), x = i3(b, i, j, k);
  const double rho = qb[x], pp = qb[x + 4 * n3];
  unsigned long long kind = 0;
  if (pp <= p.qmin5) kind = 1;
  else if (rho <= p.qmin1) kind = 2;
  if (!kind) return;
  const unsigned long long seq = stage_seq < 511 ? (unsigned long long)stage_seq : 511ull;
  const unsigned long long pos = ((unsigned long long)t.block * p.nsnap + n) * 5 + reg;
  const unsigned long long key = (seq << 55) | (pos << 29) | ((unsigned long long)(loc & 0x7FFFFFF) << 2) | kind;
  atomicMin(flag, key);
}

// ---- boundary conditions (src/bc.f90) ----
__device__ __forceinline__ void bc_ijk(int ic1, int ic2, int ic3, int a, int i2_, int i3_, int& i, int& j, int& k) {
  i = (ic1 == 1 ? a : 0) + (ic2 == 1 ? i2_ : 0) + (ic3 == 1 ? i3_ : 0);
  j = (ic1 == 2 ? a : 0) + (ic2 == 2 ? i2_ : 0) + (ic3 == 2 ? i3_ : 0);
  k = (ic1 == 3 ? a : 0) + (ic2 == 3 ? i2_ : 0) + (ic3 == 3 ? i3_ : 0);
}
__device__ __forceinline__ int cycd(int d, int m) { return (d - 1 + m - 1) % 3 + 1; }

__global__ void __launch_bounds__(256) k_bc(Par p, const Blk* __restrict__ blks, const PatchRec* __restrict__ pr,
                                            const Tile* __restrict__ tiles, double* __restrict__ q,
                                            const double* __restrict__ si, const double* __restrict__ sj,
                                            const double* __restrict__ sk, const double* __restrict__ dist,
                                            const double* __restrict__ qfb) {
  const Tile tl = tiles[blockIdx.x];
  const PatchRec rec = pr[tl.block];
  const int item = tl.base + threadIdx.x;
  if (item >= rec.count * p.nsnap) return;
  const Blk b = blks[rec.block];
  const int n = item / rec.count;
  const int loc = item - n * rec.count;
  const int idir = rec.t[1], inrout = rec.t[2];
  const int ic1 = cycd(idir, 1), ic2 = cycd(idir, 2), ic3 = cycd(idir, 3);
  // every face point is set from the interior alone, so the order of the
  // points over the threads is free: i varies fastest wherever it is a
  // transverse index (j faces cycle to (k, i), whose k-fastest order would
  // put consecutive lanes a plane apart)
  int i2_, i3_;
  if (ic3 == 1) {
    const int n3e = rec.count / rec.n2;
    i3_ = rec.t[2 * ic3 + 1] + loc % n3e;
    i2_ = rec.t[2 * ic2 + 1] + loc / n3e;
  } else {
    i2_ = rec.t[2 * ic2 + 1] + loc % rec.n2;
    i3_ = rec.t[2 * ic3 + 1] + loc / rec.n2;
  }
  const int ijk[4] = {0, b.imax, b.jmax, b.kmax};
  int ibcpt, ibcpt2, ibcn, ibcn2, ibcm;
  double sgnm;
  if (inrout == 1) {
    ibcpt = 0; ibcpt2 = -1; ibcn = 1; ibcn2 = 2; ibcm = 1; sgnm = -1.0;
  } else {
    ibcpt = ijk[idir]; ibcpt2 = ijk[idir] + 1; ibcn = ijk[idir] - 1; ibcn2 = ijk[idir] - 2; ibcm = ijk[idir];
    sgnm = 1.0;
  }
  int ib, jb, kb, ib2, jb2, kb2, in, jn, kn;
  bc_ijk(ic1, ic2, ic3, ibcpt, i2_, i3_, ib, jb, kb);
  bc_ijk(ic1, ic2, ic3, ibcpt2, i2_, i3_, ib2, jb2, kb2);
  bc_ijk(ic1, ic2, ic3, ibcn, i2_, i3_, in, jn, kn);
  double* qb = qblk(p, b, q, n);
  const ll n3 = n3of(b);
  const ll xb = i3(b, ib, jb, kb), xb2 = i3(b, ib2, jb2, kb2), xn = i3(b, in, jn, kn);
  const int type = rec.type;
  if (type == 80) {   // bc_sym (src/bc.f90:2595-2710)
    int in2, jn2, kn2;
    bc_ijk(ic1, ic2, ic3, ibcn2, i2_, i3_, in2, jn2, kn2);
    const ll xn2 = i3(b, in2, jn2, kn2);
    for (int v = 1; v <= p.npde; v++) {
      const double sg = (v == idir + 1) ? -1.0 : 1.0;
      qb[xb + (v - 1) * n3] = sg * qb[xn + (v - 1) * n3];
      qb[xb2 + (v - 1) * n3] = sg * qb[xn2 + (v - 1) * n3];
    }
    return;
  }
  if (type == 4) {    // bc_extrapol (src/bc.f90:2494-2592)
    for (int v = 0; v < p.npde; v++) {
      const double val = qb[xn + v * n3];
      qb[xb + v * n3] = val;
      qb[xb2 + v * n3] = val;
    }
    return;
  }
  int im_, jm_, km_;
  bc_ijk(ic1, ic2, ic3, ibcm, i2_, i3_, im_, jm_, km_);
  const double* s = idir == 1 ? si : (idir == 2 ? sj : sk);
  const int nh = n * p.hbmove;
  const double* sm = s + b.o2 * p.lmet * p.nsnap_met + ((ll)nh * n2of(b) + i2(b, im_, jm_, km_)) * p.lmet;
  double hv[5];
  bool set = false;
  if (type / 100 == 15) {   // bc_walvis_15 (src/bc.f90:2144-2492)
    int in2, jn2, kn2;
    bc_ijk(ic1, ic2, ic3, ibcn2, i2_, i3_, in2, jn2, kn2);
    const ll xn2 = i3(b, in2, jn2, kn2);
    double uw = 0.0, vw = 0.0, ww = 0.0;
    if (p.moving) {
      uw = sm[4];
      vw = sm[5];
      ww = sm[6];
    }
    const double p1 = qb[xn + 4 * n3], p2 = qb[xn2 + 4 * n3];
    const double pw = 0.5 * (3 * p1 - p2);
    const double rhow = pow(pw / p1, 1 / p.gamma) * qb[xn];
    hv[0] = rhow;
    hv[1] = uw;
    hv[2] = vw;
    hv[3] = ww;
    hv[4] = pw;
    for (int v = 0; v < 5; v++) {
      qb[xb + v * n3] = hv[v];
      qb[xb2 + v * n3] = hv[v];
    }
    if (p.kom) {
      const double tw = pw * p.gamma / rhow;
      const double muw = (p.stemp + 1) / (tw + p.stemp) * pow(tw, 1.5);
      const double dn = dist[b.o1 + i1(b, in, jn, kn)];
      double omew;
      if (p.wallwilc) {
        const double mor = p.machfs / p.reyno;
        const double nx = sm[0], ny = sm[1], nz = sm[2];
        const double uc = qb[xn + n3] - uw, vc = qb[xn + 2 * n3] - vw, wc = qb[xn + 3 * n3] - ww;
        const double uorth = nx * uc + ny * vc + nz * wc;
        const double a1 = uc - uorth * nx, a2 = vc - uorth * ny, a3 = wc - uorth * nz;
        const double upar = sqrt(a1 * a1 + a2 * a2 + a3 * a3);
        const double dupardn = upar / (dn + 1.e-12);
        const double tauw = muw * fabs(dupardn);
        double rkrpls = 1.0;
        if (p.roughk > 0.0) rkrpls = fmax(1.0, sqrt(rhow * tauw / mor) * p.roughk / muw);
        const double sr = rkrpls >= 25.0 ? 100.0 / rkrpls : 2500.0 / (rkrpls * rkrpls);
        omew = fabs(dupardn) * sr;
      } else {
        omew = 60.0 * muw / (p.bkw * rhow * (dn * dn) + 1.e-24) * p.machfs / p.reyno;
      }
      qb[xb + 5 * n3] = 0.0;
      qb[xb + 6 * n3] = omew;
      qb[xb2 + 5 * n3] = 0.0;
      qb[xb2 + 6 * n3] = omew;
    }
    return;
  }
  if (type == 30) {   // bc_far0 (src/bc.f90:124-403)
    // freestream of the patch's block (blocks may be independent solves)
    const double* qfst = qfb + 5 * rec.block;
    const double gamma = p.gamma;
    const double sinf = p.pinf / pow(p.rhoinf, gamma);
    double sx = sm[0] * sgnm, sy = sm[1] * sgnm, sz = sm[2] * sgnm;
    double qext = qb[xn + n3] * sx + qb[xn + 2 * n3] * sy + qb[xn + 3 * n3] * sz;
    double qinf = qfst[1] * sx + qfst[2] * sy + qfst[3] * sz;
    double rhoext = qb[xn];
    double pext = qb[xn + 4 * n3];
    double aext = sqrt(gamma * pext / rhoext);
    double rext = qext + 2 * aext / (gamma - 1);
    double rinf = qinf - 2 * p.ainf / (gamma - 1);
    double qdotn = (rext + rinf) / 2;
    double a = (gamma - 1) * (rext - rinf) / 4;
    double mach = qdotn / a;
    double sext = pext / pow(rhoext, gamma);
    if (qdotn <= 0.0) {
      if (fabs(mach) >= 1.0) {
        for (int v = 0; v < 5; v++) hv[v] = qfst[v];
        set = true;
      } else if (fabs(mach) < 1.0) {
        double rhobc = pow(a * a / (gamma * sinf), 1.0 / (gamma - 1));
        double pb = rhobc * a * a / gamma;
        hv[0] = rhobc;
        hv[1] = qfst[1] + (qdotn - qinf) * sx;
        hv[2] = qfst[2] + (qdotn - qinf) * sy;
        hv[3] = qfst[3] + (qdotn - qinf) * sz;
        hv[4] = pb;
        set = true;
      }
    } else if (qdotn > 0.0) {
      if (fabs(mach) >= 1.0) {
        for (int v = 0; v < 5; v++) hv[v] = qb[xn + v * n3];
        set = true;
      } else if (fabs(mach) < 1.0) {
        double rhobc = pow(a * a / (gamma * sext), 1.0 / (gamma - 1));
        double pb = rhobc * a * a / gamma;
        hv[0] = rhobc;
        hv[1] = qb[xn + n3] + (qdotn - qext) * sx;
        hv[2] = qb[xn + 2 * n3] + (qdotn - qext) * sy;
        hv[3] = qb[xn + 3 * n3] + (qdotn - qext) * sz;
        hv[4] = pb;
        set = true;
      }
    }
    if (p.kom) {   // src/bc.f90:385-479: k, omega fixed on inflow, extrapolated on outflow
      const double t6 = qdotn <= 0.0 ? p.tkefar : qb[xn + 5 * n3];
      const double t7 = qdotn <= 0.0 ? p.omefar : qb[xn + 6 * n3];
      if (qdotn <= 0.0 || qdotn > 0.0) {
        qb[xb + 5 * n3] = t6;
        qb[xb + 6 * n3] = t7;
        qb[xb2 + 5 * n3] = t6;
        qb[xb2 + 6 * n3] = t7;
      }
    }
  } else if (type == 14) {   // bc_walinv_14 (src/bc.f90:1457-1640)
    int in2, jn2, kn2;
    bc_ijk(ic1, ic2, ic3, ibcn2, i2_, i3_, in2, jn2, kn2);
    const ll xn2 = i3(b, in2, jn2, kn2);
    double nx = sm[0], ny = sm[1], nz = sm[2];
    double rho1 = qb[xn], u1 = qb[xn + n3], v1 = qb[xn + 2 * n3], w1 = qb[xn + 3 * n3], p1 = qb[xn + 4 * n3];
    double rho2 = qb[xn2], u2 = qb[xn2 + n3], v2 = qb[xn2 + 2 * n3], w2 = qb[xn2 + 3 * n3], p2 = qb[xn2 + 4 * n3];
    double rhow = (3 * rho1 - rho2) / 2;
    double uw = (3 * u1 - u2) / 2;
    double vw = (3 * v1 - v2) / 2;
    double ww = (3 * w1 - w2) / 2;
    double pw = (3 * p1 - p2) / 2;
    double un1 = u1 * nx + v1 * ny + w1 * nz;
    double un2 = u2 * nx + v2 * ny + w2 * nz;
    double unw = (3 * un1 - un2) / 2;
    if (p.moving) unw = unw - (sm[4] * nx + sm[5] * ny + sm[6] * nz);
    hv[0] = rhow;
    hv[1] = uw - unw * nx;
    hv[2] = vw - unw * ny;
    hv[3] = ww - unw * nz;
    hv[4] = pw;
    set = true;
    if (p.kom)   // src/bc.f90:1634-1670: zeroth-order extrapolation of k, omega
      for (int v = 5; v <= 6; v++) {
        qb[xb + v * n3] = qb[xn + v * n3];
        qb[xb2 + v * n3] = qb[xn + v * n3];
      }
  }
  if (set)
    for (int v = 0; v < 5; v++) {
      qb[xb + v * n3] = hv[v];
      qb[xb2 + v * n3] = hv[v];
    }
}

// ---- cut_q, same-device copy (src/bc.f90:1894-2141) ----
__global__ void __launch_bounds__(256) k_cut(Par p, const Blk* __restrict__ blks, const CutRec* __restrict__ cr,
                                             int first, double* __restrict__ q) {
  const CutRec rec = cr[first + blockIdx.y];
  const int item = blockIdx.x * blockDim.x + threadIdx.x;
  if (item >= rec.count) return;
  const Blk b1 = blks[rec.b1], b2 = blks[rec.b2];
  const int* t = rec.t;
  const int idir1 = t[1], inout1 = t[2], idir2 = t[10], inout2 = t[11];
  int istr1[4], iend1[4], istr2[4], iend2[4], iord[4];
  istr1[1] = t[3]; iend1[1] = t[4]; istr1[2] = t[5]; iend1[2] = t[6]; istr1[3] = t[7]; iend1[3] = t[8];
  istr2[1] = t[12]; iend2[1] = t[13]; istr2[2] = t[14]; iend2[2] = t[15]; istr2[3] = t[16]; iend2[3] = t[17];
  iord[1] = t[18]; iord[2] = t[19]; iord[3] = t[20];
  const int ijk1[4] = {0, b1.imax, b1.jmax, b1.kmax}, ijk2[4] = {0, b2.imax, b2.jmax, b2.kmax};
  int ibcpt, ibcinc, inr, inrinc;
  if (inout1 == 1) { ibcpt = 0; ibcinc = -1; } else { ibcpt = ijk1[idir1]; ibcinc = 1; }
  if (inout2 == 1) { inr = 1; inrinc = 1; } else { inr = ijk2[idir2] - 1; inrinc = -1; }
  const int ic1 = cycd(idir1, 1), ic2 = cycd(idir1, 2), ic3 = cycd(idir1, 3);
  const int jc1 = iord[ic1], jc2 = iord[ic2], jc3 = iord[ic3];
  const int i2_ = item % rec.n2, i3_ = item / rec.n2;
  const int sg1c2 = iend1[ic2] > istr1[ic2] ? 1 : -1, sg1c3 = iend1[ic3] > istr1[ic3] ? 1 : -1;
  const int sg2c2 = iend2[jc2] > istr2[jc2] ? 1 : -1, sg2c3 = iend2[jc3] > istr2[jc3] ? 1 : -1;
  int a2 = istr1[ic2] + sg1c2 * i2_, a3 = istr1[ic3] + sg1c3 * i3_;
  int ii1, jj1, kk1;
  bc_ijk(ic1, ic2, ic3, ibcpt, a2, a3, ii1, jj1, kk1);
  const int ii2 = ii1 + ibcinc * (ic1 == 1), jj2 = jj1 + ibcinc * (ic1 == 2), kk2 = kk1 + ibcinc * (ic1 == 3);
  int c2 = istr2[jc2] + sg2c2 * i2_, c3 = istr2[jc3] + sg2c3 * i3_;
  int in1, jn1, kn1;
  bc_ijk(jc1, jc2, jc3, inr, c2, c3, in1, jn1, kn1);
  const int in2 = in1 + inrinc * (jc1 == 1), jn2 = jn1 + inrinc * (jc1 == 2), kn2 = kn1 + inrinc * (jc1 == 3);
  const ll n31 = n3of(b1), n32 = n3of(b2);
  const ll d1 = i3(b1, ii1, jj1, kk1), d2 = i3(b1, ii2, jj2, kk2);
  const ll s1 = i3(b2, in1, jn1, kn1), s2 = i3(b2, in2, jn2, kn2);
  for (int n = 0; n < p.nsnap; n++) {
    double* q1 = qblk(p, b1, q, n);
    const double* q2 = qblk(p, b2, q, n);
    for (int v = 0; v < p.npde; v++) {
      q1[d1 + v * n31] = q2[s1 + v * n32];
      q1[d2 + v * n31] = q2[s2 + v * n32];
    }
  }
}

// ---- q_bedges / q_bcorns (src/flux.f90:3914-4263, 3156-3471) ----
__global__ void __launch_bounds__(256) k_edges(Par p, const Blk* __restrict__ blks, double* __restrict__ q,
                                               float* __restrict__ q32) {
  const Blk b = blks[blockIdx.y];
  const int im = b.imax - 1, jm = b.jmax - 1, km = b.kmax - 1;
  const int per = 4 * (km + jm + im);
  const int item = blockIdx.x * blockDim.x + threadIdx.x;
  if (item >= per * p.nsnap) return;
  const int n = item / per;
  int x = item - n * per;
  int grp, e, s;
  if (x < 4 * km) { grp = 0; e = x / km; s = x % km + 1; }
  else if (x < 4 * (km + jm)) { x -= 4 * km; grp = 1; e = x / jm; s = x % jm + 1; }
  else { x -= 4 * (km + jm); grp = 2; e = x / im; s = x % im + 1; }
  // first / second coordinate extents of the group's plane
  const int A = grp == 2 ? b.jmax : b.imax;
  const int B = grp == 0 ? b.jmax : b.kmax;
  int xs[4], ys[4];
  const int ex[4][4] = {{0, 0, 1, 1}, {A, A, A - 1, A - 1}, {A, A, A - 1, A - 1}, {0, 0, 1, 1}};
  const int ey[4][4] = {{0, 1, 0, 1}, {0, 1, 0, 1}, {B, B - 1, B, B - 1}, {B, B - 1, B, B - 1}};
  for (int m = 0; m < 4; m++) { xs[m] = ex[e][m]; ys[m] = ey[e][m]; }
  double* qb = qblk(p, b, q, n);
  const ll n3 = n3of(b);
  ll idx[4];
  for (int m = 0; m < 4; m++)
    idx[m] = grp == 0 ? i3(b, xs[m], ys[m], s) : (grp == 1 ? i3(b, xs[m], s, ys[m]) : i3(b, s, xs[m], ys[m]));
  for (int v = 0; v < p.npde; v++) {
    const double q1 = qb[idx[1] + v * n3], q2 = qb[idx[2] + v * n3], q3 = qb[idx[3] + v * n3];
    double bb = ((q3 - q1) * (xs[2] - xs[1]) - (q2 - q1) * (xs[3] - xs[1])) /
                ((ys[3] - ys[1]) * (xs[2] - xs[1]) - (ys[2] - ys[1]) * (xs[3] - xs[1]));
    double aa = (q2 - q1 - bb * (ys[2] - ys[1])) / (xs[2] - xs[1]);
    double cc = q1 - aa * xs[1] - bb * ys[1];
    const double e = aa * xs[0] + bb * ys[0] + cc;
    qb[idx[0] + v * n3] = e;
    if (q32) q32[(qb - q) + idx[0] + v * n3] = (float)e;
  }
}

__global__ void __launch_bounds__(64) k_corners(Par p, const Blk* __restrict__ blks, double* __restrict__ q) {
  const Blk b = blks[blockIdx.y];
  const int item = blockIdx.x * blockDim.x + threadIdx.x;
  if (item >= 8 * p.nsnap) return;
  const int n = item / 8, ic = item % 8;
  const int I = b.imax, J = b.jmax, K = b.kmax;
  const int pc[8][5][3] = {
      {{0, 0, 0}, {0, 1, 1}, {1, 1, 0}, {1, 0, 1}, {1, 1, 1}},
      {{I, 0, 0}, {I, 1, 1}, {I - 1, 1, 0}, {I - 1, 0, 1}, {I - 1, 1, 1}},
      {{I, J, 0}, {I, J - 1, 1}, {I - 1, J - 1, 0}, {I - 1, J, 1}, {I - 1, J - 1, 1}},
      {{0, J, 0}, {0, J - 1, 1}, {1, J - 1, 0}, {1, J, 1}, {1, J - 1, 1}},
      {{0, 0, K}, {0, 1, K - 1}, {1, 1, K}, {1, 0, K - 1}, {1, 1, K - 1}},
      {{I, 0, K}, {I, 1, K - 1}, {I - 1, 1, K}, {I - 1, 0, K - 1}, {I - 1, 1, K - 1}},
      {{I, J, K}, {I, J - 1, K - 1}, {I - 1, J - 1, K}, {I - 1, J, K - 1}, {I - 1, J - 1, K - 1}},
      {{0, J, K}, {0, J - 1, K - 1}, {1, J - 1, K}, {1, J, K - 1}, {1, J - 1, K - 1}}};
  int X[5][4];
  for (int m = 0; m < 5; m++)
    for (int d = 0; d < 3; d++) X[m][d + 1] = pc[ic][m][d];
  const int c1 = ((X[3][3] - X[1][3]) * (X[4][2] - X[1][2]) - (X[4][3] - X[1][3]) * (X[3][2] - X[1][2])) /
                 (X[2][3] - X[1][3]);
  const int c2 = ((X[2][3] - X[1][3]) * (X[1][2] - X[4][2]) + (X[4][3] - X[1][3]) * (X[2][2] - X[1][2])) /
                 (X[2][3] - X[1][3]);
  const int c3 = ((X[2][3] - X[1][3]) * (X[3][2] - X[1][2]) + (X[3][3] - X[1][3]) * (X[1][2] - X[2][2])) /
                 (X[2][3] - X[1][3]);
  const int dn = ((X[4][1] - X[1][1]) * (X[3][2] - X[1][2]) + (X[1][1] - X[3][1]) * (X[4][2] - X[1][2])) +
                 ((X[4][1] - X[1][1]) * (X[1][2] - X[2][2]) + (X[2][1] - X[1][1]) * (X[4][2] - X[1][2])) *
                     ((X[3][3] - X[1][3])) / ((X[2][3] - X[1][3])) +
                 ((X[1][1] - X[2][1]) * (X[3][2] - X[1][2]) + (X[1][1] - X[3][1]) * (X[1][2] - X[2][2])) *
                     ((X[4][3] - X[1][3])) / ((X[2][3] - X[1][3]));
  const int e1 = (X[1][1] - X[3][1]) + (X[2][1] - X[1][1]) * (X[3][3] - X[1][3]) / (X[2][3] - X[1][3]);
  const int e2 = (X[3][3] - X[1][3]) / (X[2][3] - X[1][3]);
  const int e3 = (X[3][2] - X[1][2]) + (X[1][2] - X[2][2]) * (X[3][3] - X[1][3]) / (X[2][3] - X[1][3]);
  double* qb = qblk(p, b, q, n);
  const ll n3 = n3of(b);
  ll idx[5];
  for (int m = 0; m < 5; m++) idx[m] = i3(b, X[m][1], X[m][2], X[m][3]);
  for (int v = 0; v < p.npde; v++) {
    const double Q1 = qb[idx[1] + v * n3], Q2 = qb[idx[2] + v * n3], Q3 = qb[idx[3] + v * n3],
                 Q4 = qb[idx[4] + v * n3];
    double a = (c1 * (Q2 - Q1) + c2 * (Q3 - Q1) + c3 * (Q4 - Q1)) / dn;
    double bb = (e1 * a + e2 * (Q1 - Q2) + (Q3 - Q1)) / e3;
    double c = ((X[1][1] - X[2][1]) * a + (X[1][2] - X[2][2]) * bb + (Q2 - Q1)) / (X[2][3] - X[1][3]);
    double d = Q1 - a * X[1][1] - bb * X[1][2] - c * X[1][3];
    qb[idx[0] + v * n3] = a * X[0][1] + bb * X[0][2] + c * X[0][3] + d;
  }
}

// ---- rest1_b / rest2_*_b (src/cosa.f90:2788-3011), one thread per coarse cell ----
__global__ void __launch_bounds__(256) k_rest(Par p, const Blk* __restrict__ b1s, const Blk* __restrict__ b2s,
                                              const Tile* __restrict__ tiles, int mode, int ho,
                                              const double* __restrict__ src1, double* __restrict__ dst2,
                                              const double* __restrict__ vol1, const double* __restrict__ vol2) {
  const Tile t = tiles[blockIdx.x];
  const Blk b1 = b1s[t.block], b2 = b2s[t.block];
  const int im = b2.imax - 1, jm = b2.jmax - 1, km = b2.kmax - 1;
  const int nc = im * jm * km;
  const int item = t.base + threadIdx.x;
  if (item >= nc * p.nsnap) return;
  const int n = item / nc;
  int rem = item - n * nc;
  const int i2_ = rem % im + 1;
  rem /= im;
  const int j2_ = rem % jm + 1, k2_ = rem / jm + 1;
  const int i1_ = 2 * i2_ - 1, j1_ = 2 * j2_ - 1, k1_ = 2 * k2_ - 1;
  const double* s = qblk(p, b1, src1, n);
  double* d = qblk(p, b2, dst2, n);
  const ll n31 = n3of(b1), n32 = n3of(b2);
  const ll x2 = i3(b2, i2_, j2_, k2_);
  for (int v = 0; v < p.npde; v++) {
    const double* sv = s + v * n31;
#define RR(di, dj, dk) sv[i3(b1, i1_ + (di), j1_ + (dj), k1_ + (dk))]
    if (mode == 1) {
      const double* V = vol1 + b1.o1;
#define R1(di, dj, dk) (V[i1(b1, i1_ + (di), j1_ + (dj), k1_ + (dk))] * RR(di, dj, dk))
      d[x2 + v * n32] = (R1(0, 0, 0) + R1(0, 1, 0) + R1(1, 0, 0) + R1(1, 1, 0) + R1(0, 0, 1) + R1(0, 1, 1) +
                         R1(1, 0, 1) + R1(1, 1, 1)) /
                        vol2[b2.o1 + i1(b2, i2_, j2_, k2_)];
#undef R1
    } else if (!ho) {
      d[x2 + v * n32] = -(RR(0, 0, 0) + RR(0, 1, 0) + RR(1, 0, 0) + RR(1, 1, 0) + RR(0, 0, 1) + RR(0, 1, 1) +
                          RR(1, 0, 1) + RR(1, 1, 1));
    } else {
      double s27 = RR(0, 0, 0) + RR(0, 1, 0) + RR(1, 0, 0) + RR(1, 1, 0) + RR(0, 0, 1) + RR(0, 1, 1) + RR(1, 0, 1) +
                   RR(1, 1, 1);
      double s9 = RR(-1, 0, 0) + RR(-1, 1, 0) + RR(2, 0, 0) + RR(2, 1, 0) + RR(-1, 0, 1) + RR(-1, 1, 1) +
                  RR(2, 0, 1) + RR(2, 1, 1) + RR(0, -1, 0) + RR(1, -1, 0) + RR(0, 2, 0) + RR(1, 2, 0) + RR(0, -1, 1) +
                  RR(1, -1, 1) + RR(0, 2, 1) + RR(1, 2, 1) + RR(0, 0, -1) + RR(0, 1, -1) + RR(0, 0, 2) + RR(0, 1, 2) +
                  RR(1, 0, -1) + RR(1, 1, -1) + RR(1, 0, 2) + RR(1, 1, 2);
      double s3 = RR(0, -1, -1) + RR(1, -1, -1) + RR(0, 2, -1) + RR(1, 2, -1) + RR(-1, 0, -1) + RR(-1, 1, -1) +
                  RR(2, 0, -1) + RR(2, 1, -1) + RR(0, -1, 2) + RR(1, -1, 2) + RR(0, 2, 2) + RR(1, 2, 2) +
                  RR(-1, 0, 2) + RR(-1, 1, 2) + RR(2, 0, 2) + RR(2, 1, 2) + RR(-1, -1, 0) + RR(-1, -1, 1) +
                  RR(-1, 2, 0) + RR(-1, 2, 1) + RR(2, -1, 0) + RR(2, -1, 1) + RR(2, 2, 0) + RR(2, 2, 1);
      double s1 = RR(-1, -1, -1) + RR(2, -1, -1) + RR(-1, 2, -1) + RR(2, 2, -1) + RR(-1, -1, 2) + RR(2, -1, 2) +
                  RR(-1, 2, 2) + RR(2, 2, 2);
      d[x2 + v * n32] = -(27 * s27 + 9 * s9 + 3 * s3 + s1) / 64;
    }
#undef RR
  }
}

// ---- prol_bln_b / prol_lin_b (src/cosa.f90:3315-3887) ----
__global__ void __launch_bounds__(256) k_prol(Par p, const Blk* __restrict__ b1s, const Blk* __restrict__ b2s,
                                              const Tile* __restrict__ tiles, int bln, double* __restrict__ q1,
                                              const double* __restrict__ q2, const double* __restrict__ vol1) {
  const Tile t = tiles[blockIdx.x];
  const Blk b1 = b1s[t.block], b2 = b2s[t.block];
  const int gi = (b1.imax + 1) / 2, gj = (b1.jmax + 1) / 2, gk = (b1.kmax + 1) / 2;   // i1 = 0,2,..,imax-1
  const int ng = gi * gj * gk;
  const int item = t.base + threadIdx.x;
  if (item >= ng * p.nsnap) return;
  const int n = item / ng;
  int rem = item - n * ng;
  const int i2_ = rem % gi;
  rem /= gi;
  const int j2_ = rem % gj, k2_ = rem / gj;
  const int i1_ = 2 * i2_, j1_ = 2 * j2_, k1_ = 2 * k2_;
  double* f = qblk(p, b1, q1, n);
  const double* c = qblk(p, b2, q2, n);
  const ll n31 = n3of(b1), n32 = n3of(b2);
  const double* V = vol1 + b1.o1;
  const int nv = bln ? 5 : p.npde;
  for (int v = 0; v < nv; v++) {
    const double* cv = c + v * n32;
    double* fv = f + v * n31;
#define C2(di, dj, dk) cv[i3(b2, i2_ + (di), j2_ + (dj), k2_ + (dk))]
#define QF(di, dj, dk) fv[i3(b1, i1_ + (di), j1_ + (dj), k1_ + (dk))]
#define VV(di, dj, dk) V[i1(b1, i1_ + (di), j1_ + (dj), k1_ + (dk))]
    if (bln) {
      double cnec_l = C2(1, 1, 0), cnwc_l = C2(0, 1, 0), csec_l = C2(1, 0, 0), cswc_l = C2(0, 0, 0);
      double cnec_u = C2(1, 1, 1), cnwc_u = C2(0, 1, 1), csec_u = C2(1, 0, 1), cswc_u = C2(0, 0, 1);
      double cswf_l = (27 * cswc_l + 9 * csec_l + 9 * cnwc_l + 3 * cnec_l + 9 * cswc_u + 3 * csec_u + 3 * cnwc_u + cnec_u) / 64;
      double csef_l = (27 * csec_l + 9 * cswc_l + 9 * cnec_l + 3 * cnwc_l + 9 * csec_u + 3 * cswc_u + 3 * cnec_u + cnwc_u) / 64;
      double cnef_l = (27 * cnec_l + 9 * cnwc_l + 9 * csec_l + 3 * cswc_l + 9 * cnec_u + 3 * cnwc_u + 3 * csec_u + cswc_u) / 64;
      double cnwf_l = (27 * cnwc_l + 9 * cnec_l + 9 * cswc_l + 3 * csec_l + 9 * cnwc_u + 3 * cnec_u + 3 * cswc_u + csec_u) / 64;
      double cswf_u = (27 * cswc_u + 9 * csec_u + 9 * cnwc_u + 3 * cnec_u + 9 * cswc_l + 3 * csec_l + 3 * cnwc_l + cnec_l) / 64;
      double csef_u = (27 * csec_u + 9 * cswc_u + 9 * cnec_u + 3 * cnwc_u + 9 * csec_l + 3 * cswc_l + 3 * cnec_l + cnwc_l) / 64;
      double cnef_u = (27 * cnec_u + 9 * cnwc_u + 9 * csec_u + 3 * cswc_u + 9 * cnec_l + 3 * cnwc_l + 3 * csec_l + cswc_l) / 64;
      double cnwf_u = (27 * cnwc_u + 9 * cnec_u + 9 * cswc_u + 3 * csec_u + 9 * cnwc_l + 3 * cnec_l + 3 * cswc_l + csec_l) / 64;
      QF(0, 0, 0) = QF(0, 0, 0) + cswf_l;
      QF(1, 0, 0) = QF(1, 0, 0) + csef_l;
      QF(1, 1, 0) = QF(1, 1, 0) + cnef_l;
      QF(0, 1, 0) = QF(0, 1, 0) + cnwf_l;
      QF(0, 0, 1) = QF(0, 0, 1) + cswf_u;
      QF(1, 0, 1) = QF(1, 0, 1) + csef_u;
      QF(1, 1, 1) = QF(1, 1, 1) + cnef_u;
      QF(0, 1, 1) = QF(0, 1, 1) + cnwf_u;
    } else {
      double fa[2][4];
      for (int l = 0; l < 2; l++) {
        double cnec = C2(1, 1, l), cnwc = C2(0, 1, l), csec = C2(1, 0, l), cswc = C2(0, 0, l);
        double dcnc = cnec - cnwc, dcsc = csec - cswc, dcec = cnec - csec, dcwc = cnwc - cswc;
        fa[l][0] = cswc + dcsc / 2 * VV(0, 0, l) / (VV(0, 0, l) + VV(1, 0, l)) +
                   dcwc / 2 * VV(0, 0, l) / (VV(0, 0, l) + VV(0, 1, l));
        fa[l][1] = csec - dcsc / 2 * VV(1, 0, l) / (VV(0, 0, l) + VV(1, 0, l)) +
                   dcec / 2 * VV(1, 0, l) / (VV(1, 0, l) + VV(1, 1, l));
        fa[l][3] = cnwc + dcnc / 2 * VV(0, 1, l) / (VV(0, 1, l) + VV(1, 1, l)) -
                   dcwc / 2 * VV(0, 1, l) / (VV(0, 0, l) + VV(0, 1, l));
        fa[l][2] = cnec - dcnc / 2 * VV(1, 1, l) / (VV(1, 1, l) + VV(0, 1, l)) -
                   dcec / 2 * VV(1, 1, l) / (VV(1, 1, l) + VV(1, 0, l));
      }
      const int di[4] = {0, 1, 1, 0}, dj[4] = {0, 0, 1, 1};
      for (int cc = 0; cc < 4; cc++) {
        double fact = VV(di[cc], dj[cc], 0) / 2 / (VV(di[cc], dj[cc], 0) + VV(di[cc], dj[cc], 1));
        double lo = fa[0][cc], up = fa[1][cc];
        QF(di[cc], dj[cc], 0) = QF(di[cc], dj[cc], 0) + lo + fact / 2 * (up - lo);
        QF(di[cc], dj[cc], 1) = QF(di[cc], dj[cc], 1) + up - (1 - fact) / 2 * (up - lo);
      }
    }
#undef C2
#undef QF
#undef VV
  }
}

// ------------------------------------------------------------------------
// launchers

hipError_t launch_cons2prim(Solve& s, Level& L, double* q, bool shadow, double* qold) {
  float* q32 = shadow && s.fp32() && q == L.f[COSAHIP_Q] ? L.q32 : nullptr;
  if (L.ntile[TK_CELLS])
    hipLaunchKernelGGL(k_cons2prim, grid1(L.ntile[TK_CELLS]), dim3(TS), 0, s.stream, s.par, L.dblk, L.tile[TK_CELLS], q,
                       q32, qold);
  return hipGetLastError();
}
hipError_t launch_halo_f32(Solve& s, Level& L) {
  if (!s.fp32() || !L.q32) return hipSuccess;
  if (L.ntile[TK_HALO])
    hipLaunchKernelGGL(k_halo_f32, grid1(L.ntile[TK_HALO]), dim3(TS), 0, s.stream, s.par, L.dblk, L.tile[TK_HALO],
                       L.f[COSAHIP_Q], L.q32);
  return hipGetLastError();
}
hipError_t launch_to_f32(Solve& s, const double* src, float* dst, ll n) {
  if (n <= 0) return hipSuccess;
  const ll nb = std::min<ll>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(k_to_f32, dim3((unsigned)nb), dim3(256), 0, s.stream, src, dst, n);
  return hipGetLastError();
}
hipError_t launch_prim2cons(Solve& s, Level& L, double* q, bool halo_only) {
  if (L.ntile[TK_HALO])
    hipLaunchKernelGGL(k_prim2cons, grid1(L.ntile[TK_HALO]), dim3(TS), 0, s.stream, s.par, L.dblk, L.tile[TK_HALO], q,
                       halo_only ? 1 : 0);
  return hipGetLastError();
}
hipError_t launch_copy_interior(Solve& s, Level& L, double* dst, const double* src) {
  if (L.ntile[TK_CELLS])
    hipLaunchKernelGGL(k_copy_int, grid1(L.ntile[TK_CELLS]), dim3(TS), 0, s.stream, s.par, L.dblk, L.tile[TK_CELLS],
                       dst, src);
  return hipGetLastError();
}
hipError_t launch_add_interior(Solve& s, Level& L, double* dst, const double* src, double fact) {
  if (L.ntile[TK_CELLS])
    hipLaunchKernelGGL(k_add_int, grid1(L.ntile[TK_CELLS]), dim3(TS), 0, s.stream, s.par, L.dblk, L.tile[TK_CELLS],
                       dst, src, fact);
  return hipGetLastError();
}
hipError_t launch_zero_res(Solve& s, Level& L, double* res) {
  if (L.ntile[TK_CELLS_P1])
    hipLaunchKernelGGL(k_zero_box, grid1(L.ntile[TK_CELLS_P1]), dim3(TS), 0, s.stream, s.par, L.dblk,
                       L.tile[TK_CELLS_P1], res);
  return hipGetLastError();
}
hipError_t launch_zero_res_shell(Solve& s, Level& L, double* res) {
  if (L.ntile[TK_SHELL])
    hipLaunchKernelGGL(k_zero_shell, grid1(L.ntile[TK_SHELL]), dim3(TS), 0, s.stream, s.par, L.dblk, L.tile[TK_SHELL],
                       res);
  return hipGetLastError();
}
hipError_t launch_sub_box(Solve& s, Level& L, double* dst, const double* src) {
  if (L.ntile[TK_CELLS_P1])
    hipLaunchKernelGGL(k_sub_box, grid1(L.ntile[TK_CELLS_P1]), dim3(TS), 0, s.stream, s.par, L.dblk,
                       L.tile[TK_CELLS_P1], dst, src);
  return hipGetLastError();
}
// threads per work-group of the rkim kernels: systems of m x m + m x 5
// doubles per thread within 64 KiB of LDS
int rkim_group(int m) {
  const ll per = (ll)m * m + 5LL * m;
  for (int cb = 64; cb >= 1; cb >>= 1)
    if ((ll)cb * per * 8 <= 65536) return cb;
  return 0;
}
static hipError_t launch_update_rkim(Solve& s, Level& L, double ar) {
  const int m = s.par.nsnap, cb = rkim_group(m);
  const int sub = TS / cb;
  const int plain = s.cfg.irsop == 0 && s.cfg.nlevel == 1;
  const size_t lds = (size_t)cb * ((size_t)m * m + 5 * (size_t)m) * sizeof(double);
  if (L.ntile[TK_CELLS])
    hipLaunchKernelGGL(k_update_rkim, dim3(L.ntile[TK_CELLS] * sub), dim3(cb), lds, s.stream, s.par, L.dblk,
                       L.tile[TK_CELLS], sub, cb, L.f[COSAHIP_RES], L.f[COSAHIP_Q], L.f[COSAHIP_QOLD],
                       L.f[COSAHIP_DTVOL], ar, L.f[COSAHIP_VOL], L.f[COSAHIP_DTVOLT], L.f[COSAHIP_DELPLS],
                       L.f[COSAHIP_PRD], L.dq, s.ddhb, plain, L.f[COSAHIP_TTRM]);
  return hipGetLastError();
}
hipError_t launch_lim_res_ome_rkim(Solve& s, Level& L) {
  const int m = s.par.nsnap, cb = rkim_group(m);
  const int sub = TS / cb;
  const size_t lds = (size_t)cb * ((size_t)m * m + 5 * (size_t)m) * sizeof(double);
  if (L.ntile[TK_CELLS])
    hipLaunchKernelGGL(k_lim_res_ome_rkim, dim3(L.ntile[TK_CELLS] * sub), dim3(cb), lds, s.stream, s.par, L.dblk,
                       L.tile[TK_CELLS], sub, cb, L.f[COSAHIP_RES], L.f[COSAHIP_Q], L.f[COSAHIP_QOLD],
                       L.f[COSAHIP_VOL], L.f[COSAHIP_DTVOLT], L.f[COSAHIP_DELPLS], L.f[COSAHIP_PRD], s.ddhb,
                       L.f[COSAHIP_TTRM]);
  return hipGetLastError();
}
hipError_t launch_rms_hb_term(Solve& s, Level& L, double* work1) {
  if (L.ntile[TK_CELLS])
    hipLaunchKernelGGL(k_rms_hb_term, grid1(L.ntile[TK_CELLS]), dim3(TS), 0, s.stream, s.par, L.dblk,
                       L.tile[TK_CELLS], work1, L.f[COSAHIP_Q], L.f[COSAHIP_VOL], s.ddhb);
  return hipGetLastError();
}
hipError_t launch_src_uns(Solve& s, Level& L, int sub_rhs) {
  if (L.ntile[TK_CELLS])
    hipLaunchKernelGGL(k_src_uns, grid1(L.ntile[TK_CELLS]), dim3(TS), 0, s.stream, s.par, L.dblk, L.tile[TK_CELLS],
                       L.f[COSAHIP_Q], L.f[COSAHIP_VOL], L.f[COSAHIP_RES], L.f[COSAHIP_RHS], sub_rhs);
  return hipGetLastError();
}
hipError_t launch_rhs_uns(Solve& s, Level& L) {
  if (L.ntile[TK_CELLS])
    hipLaunchKernelGGL(k_rhs_uns, grid1(L.ntile[TK_CELLS]), dim3(TS), 0, s.stream, s.par, L.dblk, L.tile[TK_CELLS],
                       L.f[COSAHIP_Q], L.f[COSAHIP_Q1], L.f[COSAHIP_Q2], L.f[COSAHIP_VOL], L.f[COSAHIP_RHS]);
  return hipGetLastError();
}
hipError_t launch_rms_dual_term(Solve& s, Level& L, double* work1) {
  if (L.ntile[TK_CELLS])
    hipLaunchKernelGGL(k_rms_dual_term, grid1(L.ntile[TK_CELLS]), dim3(TS), 0, s.stream, s.par, L.dblk,
                       L.tile[TK_CELLS], L.f[COSAHIP_Q], L.f[COSAHIP_RES], L.f[COSAHIP_RHS], L.f[COSAHIP_VOL],
                       work1);
  return hipGetLastError();
}
hipError_t launch_update(Solve& s, Level& L, double ar) {
  if (s.cfg.harbal && s.cfg.rkimp) return launch_update_rkim(s, L, ar);
  if (s.par.dual_rkimp) {
    if (L.ntile[TK_CELLS])
      hipLaunchKernelGGL(k_update_dual, grid1(L.ntile[TK_CELLS]), dim3(TS), 0, s.stream, s.par, L.dblk,
                         L.tile[TK_CELLS], L.f[COSAHIP_RES], L.f[COSAHIP_Q], L.f[COSAHIP_QOLD], L.f[COSAHIP_DTVOL],
                         ar, L.f[COSAHIP_VOL], L.f[COSAHIP_DTVOLT], L.f[COSAHIP_DELPLS], L.f[COSAHIP_PRD],
                         L.f[COSAHIP_TTRM]);
    return hipGetLastError();
  }
  if (L.ntile[TK_CELLS])
    hipLaunchKernelGGL(k_update, grid1(L.ntile[TK_CELLS]), dim3(TS), 0, s.stream, s.par, L.dblk, L.tile[TK_CELLS],
                       L.f[COSAHIP_RES], L.f[COSAHIP_Q], L.f[COSAHIP_QOLD], L.f[COSAHIP_DTVOL], ar,
                       L.f[COSAHIP_VOL], L.f[COSAHIP_DTVOLT], L.f[COSAHIP_DELPLS], L.f[COSAHIP_PRD],
                       L.f[COSAHIP_TTRM]);
  return hipGetLastError();
}
hipError_t launch_src_hb(Solve& s, Level& L) {
  // one work item per interior (cell, snapshot)
  if (L.ntile[TK_CELLS])
    hipLaunchKernelGGL(k_src_hb, grid1(L.ntile[TK_CELLS]), dim3(TS), 0, s.stream, s.par, L.dblk, L.tile[TK_CELLS],
                       L.f[COSAHIP_Q], L.f[COSAHIP_VOL], L.f[COSAHIP_RES], L.dq, s.ddhb);
  return hipGetLastError();
}
hipError_t launch_rms(Solve& s, Level& L, const double* res, const double* rhs, int use_rhs) {
  const int nt = L.ntile[TK_CELLS];
  if (!nt) return hipSuccess;
  hipLaunchKernelGGL(k_rms_part, grid1(nt), dim3(256), 0, s.stream, s.par, L.dblk, L.tile[TK_CELLS], res, rhs,
                     use_rhs, L.rms_part);
  hipLaunchKernelGGL(k_rms_final, dim3(1), dim3(256), 0, s.stream, L.rms_part, nt, s.par.npde,
                     L.rms_part + (ll)nt * 8);
  return hipGetLastError();
}
hipError_t launch_prtest(Solve& s, Level& L, const double* q, int stage_seq) {
  if (L.ntile[TK_HALO])
    hipLaunchKernelGGL(k_prtest, grid1(L.ntile[TK_HALO]), dim3(TS), 0, s.stream, s.par, L.dblk, L.tile[TK_HALO], q,
                       stage_seq, s.dflag + (L.nl - 1));
  return hipGetLastError();
}
hipError_t launch_bc(Solve& s, Level& L, double* q) {
  const int np = (int)L.hpatch.size();
  if (!np) return hipSuccess;
  if (!L.thin) {
    const int nt = L.ptile_first[np];
    if (nt)
      hipLaunchKernelGGL(k_bc, dim3(nt), dim3(TS), 0, s.stream, s.par, L.dblk, L.dpatch, L.ptile, q,
                         L.f[COSAHIP_SI], L.f[COSAHIP_SJ], L.f[COSAHIP_SK], L.f[COSAHIP_DIST], s.dqfst);
  } else {
    // a block one cell thick: patches read halos written by earlier
    // patches, so keep the reference's patch order (one launch per patch)
    for (int ip = 0; ip < np; ip++) {
      const int nt = L.ptile_first[ip + 1] - L.ptile_first[ip];
      if (nt)
        hipLaunchKernelGGL(k_bc, dim3(nt), dim3(TS), 0, s.stream, s.par, L.dblk, L.dpatch,
                           L.ptile + L.ptile_first[ip], q, L.f[COSAHIP_SI], L.f[COSAHIP_SJ], L.f[COSAHIP_SK],
                           L.f[COSAHIP_DIST], s.dqfst);
    }
  }
  return hipGetLastError();
}
hipError_t launch_cut(Solve& s, Level& L, double* q) {
  const int nc = (int)L.hcut.size();
  if (!nc) return hipSuccess;
  if (!L.thin) {
    dim3 g((L.maxcut + TS - 1) / TS, nc);
    hipLaunchKernelGGL(k_cut, g, dim3(TS), 0, s.stream, s.par, L.dblk, L.dcut, 0, q);
  } else {
    for (int ic = 0; ic < nc; ic++) {
      dim3 g((L.hcut[ic].count + TS - 1) / TS, 1);
      hipLaunchKernelGGL(k_cut, g, dim3(TS), 0, s.stream, s.par, L.dblk, L.dcut, ic, q);
    }
  }
  return hipGetLastError();
}
hipError_t launch_edges_corners(Solve& s, Level& L, double* q) {
  int mx = 0;
  for (auto& b : L.hblk) mx = std::max(mx, 4 * ((b.imax - 1) + (b.jmax - 1) + (b.kmax - 1)));
  dim3 g((mx * s.par.nsnap + TS - 1) / TS, L.nblocks);
  hipLaunchKernelGGL(k_edges, g, dim3(TS), 0, s.stream, s.par, L.dblk, q, nullptr);
  dim3 g2((8 * s.par.nsnap + 63) / 64, L.nblocks);
  hipLaunchKernelGGL(k_corners, g2, dim3(64), 0, s.stream, s.par, L.dblk, q);
  return hipGetLastError();
}
hipError_t launch_edges(Solve& s, Level& L, double* q, float* q32) {
  int mx = 0;
  for (auto& b : L.hblk) mx = std::max(mx, 4 * ((b.imax - 1) + (b.jmax - 1) + (b.kmax - 1)));
  dim3 g((mx * s.par.nsnap + TS - 1) / TS, L.nblocks);
  hipLaunchKernelGGL(k_edges, g, dim3(TS), 0, s.stream, s.par, L.dblk, q, q32);
  return hipGetLastError();
}
// cut copies of a P3 array with ncomp variables per snapshot (mut: 1)
hipError_t launch_cut_ncomp(Solve& s, Level& L, double* a, int ncomp) {
  const int nc = (int)L.hcut.size();
  if (!nc) return hipSuccess;
  Par p = s.par;
  p.npde = ncomp;
  if (!L.thin) {
    dim3 g((L.maxcut + TS - 1) / TS, nc);
    hipLaunchKernelGGL(k_cut, g, dim3(TS), 0, s.stream, p, L.dblk, L.dcut, 0, a);
  } else {
    for (int ic = 0; ic < nc; ic++) {
      dim3 g((L.hcut[ic].count + TS - 1) / TS, 1);
      hipLaunchKernelGGL(k_cut, g, dim3(TS), 0, s.stream, p, L.dblk, L.dcut, ic, a);
    }
  }
  return hipGetLastError();
}
// rest4_b (src/cosa.f90:3122-3164): rest1 of a scalar P3 field
hipError_t launch_rest_scalar(Solve& s, Level& L1, Level& L2, const double* a1, double* a2) {
  Par p = s.par;
  p.npde = 1;
  if (L1.ntile[TK_COARSE])
    hipLaunchKernelGGL(k_rest, grid1(L1.ntile[TK_COARSE]), dim3(TS), 0, s.stream, p, L1.dblk, L2.dblk,
                       L1.tile[TK_COARSE], 1, 0, a1, a2, L1.f[COSAHIP_VOL], L2.f[COSAHIP_VOL]);
  return hipGetLastError();
}
hipError_t launch_rest1(Solve& s, Level& L1, Level& L2, const double* q1, double* q2) {
  if (L1.ntile[TK_COARSE])
    hipLaunchKernelGGL(k_rest, grid1(L1.ntile[TK_COARSE]), dim3(TS), 0, s.stream, s.par, L1.dblk, L2.dblk,
                       L1.tile[TK_COARSE], 1, 0, q1, q2, L1.f[COSAHIP_VOL], L2.f[COSAHIP_VOL]);
  return hipGetLastError();
}
hipError_t launch_rest2(Solve& s, Level& L1, Level& L2, double* rhs2, const double* res1) {
  if (L1.ntile[TK_COARSE])
    hipLaunchKernelGGL(k_rest, grid1(L1.ntile[TK_COARSE]), dim3(TS), 0, s.stream, s.par, L1.dblk, L2.dblk,
                       L1.tile[TK_COARSE], 2, s.cfg.ho_rest, res1, rhs2, L1.f[COSAHIP_VOL], L2.f[COSAHIP_VOL]);
  return hipGetLastError();
}
hipError_t launch_prol(Solve& s, Level& L1, Level& L2, double* q1, const double* q2) {
  if (L1.ntile[TK_PROL])
    hipLaunchKernelGGL(k_prol, grid1(L1.ntile[TK_PROL]), dim3(TS), 0, s.stream, s.par, L1.dblk, L2.dblk,
                       L1.tile[TK_PROL], s.cfg.bln_prol, q1, q2, L1.f[COSAHIP_VOL]);
  return hipGetLastError();
}

}  // namespace cosahip
