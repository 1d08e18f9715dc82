// k_flux.hip -- spectral radii / local time step, the MUSCL + Roe residual
// and the implicit residual smoothing of COSA's pseudo-time stage, for
// gfx950.  Arithmetic follows the reference routines operation by
// operation (the library is built with -ffp-contract=off), so FP64 results
// match the CPU path to rounding of the few library calls (pow).
//
//   lambda + deltat : src/cosa.f90:1248-1346, 4183-4237, 4733-5203
//   muscl + limit   : src/flux.f90:1221-1812, 3472-3888
//   roflux          : src/flux.f90:45-260
//   resid           : src/cosa.f90:1149-1211
//   cirs_b + tridi  : src/cosa.f90:5238-5623, 2418-2443
#include "flux_face.h"
#include "kernels.h"
#include "viscous.h"

namespace cosahip {


// ------------------------------------------------------------------------
// i direction: one wavefront per row segment of ROWSEG cells; lane l
// computes the flux of face i0+l and the cell flux difference takes the
// neighbouring lane's face flux through a cross-lane shuffle.  NV = 7 adds
// the k-omega transport (h0tab: stale h0l/h0r of roflux_b, see roe_turb).
template <int NV, bool VIS, typename T>
__global__ void __launch_bounds__(256) k_flux_i(Par p, int nl, const Blk* __restrict__ blks,
                                                const Tile* __restrict__ tiles, const T* __restrict__ q,
                                                const T* __restrict__ si, double* __restrict__ res,
                                                const unsigned char* __restrict__ sbc, const ll* __restrict__ sbcoff,
                                                const double* __restrict__ h0tab, const T* __restrict__ mut,
                                                const T* __restrict__ xd, const T* __restrict__ ed,
                                                const T* __restrict__ zd, const double* __restrict__ dist) {
  const Tile t = tiles[blockIdx.x];
  const Blk b = blks[t.block];
  const int lane = threadIdx.x & 63;
  const int item = t.base + (threadIdx.x >> 6);
  const int im = b.imax - 1, jm = b.jmax - 1, km = b.kmax - 1;
  const int nseg = (im + ROWSEG - 1) / ROWSEG;
  const int nrows = jm * km;
  if (item >= nseg * nrows * p.nsnap) return;   // whole wavefront exits
  const int n = item / (nseg * nrows);
  const int rem = item - n * nseg * nrows;
  const int row = rem / nseg, seg = rem - (rem / nseg) * nseg;
  const int j = row % jm + 1, k = row / jm + 1;
  const int a = seg * ROWSEG + 1 + lane;   // face index
  const int L = b.imax;
  const ll n3 = n3of(b);
  const T* qb = q + b.o3 * p.npde * p.nsnap + n3 * (ll)p.npde * n;
  double* rb = res + b.o3 * p.npde * p.nsnap + n3 * (ll)p.npde * n;
  T F[NV], Fv[NV];
#pragma unroll
  for (int v = 0; v < NV; v++) F[v] = Fv[v] = 0;
  if (a <= L) {
    const unsigned char* fl = sbc + sbcoff[t.block];
    const ll pl = (ll)(b.jmax + 1) * (b.kmax + 1);
    bool fin = fl[j + (ll)(b.jmax + 1) * k] != 0;
    bool fout = fl[pl + j + (ll)(b.jmax + 1) * k] != 0;
    T w[4][NV];
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
      for (int v = 0; v < NV; v++) w[m][v] = qb[i3(b, a - 2 + m, j, k) + n3 * v];
    T QR[NV], QL[NV];
    muscl_face<NV, T>(p, nl, a, L, fin, fout, w[0], w[1], w[2], w[3], QR, QL);
    const int nh = n * p.hbmove;
    const T* s = si + b.o2 * p.lmet * p.nsnap_met + ((ll)nh * n2of(b) + i2(b, a, j, k)) * p.lmet;
    T kx, ky, kz, ds, ubn;
    face_metric(p, s, kx, ky, kz, ds, ubn);
    roe_face<NV, T>(p, QR, QL, kx, ky, kz, ds, ubn, F);
    if (NV > 5) {
      const double* h = h0tab + ((ll)(t.block * p.nsnap + n) * 3 + 0) * 2;
      roe_turb<T>(p, QR, QL, kx, ky, kz, ds, ubn, (T)h[0], (T)h[1], F);
    }
    if (VIS) viscous_face<1, NV, T>(p, b, a, j, k, n, q, mut, si, xd, ed, zd, fl, dist, Fv);
  }
  T Fn[NV], Fvn[NV];
#pragma unroll
  for (int v = 0; v < NV; v++) {
    Fn[v] = __shfl_down(F[v], 1, 64);
    if (VIS) Fvn[v] = __shfl_down(Fv[v], 1, 64);
  }
  const int c = a;   // cell c = face a, difference F(c+1) - F(c)
  if (lane < ROWSEG && c <= im) {
#pragma unroll
    for (int v = 0; v < NV; v++) {
      double r = (0.0 + (double)Fn[v]) - (double)F[v];
      if (VIS) r = (r + Fvn[v]) - Fv[v];   // then the viscous resid of this direction
      rb[i3(b, c, j, k) + n3 * v] = r;
    }
  }
}

// j or k direction: one thread per MSEG-cell segment of a line, marching
// its faces with a four-cell register window; res(c) += F(c+1) - F(c) in
// the reference order.  A face between two segments is computed by both
// (identically); every cell is updated by exactly one segment.
template <int DIR, int NV, bool VIS, typename T>
__global__ void __launch_bounds__(256) k_flux_march(Par p, int nl, const Blk* __restrict__ blks,
                                                    const Tile* __restrict__ tiles, const T* __restrict__ q,
                                                    const T* __restrict__ sm, double* __restrict__ res,
                                                    const unsigned char* __restrict__ sbc,
                                                    const ll* __restrict__ sbcoff, const double* __restrict__ h0tab,
                                                    const T* __restrict__ mut, const T* __restrict__ xd,
                                                    const T* __restrict__ ed, const T* __restrict__ zd,
                                                    const T* __restrict__ vpre, const double* __restrict__ dist) {
  const Tile t = tiles[blockIdx.x];
  const Blk b = blks[t.block];
  const int item = t.base + threadIdx.x;
  const int im = b.imax - 1;
  const int t2n = DIR == 2 ? b.kmax - 1 : b.jmax - 1;
  const int L = DIR == 2 ? b.jmax : b.kmax;
  const int nseg = (L - 1 + MSEG - 1) / MSEG;
  const int nlines = im * t2n;
  if (item >= nlines * nseg * p.nsnap) return;
  const int n = item / (nlines * nseg);
  int rem = item - n * nlines * nseg;
  const int sg = rem / nlines;
  rem -= sg * nlines;
  const int i = rem % im + 1, t2 = rem / im + 1;
  const int a0 = 1 + sg * MSEG, a1 = min(L, a0 + MSEG);   // faces a0..a1, cells a0..a1-1
  const ll n3 = n3of(b);
  const T* qb = q + b.o3 * p.npde * p.nsnap + n3 * (ll)p.npde * n;
  double* rb = res + b.o3 * p.npde * p.nsnap + n3 * (ll)p.npde * n;
  const ll step = DIR == 2 ? (ll)(b.imax + 3) : (ll)(b.imax + 3) * (b.jmax + 3);
  // point index of a = 0 on this line
  const ll base = DIR == 2 ? i3(b, i, 0, t2) : i3(b, i, t2, 0);
  const unsigned char* fl = sbc + sbcoff[t.block];
  bool fin, fout;
  if (DIR == 2) {
    const ll off = 2 * (ll)(b.jmax + 1) * (b.kmax + 1);
    const ll pl = (ll)(b.imax + 1) * (b.kmax + 1);
    fin = fl[off + i + (ll)(b.imax + 1) * t2] != 0;
    fout = fl[off + pl + i + (ll)(b.imax + 1) * t2] != 0;
  } else {
    const ll off = 2 * (ll)(b.jmax + 1) * (b.kmax + 1) + 2 * (ll)(b.imax + 1) * (b.kmax + 1);
    const ll pl = (ll)(b.imax + 1) * (b.jmax + 1);
    fin = fl[off + i + (ll)(b.imax + 1) * t2] != 0;
    fout = fl[off + pl + i + (ll)(b.imax + 1) * t2] != 0;
  }
  const int nh = n * p.hbmove;
  const T* sb = sm + b.o2 * p.lmet * p.nsnap_met + (ll)nh * n2of(b) * p.lmet;
  T h0l = 0, h0r = 0;
  if (NV > 5) {
    const double* h = h0tab + ((ll)(t.block * p.nsnap + n) * 3 + (DIR - 1)) * 2;
    h0l = (T)h[0];
    h0r = (T)h[1];
  }
  T w0[NV], w1[NV], w2[NV], w3[NV];
#pragma unroll
  for (int v = 0; v < NV; v++) {
    w0[v] = qb[base + (ll)(a0 - 2) * step + n3 * v];
    w1[v] = qb[base + (ll)(a0 - 1) * step + n3 * v];
    w2[v] = qb[base + (ll)a0 * step + n3 * v];
    w3[v] = qb[base + (ll)(a0 + 1) * step + n3 * v];
  }
  T Fp[NV], Fvp[NV];
  for (int a = a0; a <= a1; a++) {
    T QR[NV], QL[NV], F[NV], Fv[NV];
    muscl_face<NV, T>(p, nl, a, L, fin, fout, w0, w1, w2, w3, QR, QL);
    const ll sidx = DIR == 2 ? i2(b, i, a, t2) : i2(b, i, t2, a);
    T kx, ky, kz, ds, ubn;
    face_metric(p, sb + sidx * p.lmet, kx, ky, kz, ds, ubn);
    roe_face<NV, T>(p, QR, QL, kx, ky, kz, ds, ubn, F);
    if (NV > 5) roe_turb<T>(p, QR, QL, kx, ky, kz, ds, ubn, h0l, h0r, F);
    if (VIS) {
      if (DIR == 2) viscous_face<2, NV, T>(p, b, i, a, t2, n, q, mut, sm, xd, ed, zd, fl, dist, Fv);
      else viscous_face<3, NV, T>(p, b, i, t2, a, n, q, mut, sm, xd, ed, zd, fl, dist, Fv);
    }
    if (a > a0) {
      const ll c = base + (ll)(a - 1) * step;
      // vpre: first the viscous flux difference of the previous direction
      if (vpre) {
        if (DIR == 2) add_vdiff<NV, T>(p, b, vpre, 1, i, a - 1, t2, n, rb + c, n3);
        else add_vdiff<NV, T>(p, b, vpre, 2, i, t2, a - 1, n, rb + c, n3);
      }
#pragma unroll
      for (int v = 0; v < NV; v++) {
        double r = (rb[c + n3 * v] + F[v]) - Fp[v];
        if (VIS) r = (r + Fv[v]) - Fvp[v];   // then the viscous resid of this direction
        rb[c + n3 * v] = r;
      }
    }
#pragma unroll
    for (int v = 0; v < NV; v++) {
      if (VIS) Fvp[v] = Fv[v];
      Fp[v] = F[v];
      w0[v] = w1[v];
      w1[v] = w2[v];
      w2[v] = w3[v];
    }
    if (a + 2 <= L + 1) {
#pragma unroll
      for (int v = 0; v < NV; v++) w3[v] = qb[base + (ll)(a + 2) * step + n3 * v];
    }
  }
}

// h0l / h0r of the last face (ic, jc, kc) of each block, snapshot and
// direction: the values roflux_b's k-omega loop finds left over from the
// mean-flow loop (src/flux.f90:121-141, 287-302).  One thread per
// (block, snapshot, direction).
template <typename T>
__global__ void __launch_bounds__(64) k_h0_last(Par p, int nl, int nblocks, const Blk* __restrict__ blks,
                                                const T* __restrict__ q, const T* __restrict__ si,
                                                const T* __restrict__ sj, const T* __restrict__ sk,
                                                const unsigned char* __restrict__ sbc, const ll* __restrict__ sbcoff,
                                                double* __restrict__ h0tab) {
  const int item = blockIdx.x * blockDim.x + threadIdx.x;
  if (item >= nblocks * p.nsnap * 3) return;
  const int dir = item % 3 + 1;
  const int bn = item / 3;
  const int blk = bn / p.nsnap, n = bn % p.nsnap;
  const Blk b = blks[blk];
  const int L = dir == 1 ? b.imax : (dir == 2 ? b.jmax : b.kmax);
  // last face: a = L along dir, transverse indices at their maxima
  const int i = dir == 1 ? L : b.imax - 1, j = dir == 2 ? L : b.jmax - 1, k = dir == 3 ? L : b.kmax - 1;
  const ll n3 = n3of(b);
  const T* qb = q + b.o3 * p.npde * p.nsnap + n3 * (ll)p.npde * n;
  const unsigned char* fl = sbc + sbcoff[blk];
  bool fin, fout;
  if (dir == 1) {
    const ll pl = (ll)(b.jmax + 1) * (b.kmax + 1);
    fin = fl[j + (ll)(b.jmax + 1) * k] != 0;
    fout = fl[pl + j + (ll)(b.jmax + 1) * k] != 0;
  } else if (dir == 2) {
    const ll off = 2 * (ll)(b.jmax + 1) * (b.kmax + 1);
    const ll pl = (ll)(b.imax + 1) * (b.kmax + 1);
    fin = fl[off + i + (ll)(b.imax + 1) * k] != 0;
    fout = fl[off + pl + i + (ll)(b.imax + 1) * k] != 0;
  } else {
    const ll off = 2 * (ll)(b.jmax + 1) * (b.kmax + 1) + 2 * (ll)(b.imax + 1) * (b.kmax + 1);
    const ll pl = (ll)(b.imax + 1) * (b.jmax + 1);
    fin = fl[off + i + (ll)(b.imax + 1) * j] != 0;
    fout = fl[off + pl + i + (ll)(b.imax + 1) * j] != 0;
  }
  T w[4][7];
  for (int m = 0; m < 4; m++) {
    const int di = dir == 1 ? m - 2 : 0, dj = dir == 2 ? m - 2 : 0, dk = dir == 3 ? m - 2 : 0;
    for (int v = 0; v < 7; v++) w[m][v] = qb[i3(b, i + di, j + dj, k + dk) + n3 * v];
  }
  T QR[7], QL[7], F[7], h[2];
  muscl_face<7, T>(p, nl, L, L, fin, fout, w[0], w[1], w[2], w[3], QR, QL);
  const T* sm = dir == 1 ? si : (dir == 2 ? sj : sk);
  const int nh = n * p.hbmove;
  const T* s = sm + b.o2 * p.lmet * p.nsnap_met + ((ll)nh * n2of(b) + i2(b, i, j, k)) * p.lmet;
  T kx, ky, kz, ds, ubn;
  face_metric(p, s, kx, ky, kz, ds, ubn);
  roe_face<7, T>(p, QR, QL, kx, ky, kz, ds, ubn, F, h);
  h0tab[(ll)item * 2] = h[0];
  h0tab[(ll)item * 2 + 1] = h[1];
}

// ------------------------------------------------------------------------
// Fused residual, all three directions in one pass (2.5-D blocking).
// A workgroup of 32 x 8 threads owns FTI x FTJ = 31 x 7 cells of an (i,j)
// tile and marches k through a segment of up to KSEG planes.  Per plane:
//   - the plane (primitive q, i0-2..i0+FTI+1 x j0-2..j0+FTJ+1, no corners)
//     is staged in LDS: cell threads deposit their own value from their
//     k-window, the halo ring is loaded from global memory;
//   - lane (tx,ty) computes i-face i0+tx of row j0+ty and j-face j0+ty of
//     column i0+tx from LDS, and cell threads compute k-face k+1 from their
//     register window q(k-1..k+2) (face k is kept from the previous plane);
//   - res(c) = ((((0 + Fi(c+1)) - Fi(c)) + Fj(c+1)) - Fj(c)) + Fk(c+1)) - Fk(c),
//     the reference's accumulation order (zero, then resid for idir = 1,2,3).
// Every face flux is computed from the same operands as in the split
// kernels, so the results are identical.
constexpr int FTI = 31, FTJ = 7, KSEG = 32;
constexpr int PW = FTI + 4, PH = FTJ + 4;

template <typename T>
__global__ void __launch_bounds__(256) k_residual_fused(Par p, int nl, const Blk* __restrict__ blks,
                                                        const Tile* __restrict__ tiles, const T* __restrict__ q,
                                                        const T* __restrict__ si, const T* __restrict__ sj,
                                                        const T* __restrict__ sk, double* __restrict__ res,
                                                        const unsigned char* __restrict__ sbc,
                                                        const ll* __restrict__ sbcoff) {
  __shared__ T pl[5][PH][PW];
  __shared__ T fi[5][FTJ][FTI + 1];
  __shared__ T fj[5][FTJ + 1][FTI];
  const Tile t = tiles[blockIdx.x];
  const Blk b = blks[t.block];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int im = b.imax - 1, jm = b.jmax - 1, km = b.kmax - 1;
  const int ntx = (im + FTI - 1) / FTI, nty = (jm + FTJ - 1) / FTJ, nks = (km + KSEG - 1) / KSEG;
  int item = t.base;
  const int n = item / (ntx * nty * nks);
  item -= n * (ntx * nty * nks);
  const int ks_i = item / (ntx * nty);
  item -= ks_i * (ntx * nty);
  const int ty_i = item / ntx, tx_i = item - (item / ntx) * ntx;
  const int i0 = 1 + tx_i * FTI, j0 = 1 + ty_i * FTJ;
  const int ks = 1 + ks_i * KSEG, ke = min(km, ks + KSEG - 1);
  const ll n3 = n3of(b);
  const T* qb = q + b.o3 * p.npde * p.nsnap + n3 * (ll)p.npde * n;
  double* rb = res + b.o3 * p.npde * p.nsnap + n3 * (ll)p.npde * n;
  const int nh = n * p.hbmove;
  const ll mo = b.o2 * p.lmet * p.nsnap_met + (ll)nh * n2of(b) * p.lmet;
  const unsigned char* fl = sbc + sbcoff[t.block];
  const ll plI = (ll)(b.jmax + 1) * (b.kmax + 1);                  // i-face flag planes
  const ll offJ = 2 * plI, plJ = (ll)(b.imax + 1) * (b.kmax + 1);  // j-face flag planes
  const ll offK = offJ + 2 * plJ, plK = (ll)(b.imax + 1) * (b.jmax + 1);
  const int ci = i0 + tx, cj = j0 + ty;
  const bool cell = tx < FTI && ty < FTJ && ci <= im && cj <= jm;
  const ll kstep = (ll)(b.imax + 3) * (b.jmax + 3);
  // k window of this thread's column: w0..w3 = q(k-1..k+2) at plane k
  T w0[5], w1[5], w2[5], w3[5], Fkp[5];
  bool kfin = false, kfout = false;
  if (cell) {
    kfin = fl[offK + ci + (ll)(b.imax + 1) * cj] != 0;
    kfout = fl[offK + plK + ci + (ll)(b.imax + 1) * cj] != 0;
    const ll c0 = i3(b, ci, cj, ks);
#pragma unroll
    for (int v = 0; v < 5; v++) {
      const T* qv = qb + n3 * v;
      const T a_m2 = qv[c0 - 2 * kstep], a_m1 = qv[c0 - kstep], a_0 = qv[c0], a_p1 = qv[c0 + kstep];
      w0[v] = a_m2;   // temporarily q(ks-2..ks+1) for face ks
      w1[v] = a_m1;
      w2[v] = a_0;
      w3[v] = a_p1;
    }
    T QR[5], QL[5], kx, ky, kz, ds, ubn;
    muscl_face<5, T>(p, nl, ks, b.kmax, kfin, kfout, w0, w1, w2, w3, QR, QL);
    face_metric<T>(p, sk + mo + i2(b, ci, cj, ks) * p.lmet, kx, ky, kz, ds, ubn);
    roe_face<5, T>(p, QR, QL, kx, ky, kz, ds, ubn, Fkp);
    // shift to the window of plane ks: q(ks-1..ks+2)
#pragma unroll
    for (int v = 0; v < 5; v++) {
      w0[v] = w1[v];
      w1[v] = w2[v];
      w2[v] = w3[v];
      w3[v] = qb[n3 * v + c0 + 2 * kstep];
    }
  }
  // per-row / per-column strong-BC flags (constant over k only for the face
  // planes they index by k, so looked up per plane below)
  for (int k = ks; k <= ke; k++) {
    // (a) stage plane k: halo ring from memory, cells from registers
    for (int pidx = threadIdx.x; pidx < PW * PH; pidx += 256) {
      const int ii = pidx % PW, jj = pidx / PW;
      const bool in_i = ii >= 2 && ii <= FTI + 1, in_j = jj >= 2 && jj <= FTJ + 1;
      if (!in_i && !in_j) continue;   // corner: not used
      const int i = i0 - 2 + ii, j = j0 - 2 + jj;
      if (in_i && in_j && i <= im && j <= jm) continue;   // deposited by its cell thread
      if (i > b.imax + 1 || j > b.jmax + 1) continue;     // beyond the block
      const ll x = i3(b, i, j, k);
#pragma unroll
      for (int v = 0; v < 5; v++) pl[v][jj][ii] = qb[n3 * v + x];
    }
    if (cell) {
#pragma unroll
      for (int v = 0; v < 5; v++) pl[v][ty + 2][tx + 2] = w1[v];
    }
    __syncthreads();
    // (b) face fluxes
    if (ty < FTJ && cj <= jm && ci <= b.imax) {   // i-face a = ci of row cj
      const int a = ci;
      const bool fin = fl[cj + (ll)(b.jmax + 1) * k] != 0;
      const bool fout = fl[plI + cj + (ll)(b.jmax + 1) * k] != 0;
      T m2[5], m1[5], z0[5], p1[5], QR[5], QL[5], F[5], kx, ky, kz, ds, ubn;
#pragma unroll
      for (int v = 0; v < 5; v++) {
        m2[v] = pl[v][ty + 2][tx];
        m1[v] = pl[v][ty + 2][tx + 1];
        z0[v] = pl[v][ty + 2][tx + 2];
        p1[v] = pl[v][ty + 2][tx + 3];
      }
      muscl_face<5, T>(p, nl, a, b.imax, fin, fout, m2, m1, z0, p1, QR, QL);
      face_metric<T>(p, si + mo + i2(b, a, cj, k) * p.lmet, kx, ky, kz, ds, ubn);
      roe_face<5, T>(p, QR, QL, kx, ky, kz, ds, ubn, F);
#pragma unroll
      for (int v = 0; v < 5; v++) fi[v][ty][tx] = F[v];
    }
    if (tx < FTI && ci <= im && cj <= b.jmax) {   // j-face a = cj of column ci
      const int a = cj;
      const bool fin = fl[offJ + ci + (ll)(b.imax + 1) * k] != 0;
      const bool fout = fl[offJ + plJ + ci + (ll)(b.imax + 1) * k] != 0;
      T m2[5], m1[5], z0[5], p1[5], QR[5], QL[5], F[5], kx, ky, kz, ds, ubn;
#pragma unroll
      for (int v = 0; v < 5; v++) {
        m2[v] = pl[v][ty][tx + 2];
        m1[v] = pl[v][ty + 1][tx + 2];
        z0[v] = pl[v][ty + 2][tx + 2];
        p1[v] = pl[v][ty + 3][tx + 2];
      }
      muscl_face<5, T>(p, nl, a, b.jmax, fin, fout, m2, m1, z0, p1, QR, QL);
      face_metric<T>(p, sj + mo + i2(b, ci, a, k) * p.lmet, kx, ky, kz, ds, ubn);
      roe_face<5, T>(p, QR, QL, kx, ky, kz, ds, ubn, F);
#pragma unroll
      for (int v = 0; v < 5; v++) fj[v][ty][tx] = F[v];
    }
    T Fkn[5];
    if (cell) {   // k-face k+1 from the register window
      T QR[5], QL[5], kx, ky, kz, ds, ubn;
      muscl_face<5, T>(p, nl, k + 1, b.kmax, kfin, kfout, w0, w1, w2, w3, QR, QL);
      face_metric<T>(p, sk + mo + i2(b, ci, cj, k + 1) * p.lmet, kx, ky, kz, ds, ubn);
      roe_face<5, T>(p, QR, QL, kx, ky, kz, ds, ubn, Fkn);
    }
    __syncthreads();
    // (c) accumulate in the reference order and store
    if (cell) {
      const ll x = i3(b, ci, cj, k);
#pragma unroll
      for (int v = 0; v < 5; v++) {
        T r = (T(0) + fi[v][ty][tx + 1]) - fi[v][ty][tx];
        r = (r + fj[v][ty + 1][tx]) - fj[v][ty][tx];
        r = (r + Fkn[v]) - Fkp[v];
        rb[n3 * v + x] = (double)r;
        Fkp[v] = Fkn[v];
      }
      if (k < ke) {
#pragma unroll
        for (int v = 0; v < 5; v++) {
          w0[v] = w1[v];
          w1[v] = w2[v];
          w2[v] = w3[v];
          w3[v] = qb[n3 * v + x + 3 * kstep];
        }
      }
    }
  }
}

// ------------------------------------------------------------------------
// lambda (inviscid eig_eul) + deltat, one thread per (cell, snapshot)
template <bool VIS>
__global__ void __launch_bounds__(256) k_lambda_deltat(Par p, const Blk* __restrict__ blks,
                                                       const Tile* __restrict__ tiles, const double* __restrict__ q,
                                                       const double* __restrict__ si, const double* __restrict__ sj,
                                                       const double* __restrict__ sk, const double* __restrict__ xdot,
                                                       const double* __restrict__ ydot,
                                                       const double* __restrict__ zdot, double* __restrict__ dtvol,
                                                       double* __restrict__ bx, double* __restrict__ be,
                                                       double* __restrict__ bz, const double* __restrict__ vol,
                                                       const double* __restrict__ mut, double* __restrict__ dtvolt,
                                                       double* __restrict__ tbx, double* __restrict__ tbe,
                                                       double* __restrict__ tbz) {
  const Tile t = tiles[blockIdx.x];
  const Blk b = blks[t.block];
  const int item = t.base + threadIdx.x;
  const int im = b.imax - 1, jm = b.jmax - 1, km = b.kmax - 1;
  const int nc = im * jm * km;
  if (item >= nc * p.nsnap) return;
  const int n = item / nc;
  int rem = item - n * nc;
  const int i = rem % im + 1;
  rem /= im;
  const int j = rem % jm + 1, k = rem / jm + 1;
  const ll n3 = n3of(b);
  const double* qb = q + b.o3 * p.npde * p.nsnap + n3 * (ll)p.npde * n;
  const ll c = i3(b, i, j, k);
  double rho = qb[c], u = qb[c + n3], v = qb[c + 2 * n3], w = qb[c + 3 * n3], pp = qb[c + 4 * n3];
  const int nh = n * p.hbmove;
  const ll mo = b.o2 * p.lmet * p.nsnap_met + (ll)nh * n2of(b) * p.lmet;
  const double* sim = si + mo + i2(b, i, j, k) * p.lmet;
  const double* sip = si + mo + i2(b, i + 1, j, k) * p.lmet;
  const double* sjm = sj + mo + i2(b, i, j, k) * p.lmet;
  const double* sjp = sj + mo + i2(b, i, j + 1, k) * p.lmet;
  const double* skm = sk + mo + i2(b, i, j, k) * p.lmet;
  const double* skp = sk + mo + i2(b, i, j, k + 1) * p.lmet;
  double a = sqrt(p.gamma * pp / rho);
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
  if (p.moving) {
    const ll xo = b.o2 * p.nsnap_met + (ll)nh * n2of(b) + i2(b, i, j, k);
    u = u - xdot[xo];
    v = v - ydot[xo];
    w = w - zdot[xo];
  }
  double lami = fabs(u * xix + v * xiy + w * xiz) + a * axi;
  double lamj = fabs(u * etax + v * etay + w * etaz) + a * aeta;
  double lamk = fabs(u * zetax + v * zetay + w * zetaz) + a * azeta;
  const ll po = b.o1 * p.nsnap + (ll)n * n1of(b) + i1(b, i, j, k);
  double sum = lami + lamj + lamk;
  double lv[3] = {0.0, 0.0, 0.0};
  if (VIS) {
    // eig_ns viscous radii (src/cosa.f90:1350-1430, 4240-4297); mut only for
    // k-omega; the +1 neighbour's temperature divides by the -1 neighbour's
    // density, as in the reference
    const double gamma = p.gamma;
    const double t0 = gamma * pp / rho;
    const double mu = (p.stemp + 1.0) / (t0 + p.stemp) * pow(t0, 1.5);
    const double* mb = p.kom ? mut + b.o3 * p.nsnap + n3 * (ll)n : nullptr;
    const double mutc = p.kom ? mb[c] : 0.0;
    const double vc = vol[b.o1 + i1(b, i, j, k)];
    const ll st[3] = {1, (ll)(b.imax + 3), (ll)(b.imax + 3) * (b.jmax + 3)};
    const double ar_[3] = {axi, aeta, azeta};
#pragma unroll
    for (int d = 0; d < 3; d++) {
      const ll cm = c - st[d], cp = c + st[d];
      const double rhom = qb[cm], pm = qb[cm + 4 * n3];
      const double tm = gamma * pm / rhom;
      const double mum = (p.stemp + 1.0) / (tm + p.stemp) * pow(tm, 1.5);
      const double mutm = p.kom ? mb[cm] : 0.0;
      const double ppn = qb[cp + 4 * n3];
      const double tp = gamma * ppn / rhom;
      const double mup = (p.stemp + 1.0) / (tp + p.stemp) * pow(tp, 1.5);
      const double mutp = p.kom ? mb[cp] : 0.0;
      const double mud = (mum + 2 * mu + mup) / 4;
      const double mutd = (mutm + 2 * mutc + mutp) / 4;
      lv[d] = p.cdff * ar_[d] * ar_[d] * gamma * p.machfs / p.reyno * (mud / p.pranl + mutd / p.prant) / (rho * vc);
    }
    // irsop 5/6 take the viscous radii into the smoothing, not the time
    // step (src/cosa.f90:4831-4890)
    if (p.irsop != 5 && p.irsop != 6) sum = sum + lv[0] + lv[1] + lv[2];
  }
  dtvol[po] = p.cfl / sum;
  if (p.kom) dtvolt[po] = p.cflt / p.cfl * dtvol[po];
  for (int pass = 0; pass < (p.kom ? 2 : 1); pass++) {
    const double rr = pass == 0 ? p.rat : p.ratt;
    double* ox = pass == 0 ? bx : tbx;
    double* oe = pass == 0 ? be : tbe;
    double* oz = pass == 0 ? bz : tbz;
    if (p.irsop == 2 || p.irsop == 5) {
      const double zeta = 0.4;
      double phixi = 1 + pow(lamj / lami, zeta) + pow(lamk / lami, zeta);
      double phieta = 1 + pow(lami / lamj, zeta) + pow(lamk / lamj, zeta);
      double phizeta = 1 + pow(lami / lamk, zeta) + pow(lamj / lamk, zeta);
      double tt;
      tt = rr * lami * phixi / (lami + lamj + lamk);
      ox[po] = fmax(0.0, (tt * tt - 1) / 4);
      tt = rr * lamj * phieta / (lami + lamj + lamk);
      oe[po] = fmax(0.0, (tt * tt - 1) / 4);
      tt = rr * lamk * phizeta / (lami + lamj + lamk);
      oz[po] = fmax(0.0, (tt * tt - 1) / 4);
      if (p.irsop == 5) {   // cirs_v1v: at least c1/4 lambda_v / sum(lambda) (src/cosa.f90:4981-5060)
        const double c1 = 40.0;
        ox[po] = fmax(ox[po], c1 / 4 * lv[0] / (lami + lamj + lamk));
        oe[po] = fmax(oe[po], c1 / 4 * lv[1] / (lami + lamj + lamk));
        oz[po] = fmax(oz[po], c1 / 4 * lv[2] / (lami + lamj + lamk));
      }
    } else if (p.irsop == 3 || p.irsop == 6) {
      double rji = lamj / lami, rki = lamk / lami, rij = lami / lamj, rkj = lamk / lamj, rik = lami / lamk,
             rjk = lamj / lamk;
      double tt;
      tt = rr / (1 + p.psirs * (rji + rki));
      ox[po] = fmax(0.0, 0.25 * (tt * tt - 1));
      tt = rr / (1 + p.psirs * (rij + rkj));
      oe[po] = fmax(0.0, 0.25 * (tt * tt - 1));
      tt = rr / (1 + p.psirs * (rik + rjk));
      oz[po] = fmax(0.0, 0.25 * (tt * tt - 1));
      if (p.irsop == 6) {   // cirs_v2v (src/cosa.f90:4896-4975)
        const double c1 = 40.0;
        ox[po] = fmax(ox[po], 0.25 * c1 * lv[0] / (lami + lamj + lamk));
        oe[po] = fmax(oe[po], 0.25 * c1 * lv[1] / (lami + lamj + lamk));
        oz[po] = fmax(oz[po], 0.25 * c1 * lv[2] / (lami + lamj + lamk));
      }
    }
  }
}

// ------------------------------------------------------------------------
// CIRS (cirs_b, src/cosa.f90:5238-5623): tridiagonal line solves along i,
// j, k with the reference's Thomas recurrence (tridi); the multipliers r(j)
// go to a per-cell scratch, t(j) stays in place in res.  The i sweep scales
// res by the time step (dtvol, or dtvolt for k, omega), the k sweep divides
// it back; epsc is the constant coefficient of irsop = 1.

// CIRS along i, where each line is contiguous in memory: one wavefront per
// 64 lines.  Chunks of CI cells of every line (five variables, dtvol and the
// coefficients) are staged through LDS with coalesced loads and stores; the
// per-line Thomas recurrence then runs from LDS with the same arithmetic as
// k_cirs_jk (forward pass r, t; backward pass x).  This is always the first
// sweep: res is scaled by dtvol on load.
constexpr int CI = 8;     // cells per chunk
constexpr int LPAD = 65;  // padded line stride in LDS (conflict-free lane access)

template <int NV>
__global__ void __launch_bounds__(64) k_cirs_i(Par p, const Blk* __restrict__ blks, const Tile* __restrict__ tiles,
                                               double* __restrict__ res, const double* __restrict__ dtvol,
                                               const double* __restrict__ beta, double* __restrict__ rsc, int v0,
                                               double epsc) {
  __shared__ double xs[NV][CI][LPAD];
  __shared__ double rsh[CI][LPAD];
  __shared__ double dsh[CI][LPAD];
  __shared__ double bsh[CI][LPAD];
  const Tile t = tiles[blockIdx.x];
  const Blk b = blks[t.block];
  const int lane = threadIdx.x;
  const int im = b.imax - 1, jm = b.jmax - 1, km = b.kmax - 1;
  const int nlines = jm * km;
  const int total = nlines * p.nsnap;
  const int nl_here = min(64, total - t.base);
  const int L = im;
  const ll n3 = n3of(b), n1 = n1of(b);
  const bool cb_const = p.irsop == 1;
  // load slots: position ii = lane % CI of the chunk for lines lane/CI + 8 r
  const int ii_ld = lane % CI;
  ll off3[8], offr[8], off1[8];
  bool ok[8];
#pragma unroll
  for (int r = 0; r < 8; r++) {
    const int ln = lane / CI + 8 * r;
    ok[r] = ln < nl_here;
    const int gl = t.base + (ok[r] ? ln : 0);
    const int n = gl / nlines, rem = gl - (gl / nlines) * nlines;
    const int j = rem % jm + 1, k = rem / jm + 1;
    off3[r] = b.o3 * p.npde * p.nsnap + n3 * (ll)p.npde * n + n3 * v0 + i3(b, 0, j, k);
    offr[r] = b.o3 * p.nsnap + n3 * (ll)n + i3(b, 0, j, k);
    off1[r] = b.o1 * p.nsnap + n1 * (ll)n + i1(b, 0, j, k);
  }
  const bool active = lane < nl_here;
  double tp[NV];
  double rp = 0.0;
  const int nchunk = (L + CI - 1) / CI;
  // forward elimination
  for (int c = 0; c < nchunk; c++) {
    const int i0 = 1 + c * CI;
    const int cnt = min(CI, L - i0 + 1);
    if (ii_ld < cnt) {
#pragma unroll
      for (int r = 0; r < 8; r++) {
        if (!ok[r]) continue;
        const int ln = lane / CI + 8 * r;
        const int i = i0 + ii_ld;
#pragma unroll
        for (int v = 0; v < NV; v++) xs[v][ii_ld][ln] = res[off3[r] + n3 * v + i];
        dsh[ii_ld][ln] = dtvol[off1[r] + i];
        if (!cb_const) bsh[ii_ld][ln] = beta[off1[r] + i];
      }
    }
    __syncthreads();
    if (active) {
      for (int ii = 0; ii < cnt; ii++) {
        const int a = i0 + ii;
        const double eps = cb_const ? epsc : bsh[ii][lane];
        const double ca = a == 1 ? 0.0 : -eps;
        const double cb = 1 + 2 * eps;
        const double cc = a == L ? 0.0 : -eps;
        const double dtv = dsh[ii][lane];
        double r;
        if (a == 1) {
          r = cc / cb;
#pragma unroll
          for (int v = 0; v < NV; v++) tp[v] = (xs[v][ii][lane] * dtv) / cb;
        } else {
          r = cc / (cb - rp * ca);
#pragma unroll
          for (int v = 0; v < NV; v++) tp[v] = (xs[v][ii][lane] * dtv - tp[v] * ca) / (cb - rp * ca);
        }
        rp = r;
        rsh[ii][lane] = r;
#pragma unroll
        for (int v = 0; v < NV; v++) xs[v][ii][lane] = tp[v];
      }
    }
    __syncthreads();
    if (ii_ld < cnt) {
#pragma unroll
      for (int r = 0; r < 8; r++) {
        if (!ok[r]) continue;
        const int ln = lane / CI + 8 * r;
        const int i = i0 + ii_ld;
#pragma unroll
        for (int v = 0; v < NV; v++) res[off3[r] + n3 * v + i] = xs[v][ii_ld][ln];
        rsc[offr[r] + i] = rsh[ii_ld][ln];
      }
    }
    __syncthreads();
  }
  // back substitution
  double xn[NV];
  for (int c = nchunk - 1; c >= 0; c--) {
    const int i0 = 1 + c * CI;
    const int cnt = min(CI, L - i0 + 1);
    if (ii_ld < cnt) {
#pragma unroll
      for (int r = 0; r < 8; r++) {
        if (!ok[r]) continue;
        const int ln = lane / CI + 8 * r;
        const int i = i0 + ii_ld;
#pragma unroll
        for (int v = 0; v < NV; v++) xs[v][ii_ld][ln] = res[off3[r] + n3 * v + i];
        rsh[ii_ld][ln] = rsc[offr[r] + i];
      }
    }
    __syncthreads();
    if (active) {
      for (int ii = cnt - 1; ii >= 0; ii--) {
        const int a = i0 + ii;
        if (a == L) {
#pragma unroll
          for (int v = 0; v < NV; v++) xn[v] = xs[v][ii][lane];
        } else {
          const double r = rsh[ii][lane];
#pragma unroll
          for (int v = 0; v < NV; v++) xn[v] = xs[v][ii][lane] - xn[v] * r;
        }
#pragma unroll
        for (int v = 0; v < NV; v++) xs[v][ii][lane] = xn[v];
      }
    }
    __syncthreads();
    if (ii_ld < cnt) {
#pragma unroll
      for (int r = 0; r < 8; r++) {
        if (!ok[r]) continue;
        const int ln = lane / CI + 8 * r;
        const int i = i0 + ii_ld;
#pragma unroll
        for (int v = 0; v < NV; v++) res[off3[r] + n3 * v + i] = xs[v][ii_ld][ln];
      }
    }
    __syncthreads();
  }
}

// CIRS multipliers of the constant coefficient (irsop 1) in LDS: tridi's
// r(a) = c(a) / (b(a) - r(a-1) a(a)) depend only on the position along the
// line, and all lines of a work-group (one block) have the same length, so
// thread 0 computes them once instead of every line storing and reloading
// its own (same arithmetic, same values)
constexpr int RTAB_MAX = 1024;
__device__ __forceinline__ void fill_rtab(double* rtab, int L, double epsc) {
  if (threadIdx.x == 0) {
    double rp = 0.0;
    for (int a = 1; a <= L; a++) {
      const double ca = a == 1 ? 0.0 : -epsc, cb = 1 + 2 * epsc, cc = a == L ? 0.0 : -epsc;
      const double r = a == 1 ? cc / cb : cc / (cb - rp * ca);
      rtab[a] = r;
      rp = r;
    }
  }
  __syncthreads();
}

// CIRS j / k sweep of the mean flow (NM variables: dtvol, beta, rsc) and,
// k-omega, of k and omega (NT variables: dtvolt, tbeta, rsct) in one pass:
// one thread per (line, snapshot), each group with its own coefficients,
// Thomas multipliers and time step, each variable with the operations of
// tridi (src/cosa.f90:2418-2443).  first: scale by the time step on the way
// in; last: divide by it on the way out.  UPD (mean flow only, last sweep):
// instead of storing the smoothed residual, apply the explicit RK update of
// update_b to the cell at once, q = qold - ar res dtvol (src/update.f90:
// 643-651) with res = x / dtvol as the unfused path stores it; UPD = 2 then
// also converts the new state to primitive variables with cons2prim_b's
// operations (src/update.f90:2136-2188), which the next RK stage would do
// first.  The same operations on the same values: the state is the same
// bits as cirs, update and cons2prim run one after the other.
template <int DIR, int NM, int NT, int UPD>
__global__ void __launch_bounds__(256) k_cirs_jk(Par p, const Blk* __restrict__ blks, const Tile* __restrict__ tiles,
                                                 double* __restrict__ res, const double* __restrict__ dtvol,
                                                 const double* __restrict__ beta, double* __restrict__ rsc,
                                                 const double* __restrict__ dtvolt, const double* __restrict__ tbeta,
                                                 double* __restrict__ rsct, int first, int last, double epsc,
                                                 CirsUpdate up, const double* __restrict__ jkt) {
  constexpr int NV = NM + NT;
  __shared__ double rtab[RTAB_MAX + 1];
  __shared__ double ytab[RTAB_MAX + 1];
  const Tile t = tiles[blockIdx.x];
  const Blk b = blks[t.block];
  const int item = t.base + threadIdx.x;
  const int im = b.imax - 1, jm = b.jmax - 1, km = b.kmax - 1;
  const int T1 = im;
  const int T2 = DIR == 3 ? jm : km;
  const int nlines = T1 * T2;
  // irsop 1: the multipliers r(a) and reciprocal divisors 1/d(a) of the
  // block's line length, computed on the host with the recurrence's own
  // operations (correctly rounded), are loaded into LDS by the whole
  // work-group (every line of a work-group belongs to one block)
  const bool tab = p.irsop == 1 && (DIR == 2 ? jm : km) <= RTAB_MAX && jkt != nullptr;
  if (tab) {
    const int Lt = DIR == 2 ? jm : km;
    const double* tb = jkt + 2 * (ll)JK_RT * t.block;
    for (int a = 1 + threadIdx.x; a <= Lt; a += blockDim.x) {
      rtab[a] = tb[a];
      ytab[a] = tb[JK_RT + a];
    }
    __syncthreads();
  }
  if (item >= nlines * p.nsnap) return;
  const int n = item / nlines;
  const int rem = item - n * nlines;
  const int t1 = rem % T1 + 1, t2 = rem / T1 + 1;
  const int L = DIR == 2 ? jm : km;
  const ll n3 = n3of(b);
  const bool cconst = p.irsop == 1;
  double* rb = res + b.o3 * p.npde * p.nsnap + n3 * (ll)p.npde * n;
  double* rs = rsc + b.o3 * p.nsnap + n3 * (ll)n;
  double* rst = NT ? rsct + b.o3 * p.nsnap + n3 * (ll)n : nullptr;
  const ll o1 = b.o1 * p.nsnap + n1of(b) * (ll)n;
  const double* dtb = dtvol + o1;
  const double* dtt = NT ? dtvolt + o1 : nullptr;
  const double* bb = cconst ? nullptr : beta + o1;
  const double* bbt = (cconst || !NT) ? nullptr : tbeta + o1;
  double tp[NV];
  double rp = 0.0, rpt = 0.0;
  // the residual of cell a + 1 is loaded before cell a's result is stored:
  // the stores may alias the next loads as far as the compiler can tell,
  // so without it every step of the recurrence waits one memory latency
  const ll sa = DIR == 2 ? (ll)(b.imax + 3) : (ll)(b.imax + 3) * (b.jmax + 3);   // cell stride along the line
  double xnx[NV];
  {
    const ll c3a = i3(b, t1, DIR == 2 ? 1 : t2, DIR == 3 ? 1 : t2);
#pragma unroll
    for (int v = 0; v < NV; v++) xnx[v] = rb[c3a + n3 * v];
  }
#pragma unroll 1
  for (int a = 1; a <= L; a++) {
    const int i = t1, j = DIR == 2 ? a : t2, k = DIR == 3 ? a : t2;
    const ll c3 = i3(b, i, j, k), c1 = i1(b, i, j, k);
    double xa[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) xa[v] = xnx[v];
    if (a < L) {
#pragma unroll
      for (int v = 0; v < NV; v++) xnx[v] = rb[c3 + sa + n3 * v];
    }
    const double eps = cconst ? epsc : bb[c1];
    const double ca = a == 1 ? 0.0 : -eps, cb = 1 + 2 * eps, cc = a == L ? 0.0 : -eps;
    const double dtv = first ? dtb[c1] : 1.0;
    double r;
    if (tab) {
      // the forward quotient n / d by the reciprocal and one fused
      // correction: with y = RN(1/d) and q = RN(n y) faithful,
      // q + (n - q d) y rounded once is the correctly rounded n / d
      // (Markstein's theorem, IBM J. Res. Dev. 34 (1990) 111-119), the bits
      // of the division; an exact zero keeps its sign from q
      r = rtab[a];
      const double d = a == 1 ? cb : cb - rp * ca, y = ytab[a];
#pragma unroll
      for (int v = 0; v < NM; v++) {
        const double x = first ? xa[v] * dtv : xa[v];
        const double n = a == 1 ? x : x - tp[v] * ca;
        const double q = n * y;
        tp[v] = n == 0.0 ? q : fma(fma(-q, d, n), y, q);
      }
    } else if (a == 1) {
      r = cc / cb;
#pragma unroll
      for (int v = 0; v < NM; v++) {
        const double x = first ? xa[v] * dtv : xa[v];
        tp[v] = x / cb;
      }
    } else {
      r = cc / (cb - rp * ca);
#pragma unroll
      for (int v = 0; v < NM; v++) {
        const double x = first ? xa[v] * dtv : xa[v];
        tp[v] = (x - tp[v] * ca) / (cb - rp * ca);
      }
    }
    rp = r;
    if (!tab) rs[c3] = r;
    if (NT) {
      const double epst = cconst ? epsc : bbt[c1];
      const double cat = a == 1 ? 0.0 : -epst, cbt = 1 + 2 * epst, cct = a == L ? 0.0 : -epst;
      const double dtvt = first ? dtt[c1] : 1.0;
      double rt;
      if (tab) {
        // irsop 1: k and omega use the same constant coefficient
        rt = rtab[a];
        const double d = a == 1 ? cbt : cbt - rpt * cat, y = ytab[a];
#pragma unroll
        for (int v = NM; v < NV; v++) {
          const double x = first ? xa[v] * dtvt : xa[v];
          const double n = a == 1 ? x : x - tp[v] * cat;
          const double q = n * y;
          tp[v] = n == 0.0 ? q : fma(fma(-q, d, n), y, q);
        }
      } else if (a == 1) {
        rt = cct / cbt;
#pragma unroll
        for (int v = NM; v < NV; v++) {
          const double x = first ? xa[v] * dtvt : xa[v];
          tp[v] = x / cbt;
        }
      } else {
        rt = cct / (cbt - rpt * cat);
#pragma unroll
        for (int v = NM; v < NV; v++) {
          const double x = first ? xa[v] * dtvt : xa[v];
          tp[v] = (x - tp[v] * cat) / (cbt - rpt * cat);
        }
      }
      rpt = rt;
      if (!tab) rst[c3] = rt;
    }
#pragma unroll
    for (int v = 0; v < NV; v++) rb[c3 + n3 * v] = tp[v];
  }
  double xn[NV];
  {   // the forward result of cell L, loaded after its store above
    const ll c3L = i3(b, t1, DIR == 2 ? L : t2, DIR == 3 ? L : t2);
#pragma unroll
    for (int v = 0; v < NV; v++) xnx[v] = rb[c3L + n3 * v];
  }
#pragma unroll 1
  for (int a = L; a >= 1; a--) {
    const int i = t1, j = DIR == 2 ? a : t2, k = DIR == 3 ? a : t2;
    const ll c3 = i3(b, i, j, k);
    double ta[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) ta[v] = xnx[v];
    if (a > 1) {   // cell a - 1 ahead of cell a's store, as in the forward pass
#pragma unroll
      for (int v = 0; v < NV; v++) xnx[v] = rb[c3 - sa + n3 * v];
    }
    double xv[NV];
    if (a == L) {
#pragma unroll
      for (int v = 0; v < NV; v++) xv[v] = ta[v];
    } else {
      const double r = tab ? rtab[a] : rs[c3];
#pragma unroll
      for (int v = 0; v < NM; v++) xv[v] = ta[v] - xn[v] * r;
      if (NT) {
        const double rt = tab ? rtab[a] : rst[c3];
#pragma unroll
        for (int v = NM; v < NV; v++) xv[v] = ta[v] - xn[v] * rt;
      }
    }
#pragma unroll
    for (int v = 0; v < NV; v++) xn[v] = xv[v];
    if (UPD) {
      static_assert(UPD == 0 || (NM == 5 && NT == 0), "the fused update is built for the mean flow only");
      const ll c1 = i1(b, i, j, k);
      const double dtv = dtb[c1];
      const ll qo = rb - res;
      double qc[NM];
#pragma unroll
      for (int v = 0; v < NM; v++) {
        const double rv = xv[v] / dtv;
        qc[v] = up.qold[qo + c3 + n3 * v] - up.ar * rv * dtv;
      }
      double* qb = up.q + qo;
      if (UPD == 2) {
        const double rho = qc[0];
        const double u = qc[1] / rho, vv = qc[2] / rho, w = qc[3] / rho;
        const double pp = (p.gamma - 1) * (qc[4] - rho * (u * u + vv * vv + w * w) / 2 - rho * 0.0);
        qb[c3] = rho;
        qb[c3 + n3] = u;
        qb[c3 + 2 * n3] = vv;
        qb[c3 + 3 * n3] = w;
        qb[c3 + 4 * n3] = pp;
        if (up.q32) {
          float* sb = up.q32 + qo;
          sb[c3] = (float)rho;
          sb[c3 + n3] = (float)u;
          sb[c3 + 2 * n3] = (float)vv;
          sb[c3 + 3 * n3] = (float)w;
          sb[c3 + 4 * n3] = (float)pp;
        }
      } else {
#pragma unroll
        for (int v = 0; v < NM; v++) qb[c3 + n3 * v] = qc[v];
      }
    } else if (last) {
      const ll c1 = i1(b, i, j, k);
      const double dtv = dtb[c1];
#pragma unroll
      for (int v = 0; v < NM; v++) rb[c3 + n3 * v] = xv[v] / dtv;
      if (NT) {
        const double dtvt = dtt[c1];
#pragma unroll
        for (int v = NM; v < NV; v++) rb[c3 + n3 * v] = xv[v] / dtvt;
      }
    } else {
#pragma unroll
      for (int v = 0; v < NV; v++) rb[c3 + n3 * v] = xv[v];
    }
  }
}

// CIRS j / k sweep with the constant coefficient of irsop 1 (the tables of
// k_cirs_jk), the forward pass kept in chunks of C cells: the forward pass
// stores its result t(a) only at the last cell of each chunk; the backward
// pass takes the chunks from the end of the line, reloads a chunk's
// residuals, recomputes its forward values from the stored t of the cell
// before it (the same operations on the same values, so the same bits) and
// runs the back substitution through the chunk from registers.  Per cell
// the line's residual is read twice and written once (UPD: read twice, q
// written), against two reads and two writes when every forward value is
// stored and reloaded; the chunk's loads are issued together.  First-sweep
// scaling is not built (the i sweep comes first).
template <int DIR, int NM, int NT, int UPD, int C>
__global__ void __launch_bounds__(256) k_cirs_jkc(Par p, const Blk* __restrict__ blks, const Tile* __restrict__ tiles,
                                                  double* __restrict__ res, const double* __restrict__ dtvol,
                                                  const double* __restrict__ dtvolt, int last, double epsc,
                                                  CirsUpdate up, const double* __restrict__ jkt) {
  constexpr int NV = NM + NT;
  __shared__ double rtab[RTAB_MAX + 1];
  __shared__ double ytab[RTAB_MAX + 1];
  const Tile t = tiles[blockIdx.x];
  const Blk b = blks[t.block];
  const int item = t.base + threadIdx.x;
  const int im = b.imax - 1, jm = b.jmax - 1, km = b.kmax - 1;
  const int T1 = im;
  const int T2 = DIR == 3 ? jm : km;
  const int nlines = T1 * T2;
  const int L = DIR == 2 ? jm : km;   // at most RTAB_MAX (checked by the launcher)
  {
    const double* tb = jkt + 2 * (ll)JK_RT * t.block;
    for (int a = 1 + threadIdx.x; a <= L; a += blockDim.x) {
      rtab[a] = tb[a];
      ytab[a] = tb[JK_RT + a];
    }
    __syncthreads();
  }
  if (item >= nlines * p.nsnap) return;
  const int n = item / nlines;
  const int rem = item - n * nlines;
  const int t1 = rem % T1 + 1, t2 = rem / T1 + 1;
  const ll n3 = n3of(b);
  double* rb = res + b.o3 * p.npde * p.nsnap + n3 * (ll)p.npde * n;
  const ll o1 = b.o1 * p.nsnap + n1of(b) * (ll)n;
  const double* dtb = dtvol + o1;
  const double* dtt = NT ? dtvolt + o1 : nullptr;
  const ll sa = DIR == 2 ? (ll)(b.imax + 3) : (ll)(b.imax + 3) * (b.jmax + 3);   // cell stride along the line
  const ll s1 = DIR == 2 ? (ll)(b.imax + 1) : (ll)(b.imax + 1) * (b.jmax + 1);   // the same in P1 arrays
  const ll c3a = i3(b, t1, DIR == 2 ? 1 : t2, DIR == 3 ? 1 : t2);                 // cell a = 1
  const ll c1a = i1(b, t1, DIR == 2 ? 1 : t2, DIR == 3 ? 1 : t2);
  const double ca = -epsc, cb = 1 + 2 * epsc;
  // forward pass: t(a) = (x(a) - t(a-1) ca) / d(a), the quotient from the
  // reciprocal divisor with Markstein's correction as in k_cirs_jk
  {
    double tp[NV], xnx[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) xnx[v] = rb[c3a + n3 * v];
#pragma unroll 1
    for (int a = 1; a <= L; a++) {
      const ll c3 = c3a + (ll)(a - 1) * sa;
      double xa[NV];
#pragma unroll
      for (int v = 0; v < NV; v++) xa[v] = xnx[v];
      if (a < L) {
#pragma unroll
        for (int v = 0; v < NV; v++) xnx[v] = rb[c3 + sa + n3 * v];
      }
      const double d = a == 1 ? cb : cb - rtab[a - 1] * ca, y = ytab[a];
#pragma unroll
      for (int v = 0; v < NV; v++) {
        const double nn = a == 1 ? xa[v] : xa[v] - tp[v] * ca;
        const double q = nn * y;
        tp[v] = nn == 0.0 ? q : fma(fma(-q, d, nn), y, q);
      }
      if (a % C == 0 || a == L) {
#pragma unroll
        for (int v = 0; v < NV; v++) rb[c3 + n3 * v] = tp[v];
      }
    }
  }
  // backward pass, chunk by chunk from the end of the line
  const ll qo = rb - res;
  double xn[NV];
#pragma unroll 1
  for (int a0 = ((L - 1) / C) * C + 1; a0 >= 1; a0 -= C) {
    const int a1 = min(a0 + C - 1, L);
    double tc[NV], tv[C][NV], dv[C], dvt[C];
    if (a0 > 1) {   // t(a0 - 1), stored by the forward pass
#pragma unroll
      for (int v = 0; v < NV; v++) tc[v] = rb[c3a + (ll)(a0 - 2) * sa + n3 * v];
    }
#pragma unroll
    for (int u = 0; u < C; u++) {
      const int a = a0 + u;
      if (a <= a1) {   // x(a) for a < a1, t(a1) stored
        const ll c3 = c3a + (ll)(a - 1) * sa;
#pragma unroll
        for (int v = 0; v < NV; v++) tv[u][v] = rb[c3 + n3 * v];
        if (UPD || last) {
          dv[u] = dtb[c1a + (ll)(a - 1) * s1];
          if (NT) dvt[u] = dtt[c1a + (ll)(a - 1) * s1];
        }
      }
    }
#pragma unroll
    for (int u = 0; u < C; u++) {
      const int a = a0 + u;
      if (a < a1) {
        const double d = a == 1 ? cb : cb - rtab[a - 1] * ca, y = ytab[a];
#pragma unroll
        for (int v = 0; v < NV; v++) {
          const double tprev = u == 0 ? tc[v] : tv[u > 0 ? u - 1 : 0][v];
          const double nn = a == 1 ? tv[u][v] : tv[u][v] - tprev * ca;
          const double q = nn * y;
          tv[u][v] = nn == 0.0 ? q : fma(fma(-q, d, nn), y, q);
        }
      }
    }
#pragma unroll
    for (int u = C - 1; u >= 0; u--) {
      const int a = a0 + u;
      if (a > a1) continue;
      const ll c3 = c3a + (ll)(a - 1) * sa;
      double xv[NV];
      if (a == L) {
#pragma unroll
        for (int v = 0; v < NV; v++) xv[v] = tv[u][v];
      } else {
        const double r = rtab[a];
#pragma unroll
        for (int v = 0; v < NV; v++) xv[v] = tv[u][v] - xn[v] * r;
      }
#pragma unroll
      for (int v = 0; v < NV; v++) xn[v] = xv[v];
      if (UPD) {
        static_assert(UPD == 0 || (NM == 5 && NT == 0), "the fused update is built for the mean flow only");
        const double dtv = dv[u];
        double qc[NM];
#pragma unroll
        for (int v = 0; v < NM; v++) {
          const double rv = xv[v] / dtv;
          qc[v] = up.qold[qo + c3 + n3 * v] - up.ar * rv * dtv;
        }
        double* qb = up.q + qo;
        if (UPD == 2) {
          const double rho = qc[0];
          const double u_ = qc[1] / rho, vv = qc[2] / rho, w = qc[3] / rho;
          const double pp = (p.gamma - 1) * (qc[4] - rho * (u_ * u_ + vv * vv + w * w) / 2 - rho * 0.0);
          qb[c3] = rho;
          qb[c3 + n3] = u_;
          qb[c3 + 2 * n3] = vv;
          qb[c3 + 3 * n3] = w;
          qb[c3 + 4 * n3] = pp;
          if (up.q32) {
            float* sb = up.q32 + qo;
            sb[c3] = (float)rho;
            sb[c3 + n3] = (float)u_;
            sb[c3 + 2 * n3] = (float)vv;
            sb[c3 + 3 * n3] = (float)w;
            sb[c3 + 4 * n3] = (float)pp;
          }
        } else {
#pragma unroll
          for (int v = 0; v < NM; v++) qb[c3 + n3 * v] = qc[v];
        }
      } else if (last) {
#pragma unroll
        for (int v = 0; v < NM; v++) rb[c3 + n3 * v] = xv[v] / dv[u];
        if (NT) {
#pragma unroll
          for (int v = NM; v < NV; v++) rb[c3 + n3 * v] = xv[v] / dvt[u];
        }
      } else {
#pragma unroll
        for (int v = 0; v < NV; v++) rb[c3 + n3 * v] = xv[v];
      }
    }
  }
}

// CIRS i-sweep with one thread per (line, variable): a 256-thread
// workgroup takes 256 / NV lines, NV = npde variables each, so k and omega
// are smoothed in the same launch as the mean flow.  Mean-flow variables
// use dtvol and beta (or the constant coefficient of irsop 1), k and omega
// dtvolt and tbeta (irsop 1: the same constant).  Each thread walks its
// line in chunks of CI cells whose loads are issued before the recurrence
// consumes them.  The lanes of one line and group compute the same Thomas
// multipliers and store them to the same address of rsc / rsct for the
// back substitution.  Arithmetic as in k_cirs_i (tridi, src/cosa.f90:
// 2418-2443).
// With the constant coefficient of irsop 1 the multipliers depend only on
// the position along the line: every line of a work-group (one block) has
// the same, so they are computed once into LDS (RTAB_MAX cells at most)
// instead of being stored and reloaded per line and variable.
template <int NV>
__global__ void __launch_bounds__(256) k_cirs_iv(Par p, const Blk* __restrict__ blks, const Tile* __restrict__ tiles,
                                                 double* __restrict__ res, const double* __restrict__ dtvol,
                                                 const double* __restrict__ beta, const double* __restrict__ dtvolt,
                                                 const double* __restrict__ tbeta, double* __restrict__ rsc,
                                                 double* __restrict__ rsct, double epsc) {
  constexpr int LPT = 256 / NV;
  __shared__ double rtab[RTAB_MAX + 1];
  const Tile t = tiles[blockIdx.x];
  const Blk b = blks[t.block];
  const int L = b.imax - 1;
  const bool cconst = p.irsop == 1;
  const bool tab = cconst && L <= RTAB_MAX;
  if (tab) fill_rtab(rtab, L, epsc);
  const int ln = threadIdx.x / NV, v = threadIdx.x - (threadIdx.x / NV) * NV;
  if (ln >= LPT) return;
  const int jm = b.jmax - 1, km = b.kmax - 1;
  const int nlines = jm * km;
  const int gl = t.base + ln;
  if (gl >= nlines * p.nsnap) return;
  const int n = gl / nlines, rem = gl - (gl / nlines) * nlines;
  const int j = rem % jm + 1, k = rem / jm + 1;
  const ll n3 = n3of(b), n1 = n1of(b);
  const bool turb = v >= 5;
  double* x = res + b.o3 * p.npde * p.nsnap + n3 * (ll)p.npde * n + n3 * v + i3(b, 0, j, k);
  const ll o1 = b.o1 * p.nsnap + n1 * (ll)n + i1(b, 0, j, k);
  const double* dt = (turb ? dtvolt : dtvol) + o1;
  const double* be = cconst ? nullptr : (turb ? tbeta : beta) + o1;
  double* rr = (turb ? rsct : rsc) + b.o3 * p.nsnap + n3 * (ll)n + i3(b, 0, j, k);
  double tp = 0.0, rp = 0.0;
  for (int a0 = 1; a0 <= L; a0 += CI) {
    double xv[CI], dv[CI], ev[CI], rv[CI];
#pragma unroll
    for (int u = 0; u < CI; u++) {
      const int a = a0 + u;
      if (a <= L) {
        xv[u] = x[a];
        dv[u] = dt[a];
        ev[u] = cconst ? epsc : be[a];
      }
    }
#pragma unroll
    for (int u = 0; u < CI; u++) {
      const int a = a0 + u;
      if (a > L) break;
      const double eps = ev[u];
      const double ca = a == 1 ? 0.0 : -eps;
      const double cb = 1 + 2 * eps;
      const double cc = a == L ? 0.0 : -eps;
      double r;
      if (a == 1) {
        r = tab ? rtab[1] : cc / cb;
        tp = (xv[u] * dv[u]) / cb;
      } else {
        r = tab ? rtab[a] : cc / (cb - rp * ca);
        tp = (xv[u] * dv[u] - tp * ca) / (cb - rp * ca);
      }
      rp = r;
      rv[u] = r;
      xv[u] = tp;
    }
#pragma unroll
    for (int u = 0; u < CI; u++) {
      const int a = a0 + u;
      if (a <= L) {
        x[a] = xv[u];
        if (!tab) rr[a] = rv[u];
      }
    }
  }
  double xn = 0.0;
  const int nchunk = (L + CI - 1) / CI;
  for (int c = nchunk - 1; c >= 0; c--) {
    const int a0 = 1 + c * CI;
    double xv[CI], rv[CI];
#pragma unroll
    for (int u = 0; u < CI; u++) {
      const int a = a0 + u;
      if (a <= L) {
        xv[u] = x[a];
        rv[u] = tab ? rtab[a] : rr[a];
      }
    }
#pragma unroll
    for (int u = CI - 1; u >= 0; u--) {
      const int a = a0 + u;
      if (a > L) continue;
      xn = a == L ? xv[u] : xv[u] - xn * rv[u];
      xv[u] = xn;
    }
#pragma unroll
    for (int u = 0; u < CI; u++) {
      const int a = a0 + u;
      if (a <= L) x[a] = xv[u];
    }
  }
}

// One line of k_cirs_il: the forward and backward pass of the Thomas
// recurrence over x (already scaled by the time step) in LDS.  Each forward
// divisor d(a) = b - r(a-1) a(a) is known before the pass, so t(a) =
// (x(a) - t(a-1) a(a)) / d(a) is formed from its reciprocal y(a) = RN(1/d)
// as q = RN(n y), t = RN(q + (n - q d) y) with two fused multiply-adds:
// with y correctly rounded, q faithful and no underflow (the residuals
// here are far above the subnormal range), the correction gives the
// correctly rounded quotient (Markstein's theorem; P. Markstein, IBM J. Res.
// Dev. 34 (1990) 111-119), i.e. the same bits as n / d, with five
// dependent operations instead of a full division.  An exact zero n is the
// one case where the correction changes the sign (-0 + +0 rounds to +0):
// there q = n y already is the signed zero n / d, so it is taken as is.
// The restrict-qualified
// arguments let the LDS reads of the next cells and table entries run ahead
// of the stores of the recurrence.
__device__ __forceinline__ void il_thomas(double* __restrict__ x, const double* __restrict__ r,
                                          const double* __restrict__ y, int L, double ca, double cb) {
  double tp;
  {
    const double n = x[0], q = n * y[0];
    tp = n == 0.0 ? q : fma(fma(-q, cb, n), y[0], q);
  }
  x[0] = tp;
  for (int a = 1; a < L; a++) {
    const double d = cb - r[a - 1] * ca;
    const double n = x[a] - tp * ca;
    const double q = n * y[a];
    tp = n == 0.0 ? q : fma(fma(-q, d, n), y[a], q);
    x[a] = tp;
  }
  double xn = x[L - 1];
  for (int a = L - 2; a >= 0; a--) {
    xn = x[a] - xn * r[a];
    x[a] = xn;
  }
}

// CIRS i-sweep with whole lines resident in LDS (constant coefficient of
// irsop 1): a work-group takes lpw lines of one block with all NV
// variables.  The lines are read once with coalesced loads (row by row,
// consecutive threads on consecutive cells), already multiplied by the time
// step (dtvol, dtvolt for k and omega); one thread per (line, variable)
// then runs il_thomas from LDS with the block's tables (multipliers r, the
// values of fill_rtab, and the reciprocal divisors); the lines are written
// back once.  Global traffic per cell is one read and one write of res plus
// the time step, against two of each for the per-line sweeps.  Row stride S
// is odd, so the lanes of the recurrence (S doubles apart) fall in distinct
// LDS banks.  Results are those of k_cirs_iv (tridi, src/cosa.f90:
// 2418-2443), bit for bit.
template <int NV>
__global__ void __launch_bounds__(256) k_cirs_il(Par p, const Blk* __restrict__ blks, const Tile* __restrict__ tiles,
                                                 double* __restrict__ res, const double* __restrict__ dtvol,
                                                 const double* __restrict__ dtvolt, const double* __restrict__ tabs,
                                                 int lpw, int S, double epsc) {
  extern __shared__ double sm[];
  const Tile t = tiles[blockIdx.x];
  const Blk b = blks[t.block];
  const int L = b.imax - 1;
  const int jm = b.jmax - 1, km = b.kmax - 1;
  const int nlines = jm * km;
  const int nl_here = min(lpw, nlines * p.nsnap - t.base);
  const int nrow = NV * nl_here;
  double* rt = sm;                              // r(a), a = 1..L at rt[a - 1]
  double* yt = sm + S;                          // 1 / d(a) at yt[a - 1]
  double* xs = sm + 2 * S;                      // row r = line * NV + variable
  ll* base3 = reinterpret_cast<ll*>(xs + (ll)NV * lpw * S);
  ll* base1 = base3 + NV * lpw;
  const ll n3 = n3of(b), n1 = n1of(b);
  for (int r = threadIdx.x; r < nrow; r += blockDim.x) {
    const int ln = r / NV, v = r - (r / NV) * NV;
    const int gl = t.base + ln;
    const int n = gl / nlines, rem = gl - (gl / nlines) * nlines;
    const int j = rem % jm + 1, k = rem / jm + 1;
    base3[r] = b.o3 * p.npde * p.nsnap + n3 * (ll)p.npde * n + n3 * v + i3(b, 0, j, k);
    base1[r] = b.o1 * p.nsnap + n1 * (ll)n + i1(b, 0, j, k);
  }
  const double* tb = tabs + 2 * (ll)S * t.block;
  for (int a = threadIdx.x; a < L; a += blockDim.x) {
    rt[a] = tb[a];
    yt[a] = tb[S + a];
  }
  __syncthreads();
  const int tot = nrow * L;
  for (int e = threadIdx.x; e < tot; e += blockDim.x) {
    const int r = e / L, a = e - (e / L) * L + 1;
    const double* dt = r % NV >= 5 ? dtvolt : dtvol;
    xs[(ll)r * S + a - 1] = res[base3[r] + a] * dt[base1[r] + a];
  }
  __syncthreads();
  if (threadIdx.x < nrow) il_thomas(xs + (ll)threadIdx.x * S, rt, yt, L, -epsc, 1 + 2 * epsc);
  __syncthreads();
  for (int e = threadIdx.x; e < tot; e += blockDim.x) {
    const int r = e / L, a = e - (e / L) * L + 1;
    res[base3[r] + a] = xs[(ll)r * S + a - 1];
  }
}

// ------------------------------------------------------------------------
// launchers

static inline dim3 grid1(int n) { return dim3(n > 0 ? n : 1); }

hipError_t launch_lambda_deltat(Solve& s, Level& L) {
  if (!L.ntile[TK_CELLS]) return hipSuccess;
  hipLaunchKernelGGL(s.par.viscous ? k_lambda_deltat<true> : k_lambda_deltat<false>, grid1(L.ntile[TK_CELLS]),
                     dim3(TS), 0, s.stream, s.par, L.dblk,
                     L.tile[TK_CELLS], L.f[COSAHIP_Q], L.f[COSAHIP_SI], L.f[COSAHIP_SJ], L.f[COSAHIP_SK],
                     L.f[COSAHIP_XDOT], L.f[COSAHIP_YDOT], L.f[COSAHIP_ZDOT], L.f[COSAHIP_DTVOL],
                     L.f[COSAHIP_BETAXI], L.f[COSAHIP_BETAETA], L.f[COSAHIP_BETAZETA], L.f[COSAHIP_VOL],
                     L.f[COSAHIP_MUT], L.f[COSAHIP_DTVOLT], L.tbeta[0], L.tbeta[1], L.tbeta[2]);
  return hipGetLastError();
}

hipError_t launch_residual(Solve& s, Level& L, const double* q, double* res) {
  // zero (src/util.f90:238-369) over 0..imax: the residual kernels write
  // every interior cell, so only the halo shell needs it, and only the
  // cut-aware smoothing (cutcirs 1) reads the shell -- after the cut
  // exchange has filled its cut faces and with zeros on the others.  No
  // other kernel writes the shell, which stays as allocated (zero) otherwise.
  hipError_t e = hipSuccess;
  if (s.cfg.cutcirs == 1) e = launch_zero_res_shell(s, L, res);
  if (e != hipSuccess) return e;
  if (s.par.viscous || s.par.npde != 5) return launch_residual_viscous(s, L, q, res);
  if (s.fp32() && q == L.f[COSAHIP_Q] && L.q32) {
    // FP32 residual: fluxes, limiters and differences in single precision
    // from the FP32 shadows of q (kept by cons2prim and the halo pass) and
    // of the metrics (converted once per upload); res is stored in FP64.
    if (!L.m32_ok) {
      const ll nm = L.n2tot * s.par.lmet * s.par.nsnap_met;
      for (int d = 0; d < 3; d++) {
        e = launch_to_f32(s, L.f[COSAHIP_SI + d], L.m32[d], nm);
        if (e != hipSuccess) return e;
      }
      L.m32_ok = true;
    }
    const int rv = s.resid_force >= 0 ? s.resid_force : L.resid_mode;
    if (rv != RESID_GENERIC) return launch_residual_tile(s, L, rv, L.q32, L.m32[0], L.m32[1], L.m32[2], res);
    if (L.ntile[TK_FUSED])
      hipLaunchKernelGGL(k_residual_fused<float>, grid1(L.ntile[TK_FUSED]), dim3(256), 0, s.stream, s.par, L.nl, L.dblk,
                         L.tile[TK_FUSED], L.q32, L.m32[0], L.m32[1], L.m32[2], res, L.sbc, L.sbc_doff);
    return hipGetLastError();
  }
  if (!s.split_residual) {
    const int rv = s.resid_force >= 0 ? s.resid_force : L.resid_mode;
    if (s.prec_mode == 3)   // fast FP64: contracted multiply-adds
      return launch_residual_tile_contracted(s, L, rv == RESID_GENERIC ? (int)RESID_TILE7 : rv, q, L.f[COSAHIP_SI],
                                             L.f[COSAHIP_SJ], L.f[COSAHIP_SK], res);
    if (rv != RESID_GENERIC)
      return launch_residual_tile(s, L, rv, q, L.f[COSAHIP_SI], L.f[COSAHIP_SJ], L.f[COSAHIP_SK], res);
    if (L.ntile[TK_FUSED])
      hipLaunchKernelGGL(k_residual_fused<double>, grid1(L.ntile[TK_FUSED]), dim3(256), 0, s.stream, s.par, L.nl, L.dblk,
                         L.tile[TK_FUSED], q, L.f[COSAHIP_SI], L.f[COSAHIP_SJ], L.f[COSAHIP_SK], res, L.sbc,
                         L.sbc_doff);
    return hipGetLastError();
  }
  if (L.ntile[TK_ROWS_I])
    hipLaunchKernelGGL((k_flux_i<5, false, double>), grid1(L.ntile[TK_ROWS_I]), dim3(256), 0, s.stream, s.par, L.nl,
                       L.dblk, L.tile[TK_ROWS_I], q, L.f[COSAHIP_SI], res, L.sbc, L.sbc_doff, nullptr, nullptr, nullptr,
                       nullptr, nullptr, nullptr);
  if (L.ntile[TK_SEGS_J])
    hipLaunchKernelGGL((k_flux_march<2, 5, false, double>), grid1(L.ntile[TK_SEGS_J]), dim3(TS), 0, s.stream, s.par, L.nl,
                       L.dblk, L.tile[TK_SEGS_J], q, L.f[COSAHIP_SJ], res, L.sbc, L.sbc_doff, nullptr, nullptr,
                       nullptr, nullptr, nullptr, nullptr, nullptr);
  if (L.ntile[TK_SEGS_K])
    hipLaunchKernelGGL((k_flux_march<3, 5, false, double>), grid1(L.ntile[TK_SEGS_K]), dim3(TS), 0, s.stream, s.par, L.nl,
                       L.dblk, L.tile[TK_SEGS_K], q, L.f[COSAHIP_SK], res, L.sbc, L.sbc_doff, nullptr, nullptr,
                       nullptr, nullptr, nullptr, nullptr, nullptr);
  return hipGetLastError();
}

// Direction idir of the viscous / k-omega residual: inviscid Roe/MUSCL and,
// viscous, the face diffusion fluxes in the same line march (one launch).
template <int NV, bool VIS, typename T>
static void flux_dir_t(Solve& s, Level& L, int idir, const ResIn<T>& in, double* res, const T* vpre) {
  const T* xd = in.xd[3 * (idir - 1)];
  const T* ed = in.xd[3 * (idir - 1) + 1];
  const T* zd = in.xd[3 * (idir - 1) + 2];
  if (idir == 1 && L.ntile[TK_ROWS_I])
    hipLaunchKernelGGL((k_flux_i<NV, VIS, T>), grid1(L.ntile[TK_ROWS_I]), dim3(256), 0, s.stream, s.par, L.nl, L.dblk,
                       L.tile[TK_ROWS_I], in.q, in.s[0], res, L.sbc, L.sbc_doff, L.h0tab, in.mut, xd, ed, zd,
                       L.f[COSAHIP_DIST]);
  else if (idir == 2 && L.ntile[TK_SEGS_J])
    hipLaunchKernelGGL((k_flux_march<2, NV, VIS, T>), grid1(L.ntile[TK_SEGS_J]), dim3(TS), 0, s.stream, s.par, L.nl,
                       L.dblk, L.tile[TK_SEGS_J], in.q, in.s[1], res, L.sbc, L.sbc_doff, L.h0tab, in.mut, xd, ed,
                       zd, vpre, L.f[COSAHIP_DIST]);
  else if (idir == 3 && L.ntile[TK_SEGS_K])
    hipLaunchKernelGGL((k_flux_march<3, NV, VIS, T>), grid1(L.ntile[TK_SEGS_K]), dim3(TS), 0, s.stream, s.par, L.nl,
                       L.dblk, L.tile[TK_SEGS_K], in.q, in.s[2], res, L.sbc, L.sbc_doff, L.h0tab, in.mut, xd, ed,
                       zd, vpre, L.f[COSAHIP_DIST]);
}

template <typename T>
hipError_t launch_flux_dir(Solve& s, Level& L, int idir, const ResIn<T>& in, double* res, const T* vpre) {
  const bool t = s.par.npde == 7, v = s.par.viscous != 0;
  if (t && v) flux_dir_t<7, true, T>(s, L, idir, in, res, vpre);
  else if (t) flux_dir_t<7, false, T>(s, L, idir, in, res, vpre);
  else if (v) flux_dir_t<5, true, T>(s, L, idir, in, res, vpre);
  else flux_dir_t<5, false, T>(s, L, idir, in, res, vpre);
  return hipGetLastError();
}
template hipError_t launch_flux_dir<double>(Solve&, Level&, int, const ResIn<double>&, double*, const double*);
template hipError_t launch_flux_dir<float>(Solve&, Level&, int, const ResIn<float>&, double*, const float*);

template <typename T>
hipError_t launch_h0_last(Solve& s, Level& L, const ResIn<T>& in) {
  const int n = L.nblocks * s.par.nsnap * 3;
  hipLaunchKernelGGL(k_h0_last<T>, dim3((n + 63) / 64), dim3(64), 0, s.stream, s.par, L.nl, L.nblocks, L.dblk, in.q,
                     in.s[0], in.s[1], in.s[2], L.sbc, L.sbc_doff, L.h0tab);
  return hipGetLastError();
}
template hipError_t launch_h0_last<double>(Solve&, Level&, const ResIn<double>&);
template hipError_t launch_h0_last<float>(Solve&, Level&, const ResIn<float>&);

// ccirs_b sweep of direction DIR (src/cosa.f90:2067-2415, cutcirs = 1):
// thread per (line, snapshot), the variables in turn.  The xi sweep's
// right-hand side is ores (the scaled residual before smoothing), eta and
// zeta use the current res; both line ends add eps * res(halo) (the
// neighbour residual exchanged by cutman_q).  Thomas arithmetic of tridi
// (src/cosa.f90:2418-2443): multipliers in rsc, forward values in res.
template <int DIR>
__global__ void __launch_bounds__(256) k_ccirs(Par p, const Blk* __restrict__ blks, const Tile* __restrict__ tiles,
                                               double* __restrict__ res, const double* __restrict__ ores,
                                               const double* __restrict__ beta, const double* __restrict__ tbeta,
                                               double* __restrict__ rsc, double epsc) {
  const Tile t = tiles[blockIdx.x];
  const Blk b = blks[t.block];
  const int item = t.base + threadIdx.x;
  const int im = b.imax - 1, jm = b.jmax - 1, km = b.kmax - 1;
  const int T1 = DIR == 1 ? jm : im;
  const int T2 = DIR == 3 ? jm : km;
  const int nlines = T1 * T2;
  if (item >= nlines * p.nsnap) return;
  const int n = item / nlines;
  const int rem = item - n * nlines;
  const int t1 = rem % T1 + 1, t2 = rem / T1 + 1;
  const int L = DIR == 1 ? im : (DIR == 2 ? jm : km);
  const int hi = DIR == 1 ? b.imax : (DIR == 2 ? b.jmax : b.kmax);
  const ll n3 = n3of(b);
  double* rb = res + b.o3 * p.npde * p.nsnap + n3 * (ll)p.npde * n;
  const double* ob = ores + b.o3 * p.npde * p.nsnap + n3 * (ll)p.npde * n;
  double* rs = rsc + b.o3 * p.nsnap + n3 * (ll)n;
  const ll o1 = b.o1 * p.nsnap + n1of(b) * (ll)n;
  auto cell = [&](int a, ll& c3, ll& c1) {
    const int i = DIR == 1 ? a : t1, j = DIR == 2 ? a : (DIR == 1 ? t1 : t2), k = DIR == 3 ? a : t2;
    c3 = i3(b, i, j, k);
    c1 = i1(b, i, j, k);
  };
  ll lo3, hi3, tmp;
  {
    const int i = DIR == 1 ? 0 : t1, j = DIR == 2 ? 0 : (DIR == 1 ? t1 : t2), k = DIR == 3 ? 0 : t2;
    lo3 = i3(b, i, j, k);
    const int ih = DIR == 1 ? hi : t1, jh = DIR == 2 ? hi : (DIR == 1 ? t1 : t2), kh = DIR == 3 ? hi : t2;
    hi3 = i3(b, ih, jh, kh);
  }
  (void)tmp;
  for (int v = 0; v < p.npde; v++) {
    const bool turb = v >= 5;
    const double* bb = p.irsop == 1 ? nullptr : (turb ? tbeta : beta) + o1;
    const ll ov = n3 * v;
    double tp = 0.0, rp = 0.0;
    for (int a = 1; a <= L; a++) {
      ll c3, c1;
      cell(a, c3, c1);
      const double eps = bb ? bb[c1] : epsc;
      const double ca = -eps, cb = 1 + 2 * eps, cc = a == L ? 0.0 : -eps;
      double rhs = DIR == 1 ? ob[c3 + ov] : rb[c3 + ov];
      if (a == L) rhs = rhs + eps * rb[hi3 + ov];
      else if (a == 1) rhs = rhs + eps * rb[lo3 + ov];
      double r, tc;
      if (a == 1) {
        r = cc / cb;
        tc = rhs / cb;
      } else {
        r = cc / (cb - rp * ca);
        tc = (rhs - tp * ca) / (cb - rp * ca);
      }
      rs[c3] = r;
      rb[c3 + ov] = tc;
      tp = tc;
      rp = r;
    }
    double xn = 0.0;
    for (int a = L; a >= 1; a--) {
      ll c3, c1;
      cell(a, c3, c1);
      xn = a == L ? rb[c3 + ov] : rb[c3 + ov] - xn * rs[c3];
      rb[c3 + ov] = xn;
    }
  }
}

// scalef_b / scaleb_b (src/cosa.f90:5652-5773) on the interior
__global__ void __launch_bounds__(256) k_scale_dt(Par p, const Blk* __restrict__ blks, const Tile* __restrict__ tiles,
                                                  double* __restrict__ res, const double* __restrict__ dtvol,
                                                  const double* __restrict__ dtvolt, int back) {
  const Tile t = tiles[blockIdx.x];
  const Blk b = blks[t.block];
  const int im = b.imax - 1, jm = b.jmax - 1, km = b.kmax - 1;
  const int nc = im * jm * km;
  const int item = t.base + threadIdx.x;
  if (item >= nc * p.nsnap) return;
  const int n = item / nc;
  int rem = item - n * nc;
  const int i = rem % im + 1;
  rem /= im;
  const int j = rem % jm + 1, k = rem / jm + 1;
  const ll n3 = n3of(b), x = i3(b, i, j, k);
  const ll po = b.o1 * p.nsnap + n1of(b) * (ll)n + i1(b, i, j, k);
  double* rb = res + b.o3 * p.npde * p.nsnap + n3 * (ll)p.npde * n + x;
  for (int v = 0; v < p.npde; v++) {
    const double d = v < 5 ? dtvol[po] : dtvolt[po];
    rb[n3 * v] = back ? rb[n3 * v] / d : rb[n3 * v] * d;
  }
}

// CIRS with cut boundaries (src/cosa.f90:1984-1992): scalef, ores = res,
// two passes of cutman_q on res and the three ccirs sweeps, scaleb
static hipError_t launch_ccirs(Solve& s, Level& L) {
  double* res = L.f[COSAHIP_RES];
  const double epsc = 0.25 * (s.par.rat * s.par.rat - 1.0);
  const ll nq = L.n3tot * s.par.npde * s.par.nsnap;
  if (L.ntile[TK_CELLS])
    hipLaunchKernelGGL(k_scale_dt, grid1(L.ntile[TK_CELLS]), dim3(TS), 0, s.stream, s.par, L.dblk, L.tile[TK_CELLS],
                       res, L.f[COSAHIP_DTVOL], L.f[COSAHIP_DTVOLT], 0);
  hipError_t e = hipMemcpyAsync(L.work1, res, nq * sizeof(double), hipMemcpyDeviceToDevice, s.stream);
  if (e != hipSuccess) return e;
  for (int it = 0; it < 2; it++) {
    e = launch_cut(s, L, res);
    if (e != hipSuccess) return e;
    if (L.ntile[TK_LINES_I_T])
      hipLaunchKernelGGL(k_ccirs<1>, grid1(L.ntile[TK_LINES_I_T]), dim3(TS), 0, s.stream, s.par, L.dblk,
                         L.tile[TK_LINES_I_T], res, L.work1, L.f[COSAHIP_BETAXI], L.tbeta[0], L.cirs_r, epsc);
    if (L.ntile[TK_LINES_J])
      hipLaunchKernelGGL(k_ccirs<2>, grid1(L.ntile[TK_LINES_J]), dim3(TS), 0, s.stream, s.par, L.dblk,
                         L.tile[TK_LINES_J], res, L.work1, L.f[COSAHIP_BETAETA], L.tbeta[1], L.cirs_r, epsc);
    if (L.ntile[TK_LINES_K])
      hipLaunchKernelGGL(k_ccirs<3>, grid1(L.ntile[TK_LINES_K]), dim3(TS), 0, s.stream, s.par, L.dblk,
                         L.tile[TK_LINES_K], res, L.work1, L.f[COSAHIP_BETAZETA], L.tbeta[2], L.cirs_r, epsc);
  }
  if (L.ntile[TK_CELLS])
    hipLaunchKernelGGL(k_scale_dt, grid1(L.ntile[TK_CELLS]), dim3(TS), 0, s.stream, s.par, L.dblk, L.tile[TK_CELLS],
                       res, L.f[COSAHIP_DTVOL], L.f[COSAHIP_DTVOLT], 1);
  return hipGetLastError();
}

hipError_t launch_cirs(Solve& s, Level& L, const CirsUpdate* up) {
  if (s.par.irsop < 1 || s.par.irsop == 4 || s.par.irsop > 6) return hipSuccess;
  if (s.cfg.cutcirs == 1) return launch_ccirs(s, L);
  double* res = L.f[COSAHIP_RES];
  const double* dt = L.f[COSAHIP_DTVOL];
  const double epsc = 0.25 * (s.par.rat * s.par.rat - 1.0);
  ll ilines = 0;
  for (const Blk& b : L.hblk) ilines += (ll)(b.jmax - 1) * (b.kmax - 1) * s.par.nsnap;
  const int mode = s.cirs_i_mode ? s.cirs_i_mode : L.cirs_mode;
  const bool line = mode == 3 && L.ntile[TK_LINES_IL];
  const bool iv = line || mode == 2 || (mode == 0 && ilines < CIRS_IV_MAX_LINES);
  if (line) {
    const int S = L.il_stride, lpw = L.il_lpw, nv = s.par.npde;
    const size_t lds = sizeof(double) * (2 * (size_t)S + (size_t)nv * lpw * S) + 2 * sizeof(ll) * nv * lpw;
    if (nv == 7)
      hipLaunchKernelGGL(k_cirs_il<7>, grid1(L.ntile[TK_LINES_IL]), dim3(256), lds, s.stream, s.par, L.dblk,
                         L.tile[TK_LINES_IL], res, dt, L.f[COSAHIP_DTVOLT], L.il_rtab, lpw, S, epsc);
    else
      hipLaunchKernelGGL(k_cirs_il<5>, grid1(L.ntile[TK_LINES_IL]), dim3(256), lds, s.stream, s.par, L.dblk,
                         L.tile[TK_LINES_IL], res, dt, nullptr, L.il_rtab, lpw, S, epsc);
  } else if (iv && L.ntile[TK_LINES_IV]) {
    if (s.par.npde == 7)
      hipLaunchKernelGGL(k_cirs_iv<7>, grid1(L.ntile[TK_LINES_IV]), dim3(256), 0, s.stream, s.par, L.dblk,
                         L.tile[TK_LINES_IV], res, dt, L.f[COSAHIP_BETAXI], L.f[COSAHIP_DTVOLT], L.tbeta[0], L.cirs_r,
                         L.cirs_rt, epsc);
    else
      hipLaunchKernelGGL(k_cirs_iv<5>, grid1(L.ntile[TK_LINES_IV]), dim3(256), 0, s.stream, s.par, L.dblk,
                         L.tile[TK_LINES_IV], res, dt, L.f[COSAHIP_BETAXI], nullptr, nullptr, L.cirs_r, nullptr,
                         epsc);
  }
  if (!iv && L.ntile[TK_LINES_I])
    hipLaunchKernelGGL(k_cirs_i<5>, grid1(L.ntile[TK_LINES_I]), dim3(64), 0, s.stream, s.par, L.dblk,
                       L.tile[TK_LINES_I], res, dt, L.f[COSAHIP_BETAXI], L.cirs_r, 0, epsc);
  if (s.par.kom && !iv && L.ntile[TK_LINES_I])
    hipLaunchKernelGGL(k_cirs_i<2>, grid1(L.ntile[TK_LINES_I]), dim3(64), 0, s.stream, s.par, L.dblk,
                       L.tile[TK_LINES_I], res, L.f[COSAHIP_DTVOLT], L.tbeta[0], L.cirs_r, 5, epsc);
  // j and k sweeps: mean flow and k-omega in one launch per direction
  const double* dtt = L.f[COSAHIP_DTVOLT];
  const CirsUpdate none{};
  // irsop 1 with every line of the level within the LDS tables: the sweep
  // with the forward pass kept in chunks (k_cirs_jkc)
  auto chunked = [&](int dir) {
    if (s.par.irsop != 1 || s.cirs_jk_chunk == 0 || !L.jk_rtab[dir - 2]) return false;
    for (const Blk& b : L.hblk)
      if ((dir == 2 ? b.jmax : b.kmax) - 1 > RTAB_MAX) return false;
    return true;
  };
#define JKC(DIR, NT, UPD, LAST, UP, TK, RT)                                                                       \
  do {                                                                                                           \
    if (s.cirs_jk_chunk == 8)                                                                                    \
      hipLaunchKernelGGL((k_cirs_jkc<DIR, 5, NT, UPD, 8>), grid1(L.ntile[TK]), dim3(TS), 0, s.stream, s.par,     \
                         L.dblk, L.tile[TK], res, dt, NT ? dtt : nullptr, LAST, epsc, UP, RT);                   \
    else                                                                                                         \
      hipLaunchKernelGGL((k_cirs_jkc<DIR, 5, NT, UPD, 4>), grid1(L.ntile[TK]), dim3(TS), 0, s.stream, s.par,     \
                         L.dblk, L.tile[TK], res, dt, NT ? dtt : nullptr, LAST, epsc, UP, RT);                   \
  } while (0)
  if (s.par.kom) {
    if (L.ntile[TK_LINES_J]) {
      if (chunked(2)) JKC(2, 2, 0, 0, none, TK_LINES_J, L.jk_rtab[0]);
      else
        hipLaunchKernelGGL((k_cirs_jk<2, 5, 2, 0>), grid1(L.ntile[TK_LINES_J]), dim3(TS), 0, s.stream, s.par, L.dblk,
                           L.tile[TK_LINES_J], res, dt, L.f[COSAHIP_BETAETA], L.cirs_r, dtt, L.tbeta[1], L.cirs_rt, 0,
                           0, epsc, none, L.jk_rtab[0]);
    }
    if (L.ntile[TK_LINES_K]) {
      if (chunked(3)) JKC(3, 2, 0, 1, none, TK_LINES_K, L.jk_rtab[1]);
      else
        hipLaunchKernelGGL((k_cirs_jk<3, 5, 2, 0>), grid1(L.ntile[TK_LINES_K]), dim3(TS), 0, s.stream, s.par, L.dblk,
                           L.tile[TK_LINES_K], res, dt, L.f[COSAHIP_BETAZETA], L.cirs_r, dtt, L.tbeta[2], L.cirs_rt, 0,
                           1, epsc, none, L.jk_rtab[1]);
    }
  } else {
    if (L.ntile[TK_LINES_J]) {
      if (chunked(2)) JKC(2, 0, 0, 0, none, TK_LINES_J, L.jk_rtab[0]);
      else
        hipLaunchKernelGGL((k_cirs_jk<2, 5, 0, 0>), grid1(L.ntile[TK_LINES_J]), dim3(TS), 0, s.stream, s.par, L.dblk,
                           L.tile[TK_LINES_J], res, dt, L.f[COSAHIP_BETAETA], L.cirs_r, nullptr, nullptr, nullptr, 0,
                           0, epsc, none, L.jk_rtab[0]);
    }
    if (L.ntile[TK_LINES_K]) {
      const bool ch = chunked(3);
      if (!up) {
        if (ch) JKC(3, 0, 0, 1, none, TK_LINES_K, L.jk_rtab[1]);
        else
          hipLaunchKernelGGL((k_cirs_jk<3, 5, 0, 0>), grid1(L.ntile[TK_LINES_K]), dim3(TS), 0, s.stream, s.par,
                             L.dblk, L.tile[TK_LINES_K], res, dt, L.f[COSAHIP_BETAZETA], L.cirs_r, nullptr, nullptr,
                             nullptr, 0, 1, epsc, none, L.jk_rtab[1]);
      } else if (up->prim) {
        if (ch) JKC(3, 0, 2, 1, *up, TK_LINES_K, L.jk_rtab[1]);
        else
          hipLaunchKernelGGL((k_cirs_jk<3, 5, 0, 2>), grid1(L.ntile[TK_LINES_K]), dim3(TS), 0, s.stream, s.par,
                             L.dblk, L.tile[TK_LINES_K], res, dt, L.f[COSAHIP_BETAZETA], L.cirs_r, nullptr, nullptr,
                             nullptr, 0, 1, epsc, *up, L.jk_rtab[1]);
      } else {
        if (ch) JKC(3, 0, 1, 1, *up, TK_LINES_K, L.jk_rtab[1]);
        else
          hipLaunchKernelGGL((k_cirs_jk<3, 5, 0, 1>), grid1(L.ntile[TK_LINES_K]), dim3(TS), 0, s.stream, s.par,
                             L.dblk, L.tile[TK_LINES_K], res, dt, L.f[COSAHIP_BETAZETA], L.cirs_r, nullptr, nullptr,
                             nullptr, 0, 1, epsc, *up, L.jk_rtab[1]);
      }
    }
  }
#undef JKC
  return hipGetLastError();
}

}  // namespace cosahip
