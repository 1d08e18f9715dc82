// k_resid.hip -- the inviscid (Euler) residual in one pass over all three
// directions, specialised at compile time on the scheme options, for gfx950.
//
// Same arithmetic as muscl_b* + roflux_b + resid (src/flux.f90:1221-1812,
// 45-260; src/cosa.f90:1149-1211) face by face, so the results are the same
// bits as the split and generic fused kernels (k_flux.hip).  What differs is
// how the work maps onto the machine:
//
//   * one kernel instance per (limiter, first/second order, entropy fix on
//     or off): the face code holds one limiter path, not five, and no
//     entropy-fix or eigenvalue-floor branches when those are off;
//   * KW = 0: a thread's k-window q(k-1..k+2) is not carried in registers
//     across planes; q(k) comes from the staged LDS plane and q(k-1),
//     q(k+1), q(k+2) from global memory (L2-resident: each was touched one
//     or two planes before).  Only the k-face flux of the previous plane
//     stays live, which frees ~40 VGPRs and lifts the occupancy;
//     KW = 1 keeps the register window (fewer loads, fewer waves);
//   * TJ: cell rows per tile (tile 31 x TJ, work-group 32 x (TJ + 1));
//     larger TJ re-reads fewer halo rows of q and the j-face metrics;
//   * MINW: waves per SIMD the register allocation must allow (4 caps the
//     kernel at 128 VGPRs, with a few spilled).
//
// Per k-plane of a tile:
//   (a) q(k) of the tile and its 2-cell ring (no corners) is staged in LDS;
//   (b) lane (tx, ty) computes i-face i0+tx of row j0+ty and j-face j0+ty of
//       column i0+tx from LDS, and cell threads compute k-face k+1;
//   (c) res(c) = ((((0 + Fi(c+1)) - Fi(c)) + Fj(c+1)) - Fj(c)) + Fk(c+1)) - Fk(c),
//       the reference's accumulation order (zero, then resid for idir = 1, 2, 3).
#include "flux_face.h"
#include "kernels.h"

namespace cosahip {

// This file is compiled twice: as itself (no multiply-add contraction, the
// reference's rounding) and from k_resid_contracted.hip with
// COSAHIP_CONTRACTED defined and contraction on (the fast FP64 mode); the
// kernels live in a namespace of their own in each build.
#ifdef COSAHIP_CONTRACTED
namespace contracted {
#else
namespace exact {
#endif

template <typename T, int TI, int TJ, int ILIM, int FO, int EF, int KW, int MINW>
__global__ void __launch_bounds__(32 * (TJ + 1), MINW) k_residual_tile(Par p, int nl, const Blk* __restrict__ blks,
                                                                 const Tile* __restrict__ tiles,
                                                                 const T* __restrict__ q, const T* __restrict__ si,
                                                                 const T* __restrict__ sj, const T* __restrict__ sk,
                                                                 double* __restrict__ res,
                                                                 const unsigned char* __restrict__ sbc,
                                                                 const ll* __restrict__ sbcoff) {
  constexpr int NT = 32 * (TJ + 1), PW = TI + 4, PH = TJ + 4;
  static_assert(TI == 31 || TI == 32, "tiles are 31 or 32 cells wide");
  __shared__ T pl[5][PH][PW];
  __shared__ T fi[5][TJ][TI + 1];
  __shared__ T fj[5][TJ + 1][TI];
  const Tile t = tiles[blockIdx.x];
  const Blk b = blks[t.block];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int im = b.imax - 1, jm = b.jmax - 1, km = b.kmax - 1;
  const int ntx = (im + TI - 1) / TI, nty = (jm + TJ - 1) / TJ, nks = (km + RT_KSEG - 1) / RT_KSEG;
  int item = t.base;
  const int n = item / (ntx * nty * nks);
  item -= n * (ntx * nty * nks);
  const int ks_i = item / (ntx * nty);
  item -= ks_i * (ntx * nty);
  const int ty_i = item / ntx, tx_i = item - ty_i * ntx;
  const int i0 = 1 + tx_i * TI, j0 = 1 + ty_i * TJ;
  const int ks = 1 + ks_i * RT_KSEG, ke = min(km, ks + RT_KSEG - 1);
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
  const bool cell = tx < TI && ty < TJ && ci <= im && cj <= jm;
  // i-face thread: face i0 + fc of row j0 + fr, TI + 1 faces per row (for
  // 31-cell tiles the lane (tx, ty) itself)
  const int fr = threadIdx.x / (TI + 1), fc = threadIdx.x - fr * (TI + 1);
  const int fa = i0 + fc, fj0 = j0 + fr;
  const ll kstep = (ll)(b.imax + 3) * (b.jmax + 3);
  T w0[KW ? 5 : 1], w1[KW ? 5 : 1], w2[KW ? 5 : 1], w3[KW ? 5 : 1];
  T Fkp[5];
  bool kfin = false, kfout = false;
  if (cell) {
    kfin = fl[offK + ci + (ll)(b.imax + 1) * cj] != 0;
    kfout = fl[offK + plK + ci + (ll)(b.imax + 1) * cj] != 0;
    const ll c0 = i3(b, ci, cj, ks);
    T a_m2[5], a_m1[5], a_0[5], a_p1[5], QR[5], QL[5], kx, ky, kz, ds, ubn;
#pragma unroll
    for (int v = 0; v < 5; v++) {
      const T* qv = qb + n3 * v;
      a_m2[v] = qv[c0 - 2 * kstep];
      a_m1[v] = qv[c0 - kstep];
      a_0[v] = qv[c0];
      a_p1[v] = qv[c0 + kstep];
    }
    muscl_face<5, T, ILIM, FO>(p, nl, ks, b.kmax, kfin, kfout, a_m2, a_m1, a_0, a_p1, QR, QL);
    face_metric<T>(p, sk + mo + i2(b, ci, cj, ks) * p.lmet, kx, ky, kz, ds, ubn);
    roe_face<5, T, EF>(p, QR, QL, kx, ky, kz, ds, ubn, Fkp);
    if constexpr (KW != 0) {
#pragma unroll
      for (int v = 0; v < 5; v++) {
        w0[v] = a_m1[v];
        w1[v] = a_0[v];
        w2[v] = a_p1[v];
        w3[v] = qb[n3 * v + c0 + 2 * kstep];
      }
    }
  }
  for (int k = ks; k <= ke; k++) {
    // (a) stage plane k: the ring from memory, cells from registers (KW)
    //     or from memory
    for (int pidx = threadIdx.x; pidx < PW * PH; pidx += NT) {
      const int ii = pidx % PW, jj = pidx / PW;
      const bool in_i = ii >= 2 && ii <= TI + 1, in_j = jj >= 2 && jj <= TJ + 1;
      if (!in_i && !in_j) continue;   // corner: not used
      const int i = i0 - 2 + ii, j = j0 - 2 + jj;
      if (KW != 0 && in_i && in_j && i <= im && j <= jm) continue;   // deposited by its cell thread
      if (i > b.imax + 1 || j > b.jmax + 1) continue;               // beyond the block
      const ll x = i3(b, i, j, k);
#pragma unroll
      for (int v = 0; v < 5; v++) pl[v][jj][ii] = qb[n3 * v + x];
    }
    if constexpr (KW != 0) {
      if (cell) {
#pragma unroll
        for (int v = 0; v < 5; v++) pl[v][ty + 2][tx + 2] = w1[v];
      }
    }
    __syncthreads();
    // (b) face fluxes
    if (fr < TJ && fj0 <= jm && fa <= b.imax) {   // i-face a = fa of row fj0
      const int a = fa;
      const bool fin = fl[fj0 + (ll)(b.jmax + 1) * k] != 0;
      const bool fout = fl[plI + fj0 + (ll)(b.jmax + 1) * k] != 0;
      T m2[5], m1[5], z0[5], p1[5], QR[5], QL[5], F[5], kx, ky, kz, ds, ubn;
#pragma unroll
      for (int v = 0; v < 5; v++) {
        m2[v] = pl[v][fr + 2][fc];
        m1[v] = pl[v][fr + 2][fc + 1];
        z0[v] = pl[v][fr + 2][fc + 2];
        p1[v] = pl[v][fr + 2][fc + 3];
      }
      muscl_face<5, T, ILIM, FO>(p, nl, a, b.imax, fin, fout, m2, m1, z0, p1, QR, QL);
      face_metric<T>(p, si + mo + i2(b, a, fj0, k) * p.lmet, kx, ky, kz, ds, ubn);
      roe_face<5, T, EF>(p, QR, QL, kx, ky, kz, ds, ubn, F);
#pragma unroll
      for (int v = 0; v < 5; v++) fi[v][fr][fc] = F[v];
    }
    if (tx < TI && ci <= im && cj <= b.jmax) {   // j-face a = cj of column ci
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
      muscl_face<5, T, ILIM, FO>(p, nl, a, b.jmax, fin, fout, m2, m1, z0, p1, QR, QL);
      face_metric<T>(p, sj + mo + i2(b, ci, a, k) * p.lmet, kx, ky, kz, ds, ubn);
      roe_face<5, T, EF>(p, QR, QL, kx, ky, kz, ds, ubn, F);
#pragma unroll
      for (int v = 0; v < 5; v++) fj[v][ty][tx] = F[v];
    }
    T Fkn[5];
    if (cell) {   // k-face k+1 from q(k-1 .. k+2) of the column
      const ll x = i3(b, ci, cj, k);
      T QR[5], QL[5], kx, ky, kz, ds, ubn;
      if constexpr (KW != 0) {
        muscl_face<5, T, ILIM, FO>(p, nl, k + 1, b.kmax, kfin, kfout, w0, w1, w2, w3, QR, QL);
      } else {
        T a_m1[5], a_0[5], a_p1[5], a_p2[5];
#pragma unroll
        for (int v = 0; v < 5; v++) {
          const T* qv = qb + n3 * v;
          a_m1[v] = qv[x - kstep];
          a_0[v] = pl[v][ty + 2][tx + 2];
          a_p1[v] = qv[x + kstep];
          a_p2[v] = qv[x + 2 * kstep];
        }
        muscl_face<5, T, ILIM, FO>(p, nl, k + 1, b.kmax, kfin, kfout, a_m1, a_0, a_p1, a_p2, QR, QL);
      }
      face_metric<T>(p, sk + mo + i2(b, ci, cj, k + 1) * p.lmet, kx, ky, kz, ds, ubn);
      roe_face<5, T, EF>(p, QR, QL, kx, ky, kz, ds, ubn, Fkn);
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
      if constexpr (KW != 0) {
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
}

// Van Albada, second order, no entropy fix or eigenvalue floor: the
// reference's limit_i/j/k compute one slope per cell, dqmx(c) from dq(c+1)
// and dq(c), and hand it to both faces of the cell, dqp(c) = -dqmx/2 and
// dqm(c+1) = dqmx/2 (src/flux.f90:3584-3604, 3723-3743, 3862-3882).  This
// kernel keeps that sharing, where k_residual_tile evaluates both slopes of
// every face (half its limiter divisions are repeats):
//   (a) q(k) of the tile and its ring is staged in LDS, as in k_residual_tile;
//   (b) the half-slopes h = dqmx/2 of cells i0-1 .. i0+TI of every tile row
//       and of cells j0-1 .. j0+TJ of every tile column go to LDS;
//   (c) face a takes QR = q(a) - h(a), QL = q(a-1) + h(a-1), then the
//       strong-BC overrides of muscl_b*; the k-face carries h(k) of its
//       column from one plane to the next;
//   (d) the face fluxes replace the slopes in the same LDS (one more
//       barrier) and the cell sums them in the reference order.
template <typename T, int TI, int TJ, int MINW>
__global__ void __launch_bounds__(32 * (TJ + 1), MINW) k_residual_sva(Par p, int nl, const Blk* __restrict__ blks,
                                                                const Tile* __restrict__ tiles,
                                                                const T* __restrict__ q, const T* __restrict__ si,
                                                                const T* __restrict__ sj, const T* __restrict__ sk,
                                                                double* __restrict__ res,
                                                                const unsigned char* __restrict__ sbc,
                                                                const ll* __restrict__ sbcoff) {
  constexpr int NT = 32 * (TJ + 1), PW = TI + 4, PH = TJ + 4;
  static_assert(TI == 31 || TI == 32, "tiles are 31 or 32 cells wide");
  constexpr int NI = TJ * (TI + 2), NJ = (TJ + 2) * TI;               // slope cells along i / along j
  constexpr int NF = 5 * TJ * (TI + 1) + 5 * (TJ + 1) * TI, NH = 5 * (NI + NJ);
  __shared__ T pl[5][PH][PW];
  __shared__ T buf[NF > NH ? NF : NH];
  T(*fi)[TJ][TI + 1] = reinterpret_cast<T(*)[TJ][TI + 1]>(buf);
  T(*fj)[TJ + 1][TI] = reinterpret_cast<T(*)[TJ + 1][TI]>(buf + 5 * TJ * (TI + 1));
  const T* hi = buf;            // h of row r, cell i0-1+c: hi[v * NI + r * (TI + 2) + c]
  const T* hj = buf + 5 * NI;   // h of column c, cell j0-1+r: hj[v * NJ + r * TI + c]
  const Tile t = tiles[blockIdx.x];
  const Blk b = blks[t.block];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int im = b.imax - 1, jm = b.jmax - 1, km = b.kmax - 1;
  const int ntx = (im + TI - 1) / TI, nty = (jm + TJ - 1) / TJ, nks = (km + RT_KSEG - 1) / RT_KSEG;
  int item = t.base;
  const int n = item / (ntx * nty * nks);
  item -= n * (ntx * nty * nks);
  const int ks_i = item / (ntx * nty);
  item -= ks_i * (ntx * nty);
  const int ty_i = item / ntx, tx_i = item - ty_i * ntx;
  const int i0 = 1 + tx_i * TI, j0 = 1 + ty_i * TJ;
  const int ks = 1 + ks_i * RT_KSEG, ke = min(km, ks + RT_KSEG - 1);
  const ll n3 = n3of(b);
  const T* qb = q + b.o3 * p.npde * p.nsnap + n3 * (ll)p.npde * n;
  double* rb = res + b.o3 * p.npde * p.nsnap + n3 * (ll)p.npde * n;
  const int nh = n * p.hbmove;
  const ll mo = b.o2 * p.lmet * p.nsnap_met + (ll)nh * n2of(b) * p.lmet;
  const unsigned char* fl = sbc + sbcoff[t.block];
  const ll plI = (ll)(b.jmax + 1) * (b.kmax + 1);
  const ll offJ = 2 * plI, plJ = (ll)(b.imax + 1) * (b.kmax + 1);
  const ll offK = offJ + 2 * plJ, plK = (ll)(b.imax + 1) * (b.jmax + 1);
  const int ci = i0 + tx, cj = j0 + ty;
  const bool cell = tx < TI && ty < TJ && ci <= im && cj <= jm;
  // i-face thread: face i0 + fc of row j0 + fr, TI + 1 faces per row (for
  // 31-cell tiles the lane (tx, ty) itself)
  const int fr = threadIdx.x / (TI + 1), fc = threadIdx.x - fr * (TI + 1);
  const int fa = i0 + fc, fj0 = j0 + fr;
  const ll kstep = (ll)(b.imax + 3) * (b.jmax + 3);
  const T epsq = (T)p.epslimf * (T)p.epslimf;
  T Fkp[5], hk[5];
  bool kfin = false, kfout = false;
  if (cell) {   // k-face ks from q(ks-2 .. ks+1): h(ks-1), h(ks)
    kfin = fl[offK + ci + (ll)(b.imax + 1) * cj] != 0;
    kfout = fl[offK + plK + ci + (ll)(b.imax + 1) * cj] != 0;
    const ll c0 = i3(b, ci, cj, ks);
    const int a = ks;
    const bool sm1 = (a - 1 == 1 && kfin) || (a - 1 == b.kmax && kfout);
    const bool s0 = (a == 1 && kfin) || (a == b.kmax && kfout);
    const bool sp1 = (a + 1 == 1 && kfin) || (a + 1 == b.kmax && kfout);
    T QR[5], QL[5], kx, ky, kz, ds, ubn;
#pragma unroll
    for (int v = 0; v < 5; v++) {
      const T* qv = qb + n3 * v;
      const T qm2 = qv[c0 - 2 * kstep], qm1 = qv[c0 - kstep], q0 = qv[c0], qp1 = qv[c0 + kstep];
      T dm1 = qm1 - qm2, d0 = q0 - qm1, dp1 = qp1 - q0;
      if (sm1) dm1 = dm1 * 2;
      if (s0) d0 = d0 * 2;
      if (sp1) dp1 = dp1 * 2;
      hk[v] = va_slope<T>(epsq, dp1, d0) / 2;
      QR[v] = q0 - hk[v];
      QL[v] = qm1 + va_slope<T>(epsq, d0, dm1) / 2;
      if (a == 1 && kfin) QR[v] = QL[v] = qm1;
      if (a == b.kmax && kfout) QR[v] = QL[v] = q0;
    }
    face_metric<T>(p, sk + mo + i2(b, ci, cj, ks) * p.lmet, kx, ky, kz, ds, ubn);
    roe_face<5, T, 0>(p, QR, QL, kx, ky, kz, ds, ubn, Fkp);
  }
  for (int k = ks; k <= ke; k++) {
    // (a) stage plane k
    for (int pidx = threadIdx.x; pidx < PW * PH; pidx += NT) {
      const int ii = pidx % PW, jj = pidx / PW;
      const bool in_i = ii >= 2 && ii <= TI + 1, in_j = jj >= 2 && jj <= TJ + 1;
      if (!in_i && !in_j) continue;   // corner: not used
      const int i = i0 - 2 + ii, j = j0 - 2 + jj;
      if (i > b.imax + 1 || j > b.jmax + 1) continue;   // beyond the block
      const ll x = i3(b, i, j, k);
#pragma unroll
      for (int v = 0; v < 5; v++) pl[v][jj][ii] = qb[n3 * v + x];
    }
    __syncthreads();
    // (b) half-slopes of the row cells (along i) and column cells (along j);
    //     cells beyond the block compute on unused LDS and are never read
    for (int it = threadIdx.x; it < NI + NJ; it += NT) {
      int c0, st, a, L, ho, hs;
      bool fin, fout;
      if (it < NI) {
        const int r = it / (TI + 2), c = it - r * (TI + 2);
        c0 = (r + 2) * PW + c + 1;
        st = 1;
        a = i0 - 1 + c;
        L = b.imax;
        const ll fx = min(j0 + r, b.jmax) + (ll)(b.jmax + 1) * k;
        fin = fl[fx] != 0;
        fout = fl[plI + fx] != 0;
        ho = it;
        hs = NI;
      } else {
        const int jt = it - NI, r = jt / TI, c = jt - r * TI;
        c0 = (r + 1) * PW + c + 2;
        st = PW;
        a = j0 - 1 + r;
        L = b.jmax;
        const ll fx = offJ + min(i0 + c, b.imax) + (ll)(b.imax + 1) * k;
        fin = fl[fx] != 0;
        fout = fl[plJ + fx] != 0;
        ho = 5 * NI + jt;
        hs = NJ;
      }
      const bool s0 = (a == 1 && fin) || (a == L && fout);            // face a (below the cell)
      const bool s1 = (a + 1 == 1 && fin) || (a + 1 == L && fout);    // face a+1 (above)
#pragma unroll
      for (int v = 0; v < 5; v++) {
        const T* pv = &pl[v][0][0];
        const T qm = pv[c0 - st], qc = pv[c0], qp = pv[c0 + st];
        T d2 = qc - qm, d1 = qp - qc;
        if (s0) d2 = d2 * 2;
        if (s1) d1 = d1 * 2;
        buf[ho + v * hs] = va_slope<T>(epsq, d1, d2) / 2;
      }
    }
    __syncthreads();
    // (c) face fluxes into registers
    T Fi[5], Fj[5], Fkn[5];
    const bool iface = fr < TJ && fj0 <= jm && fa <= b.imax;
    const bool jface = tx < TI && ci <= im && cj <= b.jmax;
    if (iface) {   // i-face a = fa of row fj0
      const int a = fa;
      const bool fin = fl[fj0 + (ll)(b.jmax + 1) * k] != 0;
      const bool fout = fl[plI + fj0 + (ll)(b.jmax + 1) * k] != 0;
      T QR[5], QL[5], kx, ky, kz, ds, ubn;
#pragma unroll
      for (int v = 0; v < 5; v++) {
        const T qa = pl[v][fr + 2][fc + 2], qm = pl[v][fr + 2][fc + 1];
        QR[v] = qa - hi[v * NI + fr * (TI + 2) + fc + 1];
        QL[v] = qm + hi[v * NI + fr * (TI + 2) + fc];
        if (a == 1 && fin) QR[v] = QL[v] = qm;
        if (a == b.imax && fout) QR[v] = QL[v] = qa;
      }
      face_metric<T>(p, si + mo + i2(b, a, fj0, k) * p.lmet, kx, ky, kz, ds, ubn);
      roe_face<5, T, 0>(p, QR, QL, kx, ky, kz, ds, ubn, Fi);
    }
    if (jface) {   // j-face a = cj of column ci
      const int a = cj;
      const bool fin = fl[offJ + ci + (ll)(b.imax + 1) * k] != 0;
      const bool fout = fl[offJ + plJ + ci + (ll)(b.imax + 1) * k] != 0;
      T QR[5], QL[5], kx, ky, kz, ds, ubn;
#pragma unroll
      for (int v = 0; v < 5; v++) {
        const T qa = pl[v][ty + 2][tx + 2], qm = pl[v][ty + 1][tx + 2];
        QR[v] = qa - hj[v * NJ + (ty + 1) * TI + tx];
        QL[v] = qm + hj[v * NJ + ty * TI + tx];
        if (a == 1 && fin) QR[v] = QL[v] = qm;
        if (a == b.jmax && fout) QR[v] = QL[v] = qa;
      }
      face_metric<T>(p, sj + mo + i2(b, ci, a, k) * p.lmet, kx, ky, kz, ds, ubn);
      roe_face<5, T, 0>(p, QR, QL, kx, ky, kz, ds, ubn, Fj);
    }
    if (cell) {   // k-face k+1: h(k) carried, h(k+1) from q(k .. k+2)
      const ll x = i3(b, ci, cj, k);
      const int a = k + 1;
      const bool s0 = (a == 1 && kfin) || (a == b.kmax && kfout);
      const bool s1 = (a + 1 == 1 && kfin) || (a + 1 == b.kmax && kfout);
      T QR[5], QL[5], kx, ky, kz, ds, ubn;
#pragma unroll
      for (int v = 0; v < 5; v++) {
        const T* qv = qb + n3 * v;
        const T q0 = pl[v][ty + 2][tx + 2], q1 = qv[x + kstep], q2 = qv[x + 2 * kstep];
        T d2 = q1 - q0, d1 = q2 - q1;
        if (s0) d2 = d2 * 2;
        if (s1) d1 = d1 * 2;
        const T h1 = va_slope<T>(epsq, d1, d2) / 2;
        QR[v] = q1 - h1;
        QL[v] = q0 + hk[v];
        hk[v] = h1;
        if (a == 1 && kfin) QR[v] = QL[v] = q0;
        if (a == b.kmax && kfout) QR[v] = QL[v] = q1;
      }
      face_metric<T>(p, sk + mo + i2(b, ci, cj, a) * p.lmet, kx, ky, kz, ds, ubn);
      roe_face<5, T, 0>(p, QR, QL, kx, ky, kz, ds, ubn, Fkn);
    }
    __syncthreads();   // every slope read: the fluxes take their place
    if (iface) {
#pragma unroll
      for (int v = 0; v < 5; v++) fi[v][fr][fc] = Fi[v];
    }
    if (jface) {
#pragma unroll
      for (int v = 0; v < 5; v++) fj[v][ty][tx] = Fj[v];
    }
    __syncthreads();
    // (d) accumulate in the reference order and store
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
    }
  }
}

template <typename T, int TJ, int ILIM, int FO, int EF, int KW, int MINW = 1, int TI = 31>
static void launch_tile(Solve& s, Level& L, int ntile, const Tile* tiles, const T* q, const T* si, const T* sj,
                        const T* sk, double* res) {
  hipLaunchKernelGGL((k_residual_tile<T, TI, TJ, ILIM, FO, EF, KW, MINW>), dim3(ntile), dim3(32 * (TJ + 1)), 0,
                     s.stream, s.par, L.nl, L.dblk, tiles, q, si, sj, sk, res, L.sbc, L.sbc_doff);
}

template <typename T, int TJ, int MINW = 1, int TI = 31>
static void launch_sva(Solve& s, Level& L, int ntile, const Tile* tiles, const T* q, const T* si, const T* sj,
                       const T* sk, double* res) {
  hipLaunchKernelGGL((k_residual_sva<T, TI, TJ, MINW>), dim3(ntile), dim3(32 * (TJ + 1)), 0, s.stream, s.par, L.nl,
                     L.dblk, tiles, q, si, sj, sk, res, L.sbc, L.sbc_doff);
}

// The instances built: van Albada (the default limiter) specialised, every
// other option through the run-time flags.  Returns false when `variant`
// names no instance for this level (the caller uses the generic kernel).
template <typename T>
static bool dispatch(Solve& s, Level& L, int variant, const T* q, const T* si, const T* sj, const T* sk, double* res) {
  const Par& p = s.par;
  const bool w32 = variant == RESID_TILE32_W4 || variant == RESID_SVA32;
  const int tk = w32 ? TK_T32 : (variant == RESID_TILE15 || variant == RESID_SVA15 ? TK_FUSED15 : TK_FUSED);
  const int nt = L.ntile[tk];
  if (!nt) return true;
  const Tile* tl = L.tile[tk];
  const bool fo = L.nl >= p.nl_crs;
  const bool plain = p.ilimf == 4 && !p.efx_setup_f && !p.eig_limit_f;
  if (variant == RESID_TILE7) {
    if (!plain) launch_tile<T, 7, -1, -1, -1, 0>(s, L, nt, tl, q, si, sj, sk, res);
    else if (fo) launch_tile<T, 7, 4, 1, 0, 0>(s, L, nt, tl, q, si, sj, sk, res);
    else launch_tile<T, 7, 4, 0, 0, 0>(s, L, nt, tl, q, si, sj, sk, res);
  } else if (variant == RESID_TILE7_REG) {
    if (!plain) launch_tile<T, 7, -1, -1, -1, 1>(s, L, nt, tl, q, si, sj, sk, res);
    else if (fo) launch_tile<T, 7, 4, 1, 0, 1>(s, L, nt, tl, q, si, sj, sk, res);
    else launch_tile<T, 7, 4, 0, 0, 1>(s, L, nt, tl, q, si, sj, sk, res);
  } else if (variant == RESID_TILE7_W4) {
    if (!plain) launch_tile<T, 7, -1, -1, -1, 0, 4>(s, L, nt, tl, q, si, sj, sk, res);
    else if (fo) launch_tile<T, 7, 4, 1, 0, 0, 4>(s, L, nt, tl, q, si, sj, sk, res);
    else launch_tile<T, 7, 4, 0, 0, 0, 4>(s, L, nt, tl, q, si, sj, sk, res);
  } else if (variant == RESID_TILE32_W4) {
    if (!plain) launch_tile<T, 7, -1, -1, -1, 0, 4, 32>(s, L, nt, tl, q, si, sj, sk, res);
    else if (fo) launch_tile<T, 7, 4, 1, 0, 0, 4, 32>(s, L, nt, tl, q, si, sj, sk, res);
    else launch_tile<T, 7, 4, 0, 0, 0, 4, 32>(s, L, nt, tl, q, si, sj, sk, res);
  } else if (variant == RESID_SVA32) {
    if (plain && !fo) launch_sva<T, 7, 1, 32>(s, L, nt, tl, q, si, sj, sk, res);
    else if (!plain) launch_tile<T, 7, -1, -1, -1, 0, 4, 32>(s, L, nt, tl, q, si, sj, sk, res);
    else launch_tile<T, 7, 4, 1, 0, 0, 4, 32>(s, L, nt, tl, q, si, sj, sk, res);
  } else if (variant == RESID_SVA7 || variant == RESID_SVA15 || variant == RESID_SVA7_W4) {
    const bool t15 = variant == RESID_SVA15;
    if (plain && !fo) {
      if (t15) launch_sva<T, 15>(s, L, nt, tl, q, si, sj, sk, res);
      else if (variant == RESID_SVA7_W4) launch_sva<T, 7, 4>(s, L, nt, tl, q, si, sj, sk, res);
      else launch_sva<T, 7>(s, L, nt, tl, q, si, sj, sk, res);
    } else if (t15) {
      if (!plain) launch_tile<T, 15, -1, -1, -1, 0>(s, L, nt, tl, q, si, sj, sk, res);
      else launch_tile<T, 15, 4, 1, 0, 0>(s, L, nt, tl, q, si, sj, sk, res);
    } else {
      if (!plain) launch_tile<T, 7, -1, -1, -1, 0>(s, L, nt, tl, q, si, sj, sk, res);
      else launch_tile<T, 7, 4, 1, 0, 0>(s, L, nt, tl, q, si, sj, sk, res);
    }
  } else if (variant == RESID_TILE15) {
    if (!plain) launch_tile<T, 15, -1, -1, -1, 0>(s, L, nt, tl, q, si, sj, sk, res);
    else if (fo) launch_tile<T, 15, 4, 1, 0, 0>(s, L, nt, tl, q, si, sj, sk, res);
    else launch_tile<T, 15, 4, 0, 0, 0>(s, L, nt, tl, q, si, sj, sk, res);
  } else {
    return false;
  }
  return true;
}

}  // namespace exact / contracted

#ifdef COSAHIP_CONTRACTED
hipError_t launch_residual_tile_contracted(Solve& s, Level& L, int variant, const double* q, const double* si,
                                           const double* sj, const double* sk, double* res) {
  if (!contracted::dispatch<double>(s, L, variant, q, si, sj, sk, res)) return hipErrorInvalidValue;
  return hipGetLastError();
}
#else
// Work items of one block for a TI x TJ tile table: (i tiles) x (j tiles) x
// (k segments) x snapshots, i tiles fastest.
ll residual_tile_items(int TI, int TJ, int imax, int jmax, int kmax, int nsnap) {
  const ll im = imax - 1, jm = jmax - 1, km = kmax - 1;
  return (ll)nsnap * ((im + TI - 1) / TI) * ((jm + TJ - 1) / TJ) * ((km + RT_KSEG - 1) / RT_KSEG);
}

hipError_t launch_residual_tile(Solve& s, Level& L, int variant, const double* q, const double* si,
                                const double* sj, const double* sk, double* res) {
  if (!exact::dispatch<double>(s, L, variant, q, si, sj, sk, res)) return hipErrorInvalidValue;
  return hipGetLastError();
}
hipError_t launch_residual_tile(Solve& s, Level& L, int variant, const float* q, const float* si, const float* sj,
                                const float* sk, double* res) {
  if (!exact::dispatch<float>(s, L, variant, q, si, sj, sk, res)) return hipErrorInvalidValue;
  return hipGetLastError();
}
#endif

}  // namespace cosahip
