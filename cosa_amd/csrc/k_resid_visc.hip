// k_resid_visc.hip -- the inviscid part of the viscous / k-omega residual in
// one pass over all three directions, with the stored viscous face fluxes of
// every direction differenced in the same pass, for gfx950.
//
// The reference's viscous residual adds, for idir = 1, 2, 3 in turn, the
// inviscid Roe/MUSCL flux difference of the direction and then its viscous
// flux difference (roflux_b + vflux + resid, src/flux.f90:45-387, 878-1182;
// src/cosa.f90:1149-1211), so per cell and variable
//   res = ((((((((((0 + Fi(c+1)) - Fi(c)) + Vi(c+1)) - Vi(c)) + Fj(c+1)) - Fj(c))
//         + Vj(c+1)) - Vj(c)) + Fk(c+1)) - Fk(c)) + Vk(c+1)) - Vk(c)
// in FP64.  k_vflux stores the viscous face fluxes V of each direction in a
// slot of its own; this kernel forms the inviscid face fluxes F of a TI x TJ
// tile of cells plane by plane (q(k) staged in LDS with its 2-cell ring, the
// i and j faces from LDS by the tile's threads -- the TI + 1 faces of a row
// spread over the work-group --, the k face by the cell's own
// thread from q(k-1 .. k+2)), with the k-omega transport of roflux_b
// (roe_turb, fed the block's left-over h0l/h0r of the direction from h0tab),
// and sums F and V in the order above.  The per-direction kernels (k_flux_i,
// k_flux_march + add_vdiff, k_turb_src) do the same operations on the same
// operands, so the residuals are the same bits; what this arrangement saves
// is the residual's read-modify-write per direction and two of the three
// passes over q and its halo.
#include "flux_face.h"
#include "kernels.h"

namespace cosahip {

template <typename T, int NV, bool VIS, int TI, int TJ, int ILIM, int FO, int EF>
__global__ void __launch_bounds__(32 * (TJ + 1)) k_residual_tile_v(
    Par p, int nl, const Blk* __restrict__ blks, const Tile* __restrict__ tiles, const T* __restrict__ q,
    const T* __restrict__ si, const T* __restrict__ sj, const T* __restrict__ sk, double* __restrict__ res,
    const unsigned char* __restrict__ sbc, const ll* __restrict__ sbcoff, const double* __restrict__ h0tab,
    const T* __restrict__ vf, ll vslot) {
  constexpr int NT = 32 * (TJ + 1), PW = TI + 4, PH = TJ + 4;
  static_assert(TI == 31 || TI == 32, "tiles are 31 or 32 cells wide");
  __shared__ T pl[NV][PH][PW];
  __shared__ T fi[NV][TJ][TI + 1];
  __shared__ T fj[NV][TJ + 1][TI];
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
  // i-face thread: face i0 + fc of row j0 + fr, TI + 1 faces per row
  const int fr = threadIdx.x / (TI + 1), fc = threadIdx.x - fr * (TI + 1);
  const int fa = i0 + fc, fj0 = j0 + fr;
  const ll kstep = (ll)(b.imax + 3) * (b.jmax + 3);
  // h0l / h0r the k-omega flux of each direction takes (see roe_turb)
  T h0[3][2] = {{0, 0}, {0, 0}, {0, 0}};
  if constexpr (NV > 5) {
    const double* h = h0tab + (ll)(t.block * p.nsnap + n) * 6;
#pragma unroll
    for (int d = 0; d < 3; d++) {
      h0[d][0] = (T)h[2 * d];
      h0[d][1] = (T)h[2 * d + 1];
    }
  }
  // viscous face fluxes vflux(imax, jmax, kmax, npde, 0:2N) of direction d
  // in slot d - 1
  const ll fst = (ll)b.imax * b.jmax * b.kmax;
  const T* vb = VIS ? vf + b.o0 * p.npde * p.nsnap + fst * (ll)p.npde * n : nullptr;
  T Fkp[NV];
  bool kfin = false, kfout = false;
  if (cell) {   // k-face ks
    kfin = fl[offK + ci + (ll)(b.imax + 1) * cj] != 0;
    kfout = fl[offK + plK + ci + (ll)(b.imax + 1) * cj] != 0;
    const ll c0 = i3(b, ci, cj, ks);
    T a_m2[NV], a_m1[NV], a_0[NV], a_p1[NV], QR[NV], QL[NV], kx, ky, kz, ds, ubn;
#pragma unroll
    for (int v = 0; v < NV; v++) {
      const T* qv = qb + n3 * v;
      a_m2[v] = qv[c0 - 2 * kstep];
      a_m1[v] = qv[c0 - kstep];
      a_0[v] = qv[c0];
      a_p1[v] = qv[c0 + kstep];
    }
    muscl_face<NV, T, ILIM, FO>(p, nl, ks, b.kmax, kfin, kfout, a_m2, a_m1, a_0, a_p1, QR, QL);
    face_metric<T>(p, sk + mo + i2(b, ci, cj, ks) * p.lmet, kx, ky, kz, ds, ubn);
    roe_face<NV, T, EF>(p, QR, QL, kx, ky, kz, ds, ubn, Fkp);
    if constexpr (NV > 5) roe_turb<T>(p, QR, QL, kx, ky, kz, ds, ubn, h0[2][0], h0[2][1], Fkp);
  }
  for (int k = ks; k <= ke; k++) {
    // (a) stage plane k with its ring (no corners)
    for (int pidx = threadIdx.x; pidx < PW * PH; pidx += NT) {
      const int ii = pidx % PW, jj = pidx / PW;
      const bool in_i = ii >= 2 && ii <= TI + 1, in_j = jj >= 2 && jj <= TJ + 1;
      if (!in_i && !in_j) continue;   // corner: not used
      const int i = i0 - 2 + ii, j = j0 - 2 + jj;
      if (i > b.imax + 1 || j > b.jmax + 1) continue;   // beyond the block
      const ll x = i3(b, i, j, k);
#pragma unroll
      for (int v = 0; v < NV; v++) pl[v][jj][ii] = qb[n3 * v + x];
    }
    __syncthreads();
    // (b) face fluxes: i-face ci of row cj, j-face cj of column ci, and the
    //     cell's k-face k + 1
    if (fr < TJ && fj0 <= jm && fa <= b.imax) {
      const int a = fa;
      const bool fin = fl[fj0 + (ll)(b.jmax + 1) * k] != 0;
      const bool fout = fl[plI + fj0 + (ll)(b.jmax + 1) * k] != 0;
      T m2[NV], m1[NV], z0[NV], p1[NV], QR[NV], QL[NV], F[NV], kx, ky, kz, ds, ubn;
#pragma unroll
      for (int v = 0; v < NV; v++) {
        m2[v] = pl[v][fr + 2][fc];
        m1[v] = pl[v][fr + 2][fc + 1];
        z0[v] = pl[v][fr + 2][fc + 2];
        p1[v] = pl[v][fr + 2][fc + 3];
      }
      muscl_face<NV, T, ILIM, FO>(p, nl, a, b.imax, fin, fout, m2, m1, z0, p1, QR, QL);
      face_metric<T>(p, si + mo + i2(b, a, fj0, k) * p.lmet, kx, ky, kz, ds, ubn);
      roe_face<NV, T, EF>(p, QR, QL, kx, ky, kz, ds, ubn, F);
      if constexpr (NV > 5) roe_turb<T>(p, QR, QL, kx, ky, kz, ds, ubn, h0[0][0], h0[0][1], F);
#pragma unroll
      for (int v = 0; v < NV; v++) fi[v][fr][fc] = F[v];
    }
    if (tx < TI && ci <= im && cj <= b.jmax) {
      const int a = cj;
      const bool fin = fl[offJ + ci + (ll)(b.imax + 1) * k] != 0;
      const bool fout = fl[offJ + plJ + ci + (ll)(b.imax + 1) * k] != 0;
      T m2[NV], m1[NV], z0[NV], p1[NV], QR[NV], QL[NV], F[NV], kx, ky, kz, ds, ubn;
#pragma unroll
      for (int v = 0; v < NV; v++) {
        m2[v] = pl[v][ty][tx + 2];
        m1[v] = pl[v][ty + 1][tx + 2];
        z0[v] = pl[v][ty + 2][tx + 2];
        p1[v] = pl[v][ty + 3][tx + 2];
      }
      muscl_face<NV, T, ILIM, FO>(p, nl, a, b.jmax, fin, fout, m2, m1, z0, p1, QR, QL);
      face_metric<T>(p, sj + mo + i2(b, ci, a, k) * p.lmet, kx, ky, kz, ds, ubn);
      roe_face<NV, T, EF>(p, QR, QL, kx, ky, kz, ds, ubn, F);
      if constexpr (NV > 5) roe_turb<T>(p, QR, QL, kx, ky, kz, ds, ubn, h0[1][0], h0[1][1], F);
#pragma unroll
      for (int v = 0; v < NV; v++) fj[v][ty][tx] = F[v];
    }
    T Fkn[NV];
    if (cell) {
      const ll x = i3(b, ci, cj, k);
      T a_m1[NV], a_0[NV], a_p1[NV], a_p2[NV], QR[NV], QL[NV], kx, ky, kz, ds, ubn;
#pragma unroll
      for (int v = 0; v < NV; v++) {
        const T* qv = qb + n3 * v;
        a_m1[v] = qv[x - kstep];
        a_0[v] = pl[v][ty + 2][tx + 2];
        a_p1[v] = qv[x + kstep];
        a_p2[v] = qv[x + 2 * kstep];
      }
      muscl_face<NV, T, ILIM, FO>(p, nl, k + 1, b.kmax, kfin, kfout, a_m1, a_0, a_p1, a_p2, QR, QL);
      face_metric<T>(p, sk + mo + i2(b, ci, cj, k + 1) * p.lmet, kx, ky, kz, ds, ubn);
      roe_face<NV, T, EF>(p, QR, QL, kx, ky, kz, ds, ubn, Fkn);
      if constexpr (NV > 5) roe_turb<T>(p, QR, QL, kx, ky, kz, ds, ubn, h0[2][0], h0[2][1], Fkn);
    }
    __syncthreads();
    // (c) sum in the reference order and store
    if (cell) {
      const ll x = i3(b, ci, cj, k);
      const ll f = (ll)(ci - 1) + (ll)b.imax * ((ll)(cj - 1) + (ll)b.jmax * (ll)(k - 1));
#pragma unroll
      for (int v = 0; v < NV; v++) {
        double r = (0.0 + (double)fi[v][ty][tx + 1]) - (double)fi[v][ty][tx];
        if constexpr (VIS) {
          const T* v1 = vb + fst * v;
          r = (r + (double)v1[f + 1]) - (double)v1[f];
        }
        r = (r + (double)fj[v][ty + 1][tx]) - (double)fj[v][ty][tx];
        if constexpr (VIS) {
          const T* v2 = vb + vslot + fst * v;
          r = (r + (double)v2[f + b.imax]) - (double)v2[f];
        }
        r = (r + (double)Fkn[v]) - (double)Fkp[v];
        if constexpr (VIS) {
          const T* v3 = vb + 2 * vslot + fst * v;
          r = (r + (double)v3[f + (ll)b.imax * b.jmax]) - (double)v3[f];
        }
        rb[n3 * v + x] = r;
        Fkp[v] = Fkn[v];
      }
    }
  }
}

template <typename T, int NV, bool VIS, int ILIM, int FO, int EF>
static void launch_v(Solve& s, Level& L, const T* q, const T* si, const T* sj, const T* sk, double* res,
                     const T* vf, ll vslot) {
  hipLaunchKernelGGL((k_residual_tile_v<T, NV, VIS, 32, 7, ILIM, FO, EF>), dim3(L.ntile[TK_T32]), dim3(256), 0,
                     s.stream, s.par, L.nl, L.dblk, L.tile[TK_T32], q, si, sj, sk, res, L.sbc, L.sbc_doff, L.h0tab,
                     vf, vslot);
}

template <typename T, int NV, bool VIS>
static void dispatch_v(Solve& s, Level& L, const T* q, const T* si, const T* sj, const T* sk, double* res,
                       const T* vf, ll vslot) {
  const Par& p = s.par;
  const bool fo = L.nl >= p.nl_crs;
  // van Albada with no entropy fix or eigenvalue floor in either flux:
  // specialised; everything else through the run-time flags
  const bool plain = p.ilimf == 4 && !p.efx_setup_f && !p.eig_limit_f;
  if (!plain) launch_v<T, NV, VIS, -1, -1, -1>(s, L, q, si, sj, sk, res, vf, vslot);
  else if (fo) launch_v<T, NV, VIS, 4, 1, 0>(s, L, q, si, sj, sk, res, vf, vslot);
  else launch_v<T, NV, VIS, 4, 0, 0>(s, L, q, si, sj, sk, res, vf, vslot);
}

template <typename T>
hipError_t launch_residual_tile_visc(Solve& s, Level& L, const ResIn<T>& in, double* res, const T* vf, ll vslot) {
  if (!L.ntile[TK_T32]) return hipSuccess;
  const bool t = s.par.npde == 7;
  if (vf) {
    if (t) dispatch_v<T, 7, true>(s, L, in.q, in.s[0], in.s[1], in.s[2], res, vf, vslot);
    else dispatch_v<T, 5, true>(s, L, in.q, in.s[0], in.s[1], in.s[2], res, vf, vslot);
  } else {
    if (t) dispatch_v<T, 7, false>(s, L, in.q, in.s[0], in.s[1], in.s[2], res, vf, vslot);
    else dispatch_v<T, 5, false>(s, L, in.q, in.s[0], in.s[1], in.s[2], res, vf, vslot);
  }
  return hipGetLastError();
}
template hipError_t launch_residual_tile_visc<double>(Solve&, Level&, const ResIn<double>&, double*, const double*,
                                                      ll);
template hipError_t launch_residual_tile_visc<float>(Solve&, Level&, const ResIn<float>&, double*, const float*,
                                                     ll);

}  // namespace cosahip
