// k_turb.hip -- viscous fluxes and the Wilcox k-omega model of COSA's
// pseudo-time stage for gfx950: face-centred viscous/turbulent diffusion
// fluxes (q_face + q_der + rtst + vflux fused per face), their flux
// difference, the k-omega sources (q_face + q_der_c + turprod + srckw
// fused per cell), mut, the k/omega floors and the restriction-stage
// limiters.  Operation order follows the reference routines so results
// match the CPU path to the rounding of the math library's pow.
//
//   q_face / q_der  : src/flux.f90:2391-2801, 1813-2390
//   rtst / vflux    : src/flux.f90:2802-2948, 831-1182
//   q_der_c         : src/flux.f90:2949-3131
//   turprod / srckw : src/src.f90:2-272
//   tur_vis / fixq  : src/update.f90:2191-2337
//   lim_res_ome     : src/update.f90:2527-2835
//   lim_prod_mg     : src/src.f90:674-757
#include "kernels.h"
#include "viscous.h"

namespace cosahip {

// ------------------------------------------------------------------------
// Viscous fluxes of the faces of direction DIR, one thread per face:
// face average (q_face), xi/eta/zeta differences (q_der), Reynolds
// stresses (rtst, k-omega) and the flux (vflux) into vflux(imax,jmax,kmax,
// npde,0:2N).  Wall-like faces (strong-BC flag) take the halo value and
// one-sided differences as bq_face / bq_der do.
template <int DIR, int NV, typename T>
__global__ void __launch_bounds__(256) k_vflux(Par p, const Blk* __restrict__ blks, const Tile* __restrict__ tiles,
                                               const T* __restrict__ q, const T* __restrict__ mut,
                                               const T* __restrict__ sm, const T* __restrict__ xd,
                                               const T* __restrict__ ed, const T* __restrict__ zd,
                                               const unsigned char* __restrict__ sbc,
                                               const ll* __restrict__ sbcoff, T* __restrict__ vf,
                                               const double* __restrict__ dist) {
  const Tile t = tiles[blockIdx.x];
  const Blk b = blks[t.block];
  const int ic = DIR == 1 ? b.imax : b.imax - 1, jc = DIR == 2 ? b.jmax : b.jmax - 1,
            kc = DIR == 3 ? b.kmax : b.kmax - 1;
  const int nf = ic * jc * kc;
  const int item = t.base + threadIdx.x;
  if (item >= nf * p.nsnap) return;
  const int n = item / nf;
  int rem = item - n * nf;
  const int i = rem % ic + 1;
  rem /= ic;
  const int j = rem % jc + 1, k = rem / jc + 1;
  T F[NV];
  viscous_face<DIR, NV, T>(p, b, i, j, k, n, q, mut, sm, xd, ed, zd, sbc + sbcoff[t.block], dist, F);
  T* fb = vf + b.o0 * p.npde * p.nsnap;
  const ll fo = (ll)(i - 1) + (ll)b.imax * ((ll)(j - 1) + (ll)b.jmax * (ll)(k - 1));
  const ll fst = (ll)b.imax * b.jmax * b.kmax;
#pragma unroll
  for (int m = 0; m < NV; m++) fb[fo + fst * (m + (ll)p.npde * n)] = F[m];
}

// res(c) += F(c+1) - F(c) of the viscous face fluxes of direction DIR
template <int DIR, int NV, typename T>
__global__ void __launch_bounds__(256) k_vresid(Par p, const Blk* __restrict__ blks, const Tile* __restrict__ tiles,
                                                const T* __restrict__ vf, double* __restrict__ res) {
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
  const ll n3 = n3of(b);
  double* rb = res + b.o3 * p.npde * p.nsnap + n3 * (ll)p.npde * n + i3(b, i, j, k);
  const T* fb = vf + b.o0 * p.npde * p.nsnap;
  const ll fst = (ll)b.imax * b.jmax * b.kmax;
  const ll f0 = (ll)(i - 1) + (ll)b.imax * ((ll)(j - 1) + (ll)b.jmax * (ll)(k - 1));
  const ll f1 = f0 + (DIR == 1 ? 1 : (DIR == 2 ? (ll)b.imax : (ll)b.imax * b.jmax));
#pragma unroll
  for (int m = 0; m < NV; m++) {
    const ll o = fst * (m + (ll)p.npde * n);
    rb[n3 * m] = rb[n3 * m] + fb[f1 + o] - fb[f0 + o];
  }
}

// Viscous fluxes of direction 1 and their difference in one pass: one
// wavefront per i-row segment, lane l computes the viscous flux of face
// i0 + l and cell c takes F(c+1) from the neighbouring lane.  Same
// arithmetic as k_vflux<1> followed by k_vresid<1>.
template <int NV, typename T>
__global__ void __launch_bounds__(256) k_vres_i(Par p, const Blk* __restrict__ blks, const Tile* __restrict__ tiles,
                                                const T* __restrict__ q, const T* __restrict__ mut,
                                                const T* __restrict__ sm, const T* __restrict__ xd,
                                                const T* __restrict__ ed, const T* __restrict__ zd,
                                                const unsigned char* __restrict__ sbc, const ll* __restrict__ sbcoff,
                                                double* __restrict__ res, const double* __restrict__ dist) {
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
  T F[NV];
#pragma unroll
  for (int v = 0; v < NV; v++) F[v] = 0;
  if (a <= b.imax)
    viscous_face<1, NV, T>(p, b, a, j, k, n, q, mut, sm, xd, ed, zd, sbc + sbcoff[t.block], dist, F);
  T Fn[NV];
#pragma unroll
  for (int v = 0; v < NV; v++) Fn[v] = __shfl_down(F[v], 1, 64);
  const int c = a;
  if (lane < ROWSEG && c <= im) {
    const ll n3 = n3of(b);
    double* rb = res + b.o3 * p.npde * p.nsnap + n3 * (ll)p.npde * n + i3(b, c, j, k);
#pragma unroll
    for (int v = 0; v < NV; v++) rb[n3 * v] = rb[n3 * v] + (double)Fn[v] - (double)F[v];
  }
}

// Viscous fluxes of direction DIR = 2 or 3 and their difference in one
// pass: one thread per MSEG-cell segment of a line, marching its faces and
// carrying the previous face's flux.  Same arithmetic as k_vflux<DIR> +
// k_vresid<DIR>.
template <int DIR, int NV, typename T>
__global__ void __launch_bounds__(256) k_vres_march(Par p, const Blk* __restrict__ blks,
                                                    const Tile* __restrict__ tiles, const T* __restrict__ q,
                                                    const T* __restrict__ mut, const T* __restrict__ sm,
                                                    const T* __restrict__ xd, const T* __restrict__ ed,
                                                    const T* __restrict__ zd, const unsigned char* __restrict__ sbc,
                                                    const ll* __restrict__ sbcoff, double* __restrict__ res,
                                                    const double* __restrict__ dist) {
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
  double* rb = res + b.o3 * p.npde * p.nsnap + n3 * (ll)p.npde * n;
  const ll step = DIR == 2 ? (ll)(b.imax + 3) : (ll)(b.imax + 3) * (b.jmax + 3);
  const ll base = DIR == 2 ? i3(b, i, 0, t2) : i3(b, i, t2, 0);
  const unsigned char* fl = sbc + sbcoff[t.block];
  T Fp[NV];
  for (int a = a0; a <= a1; a++) {
    T F[NV];
    if (DIR == 2) viscous_face<2, NV, T>(p, b, i, a, t2, n, q, mut, sm, xd, ed, zd, fl, dist, F);
    else viscous_face<3, NV, T>(p, b, i, t2, a, n, q, mut, sm, xd, ed, zd, fl, dist, F);
    if (a > a0) {
      const ll c = base + (ll)(a - 1) * step;
#pragma unroll
      for (int v = 0; v < NV; v++) rb[c + n3 * v] = rb[c + n3 * v] + (double)F[v] - (double)Fp[v];
    }
#pragma unroll
    for (int v = 0; v < NV; v++) Fp[v] = F[v];
  }
}

// ------------------------------------------------------------------------
// Cell-centre averages of the metric derivatives used by the k-omega
// sources (src/cosa.f90 srckw_b / update_ttrm_b: xix = (xideri(1,i+1) +
// xideri(1,i)) / 2, xiy from xiderj(2, j, j+1), xiz from xiderk(3, k, k+1),
// and the same for eta and zeta), stored (P3, 9, nsnap_met) per block and
// metric snapshot: xix, xiy, xiz, etax, etay, etaz, zetax, zetay, zetaz.
// The same operations as when k_turb_src formed them itself, once per
// metric upload instead of every stage, and read as nine coalesced values
// instead of gathers from nine triples of which one component is used.
template <typename T>
__global__ void __launch_bounds__(256) k_cell_metrics(Par p, const Blk* __restrict__ blks,
                                                      const Tile* __restrict__ tiles, const T* const* __restrict__ xdp,
                                                      T* __restrict__ cmet) {
  const Tile t = tiles[blockIdx.x];
  const Blk b = blks[t.block];
  const int im = b.imax - 1, jm = b.jmax - 1, km = b.kmax - 1;
  const int nc = im * jm * km;
  const int item = t.base + threadIdx.x;
  if (item >= nc * p.nsnap) return;
  const int n = item / nc;
  if (n >= p.nsnap_met) return;
  int rem = item - n * nc;
  const int i = rem % im + 1;
  rem /= im;
  const int j = rem % jm + 1, k = rem / jm + 1;
  const ll n3 = n3of(b);
  const int nh = n;   // nsnap_met > 1 only with hbmove
  T* cm = cmet + b.o3 * 9 * p.nsnap_met + n3 * 9 * (ll)nh + i3(b, i, j, k);
  for (int e = 0; e < 3; e++) {      // xi, eta, zeta
    cm[(3 * e + 0) * n3] = (mder<T>(p, b, xdp[e], i + 1, j, k, nh)[0] + mder<T>(p, b, xdp[e], i, j, k, nh)[0]) / 2;
    cm[(3 * e + 1) * n3] = (mder<T>(p, b, xdp[3 + e], i, j + 1, k, nh)[1] + mder<T>(p, b, xdp[3 + e], i, j, k, nh)[1]) / 2;
    cm[(3 * e + 2) * n3] = (mder<T>(p, b, xdp[6 + e], i, j, k + 1, nh)[2] + mder<T>(p, b, xdp[6 + e], i, j, k, nh)[2]) / 2;
  }
}

// ------------------------------------------------------------------------
// k-omega sources at cell centres: the differences of the face averages
// (q_face + q_der_c, three directions; u, v, w, and for BSL/SST k and
// omega), turprod (finest level), then srckw (Wilcox) or update_ttrm +
// srckw_sst (BSL/SST, src/cosa.f90:1894-1900) into res(6:7).  SST keeps
// the velocity differences (dqc) for tur_vis_sst after the update.
// q primitive.
template <typename T>
__global__ void __launch_bounds__(256) k_turb_src(Par p, int nl, const Blk* __restrict__ blks,
                                                  const Tile* __restrict__ tiles, const T* __restrict__ q,
                                                  const T* __restrict__ mut, double* __restrict__ prd,
                                                  double* __restrict__ divvel, double* __restrict__ delpls,
                                                  double* __restrict__ res, const double* __restrict__ vol,
                                                  const T* const* __restrict__ xdp,
                                                  const unsigned char* __restrict__ sbc,
                                                  const ll* __restrict__ sbcoff, const T* __restrict__ vpre,
                                                  double* __restrict__ ttrm, double* __restrict__ dqc,
                                                  const double* __restrict__ dist, const T* __restrict__ cmet) {
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
  const ll n3 = n3of(b);
  const T* qb = q + b.o3 * p.npde * p.nsnap + n3 * (ll)p.npde * n;
  const ll c = i3(b, i, j, k);
  const ll st[4] = {0, 1, (ll)(b.imax + 3), (ll)(b.imax + 3) * (b.jmax + 3)};
  const unsigned char* fl = sbc + sbcoff[t.block];
  const ll so = b.o3 * p.nsnap + n3 * (ll)n + c;
  const bool prodlvl = nl == 1 || p.prd_all_lvls;
  T cdif = 0, gav = (T)p.gkw, bav = (T)p.bkw;
  if (prodlvl || p.sst) {
    // dq[d][s] = qf(face c+1) - qf(face c), s = u, v, w, k, omega
    T dq[3][5];
    const int nvar = p.sst ? 5 : 3;
    for (int d = 1; d <= 3; d++) {
      const int a = d == 1 ? i : (d == 2 ? j : k);
      const int L = d == 1 ? b.imax : (d == 2 ? b.jmax : b.kmax);
      const int t1 = d == 1 ? j : i, t2 = d == 3 ? j : k;
      // face a (cells c-1, c) and face a+1 (cells c, c+1)
      bool wlo = false, whi = false;
      if (a == 1) {
        const unsigned char f = d == 1 ? face_flag<1>(b, fl, 0, t1, t2)
                                       : (d == 2 ? face_flag<2>(b, fl, 0, t1, t2) : face_flag<3>(b, fl, 0, t1, t2));
        wlo = (f & 1) != 0;
      }
      if (a + 1 == L) {
        const unsigned char f = d == 1 ? face_flag<1>(b, fl, 1, t1, t2)
                                       : (d == 2 ? face_flag<2>(b, fl, 1, t1, t2) : face_flag<3>(b, fl, 1, t1, t2));
        whi = (f & 1) != 0;
      }
      for (int sv = 0; sv < nvar; sv++) {
        const T* qv = qb + n3 * (sv < 3 ? sv + 1 : sv + 2);
        const T flo = wlo ? qv[c - st[d]] : (qv[c - st[d]] + qv[c]) / 2;
        const T fhi = whi ? qv[c + st[d]] : (qv[c] + qv[c + st[d]]) / 2;
        dq[d - 1][sv] = fhi - flo;
      }
    }
    // cell-centre metric derivatives, averaged once per grid by k_cell_metrics
    const T* cm = cmet + b.o3 * 9 * p.nsnap_met + n3 * 9 * (ll)(n * p.hbmove) + c;
    const T xix = cm[0], xiy = cm[n3], xiz = cm[2 * n3];
    const T etax = cm[3 * n3], etay = cm[4 * n3], etaz = cm[5 * n3];
    const T zetax = cm[6 * n3], zetay = cm[7 * n3], zetaz = cm[8 * n3];
    if (prodlvl) {
      const T uxi = dq[0][0], ueta = dq[1][0], uzeta = dq[2][0];
      const T vxi = dq[0][1], veta = dq[1][1], vzeta = dq[2][1];
      const T wxi = dq[0][2], weta = dq[1][2], wzeta = dq[2][2];
      const T dudx = uxi * xix + ueta * etax + uzeta * zetax;
      const T dudy = uxi * xiy + ueta * etay + uzeta * zetay;
      const T dudz = uxi * xiz + ueta * etaz + uzeta * zetaz;
      const T dvdx = vxi * xix + veta * etax + vzeta * zetax;
      const T dvdy = vxi * xiy + veta * etay + vzeta * zetay;
      const T dvdz = vxi * xiz + veta * etaz + vzeta * zetaz;
      const T dwdx = wxi * xix + weta * etax + wzeta * zetax;
      const T dwdy = wxi * xiy + weta * etay + wzeta * zetay;
      const T dwdz = wxi * xiz + weta * etaz + wzeta * zetaz;
      const T divv = (dudx + dvdy + dwdz) * (T)p.turcmp;
      const T txx = dudx, txy = (dudy + dvdx) / 2, txz = (dudz + dwdx) / 2, tyy = dvdy, tyz = (dvdz + dwdy) / 2,
              tzz = dwdz;
      prd[so] = 2 * (txx * txx + tyy * tyy + tzz * tzz + 2 * (txy * txy + txz * txz + tyz * tyz) - divv * divv / 3);
      divvel[so] = divv;
      delpls[so] = T(2.0 / 3.0) * fmax(divv, T(0));
    }
    if (p.sst == 2) {
      // u, v, w differences along xi, eta, zeta for tur_vis_sst
      double* db = dqc + b.o3 * 9 * p.nsnap + n3 * 9 * (ll)n + c;
#pragma unroll
      for (int sv = 0; sv < 3; sv++)
#pragma unroll
        for (int d = 0; d < 3; d++) db[n3 * (3 * sv + d)] = (double)dq[d][sv];
    }
    if (p.sst) {
      // update_ttrm_b (src/update.f90:3083-3155): F1 blending
      const T rho = qb[c], tke = qb[c + 5 * n3], ome = qb[c + 6 * n3], pr = qb[c + 4 * n3];
      const T tt = (T)p.gamma * pr / rho;
      const T mu = suth<T>(p, tt);
      const T dtkedx = dq[0][3] * xix + dq[1][3] * etax + dq[2][3] * zetax;
      const T dtkedy = dq[0][3] * xiy + dq[1][3] * etay + dq[2][3] * zetay;
      const T dtkedz = dq[0][3] * xiz + dq[1][3] * etaz + dq[2][3] * zetaz;
      const T domedx = dq[0][4] * xix + dq[1][4] * etax + dq[2][4] * zetax;
      const T domedy = dq[0][4] * xiy + dq[1][4] * etay + dq[2][4] * zetay;
      const T domedz = dq[0][4] * xiz + dq[1][4] * etaz + dq[2][4] * zetaz;
      const T de = (T)dist[b.o1 + i1(b, i, j, k)] + (T)p.epst;
      T arg1 = fmax(sqrt(tke) / (T)0.09 / ome / de, (T)500 * mu / rho / ome / (de * de) * (T)p.machfs / (T)p.reyno);
      const T cdt = 2 * rho * (T)p.sigw2 / ome * (dtkedx * domedx + dtkedy * domedy + dtkedz * domedz);
      const T acdt = fmax(cdt, (T)1e-20);
      arg1 = fmin(arg1, 4 * rho * (T)p.sigw2 * tke / acdt / (de * de));
      arg1 = arg1 * arg1 * arg1 * arg1;   // arg1**4 as ((a a) a) a
      arg1 = fmin(arg1, (T)1e2);
      // FP32: exp(2 arg1) overflows above arg1 = 44; F1 is 1 to float precision from arg1 = 9 on
      const T e2 = sizeof(T) == 4 ? exp(fmin(2 * arg1, (T)80)) : exp(2 * arg1);
      const T f1 = (e2 - 1) / (e2 + 1);
      cdif = cdt * (1 - f1);
      gav = f1 * (T)p.g1 + (1 - f1) * (T)p.g2;
      bav = f1 * (T)p.b1 + (1 - f1) * (T)p.b2;
      double* tb = ttrm + b.o3 * 3 * p.nsnap + n3 * 3 * (ll)n + c;
      tb[0] = (double)cdif;
      tb[n3] = (double)gav;
      tb[2 * n3] = (double)bav;
    }
  }
  // srckw_b / srckw_sst_b
  const T fact = T(2.0 / 3.0), mor = (T)(p.machfs / p.reyno);
  const T rho = qb[c], tke = qb[c + 5 * n3], ome = qb[c + 6 * n3];
  const T mu_t = mut[so];
  const T divv = (T)divvel[so];
  T prod = (T)prd[so], prodw = (T)prd[so];
  if (p.lim_prodtke && prodlvl) {
    const T tstdis = (T)p.prdlim * (T)p.bstrkw * rho * tke * ome;
    if (mor * mu_t * prod > tstdis) {
      prod = tstdis / mor / mu_t;
      if (p.lim_prodomega) prodw = prod;
    }
  }
  const T vv = (T)vol[b.o1 + i1(b, i, j, k)];
  double* rb = res + b.o3 * p.npde * p.nsnap + n3 * (ll)p.npde * n + c;
  // vpre: first the difference of the direction-3 viscous face fluxes
  if (vpre) add_vdiff<7, T>(p, b, vpre, 3, i, j, k, n, rb, n3);
  rb[5 * n3] = rb[5 * n3] - vv * (mor * mu_t * prod - fact * divv * rho * tke - (T)p.bstrkw * rho * tke * ome);
  if (p.sst)
    rb[6 * n3] = rb[6 * n3] - vv * (gav * (T)p.cmu * rho * prodw - gav * fact * divv * rho * ome -
                                    bav * rho * (ome * ome) + cdif);
  else
    rb[6 * n3] = rb[6 * n3] - vv * ((T)p.gkw * (T)p.cmu * rho * prodw - (T)p.gkw * fact * divv * rho * ome -
                                    (T)p.bkw * rho * (ome * ome));
}

// tur_vis_sst_b interior (src/update.f90:2422-2486): SST eddy viscosity
// with the F2 limiter on the strain rate, from the velocity differences dqc
// of the stage; q conservative
__global__ void __launch_bounds__(256) k_tur_vis_sst(Par p, const Blk* __restrict__ blks,
                                                     const Tile* __restrict__ tiles, const double* __restrict__ q,
                                                     double* __restrict__ mut, const double* __restrict__ dqc,
                                                     const double* const* __restrict__ xdp,
                                                     const double* __restrict__ dist) {
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
  const ll n3 = n3of(b), c = i3(b, i, j, k);
  const double* qb = q + b.o3 * p.npde * p.nsnap + n3 * (ll)p.npde * n + c;
  const double* db = dqc + b.o3 * 9 * p.nsnap + n3 * 9 * (ll)n + c;
  const int nh = n * p.hbmove;
  const double q1 = qb[0], q2 = qb[n3], q3 = qb[2 * n3], q4 = qb[3 * n3];
  const double pr = (p.gamma - 1) * (qb[4 * n3] - (q2 * q2 + q3 * q3 + q4 * q4) / q1 / 2 - qb[5 * n3]);
  const double rho = q1;
  const double tke = qb[5 * n3] / rho;
  const double ome = qb[6 * n3] / rho;
  const double tt = p.gamma * pr / rho;
  const double mu = suth<double>(p, tt);
  const double xix = (mder<double>(p, b, xdp[0], i + 1, j, k, nh)[0] + mder<double>(p, b, xdp[0], i, j, k, nh)[0]) / 2;
  const double xiy = (mder<double>(p, b, xdp[3], i, j + 1, k, nh)[1] + mder<double>(p, b, xdp[3], i, j, k, nh)[1]) / 2;
  const double xiz = (mder<double>(p, b, xdp[6], i, j, k + 1, nh)[2] + mder<double>(p, b, xdp[6], i, j, k, nh)[2]) / 2;
  const double etax = (mder<double>(p, b, xdp[1], i + 1, j, k, nh)[0] + mder<double>(p, b, xdp[1], i, j, k, nh)[0]) / 2;
  const double etay = (mder<double>(p, b, xdp[4], i, j + 1, k, nh)[1] + mder<double>(p, b, xdp[4], i, j, k, nh)[1]) / 2;
  const double etaz = (mder<double>(p, b, xdp[7], i, j, k + 1, nh)[2] + mder<double>(p, b, xdp[7], i, j, k, nh)[2]) / 2;
  const double zetax =
      (mder<double>(p, b, xdp[2], i + 1, j, k, nh)[0] + mder<double>(p, b, xdp[2], i, j, k, nh)[0]) / 2;
  const double zetay =
      (mder<double>(p, b, xdp[5], i, j + 1, k, nh)[1] + mder<double>(p, b, xdp[5], i, j, k, nh)[1]) / 2;
  const double zetaz =
      (mder<double>(p, b, xdp[8], i, j, k + 1, nh)[2] + mder<double>(p, b, xdp[8], i, j, k, nh)[2]) / 2;
  const double uxi = db[0], ueta = db[n3], uzeta = db[2 * n3];
  const double vxi = db[3 * n3], veta = db[4 * n3], vzeta = db[5 * n3];
  const double wxi = db[6 * n3], weta = db[7 * n3], wzeta = db[8 * n3];
  const double dudx = uxi * xix + ueta * etax + uzeta * zetax;
  const double dudy = uxi * xiy + ueta * etay + uzeta * zetay;
  const double dudz = uxi * xiz + ueta * etaz + uzeta * zetaz;
  const double dvdx = vxi * xix + veta * etax + vzeta * zetax;
  const double dvdy = vxi * xiy + veta * etay + vzeta * zetay;
  const double dvdz = vxi * xiz + veta * etaz + vzeta * zetaz;
  const double dwdx = wxi * xix + weta * etax + wzeta * zetax;
  const double dwdy = wxi * xiy + weta * etay + wzeta * zetay;
  const double dwdz = wxi * xiz + weta * etaz + wzeta * zetaz;
  const double sxx = dudx, sxy = (dudy + dvdx) / 2, sxz = (dudz + dwdx) / 2, syy = dvdy, syz = (dvdz + dwdy) / 2,
               szz = dwdz;
  const double sv = sxx * sxx + syy * syy + szz * szz + 2 * (sxy * sxy + sxz * sxz + syz * syz);
  const double de = dist[b.o1 + i1(b, i, j, k)] + p.epst;
  double arg2 = fmax(2 * sqrt(tke) / 0.09 / ome / de, 500 * mu / rho / ome / (de * de) * p.machfs / p.reyno);
  arg2 = arg2 * arg2;
  arg2 = fmin(arg2, 1e2);
  const double e2 = exp(2 * arg2);
  const double f2 = (e2 - 1) / (e2 + 1);
  const double fct = sqrt(2 * sv);
  mut[b.o3 * p.nsnap + n3 * (ll)n + c] = p.sst_a1 * rho * tke / fmax(p.sst_a1 * ome, fct * f2) * p.reyno / p.machfs;
}

// tur_vis_sst_b halo faces (src/update.f90:2488-2520): mut of the face
// halos copied from the first interior layer (no edges)
__global__ void __launch_bounds__(256) k_mut_faces(Par p, const Blk* __restrict__ blks,
                                                   const Tile* __restrict__ tiles, double* __restrict__ mut) {
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
  const ll n3 = n3of(b);
  double* mb = mut + b.o3 * p.nsnap + n3 * (ll)n;
  const double v = mb[i3(b, i, j, k)];
  if (k == 1) mb[i3(b, i, j, 0)] = v;
  if (k == km) mb[i3(b, i, j, b.kmax)] = v;
  if (j == 1) mb[i3(b, i, 0, k)] = v;
  if (j == jm) mb[i3(b, i, b.jmax, k)] = v;
  if (i == 1) mb[i3(b, 0, j, k)] = v;
  if (i == im) mb[i3(b, b.imax, j, k)] = v;
}

// q_face + q_der_c of u, v, w (src/flux.f90:2391-2801, 2979-3131) into
// dqc: the first visit of a coarse level computes SST mut from them
// (src/cosa.f90:1739-1750); q primitive
__global__ void __launch_bounds__(256) k_vel_diffs(Par p, const Blk* __restrict__ blks,
                                                   const Tile* __restrict__ tiles, const double* __restrict__ q,
                                                   double* __restrict__ dqc, const unsigned char* __restrict__ sbc,
                                                   const ll* __restrict__ sbcoff) {
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
  const ll n3 = n3of(b), c = i3(b, i, j, k);
  const double* qb = q + b.o3 * p.npde * p.nsnap + n3 * (ll)p.npde * n;
  const ll st[4] = {0, 1, (ll)(b.imax + 3), (ll)(b.imax + 3) * (b.jmax + 3)};
  const unsigned char* fl = sbc + sbcoff[t.block];
  double* db = dqc + b.o3 * 9 * p.nsnap + n3 * 9 * (ll)n + c;
  for (int d = 1; d <= 3; d++) {
    const int a = d == 1 ? i : (d == 2 ? j : k);
    const int L = d == 1 ? b.imax : (d == 2 ? b.jmax : b.kmax);
    const int t1 = d == 1 ? j : i, t2 = d == 3 ? j : k;
    bool wlo = false, whi = false;
    if (a == 1) {
      const unsigned char f = d == 1 ? face_flag<1>(b, fl, 0, t1, t2)
                                     : (d == 2 ? face_flag<2>(b, fl, 0, t1, t2) : face_flag<3>(b, fl, 0, t1, t2));
      wlo = (f & 1) != 0;
    }
    if (a + 1 == L) {
      const unsigned char f = d == 1 ? face_flag<1>(b, fl, 1, t1, t2)
                                     : (d == 2 ? face_flag<2>(b, fl, 1, t1, t2) : face_flag<3>(b, fl, 1, t1, t2));
      whi = (f & 1) != 0;
    }
    for (int sv = 0; sv < 3; sv++) {
      const double* qv = qb + n3 * (sv + 1);
      const double flo = wlo ? qv[c - st[d]] : (qv[c - st[d]] + qv[c]) / 2;
      const double fhi = whi ? qv[c + st[d]] : (qv[c] + qv[c + st[d]]) / 2;
      db[n3 * (3 * sv + d - 1)] = fhi - flo;
    }
  }
}

// tur_vis_b over 0..imax x 0..jmax x 0..kmax (q primitive or conservative,
// as the caller has it)
__global__ void __launch_bounds__(256) k_tur_vis(Par p, const Blk* __restrict__ blks, const Tile* __restrict__ tiles,
                                                 const double* __restrict__ q, double* __restrict__ mut) {
  const Tile t = tiles[blockIdx.x];
  const Blk b = blks[t.block];
  const int ni = b.imax + 1, nj = b.jmax + 1, nk = b.kmax + 1;
  const int nc = ni * nj * nk;
  const int item = t.base + threadIdx.x;
  if (item >= nc * p.nsnap) return;
  const int n = item / nc;
  int rem = item - n * nc;
  const int i = rem % ni;
  rem /= ni;
  const int j = rem % nj, k = rem / nj;
  const ll n3 = n3of(b);
  const double* qb = q + b.o3 * p.npde * p.nsnap + n3 * (ll)p.npde * n;
  const ll c = i3(b, i, j, k);
  mut[b.o3 * p.nsnap + n3 * (ll)n + c] =
      p.cmu * qb[c] * qb[c + 5 * n3] / (qb[c + 6 * n3] + p.epst) * p.reyno / p.machfs;
}

// fixq_b: k and omega floors, interior, q conservative
__global__ void __launch_bounds__(256) k_fixq(Par p, const Blk* __restrict__ blks, const Tile* __restrict__ tiles,
                                              double* __restrict__ q) {
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
  const ll n3 = n3of(b);
  double* qb = q + b.o3 * p.npde * p.nsnap + n3 * (ll)p.npde * n;
  const ll c = i3(b, i, j, k);
  const double rho = qb[c];
  double tke = qb[c + 5 * n3] / rho;
  double ome = qb[c + 6 * n3] / rho;
  if (tke < p.qmin6) {
    tke = p.qmin6;
    qb[c + 5 * n3] = rho * tke;
  }
  if (ome < p.qmin7) {
    ome = p.qmin7;
    qb[c + 6 * n3] = rho * ome;
  }
}

// lim_prod_mg_b then lim_res_ome (restriction stage), interior, every
// snapshot, q conservative.  lim: 1 lim_res_ome_b (steady), 2 the explicit
// branch of lim_res_ome_bhb (src/update.f90:2880-2905: no RK coefficient,
// from qold), 3 the rkimp branch of lim_res_ome_b (dual time with IRS or
// multigrid), 0 none (HB rkim runs as its own kernel; dual-time rkimp
// without IRS on one level has no limit).
__global__ void __launch_bounds__(256) k_turb_limits(Par p, int do_prod, int lim, double ar,
                                                     const Blk* __restrict__ blks, const Tile* __restrict__ tiles,
                                                     const double* __restrict__ q, const double* __restrict__ qold,
                                                     double* __restrict__ res, const double* __restrict__ vol,
                                                     const double* __restrict__ dtvolt,
                                                     const double* __restrict__ delpls, double* __restrict__ prd,
                                                     const double* __restrict__ mut, const double* __restrict__ ttrm) {
  const Tile t = tiles[blockIdx.x];
  const Blk b = blks[t.block];
  const int im = b.imax - 1, jm = b.jmax - 1, km = b.kmax - 1;
  const int nc = im * jm * km;
  const int item = t.base + threadIdx.x;
  if (item >= nc) return;
  int rem = item;
  const int i = rem % im + 1;
  rem /= im;
  const int j = rem % jm + 1, k = rem / jm + 1;
  const ll n3 = n3of(b), n1 = n1of(b);
  const ll c = i3(b, i, j, k);
  const double mor = p.machfs / p.reyno;
  const double vv = vol[b.o1 + i1(b, i, j, k)];
  for (int n = 0; n < p.nsnap; n++) {
    // ttrm(2:3): gkw, bkw (Wilcox) or blended (BSL/SST)
    const double* tb = p.sst ? ttrm + b.o3 * 3 * p.nsnap + n3 * 3 * (ll)n + c : nullptr;
    const double gav = tb ? tb[n3] : p.gkw, bav = tb ? tb[2 * n3] : p.bkw;
    const double* qb = q + b.o3 * p.npde * p.nsnap + n3 * p.npde * n;
    double* rb = res + b.o3 * p.npde * p.nsnap + n3 * p.npde * n;
    const ll so = b.o3 * p.nsnap + n3 * n + c;
    if (do_prod) {
      const double tstdis = p.prdlim * p.bstrkw * qb[c] * qb[c + 5 * n3] * qb[c + 6 * n3];
      if (mor * mut[so] * prd[so] > tstdis) prd[so] = tstdis / mor / mut[so];
    }
    const double dtt = dtvolt[b.o1 * p.nsnap + n1 * n + i1(b, i, j, k)];
    const double dtau = dtt * vv;
    const double rho = qb[c];
    const double ome = qb[c + 6 * n3] / qb[c];
    const double rhomegamin = gav * p.cmu * rho * sqrt(prd[so]);
    if (lim == 1) {
      const double den1 = 1 + ar * dtau * (gav * delpls[so] + 2 * bav * ome);
      const double rhomega = qb[c + 6 * n3] - ar * rb[c + 6 * n3] * dtt / den1;
      if (rhomega < rhomegamin) rb[c + 6 * n3] = den1 * (qb[c + 6 * n3] - rhomegamin) / dtt / ar;
    } else if (lim == 3) {   // dual time, rkimp: src/update.f90:2753-2794 (den1 without ar)
      const double den1 = 1 + dtau * (gav * delpls[so] + 2 * bav * ome);
      const double den2 = ar * p.dual_fact * dtau / p.dual_dt;
      const double rhomega = qb[c + 6 * n3] - ar * rb[c + 6 * n3] * dtt / (den1 + den2);
      if (rhomega < rhomegamin) rb[c + 6 * n3] = ((den1 + den2) * (qb[c + 6 * n3] - rhomegamin)) / dtt / ar;
    } else if (lim == 2) {
      const double* ob = qold + b.o3 * p.npde * p.nsnap + n3 * p.npde * n;
      const double rhomega =
          ob[c + 6 * n3] - rb[c + 6 * n3] * dtt / (1 + dtau * (gav * delpls[so] + 2 * bav * ome));
      rb[c + 6 * n3] = rb[c + 6 * n3] -
                       (1 + dtau * (gav * delpls[so] + 2 * bav * ome)) * fmax(0.0, (rhomegamin - rhomega) / dtt);
    }
  }
}

// ------------------------------------------------------------------------
// launchers

static inline dim3 grid1(int n) { return dim3(n > 0 ? n : 1); }

template <typename T>
static hipError_t residual_viscous_t(Solve& s, Level& L, const ResIn<T>& in, double* res, T* vflux) {
  const Par& p = s.par;
  hipError_t e;
  if (p.kom) {
    e = launch_h0_last<T>(s, L, in);
    if (e != hipSuccess) return e;
  }
  const int fk[3] = {TK_FACES_I, TK_FACES_J, TK_FACES_K};
  if (s.fuse_viscous || !p.viscous) {
    for (int d = 1; d <= 3; d++) {
      e = launch_flux_dir<T>(s, L, d, in, res);   // inviscid + viscous fluxes of direction d
      if (e != hipSuccess) return e;
    }
  } else if (s.viscous_tile) {
    // the viscous face fluxes of direction d into slot d - 1, then one pass
    // forms the inviscid fluxes of all directions and sums both in the
    // reference's order (k_resid_visc.hip)
    const ll vslot = L.n0tot * p.npde * p.nsnap;
    for (int d = 1; d <= 3; d++) {
      const int nt = L.ntile[fk[d - 1]];
      if (!nt) continue;
      const T* xd = in.xd[3 * (d - 1)];
      const T* ed = in.xd[3 * (d - 1) + 1];
      const T* zd = in.xd[3 * (d - 1) + 2];
      T* vd = vflux + (d - 1) * vslot;
#define VF(D, NV)                                                                                                 \
  hipLaunchKernelGGL((k_vflux<D, NV, T>), grid1(nt), dim3(TS), 0, s.stream, p, L.dblk, L.tile[fk[d - 1]], in.q,    \
                     in.mut, in.s[D - 1], xd, ed, zd, L.sbc, L.sbc_doff, vd, L.f[COSAHIP_DIST])
      if (p.npde == 7) {
        if (d == 1) VF(1, 7);
        else if (d == 2) VF(2, 7);
        else VF(3, 7);
      } else {
        if (d == 1) VF(1, 5);
        else if (d == 2) VF(2, 5);
        else VF(3, 5);
      }
#undef VF
    }
    e = launch_residual_tile_visc<T>(s, L, in, res, vflux, vslot);
    if (e != hipSuccess) return e;
  } else {
    // per direction: inviscid line march, then the viscous face fluxes.
    // Default (viscous_faces): k_vflux stores the face fluxes of direction d
    // and their difference is added to res by the next kernel that reads and
    // writes res -- the inviscid march of direction d + 1, or for d = 3 the
    // k-omega sources (a difference kernel for laminar runs) -- in the
    // reference's order.  Otherwise one marching kernel per direction does
    // flux and difference.
    const Par porig = s.par;
    Par pi = porig;
    pi.viscous = 0;
    for (int d = 1; d <= 3; d++) {
      s.par = pi;
      e = launch_flux_dir<T>(s, L, d, in, res, (s.viscous_faces && d > 1) ? vflux : nullptr);
      s.par = porig;
      if (e != hipSuccess) return e;
      const T* xd = in.xd[3 * (d - 1)];
      const T* ed = in.xd[3 * (d - 1) + 1];
      const T* zd = in.xd[3 * (d - 1) + 2];
      const int nt = L.ntile[fk[d - 1]];
      const Tile* tl = L.tile[fk[d - 1]];
#define VF(D, NV)                                                                                                 \
  hipLaunchKernelGGL((k_vflux<D, NV, T>), grid1(nt), dim3(TS), 0, s.stream, p, L.dblk, tl, in.q, in.mut, in.s[D - 1], \
                     xd, ed, zd, L.sbc, L.sbc_doff, vflux, L.f[COSAHIP_DIST])
#define VI(NV)                                                                                                    \
  hipLaunchKernelGGL((k_vres_i<NV, T>), grid1(L.ntile[TK_ROWS_I]), dim3(256), 0, s.stream, p, L.dblk,             \
                     L.tile[TK_ROWS_I], in.q, in.mut, in.s[0], xd, ed, zd, L.sbc, L.sbc_doff, res, L.f[COSAHIP_DIST])
#define VM(D, NV, TK)                                                                                             \
  hipLaunchKernelGGL((k_vres_march<D, NV, T>), grid1(L.ntile[TK]), dim3(TS), 0, s.stream, p, L.dblk, L.tile[TK],   \
                     in.q, in.mut, in.s[D - 1], xd, ed, zd, L.sbc, L.sbc_doff, res, L.f[COSAHIP_DIST])
      if (s.viscous_faces) {
        if (nt) {
          if (p.npde == 7) {
            if (d == 1) VF(1, 7);
            else if (d == 2) VF(2, 7);
            else VF(3, 7);
          } else {
            if (d == 1) VF(1, 5);
            else if (d == 2) VF(2, 5);
            else VF(3, 5);
          }
        }
      } else if (p.npde == 7) {
        if (d == 1 && L.ntile[TK_ROWS_I]) VI(7);
        else if (d == 2 && L.ntile[TK_SEGS_J]) VM(2, 7, TK_SEGS_J);
        else if (d == 3 && L.ntile[TK_SEGS_K]) VM(3, 7, TK_SEGS_K);
      } else {
        if (d == 1 && L.ntile[TK_ROWS_I]) VI(5);
        else if (d == 2 && L.ntile[TK_SEGS_J]) VM(2, 5, TK_SEGS_J);
        else if (d == 3 && L.ntile[TK_SEGS_K]) VM(3, 5, TK_SEGS_K);
      }
#undef VF
#undef VI
#undef VM
    }
    if (s.viscous_faces && !p.kom && L.ntile[TK_CELLS])   // laminar: difference of the direction-3 fluxes
      hipLaunchKernelGGL((k_vresid<3, 5, T>), grid1(L.ntile[TK_CELLS]), dim3(TS), 0, s.stream, p, L.dblk,
                         L.tile[TK_CELLS], vflux, res);
  }
  const bool vpre3 = p.viscous && !s.fuse_viscous && !s.viscous_tile && s.viscous_faces;
  T* cmet = sizeof(T) == 4 ? reinterpret_cast<T*>(L.cmet32) : reinterpret_cast<T*>(L.cmet);
  bool& cmet_ok = sizeof(T) == 4 ? L.cmet32_ok : L.cmet_ok;
  if (p.kom && L.ntile[TK_CELLS] && !cmet_ok) {
    hipLaunchKernelGGL(k_cell_metrics<T>, grid1(L.ntile[TK_CELLS]), dim3(TS), 0, s.stream, p, L.dblk, L.tile[TK_CELLS],
                       in.dxd, cmet);
    cmet_ok = true;
  }
  if (p.kom && L.ntile[TK_CELLS])
    hipLaunchKernelGGL(k_turb_src<T>, grid1(L.ntile[TK_CELLS]), dim3(TS), 0, s.stream, p, L.nl, L.dblk,
                       L.tile[TK_CELLS], in.q, in.mut, L.f[COSAHIP_PRD], L.f[COSAHIP_DIVVEL], L.f[COSAHIP_DELPLS], res,
                       L.f[COSAHIP_VOL], in.dxd, L.sbc, L.sbc_doff, vpre3 ? vflux : nullptr, L.f[COSAHIP_TTRM],
                       L.dqc, L.f[COSAHIP_DIST], cmet);
  return hipGetLastError();
}

hipError_t launch_residual_viscous(Solve& s, Level& L, const double* q, double* res) {
  const Par& p = s.par;
  const bool f32 = s.fp32() && q == L.f[COSAHIP_Q] && L.q32;
  hipError_t e;
  if (p.viscous) {
    e = launch_edges(s, L, const_cast<double*>(q), f32 ? L.q32 : nullptr);
    if (e != hipSuccess) return e;
  }
  if (!f32) {
    ResIn<double> in{};
    in.q = q;
    for (int d = 0; d < 3; d++) in.s[d] = L.f[COSAHIP_SI + d];
    in.mut = L.f[COSAHIP_MUT];
    for (int m = 0; m < 9; m++) in.xd[m] = L.f[COSAHIP_XIDERI + m];
    in.dxd = L.dxd;
    return residual_viscous_t<double>(s, L, in, res, L.vflux);
  }
  // FP32 residual: single-precision copies of the inputs, converted when
  // their FP64 source has changed (metrics on upload, mut after tur_vis)
  if (!L.m32_ok) {
    const ll nm = L.n2tot * p.lmet * p.nsnap_met;
    for (int d = 0; d < 3; d++)
      if ((e = launch_to_f32(s, L.f[COSAHIP_SI + d], L.m32[d], nm)) != hipSuccess) return e;
    L.m32_ok = true;
  }
  if (p.viscous && !L.xd32_ok) {
    const ll nx = L.n0tot * 3 * p.nsnap_met;
    for (int m = 0; m < 9; m++)
      if ((e = launch_to_f32(s, L.f[COSAHIP_XIDERI + m], L.xd32[m], nx)) != hipSuccess) return e;
    L.xd32_ok = true;
  }
  if (p.kom && !L.mut32_ok) {
    if ((e = launch_to_f32(s, L.f[COSAHIP_MUT], L.mut32, L.n3tot * p.nsnap)) != hipSuccess) return e;
    L.mut32_ok = true;
  }
  ResIn<float> in{};
  in.q = L.q32;
  for (int d = 0; d < 3; d++) in.s[d] = L.m32[d];
  in.mut = L.mut32;
  for (int m = 0; m < 9; m++) in.xd[m] = L.xd32[m];
  in.dxd = L.dxd32;
  return residual_viscous_t<float>(s, L, in, res, reinterpret_cast<float*>(L.vflux));
}

__global__ void __launch_bounds__(256) k_zero_n(double* __restrict__ a, ll n) {
  const ll x = (ll)blockIdx.x * 256 + threadIdx.x;
  if (x < n) a[x] = 0.0;
}

// SST: tur_vis_sst from the velocity differences dqc of the stage's
// residual, halo faces, cut
static hipError_t launch_tur_vis_sst(Solve& s, Level& L, const double* q) {
  if (L.ntile[TK_CELLS]) {
    hipLaunchKernelGGL(k_tur_vis_sst, grid1(L.ntile[TK_CELLS]), dim3(TS), 0, s.stream, s.par, L.dblk,
                       L.tile[TK_CELLS], q, L.f[COSAHIP_MUT], L.dqc, (const double* const*)L.dxd,
                       L.f[COSAHIP_DIST]);
    hipLaunchKernelGGL(k_mut_faces, grid1(L.ntile[TK_CELLS]), dim3(TS), 0, s.stream, s.par, L.dblk, L.tile[TK_CELLS],
                       L.f[COSAHIP_MUT]);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_cut_ncomp(s, L, L.f[COSAHIP_MUT], 1);
}

hipError_t launch_sst_first_mut(Solve& s, Level& L) {
  L.mut32_ok = false;
  const ll nq = L.n3tot * s.par.npde * s.par.nsnap;
  double* res = L.f[COSAHIP_RES];
  hipError_t e = hipMemcpyAsync(res, L.f[COSAHIP_Q], nq * sizeof(double), hipMemcpyDeviceToDevice, s.stream);
  if (e != hipSuccess) return e;
  if (L.ntile[TK_CELLS])
    hipLaunchKernelGGL(k_vel_diffs, grid1(L.ntile[TK_CELLS]), dim3(TS), 0, s.stream, s.par, L.dblk, L.tile[TK_CELLS],
                       res, L.dqc, L.sbc, L.sbc_doff);
  e = launch_prim2cons(s, L, res);
  if (e != hipSuccess) return e;
  e = launch_tur_vis_sst(s, L, res);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_zero_n, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0, s.stream, res, nq);
  return hipGetLastError();
}

hipError_t launch_tur_vis(Solve& s, Level& L, const double* q) {
  L.mut32_ok = false;
  if (s.par.sst == 2) return launch_tur_vis_sst(s, L, q);
  if (L.ntile[TK_CELLS_P1])
    hipLaunchKernelGGL(k_tur_vis, grid1(L.ntile[TK_CELLS_P1]), dim3(TS), 0, s.stream, s.par, L.dblk,
                       L.tile[TK_CELLS_P1], q, L.f[COSAHIP_MUT]);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_cut_ncomp(s, L, L.f[COSAHIP_MUT], 1);
}

hipError_t launch_fixq(Solve& s, Level& L) {
  if (L.ntile[TK_CELLS])
    hipLaunchKernelGGL(k_fixq, grid1(L.ntile[TK_CELLS]), dim3(TS), 0, s.stream, s.par, L.dblk, L.tile[TK_CELLS],
                       L.f[COSAHIP_Q]);
  return hipGetLastError();
}

hipError_t launch_turb_limits(Solve& s, Level& L, double ar) {
  const Par& p = s.par;
  const int do_prod = p.nlevel > 1 && L.nl != p.nlevel && !p.prd_all_lvls && p.lim_prodtke;
  const int lim = p.dual_rkimp ? (p.dual_plain ? 0 : 3) : !s.cfg.harbal ? 1 : (s.cfg.rkimp ? 0 : 2);
  if (L.ntile[TK_CELLS])
    hipLaunchKernelGGL(k_turb_limits, grid1(L.ntile[TK_CELLS]), dim3(TS), 0, s.stream, p, do_prod, lim, ar, L.dblk,
                       L.tile[TK_CELLS], L.f[COSAHIP_Q], L.f[COSAHIP_QOLD], L.f[COSAHIP_RES], L.f[COSAHIP_VOL],
                       L.f[COSAHIP_DTVOLT], L.f[COSAHIP_DELPLS], L.f[COSAHIP_PRD], L.f[COSAHIP_MUT],
                       L.f[COSAHIP_TTRM]);
  if (s.cfg.harbal && s.cfg.rkimp) return launch_lim_res_ome_rkim(s, L);
  return hipGetLastError();
}

}  // namespace cosahip
