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

namespace cosahip {

__device__ __forceinline__ double suth(const Par& p, double t) {
  return (p.stemp + 1) / (t + p.stemp) * pow(t, 1.5);
}

// strong-BC flag byte of the face of direction DIR at position a (1 or L)
// and transverse (t1, t2) (lower-numbered transverse direction first)
template <int DIR>
__device__ __forceinline__ unsigned char face_flag(const Blk& b, const unsigned char* fl, int outer, int t1, int t2) {
  if (DIR == 1) return fl[outer * (ll)(b.jmax + 1) * (b.kmax + 1) + t1 + (ll)(b.jmax + 1) * t2];
  if (DIR == 2)
    return fl[2 * (ll)(b.jmax + 1) * (b.kmax + 1) + outer * (ll)(b.imax + 1) * (b.kmax + 1) + t1 +
              (ll)(b.imax + 1) * t2];
  return fl[2 * (ll)(b.jmax + 1) * (b.kmax + 1) + 2 * (ll)(b.imax + 1) * (b.kmax + 1) +
            outer * (ll)(b.imax + 1) * (b.jmax + 1) + t1 + (ll)(b.imax + 1) * t2];
}

// metric-derivative triple (3, imax, jmax, kmax, 0:2N*hbmove) at face (i,j,k)
__device__ __forceinline__ const double* mder(const Par& p, const Blk& b, const double* a, int i, int j, int k,
                                              int nh) {
  return a + b.o0 * 3 * p.nsnap_met +
         3 * ((ll)(i - 1) + (ll)b.imax * ((ll)(j - 1) + (ll)b.jmax * ((ll)(k - 1) + (ll)b.kmax * nh)));
}

// ------------------------------------------------------------------------
// Viscous fluxes of the faces of direction DIR, one thread per face:
// face average (q_face), xi/eta/zeta differences (q_der), Reynolds
// stresses (rtst, k-omega) and the flux (vflux) into vflux(imax,jmax,kmax,
// npde,0:2N).  Wall-like faces (strong-BC flag) take the halo value and
// one-sided differences as bq_face / bq_der do.
template <int DIR, int NV>
__global__ void __launch_bounds__(256) k_vflux(Par p, const Blk* __restrict__ blks, const Tile* __restrict__ tiles,
                                               const double* __restrict__ q, const double* __restrict__ mut,
                                               const double* __restrict__ sm, const double* __restrict__ xd,
                                               const double* __restrict__ ed, const double* __restrict__ zd,
                                               const unsigned char* __restrict__ sbc,
                                               const ll* __restrict__ sbcoff, double* __restrict__ vf) {
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
  const ll n3 = n3of(b);
  const double* qb = q + b.o3 * p.npde * p.nsnap + n3 * (ll)p.npde * n;
  const ll st[4] = {0, 1, (ll)(b.imax + 3), (ll)(b.imax + 3) * (b.jmax + 3)};
  const ll cp = i3(b, i, j, k), cm = cp - st[DIR];
  const int a = DIR == 1 ? i : (DIR == 2 ? j : k);
  const int L = DIR == 1 ? b.imax : (DIR == 2 ? b.jmax : b.kmax);
  const int t1 = DIR == 1 ? j : i, t2 = DIR == 3 ? j : k;
  unsigned char flag = 0;
  if (a == 1 || a == L) flag = face_flag<DIR>(b, sbc + sbcoff[t.block], a == L ? 1 : 0, t1, t2);
  const bool wall = (flag & 1) != 0;
  // transverse directions, lower-numbered first
  const int d1 = DIR == 1 ? 2 : 1, d2 = DIR == 3 ? 2 : 3;
  ll aux = 0, n1 = 0, n2 = 0;
  double sgnm = 1.0;
  if (wall) {
    if (a == 1) {
      aux = cm; n1 = cp; n2 = cp + st[DIR]; sgnm = 1.0;
    } else {
      aux = cp; n1 = cm; n2 = cm - st[DIR]; sgnm = -1.0;
    }
  }
  double qf[NV], dd[3][NV];
#pragma unroll
  for (int v = 0; v < NV; v++) {
    const double* qv = qb + n3 * v;
    qf[v] = wall ? qv[aux] : (qv[cm] + qv[cp]) / 2;
    for (int d = 1; d <= 3; d++) {
      double val;
      if (!wall) {
        if (d == DIR) {
          val = qv[cp] - qv[cm];
        } else {
          const ll tt = st[d];
          const double q2 = (qv[cm] + qv[cp] + qv[cp + tt] + qv[cm + tt]) / 4;
          const double q1 = (qv[cm - tt] + qv[cp - tt] + qv[cp] + qv[cm]) / 4;
          val = q2 - q1;
        }
      } else {
        if (d == DIR) {
          val = sgnm / 3 * (-qv[aux] * 8 + qv[n1] * 9 - qv[n2]);
        } else {
          const ll tt = st[d];
          const bool first = d == d1 ? (flag & 2) != 0 : (flag & 8) != 0;
          const bool last = d == d1 ? (flag & 4) != 0 : (flag & 16) != 0;
          const double qc = qv[aux], qp1 = qv[aux + tt], qm1 = qv[aux - tt];
          double q2, q1;
          if (first) {
            q2 = (qc + qp1) / 2;
            q1 = (qc * 3 - qp1) / 2;
          } else if (last) {
            q2 = (qc * 3 - qm1) / 2;
            q1 = (qc + qm1) / 2;
          } else {
            q2 = (qc + qp1) / 2;
            q1 = (qm1 + qc) / 2;
          }
          val = q2 - q1;
        }
      }
      dd[d - 1][v] = val;
    }
  }
  (void)d2;
  double mutf = 0.0;
  if (p.kom) {
    const double* mb = mut + b.o3 * p.nsnap + n3 * (ll)n;
    mutf = wall ? mb[aux] : (mb[cm] + mb[cp]) / 2;
  }
  const int nh = n * p.hbmove;
  const double* x3 = mder(p, b, xd, i, j, k, nh);
  const double* e3 = mder(p, b, ed, i, j, k, nh);
  const double* z3 = mder(p, b, zd, i, j, k, nh);
  const double xix = x3[0], xiy = x3[1], xiz = x3[2];
  const double etax = e3[0], etay = e3[1], etaz = e3[2];
  const double zetax = z3[0], zetay = z3[1], zetaz = z3[2];
  const double* s = sm + b.o2 * p.lmet * p.nsnap_met + ((ll)nh * n2of(b) + i2(b, i, j, k)) * p.lmet;
  const double kx = s[0], ky = s[1], kz = s[2], ds = s[3];
  const double gamma = p.gamma;
  const double rho = qf[0], u = qf[1], v = qf[2], w = qf[3], pr = qf[4];
  const double tmp = gamma * pr / rho;
  const double mu = suth(p, tmp);
  const double uxi = dd[0][1], ueta = dd[1][1], uzeta = dd[2][1];
  const double vxi = dd[0][2], veta = dd[1][2], vzeta = dd[2][2];
  const double wxi = dd[0][3], weta = dd[1][3], wzeta = dd[2][3];
  const double dudx = uxi * xix + ueta * etax + uzeta * zetax;
  const double dudy = uxi * xiy + ueta * etay + uzeta * zetay;
  const double dudz = uxi * xiz + ueta * etaz + uzeta * zetaz;
  const double dvdx = vxi * xix + veta * etax + vzeta * zetax;
  const double dvdy = vxi * xiy + veta * etay + vzeta * zetay;
  const double dvdz = vxi * xiz + veta * etaz + vzeta * zetaz;
  const double dwdx = wxi * xix + weta * etax + wzeta * zetax;
  const double dwdy = wxi * xiy + weta * etay + wzeta * zetay;
  const double dwdz = wxi * xiz + weta * etaz + wzeta * zetaz;
  const double mr = p.machfs / p.reyno;
  double F[NV];
#pragma unroll
  for (int m = 0; m < NV; m++) F[m] = 0.0;
  if (NV == 7) {
    // rtst: Reynolds stresses with mut at the face
    const double tke = qf[5];
    const double fact = 2.0 / 3.0;
    const double divt = (dudx + dvdy + dwdz) * (double)p.turcmp;
    const double ttxx = 2 * mutf * (dudx - divt / 3);
    const double ttxy = mutf * (dudy + dvdx);
    const double ttxz = mutf * (dudz + dwdx);
    const double ttyy = 2 * mutf * (dvdy - divt / 3);
    const double ttyz = mutf * (dvdz + dwdy);
    const double ttzz = 2 * mutf * (dwdz - divt / 3);
    const double r1 = mr * ttxx - fact * rho * tke;
    const double r2 = mr * ttxy;
    const double r3 = mr * ttxz;
    const double r4 = mr * ttyy - fact * rho * tke;
    const double r5 = mr * ttyz;
    const double r6 = mr * ttzz - fact * rho * tke;
    // vflux, k-omega branch
    const double tkexi = dd[0][5], tkeeta = dd[1][5], tkezeta = dd[2][5];
    const double omexi = dd[0][6], omeeta = dd[1][6];
    const double dtkedx = tkexi * xix + tkeeta * etax + tkezeta * zetax;
    const double dtkedy = tkexi * xiy + tkeeta * etay + tkezeta * zetay;
    const double dtkedz = tkexi * xiz + tkeeta * etaz + tkezeta * zetaz;
    const double domedx = omexi * xix + omeeta * etax + dd[2][6] * zetax;
    const double domedy = omexi * xiy + omeeta * etay + dd[2][6] * zetay;
    F[1] = -((r1 * kx + r2 * ky + r3 * kz) * ds);
    F[2] = -((r2 * kx + r4 * ky + r5 * kz) * ds);
    F[3] = -((r3 * kx + r5 * ky + r6 * kz) * ds);
    const double ctke = (mu + p.sigk * mutf) * p.machfs / p.reyno;
    const double come = (mu + p.sigw * mutf) * p.machfs / p.reyno;
    F[4] = -(ctke * (dtkedx * kx + dtkedy * ky + dtkedz * kz) * ds);
    F[5] = -(ctke * (dtkedx * kx + dtkedy * ky + dtkedz * kz) * ds);
    // as in the reference: dk/dz in place of domega/dz
    F[6] = -(come * (domedx * kx + domedy * ky + dtkedz * kz) * ds);
  }
  // laminar stresses and heat flux (common part of vflux_b)
  const double txi = tmp * (dd[0][4] / pr - dd[0][0] / rho);
  const double teta = tmp * (dd[1][4] / pr - dd[1][0] / rho);
  const double tzeta = tmp * (dd[2][4] / pr - dd[2][0] / rho);
  const double dtdx = txi * xix + teta * etax + tzeta * zetax;
  const double dtdy = txi * xiy + teta * etay + tzeta * zetay;
  const double dtdz = txi * xiz + teta * etaz + tzeta * zetaz;
  const double divv = dudx + dvdy + dwdz;
  const double txx = mr * 2 * mu * (dudx - divv / 3);
  const double txy = mr * mu * (dudy + dvdx);
  const double txz = mr * mu * (dudz + dwdx);
  const double tyy = mr * 2 * mu * (dvdy - divv / 3);
  const double tyz = mr * mu * (dvdz + dwdy);
  const double tzz = mr * 2 * mu * (dwdz - divv / 3);
  const double cq = mr * (mu / p.pranl + mutf / p.prant) / (gamma - 1);
  const double qx = -(cq * dtdx), qy = -(cq * dtdy), qz = -(cq * dtdz);
  F[0] = 0;
  F[1] = F[1] - (txx * kx + txy * ky + txz * kz) * ds;
  F[2] = F[2] - (txy * kx + tyy * ky + tyz * kz) * ds;
  F[3] = F[3] - (txz * kx + tyz * ky + tzz * kz) * ds;
  // as in the reference: the z term repeats the y-row stresses
  F[4] = F[4] - ((u * txx + v * txy + w * txz) * kx + (u * txy + v * tyy + w * tyz) * ky +
                 (u * txy + v * tyy + w * tyz) * kz) * ds +
         (qx * kx + qy * ky + qz * kz) * ds;
  double* fb = vf + b.o0 * p.npde * p.nsnap;
  const ll fo = (ll)(i - 1) + (ll)b.imax * ((ll)(j - 1) + (ll)b.jmax * (ll)(k - 1));
  const ll fst = (ll)b.imax * b.jmax * b.kmax;
#pragma unroll
  for (int m = 0; m < NV; m++) fb[fo + fst * (m + (ll)p.npde * n)] = F[m];
}

// res(c) += F(c+1) - F(c) of the viscous face fluxes of direction DIR
template <int DIR, int NV>
__global__ void __launch_bounds__(256) k_vresid(Par p, const Blk* __restrict__ blks, const Tile* __restrict__ tiles,
                                                const double* __restrict__ vf, double* __restrict__ res) {
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
  const double* fb = vf + b.o0 * p.npde * p.nsnap;
  const ll fst = (ll)b.imax * b.jmax * b.kmax;
  const ll f0 = (ll)(i - 1) + (ll)b.imax * ((ll)(j - 1) + (ll)b.jmax * (ll)(k - 1));
  const ll f1 = f0 + (DIR == 1 ? 1 : (DIR == 2 ? (ll)b.imax : (ll)b.imax * b.jmax));
#pragma unroll
  for (int m = 0; m < NV; m++) {
    const ll o = fst * (m + (ll)p.npde * n);
    rb[n3 * m] = rb[n3 * m] + fb[f1 + o] - fb[f0 + o];
  }
}

// ------------------------------------------------------------------------
// k-omega sources at cell centres: the velocity differences of the face
// averages (q_face + q_der_c, three directions), turprod (finest level)
// and srckw into res(6:7).  q primitive.
__global__ void __launch_bounds__(256) k_turb_src(Par p, int nl, const Blk* __restrict__ blks,
                                                  const Tile* __restrict__ tiles, const double* __restrict__ q,
                                                  const double* __restrict__ mut, double* __restrict__ prd,
                                                  double* __restrict__ divvel, double* __restrict__ delpls,
                                                  double* __restrict__ res, const double* __restrict__ vol,
                                                  const double* const* __restrict__ xdp,
                                                  const unsigned char* __restrict__ sbc,
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
  const ll n3 = n3of(b);
  const double* qb = q + b.o3 * p.npde * p.nsnap + n3 * (ll)p.npde * n;
  const ll c = i3(b, i, j, k);
  const ll st[4] = {0, 1, (ll)(b.imax + 3), (ll)(b.imax + 3) * (b.jmax + 3)};
  const unsigned char* fl = sbc + sbcoff[t.block];
  const ll so = b.o3 * p.nsnap + n3 * (ll)n + c;
  if (nl == 1 || p.prd_all_lvls) {
    // dq(d)[v] = qf(face c+1) - qf(face c), v = u, v, w
    double dq[3][3];
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
      for (int v = 1; v <= 3; v++) {
        const double* qv = qb + n3 * v;
        const double flo = wlo ? qv[c - st[d]] : (qv[c - st[d]] + qv[c]) / 2;
        const double fhi = whi ? qv[c + st[d]] : (qv[c] + qv[c + st[d]]) / 2;
        dq[d - 1][v - 1] = fhi - flo;
      }
    }
    const int nh = n * p.hbmove;
    // xdp: xideri, etaderi, zetaderi, xiderj, etaderj, zetaderj, xiderk, etaderk, zetaderk
    const double xix = (mder(p, b, xdp[0], i + 1, j, k, nh)[0] + mder(p, b, xdp[0], i, j, k, nh)[0]) / 2;
    const double xiy = (mder(p, b, xdp[3], i, j + 1, k, nh)[1] + mder(p, b, xdp[3], i, j, k, nh)[1]) / 2;
    const double xiz = (mder(p, b, xdp[6], i, j, k + 1, nh)[2] + mder(p, b, xdp[6], i, j, k, nh)[2]) / 2;
    const double etax = (mder(p, b, xdp[1], i + 1, j, k, nh)[0] + mder(p, b, xdp[1], i, j, k, nh)[0]) / 2;
    const double etay = (mder(p, b, xdp[4], i, j + 1, k, nh)[1] + mder(p, b, xdp[4], i, j, k, nh)[1]) / 2;
    const double etaz = (mder(p, b, xdp[7], i, j, k + 1, nh)[2] + mder(p, b, xdp[7], i, j, k, nh)[2]) / 2;
    const double zetax = (mder(p, b, xdp[2], i + 1, j, k, nh)[0] + mder(p, b, xdp[2], i, j, k, nh)[0]) / 2;
    const double zetay = (mder(p, b, xdp[5], i, j + 1, k, nh)[1] + mder(p, b, xdp[5], i, j, k, nh)[1]) / 2;
    const double zetaz = (mder(p, b, xdp[8], i, j, k + 1, nh)[2] + mder(p, b, xdp[8], i, j, k, nh)[2]) / 2;
    const double uxi = dq[0][0], ueta = dq[1][0], uzeta = dq[2][0];
    const double vxi = dq[0][1], veta = dq[1][1], vzeta = dq[2][1];
    const double wxi = dq[0][2], weta = dq[1][2], wzeta = dq[2][2];
    const double dudx = uxi * xix + ueta * etax + uzeta * zetax;
    const double dudy = uxi * xiy + ueta * etay + uzeta * zetay;
    const double dudz = uxi * xiz + ueta * etaz + uzeta * zetaz;
    const double dvdx = vxi * xix + veta * etax + vzeta * zetax;
    const double dvdy = vxi * xiy + veta * etay + vzeta * zetay;
    const double dvdz = vxi * xiz + veta * etaz + vzeta * zetaz;
    const double dwdx = wxi * xix + weta * etax + wzeta * zetax;
    const double dwdy = wxi * xiy + weta * etay + wzeta * zetay;
    const double dwdz = wxi * xiz + weta * etaz + wzeta * zetaz;
    const double divv = (dudx + dvdy + dwdz) * (double)p.turcmp;
    const double txx = dudx, txy = (dudy + dvdx) / 2, txz = (dudz + dwdx) / 2, tyy = dvdy, tyz = (dvdz + dwdy) / 2,
                 tzz = dwdz;
    prd[so] = 2 * (txx * txx + tyy * tyy + tzz * tzz + 2 * (txy * txy + txz * txz + tyz * tyz) - divv * divv / 3);
    divvel[so] = divv;
    delpls[so] = 2.0 / 3.0 * fmax(divv, 0.0);
  }
  // srckw_b
  const double fact = 2.0 / 3.0, mor = p.machfs / p.reyno;
  const double rho = qb[c], tke = qb[c + 5 * n3], ome = qb[c + 6 * n3];
  const double mu_t = mut[so];
  const double divv = divvel[so];
  double prod = prd[so], prodw = prd[so];
  if (p.lim_prodtke && (nl == 1 || p.prd_all_lvls)) {
    const double tstdis = p.prdlim * p.bstrkw * rho * tke * ome;
    if (mor * mu_t * prod > tstdis) {
      prod = tstdis / mor / mu_t;
      if (p.lim_prodomega) prodw = prod;
    }
  }
  const double vv = vol[b.o1 + i1(b, i, j, k)];
  double* rb = res + b.o3 * p.npde * p.nsnap + n3 * (ll)p.npde * n + c;
  rb[5 * n3] = rb[5 * n3] - vv * (mor * mu_t * prod - fact * divv * rho * tke - p.bstrkw * rho * tke * ome);
  rb[6 * n3] = rb[6 * n3] - vv * (p.gkw * p.cmu * rho * prodw - p.gkw * fact * divv * rho * ome -
                                  p.bkw * rho * (ome * ome));
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

// lim_prod_mg_b then lim_res_ome_b (restriction stage), interior, one
// snapshot, q conservative
__global__ void __launch_bounds__(256) k_turb_limits(Par p, int do_prod, double ar, const Blk* __restrict__ blks,
                                                     const Tile* __restrict__ tiles, const double* __restrict__ q,
                                                     double* __restrict__ res, const double* __restrict__ vol,
                                                     const double* __restrict__ dtvolt,
                                                     const double* __restrict__ delpls, double* __restrict__ prd,
                                                     const double* __restrict__ mut) {
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
  const ll n3 = n3of(b);
  const double* qb = q + b.o3 * p.npde * p.nsnap;
  double* rb = res + b.o3 * p.npde * p.nsnap;
  const ll c = i3(b, i, j, k), so = b.o3 * p.nsnap + c;
  const double mor = p.machfs / p.reyno;
  if (do_prod) {
    const double tstdis = p.prdlim * p.bstrkw * qb[c] * qb[c + 5 * n3] * qb[c + 6 * n3];
    if (mor * mut[so] * prd[so] > tstdis) prd[so] = tstdis / mor / mut[so];
  }
  const double gav = p.gkw, bav = p.bkw;
  const ll po = b.o1 * p.nsnap + i1(b, i, j, k);
  const double dtt = dtvolt[po];
  const double dtau = dtt * vol[b.o1 + i1(b, i, j, k)];
  const double rho = qb[c];
  const double ome = qb[c + 6 * n3] / qb[c];
  const double den1 = 1 + ar * dtau * (gav * delpls[so] + 2 * bav * ome);
  const double rhomega = qb[c + 6 * n3] - ar * rb[c + 6 * n3] * dtt / den1;
  const double rhomegamin = gav * p.cmu * rho * sqrt(prd[so]);
  if (rhomega < rhomegamin) rb[c + 6 * n3] = den1 * (qb[c + 6 * n3] - rhomegamin) / dtt / ar;
}

// ------------------------------------------------------------------------
// launchers

static inline dim3 grid1(int n) { return dim3(n > 0 ? n : 1); }

hipError_t launch_residual_viscous(Solve& s, Level& L, const double* q, double* res) {
  const Par& p = s.par;
  hipError_t e;
  if (p.viscous) {
    e = launch_edges(s, L, const_cast<double*>(q));
    if (e != hipSuccess) return e;
  }
  if (p.kom) {
    e = launch_h0_last(s, L, q);
    if (e != hipSuccess) return e;
  }
  const double* sm[3] = {L.f[COSAHIP_SI], L.f[COSAHIP_SJ], L.f[COSAHIP_SK]};
  const int fk[3] = {TK_FACES_I, TK_FACES_J, TK_FACES_K};
  for (int d = 1; d <= 3; d++) {
    e = launch_flux_dir(s, L, d, q, res);
    if (e != hipSuccess) return e;
    if (!p.viscous) continue;
    const double* xd = L.f[COSAHIP_XIDERI + 3 * (d - 1)];
    const double* ed = L.f[COSAHIP_XIDERI + 3 * (d - 1) + 1];
    const double* zd = L.f[COSAHIP_XIDERI + 3 * (d - 1) + 2];
    const int nt = L.ntile[fk[d - 1]];
    const Tile* tl = L.tile[fk[d - 1]];
    const double* mu = L.f[COSAHIP_MUT];
#define VF(D, NV)                                                                                                  \
  hipLaunchKernelGGL((k_vflux<D, NV>), grid1(nt), dim3(TS), 0, s.stream, p, L.dblk, tl, q, mu, sm[D - 1], xd, ed, zd, \
                     L.sbc, L.sbc_doff, L.vflux)
#define VR(D, NV)                                                                                          \
  hipLaunchKernelGGL((k_vresid<D, NV>), grid1(L.ntile[TK_CELLS]), dim3(TS), 0, s.stream, p, L.dblk,        \
                     L.tile[TK_CELLS], L.vflux, res)
    if (nt) {
      if (p.npde == 7) {
        if (d == 1) { VF(1, 7); VR(1, 7); }
        else if (d == 2) { VF(2, 7); VR(2, 7); }
        else { VF(3, 7); VR(3, 7); }
      } else {
        if (d == 1) { VF(1, 5); VR(1, 5); }
        else if (d == 2) { VF(2, 5); VR(2, 5); }
        else { VF(3, 5); VR(3, 5); }
      }
    }
#undef VF
#undef VR
  }
  if (p.kom && L.ntile[TK_CELLS])
    hipLaunchKernelGGL(k_turb_src, grid1(L.ntile[TK_CELLS]), dim3(TS), 0, s.stream, p, L.nl, L.dblk,
                       L.tile[TK_CELLS], q, L.f[COSAHIP_MUT], L.f[COSAHIP_PRD], L.f[COSAHIP_DIVVEL],
                       L.f[COSAHIP_DELPLS], res, L.f[COSAHIP_VOL], (const double* const*)L.dxd, L.sbc, L.sbc_doff);
  return hipGetLastError();
}

hipError_t launch_tur_vis(Solve& s, Level& L, const double* q) {
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
  if (L.ntile[TK_CELLS])
    hipLaunchKernelGGL(k_turb_limits, grid1(L.ntile[TK_CELLS]), dim3(TS), 0, s.stream, p, do_prod, ar, L.dblk,
                       L.tile[TK_CELLS], L.f[COSAHIP_Q], L.f[COSAHIP_RES], L.f[COSAHIP_VOL], L.f[COSAHIP_DTVOLT],
                       L.f[COSAHIP_DELPLS], L.f[COSAHIP_PRD], L.f[COSAHIP_MUT]);
  return hipGetLastError();
}

}  // namespace cosahip
