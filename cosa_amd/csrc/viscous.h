// viscous.h -- device helpers of the viscous / k-omega face fluxes, shared
// by the fused face kernels (k_turb.hip) and the line-marching residual
// kernels (k_flux.hip).
#pragma once
#include "cosahip_internal.h"

namespace cosahip {

template <typename T>
__device__ __forceinline__ T suth(const Par& p, T t) {
  return ((T)p.stemp + 1) / (t + (T)p.stemp) * pow(t, T(1.5));
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

// res(c) = (res(c) + vf(c + e_d)) - vf(c) for the NV variables of cell
// (i,j,k), snapshot n: the difference of the viscous face fluxes of
// direction d stored by k_vflux (vflux(imax,jmax,kmax,npde,0:2N), off_0),
// i.e. resid of the viscous fluxes (src/cosa.f90:1123-1211), applied inside
// the kernel that next reads and writes res.
template <int NV, typename T>
__device__ __forceinline__ void add_vdiff(const Par& p, const Blk& b, const T* __restrict__ vf, int d, int i, int j,
                                          int k, int n, double* __restrict__ rc, ll n3) {
  const T* fb = vf + b.o0 * p.npde * p.nsnap;
  const ll fst = (ll)b.imax * b.jmax * b.kmax;
  const ll f0 = (ll)(i - 1) + (ll)b.imax * ((ll)(j - 1) + (ll)b.jmax * (ll)(k - 1));
  const ll f1 = f0 + (d == 1 ? 1 : (d == 2 ? (ll)b.imax : (ll)b.imax * b.jmax));
#pragma unroll
  for (int m = 0; m < NV; m++) {
    const ll o = fst * (m + (ll)p.npde * n);
    rc[n3 * m] = rc[n3 * m] + (double)fb[f1 + o] - (double)fb[f0 + o];
  }
}

// metric-derivative triple (3, imax, jmax, kmax, 0:2N*hbmove) at face (i,j,k)
template <typename T>
__device__ __forceinline__ const T* mder(const Par& p, const Blk& b, const T* a, int i, int j, int k, int nh) {
  return a + b.o0 * 3 * p.nsnap_met +
         3 * ((ll)(i - 1) + (ll)b.imax * ((ll)(j - 1) + (ll)b.jmax * ((ll)(k - 1) + (ll)b.kmax * nh)));
}

// Viscous / turbulent diffusion flux of face (i,j,k) of direction DIR and
// snapshot n (q_face + q_der + rtst + vflux, src/flux.f90:2391-2801,
// 1813-2390, 2802-2948, 878-1182).  Wall-like faces (strong-BC flag) take
// the halo value and one-sided differences as bq_face / bq_der do.
template <int DIR, int NV, typename T = double>
__device__ __forceinline__ void viscous_face(const Par& p, const Blk& b, int i, int j, int k, int n,
                                             const T* __restrict__ q, const T* __restrict__ mut,
                                             const T* __restrict__ sm, const T* __restrict__ xd,
                                             const T* __restrict__ ed, const T* __restrict__ zd,
                                             const unsigned char* __restrict__ fl, const double* __restrict__ dist,
                                             T* F) {
  const ll n3 = n3of(b);
  const T* qb = q + b.o3 * p.npde * p.nsnap + n3 * (ll)p.npde * n;
  const ll st[4] = {0, 1, (ll)(b.imax + 3), (ll)(b.imax + 3) * (b.jmax + 3)};
  const ll cp = i3(b, i, j, k), cm = cp - st[DIR];
  const int a = DIR == 1 ? i : (DIR == 2 ? j : k);
  const int L = DIR == 1 ? b.imax : (DIR == 2 ? b.jmax : b.kmax);
  const int t1 = DIR == 1 ? j : i, t2 = DIR == 3 ? j : k;
  unsigned char flag = 0;
  if (a == 1 || a == L) flag = face_flag<DIR>(b, fl, a == L ? 1 : 0, t1, t2);
  const bool wall = (flag & 1) != 0;
  // transverse directions, lower-numbered first
  const int d1 = DIR == 1 ? 2 : 1, d2 = DIR == 3 ? 2 : 3;
  ll aux = 0, n1 = 0, n2 = 0;
  T sgnm = 1;
  if (wall) {
    if (a == 1) {
      aux = cm; n1 = cp; n2 = cp + st[DIR]; sgnm = 1;
    } else {
      aux = cp; n1 = cm; n2 = cm - st[DIR]; sgnm = -1;
    }
  }
  T qf[NV], dd[3][NV];
#pragma unroll
  for (int v = 0; v < NV; v++) {
    const T* qv = qb + n3 * v;
    qf[v] = wall ? qv[aux] : (qv[cm] + qv[cp]) / 2;
    for (int d = 1; d <= 3; d++) {
      T val;
      if (!wall) {
        if (d == DIR) {
          val = qv[cp] - qv[cm];
        } else {
          const ll tt = st[d];
          const T q2 = (qv[cm] + qv[cp] + qv[cp + tt] + qv[cm + tt]) / 4;
          const T q1 = (qv[cm - tt] + qv[cp - tt] + qv[cp] + qv[cm]) / 4;
          val = q2 - q1;
        }
      } else {
        if (d == DIR) {
          val = sgnm / 3 * (-qv[aux] * 8 + qv[n1] * 9 - qv[n2]);
        } else {
          const ll tt = st[d];
          const bool first = d == d1 ? (flag & 2) != 0 : (flag & 8) != 0;
          const bool last = d == d1 ? (flag & 4) != 0 : (flag & 16) != 0;
          const T qc = qv[aux], qp1 = qv[aux + tt], qm1 = qv[aux - tt];
          T q2, q1;
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
  T mutf = 0;
  if (NV == 7) {
    const T* mb = mut + b.o3 * p.nsnap + n3 * (ll)n;
    mutf = wall ? mb[aux] : (mb[cm] + mb[cp]) / 2;
  }
  const int nh = n * p.hbmove;
  const T* x3 = mder<T>(p, b, xd, i, j, k, nh);
  const T* e3 = mder<T>(p, b, ed, i, j, k, nh);
  const T* z3 = mder<T>(p, b, zd, i, j, k, nh);
  const T xix = x3[0], xiy = x3[1], xiz = x3[2];
  const T etax = e3[0], etay = e3[1], etaz = e3[2];
  const T zetax = z3[0], zetay = z3[1], zetaz = z3[2];
  const T* s = sm + b.o2 * p.lmet * p.nsnap_met + ((ll)nh * n2of(b) + i2(b, i, j, k)) * p.lmet;
  const T kx = s[0], ky = s[1], kz = s[2], ds = s[3];
  const T gamma = (T)p.gamma;
  const T rho = qf[0], u = qf[1], v = qf[2], w = qf[3], pr = qf[4];
  const T tmp = gamma * pr / rho;
  const T mu = suth<T>(p, tmp);
  const T uxi = dd[0][1], ueta = dd[1][1], uzeta = dd[2][1];
  const T vxi = dd[0][2], veta = dd[1][2], vzeta = dd[2][2];
  const T wxi = dd[0][3], weta = dd[1][3], wzeta = dd[2][3];
  const T dudx = uxi * xix + ueta * etax + uzeta * zetax;
  const T dudy = uxi * xiy + ueta * etay + uzeta * zetay;
  const T dudz = uxi * xiz + ueta * etaz + uzeta * zetaz;
  const T dvdx = vxi * xix + veta * etax + vzeta * zetax;
  const T dvdy = vxi * xiy + veta * etay + vzeta * zetay;
  const T dvdz = vxi * xiz + veta * etaz + vzeta * zetaz;
  const T dwdx = wxi * xix + weta * etax + wzeta * zetax;
  const T dwdy = wxi * xiy + weta * etay + wzeta * zetay;
  const T dwdz = wxi * xiz + weta * etaz + wzeta * zetaz;
  const T mr = (T)(p.machfs / p.reyno);
#pragma unroll
  for (int m = 0; m < NV; m++) F[m] = 0;
  if (NV == 7) {
    // rtst: Reynolds stresses with mut at the face
    const T tke = qf[5];
    const T fact = T(2.0 / 3.0);
    const T divt = (dudx + dvdy + dwdz) * (T)p.turcmp;
    const T ttxx = 2 * mutf * (dudx - divt / 3);
    const T ttxy = mutf * (dudy + dvdx);
    const T ttxz = mutf * (dudz + dwdx);
    const T ttyy = 2 * mutf * (dvdy - divt / 3);
    const T ttyz = mutf * (dvdz + dwdy);
    const T ttzz = 2 * mutf * (dwdz - divt / 3);
    const T r1 = mr * ttxx - fact * rho * tke;
    const T r2 = mr * ttxy;
    const T r3 = mr * ttxz;
    const T r4 = mr * ttyy - fact * rho * tke;
    const T r5 = mr * ttyz;
    const T r6 = mr * ttzz - fact * rho * tke;
    // vflux, k-omega branch
    const T tkexi = dd[0][5], tkeeta = dd[1][5], tkezeta = dd[2][5];
    const T omexi = dd[0][6], omeeta = dd[1][6];
    const T dtkedx = tkexi * xix + tkeeta * etax + tkezeta * zetax;
    const T dtkedy = tkexi * xiy + tkeeta * etay + tkezeta * zetay;
    const T dtkedz = tkexi * xiz + tkeeta * etaz + tkezeta * zetaz;
    const T domedx = omexi * xix + omeeta * etax + dd[2][6] * zetax;
    const T domedy = omexi * xiy + omeeta * etay + dd[2][6] * zetay;
    F[1] = -((r1 * kx + r2 * ky + r3 * kz) * ds);
    F[2] = -((r2 * kx + r4 * ky + r5 * kz) * ds);
    F[3] = -((r3 * kx + r5 * ky + r6 * kz) * ds);
    T sk = (T)p.sigk, sw = (T)p.sigw;
    if (p.sst) {
      // BSL/SST: sigma_k, sigma_omega blended with F1 of the face
      // (src/flux.f90:1035-1046); the face wall distance of facedist_b
      // (src/grid.f90:2834-2884) from the cell-centre distance
      const double* db = dist + b.o1;
      const ll ihi = i1(b, i, j, k);
      const ll ilo = DIR == 1 ? i1(b, i - 1, j, k) : (DIR == 2 ? i1(b, i, j - 1, k) : i1(b, i, j, k - 1));
      const T fd = a == 1 ? (T)db[ilo] : (a == L ? (T)db[ihi] : (T)((db[ilo] + db[ihi]) / 2));
      const T ome = qf[6];
      const T domedz = omexi * xiz + omeeta * etaz + dd[2][6] * zetaz;
      const T de = fd + (T)p.epst;
      T arg1 = fmax(sqrt(tke) / (T)0.09 / ome / de, (T)500 * mu / rho / ome / (de * de) * (T)p.machfs / (T)p.reyno);
      const T cdt = 2 * rho * (T)p.sigw2 / ome * (dtkedx * domedx + dtkedy * domedy + dtkedz * domedz);
      const T acdt = fmax(cdt, (T)1e-20);
      arg1 = fmin(arg1, 4 * rho * (T)p.sigw2 * tke / acdt / (de * de));
      arg1 = fmin(arg1 * arg1 * arg1 * arg1, (T)1e2);   // arg1**4 as ((a a) a) a
      // FP32: exp(2 arg1) overflows above arg1 = 44; F1 is 1 to float precision from arg1 = 9 on
      const T e2 = sizeof(T) == 4 ? exp(fmin(2 * arg1, (T)80)) : exp(2 * arg1);
      const T f1 = (e2 - 1) / (e2 + 1);
      sk = f1 * (T)p.sigk1 + (1 - f1) * (T)p.sigk2;
      sw = f1 * (T)p.sigw1 + (1 - f1) * (T)p.sigw2;
    }
    const T ctke = (mu + sk * mutf) * (T)p.machfs / (T)p.reyno;
    const T come = (mu + sw * mutf) * (T)p.machfs / (T)p.reyno;
    F[4] = -(ctke * (dtkedx * kx + dtkedy * ky + dtkedz * kz) * ds);
    F[5] = -(ctke * (dtkedx * kx + dtkedy * ky + dtkedz * kz) * ds);
    // as in the reference: dk/dz in place of domega/dz
    F[6] = -(come * (domedx * kx + domedy * ky + dtkedz * kz) * ds);
  }
  // laminar stresses and heat flux (common part of vflux_b)
  const T txi = tmp * (dd[0][4] / pr - dd[0][0] / rho);
  const T teta = tmp * (dd[1][4] / pr - dd[1][0] / rho);
  const T tzeta = tmp * (dd[2][4] / pr - dd[2][0] / rho);
  const T dtdx = txi * xix + teta * etax + tzeta * zetax;
  const T dtdy = txi * xiy + teta * etay + tzeta * zetay;
  const T dtdz = txi * xiz + teta * etaz + tzeta * zetaz;
  const T divv = dudx + dvdy + dwdz;
  const T txx = mr * 2 * mu * (dudx - divv / 3);
  const T txy = mr * mu * (dudy + dvdx);
  const T txz = mr * mu * (dudz + dwdx);
  const T tyy = mr * 2 * mu * (dvdy - divv / 3);
  const T tyz = mr * mu * (dvdz + dwdy);
  const T tzz = mr * 2 * mu * (dwdz - divv / 3);
  const T cq = mr * (mu / (T)p.pranl + mutf / (T)p.prant) / (gamma - 1);
  const T qx = -(cq * dtdx), qy = -(cq * dtdy), qz = -(cq * dtdz);
  F[0] = 0;
  F[1] = F[1] - (txx * kx + txy * ky + txz * kz) * ds;
  F[2] = F[2] - (txy * kx + tyy * ky + tyz * kz) * ds;
  F[3] = F[3] - (txz * kx + tyz * ky + tzz * kz) * ds;
  // as in the reference: the z term repeats the y-row stresses
  F[4] = F[4] - ((u * txx + v * txy + w * txz) * kx + (u * txy + v * tyy + w * tyz) * ky +
                 (u * txy + v * tyy + w * tyz) * kz) * ds +
         (qx * kx + qy * ky + qz * kz) * ds;
}

}  // namespace cosahip
