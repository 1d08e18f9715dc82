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
                                               const ll* __restrict__ sbcoff, T* __restrict__ vf) {
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
  viscous_face<DIR, NV, T>(p, b, i, j, k, n, q, mut, sm, xd, ed, zd, sbc + sbcoff[t.block], F);
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
                                                double* __restrict__ res) {
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
  if (a <= b.imax) viscous_face<1, NV, T>(p, b, a, j, k, n, q, mut, sm, xd, ed, zd, sbc + sbcoff[t.block], F);
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
                                                    const ll* __restrict__ sbcoff, double* __restrict__ res) {
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
    if (DIR == 2) viscous_face<2, NV, T>(p, b, i, a, t2, n, q, mut, sm, xd, ed, zd, fl, F);
    else viscous_face<3, NV, T>(p, b, i, t2, a, n, q, mut, sm, xd, ed, zd, fl, F);
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
// k-omega sources at cell centres: the velocity differences of the face
// averages (q_face + q_der_c, three directions), turprod (finest level)
// and srckw into res(6:7).  q primitive.
template <typename T>
__global__ void __launch_bounds__(256) k_turb_src(Par p, int nl, const Blk* __restrict__ blks,
                                                  const Tile* __restrict__ tiles, const T* __restrict__ q,
                                                  const T* __restrict__ mut, double* __restrict__ prd,
                                                  double* __restrict__ divvel, double* __restrict__ delpls,
                                                  double* __restrict__ res, const double* __restrict__ vol,
                                                  const T* const* __restrict__ xdp,
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
  const T* qb = q + b.o3 * p.npde * p.nsnap + n3 * (ll)p.npde * n;
  const ll c = i3(b, i, j, k);
  const ll st[4] = {0, 1, (ll)(b.imax + 3), (ll)(b.imax + 3) * (b.jmax + 3)};
  const unsigned char* fl = sbc + sbcoff[t.block];
  const ll so = b.o3 * p.nsnap + n3 * (ll)n + c;
  if (nl == 1 || p.prd_all_lvls) {
    // dq(d)[v] = qf(face c+1) - qf(face c), v = u, v, w
    T dq[3][3];
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
        const T* qv = qb + n3 * v;
        const T flo = wlo ? qv[c - st[d]] : (qv[c - st[d]] + qv[c]) / 2;
        const T fhi = whi ? qv[c + st[d]] : (qv[c] + qv[c + st[d]]) / 2;
        dq[d - 1][v - 1] = fhi - flo;
      }
    }
    const int nh = n * p.hbmove;
    // xdp: xideri, etaderi, zetaderi, xiderj, etaderj, zetaderj, xiderk, etaderk, zetaderk
    const T xix = (mder<T>(p, b, xdp[0], i + 1, j, k, nh)[0] + mder<T>(p, b, xdp[0], i, j, k, nh)[0]) / 2;
    const T xiy = (mder<T>(p, b, xdp[3], i, j + 1, k, nh)[1] + mder<T>(p, b, xdp[3], i, j, k, nh)[1]) / 2;
    const T xiz = (mder<T>(p, b, xdp[6], i, j, k + 1, nh)[2] + mder<T>(p, b, xdp[6], i, j, k, nh)[2]) / 2;
    const T etax = (mder<T>(p, b, xdp[1], i + 1, j, k, nh)[0] + mder<T>(p, b, xdp[1], i, j, k, nh)[0]) / 2;
    const T etay = (mder<T>(p, b, xdp[4], i, j + 1, k, nh)[1] + mder<T>(p, b, xdp[4], i, j, k, nh)[1]) / 2;
    const T etaz = (mder<T>(p, b, xdp[7], i, j, k + 1, nh)[2] + mder<T>(p, b, xdp[7], i, j, k, nh)[2]) / 2;
    const T zetax = (mder<T>(p, b, xdp[2], i + 1, j, k, nh)[0] + mder<T>(p, b, xdp[2], i, j, k, nh)[0]) / 2;
    const T zetay = (mder<T>(p, b, xdp[5], i, j + 1, k, nh)[1] + mder<T>(p, b, xdp[5], i, j, k, nh)[1]) / 2;
    const T zetaz = (mder<T>(p, b, xdp[8], i, j, k + 1, nh)[2] + mder<T>(p, b, xdp[8], i, j, k, nh)[2]) / 2;
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
  // srckw_b
  const T fact = T(2.0 / 3.0), mor = (T)(p.machfs / p.reyno);
  const T rho = qb[c], tke = qb[c + 5 * n3], ome = qb[c + 6 * n3];
  const T mu_t = mut[so];
  const T divv = (T)divvel[so];
  T prod = (T)prd[so], prodw = (T)prd[so];
  if (p.lim_prodtke && (nl == 1 || p.prd_all_lvls)) {
    const T tstdis = (T)p.prdlim * (T)p.bstrkw * rho * tke * ome;
    if (mor * mu_t * prod > tstdis) {
      prod = tstdis / mor / mu_t;
      if (p.lim_prodomega) prodw = prod;
    }
  }
  const T vv = (T)vol[b.o1 + i1(b, i, j, k)];
  double* rb = res + b.o3 * p.npde * p.nsnap + n3 * (ll)p.npde * n + c;
  rb[5 * n3] = rb[5 * n3] - vv * (mor * mu_t * prod - fact * divv * rho * tke - (T)p.bstrkw * rho * tke * ome);
  rb[6 * n3] = rb[6 * n3] - vv * ((T)p.gkw * (T)p.cmu * rho * prodw - (T)p.gkw * fact * divv * rho * ome -
                                  (T)p.bkw * rho * (ome * ome));
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

template <typename T>
static hipError_t residual_viscous_t(Solve& s, Level& L, const ResIn<T>& in, double* res, T* vflux) {
  const Par& p = s.par;
  hipError_t e;
  if (p.kom) {
    e = launch_h0_last<T>(s, L, in);
    if (e != hipSuccess) return e;
  }
  if (s.fuse_viscous || !p.viscous) {
    for (int d = 1; d <= 3; d++) {
      e = launch_flux_dir<T>(s, L, d, in, res);   // inviscid + viscous fluxes of direction d
      if (e != hipSuccess) return e;
    }
  } else {
    // per direction: inviscid line march, then the face-parallel viscous
    // fluxes into vflux and their difference (faster than fusing the
    // viscous face work into the line march, which runs out of registers)
    const Par porig = s.par;
    Par pi = porig;
    pi.viscous = 0;
    const int fk[3] = {TK_FACES_I, TK_FACES_J, TK_FACES_K};
    for (int d = 1; d <= 3; d++) {
      s.par = pi;
      e = launch_flux_dir<T>(s, L, d, in, res);
      s.par = porig;
      if (e != hipSuccess) return e;
      const T* xd = in.xd[3 * (d - 1)];
      const T* ed = in.xd[3 * (d - 1) + 1];
      const T* zd = in.xd[3 * (d - 1) + 2];
      const int nt = L.ntile[fk[d - 1]];
      const Tile* tl = L.tile[fk[d - 1]];
#define VF(D, NV)                                                                                                 \
  hipLaunchKernelGGL((k_vflux<D, NV, T>), grid1(nt), dim3(TS), 0, s.stream, p, L.dblk, tl, in.q, in.mut, in.s[D - 1], \
                     xd, ed, zd, L.sbc, L.sbc_doff, vflux)
#define VR(D, NV)                                                                                          \
  hipLaunchKernelGGL((k_vresid<D, NV, T>), grid1(L.ntile[TK_CELLS]), dim3(TS), 0, s.stream, p, L.dblk,     \
                     L.tile[TK_CELLS], vflux, res)
#define VI(NV)                                                                                                    \
  hipLaunchKernelGGL((k_vres_i<NV, T>), grid1(L.ntile[TK_ROWS_I]), dim3(256), 0, s.stream, p, L.dblk,             \
                     L.tile[TK_ROWS_I], in.q, in.mut, in.s[0], xd, ed, zd, L.sbc, L.sbc_doff, res)
#define VM(D, NV, TK)                                                                                             \
  hipLaunchKernelGGL((k_vres_march<D, NV, T>), grid1(L.ntile[TK]), dim3(TS), 0, s.stream, p, L.dblk, L.tile[TK],   \
                     in.q, in.mut, in.s[D - 1], xd, ed, zd, L.sbc, L.sbc_doff, res)
      if (s.viscous_faces) {
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
      } else if (p.npde == 7) {
        if (d == 1 && L.ntile[TK_ROWS_I]) VI(7);
        else if (d == 2 && L.ntile[TK_SEGS_J]) VM(2, 7, TK_SEGS_J);
        else if (d == 3 && L.ntile[TK_SEGS_K]) VM(3, 7, TK_SEGS_K);
      } else {
        if (d == 1 && L.ntile[TK_ROWS_I]) VI(5);
        else if (d == 2 && L.ntile[TK_SEGS_J]) VM(2, 5, TK_SEGS_J);
        else if (d == 3 && L.ntile[TK_SEGS_K]) VM(3, 5, TK_SEGS_K);
      }
#undef VR
#undef VI
#undef VM
    }
  }
  if (p.kom && L.ntile[TK_CELLS])
    hipLaunchKernelGGL(k_turb_src<T>, grid1(L.ntile[TK_CELLS]), dim3(TS), 0, s.stream, p, L.nl, L.dblk,
                       L.tile[TK_CELLS], in.q, in.mut, L.f[COSAHIP_PRD], L.f[COSAHIP_DIVVEL], L.f[COSAHIP_DELPLS], res,
                       L.f[COSAHIP_VOL], in.dxd, L.sbc, L.sbc_doff);
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

hipError_t launch_tur_vis(Solve& s, Level& L, const double* q) {
  L.mut32_ok = false;
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
