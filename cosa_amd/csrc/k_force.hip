// k_force.hip -- force and moment coefficients of wall surfaces for gfx950:
// force_b (src/output.f90:4388-4886) and the block sum of force
// (src/output.f90:4290-4330).  Three passes keep the reference's
// summation order, so the coefficients equal the sequential CPU result:
//   k_force_terms  one thread per (wall face, snapshot): the twelve terms the
//                  reference adds to fp(1:6) / fv(1:6) for that face;
//   k_force_sums   one thread per (block, snapshot, surface, component):
//                  the sum over the block's wall faces in patch, i3, i2
//                  order;
//   k_force_coef   one thread per (snapshot, surface): the coefficients of
//                  each block, then their sum over the blocks in order.
#include "kernels.h"
#include "viscous.h"

namespace cosahip {

__global__ void __launch_bounds__(256) k_force_terms(Par p, const Blk* __restrict__ blks,
                                                     const WallRun* __restrict__ runs, int nruns, int nf,
                                                     const double* __restrict__ q, const double* __restrict__ x,
                                                     const double* __restrict__ y, const double* __restrict__ z,
                                                     const double* __restrict__ xdot, const double* __restrict__ ydot,
                                                     const double* __restrict__ zdot, const double* __restrict__ si,
                                                     const double* __restrict__ sj, const double* __restrict__ sk,
                                                     const double* __restrict__ xideri,
                                                     const double* __restrict__ etaderj,
                                                     const double* __restrict__ zetaderk, const double* __restrict__ momr,
                                                     double* __restrict__ T) {
  const ll gid = (ll)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (ll)nf * p.nsnap) return;
  const int n = (int)(gid / nf), f = (int)(gid - (ll)n * nf);
  int r = 0;
  while (r + 1 < nruns && runs[r + 1].first <= f) r++;
  const WallRun w = runs[r];
  const Blk b = blks[w.block];
  const int loc = f - w.first;
  const int t2 = w.lo2 + loc % w.n2, t3 = w.lo3 + loc / w.n2;
  const int idir = w.idir;
  const int L = idir == 1 ? b.imax : (idir == 2 ? b.jmax : b.kmax);
  const int ibcpt = w.inrout == 1 ? 0 : L, ibcm = w.inrout == 1 ? 1 : L, ibcw = w.inrout == 1 ? 0 : L;
  const int sgn = w.inrout == 1 ? 1 : -1;
  // (ic1, ic2, ic3) = cyc(idir, 1:3): face normal, then the two transverse
  // directions in cyclic order
  const int ic2 = idir % 3 + 1, ic3 = (idir + 1) % 3 + 1;
  int c[3], cm[3], cw[3];
  c[idir - 1] = ibcpt;
  cm[idir - 1] = ibcm;
  cw[idir - 1] = ibcw;
  c[ic2 - 1] = cm[ic2 - 1] = cw[ic2 - 1] = t2;
  c[ic3 - 1] = cm[ic3 - 1] = cw[ic3 - 1] = t3;
  const int nh = n * p.hbmove;
  const ll n2 = n2of(b), n3 = n3of(b);
  const double* sm = idir == 1 ? si : (idir == 2 ? sj : sk);
  const double* s = sm + b.o2 * p.lmet * p.nsnap_met + ((ll)nh * n2 + i2(b, cm[0], cm[1], cm[2])) * p.lmet;
  const double nx = s[0], ny = s[1], nz = s[2], ds = s[3];
  const double kx = sgn * nx, ky = sgn * ny, kz = sgn * nz;
  const double* qb = q + b.o3 * p.npde * p.nsnap + n3 * (ll)p.npde * n;
  const ll xb = i3(b, c[0], c[1], c[2]);
  const double r1 = qb[xb], r2 = qb[xb + n3], r3 = qb[xb + 2 * n3], r4 = qb[xb + 3 * n3];
  const double gamma = p.gamma;
  const double pw = (gamma - 1) * (qb[xb + 4 * n3] - r2 * r2 / r1 / 2 - r3 * r3 / r1 / 2 - r4 * r4 / r1 / 2);
  const ll xo = b.o2 * p.nsnap_met + (ll)nh * n2;
  const ll wi = xo + i2(b, cw[0], cw[1], cw[2]);
  const double xarm = x[wi] - momr[3 * n], yarm = y[wi] - momr[3 * n + 1], zarm = z[wi] - momr[3 * n + 2];
  double* t = T + ((ll)n * nf + f) * 12;
  t[0] = pw * kx * ds;
  t[1] = pw * ky * ds;
  t[2] = pw * kz * ds;
  t[3] = (ky * zarm - kz * yarm) * pw * ds;
  t[4] = (kz * xarm - kx * zarm) * pw * ds;
  t[5] = (kx * yarm - ky * xarm) * pw * ds;
  if (!w.visc) {
#pragma unroll
    for (int m = 6; m < 12; m++) t[m] = 0.0;
    return;
  }
  // viscous wall shear from the one-sided wall-normal derivative
  const double* ed = idir == 1 ? xideri : (idir == 2 ? etaderj : zetaderk);
  const double* e3 = mder<double>(p, b, ed, cm[0], cm[1], cm[2], nh);
  const double etax = e3[0], etay = e3[1], etaz = e3[2];
  const double rhow = r1;
  const double tw = pw * gamma / rhow;
  const double muw = (p.stemp + 1) / (tw + p.stemp) * pow(tw, 1.5);
  int c1[3] = {c[0], c[1], c[2]}, c2[3] = {c[0], c[1], c[2]};
  c1[idir - 1] += sgn;
  c2[idir - 1] += 2 * sgn;
  const ll x1 = i3(b, c1[0], c1[1], c1[2]), x2 = i3(b, c2[0], c2[1], c2[2]);
  double uw = qb[xb + n3] / qb[xb], u1 = qb[x1 + n3] / qb[x1], u2 = qb[x2 + n3] / qb[x2];
  double vw = qb[xb + 2 * n3] / qb[xb], v1 = qb[x1 + 2 * n3] / qb[x1], v2 = qb[x2 + 2 * n3] / qb[x2];
  double ww = qb[xb + 3 * n3] / qb[xb], w1 = qb[x1 + 3 * n3] / qb[x1], w2 = qb[x2 + 3 * n3] / qb[x2];
  if (p.moving) {
    const ll g0 = xo + i2(b, c[0], c[1], c[2]), g1 = xo + i2(b, c1[0], c1[1], c1[2]),
             g2 = xo + i2(b, c2[0], c2[1], c2[2]);
    uw = uw - xdot[g0];
    u1 = u1 - xdot[g1];
    u2 = u2 - xdot[g2];
    vw = vw - ydot[g0];
    v1 = v1 - ydot[g1];
    v2 = v2 - ydot[g2];
    ww = ww - zdot[g0];
    w1 = w1 - zdot[g1];
    w2 = w2 - zdot[g2];
  }
  const double ueta = sgn * (-8 * uw + 9 * u1 - u2) / 3;
  const double veta = sgn * (-8 * vw + 9 * v1 - v2) / 3;
  const double weta = sgn * (-8 * ww + 9 * w1 - w2) / 3;
  const double dudx = ueta * etax, dudy = ueta * etay, dudz = ueta * etaz;
  const double dvdx = veta * etax, dvdy = veta * etay, dvdz = veta * etaz;
  const double dwdx = weta * etax, dwdy = weta * etay, dwdz = weta * etaz;
  const double divv = dudx + dvdy + dwdz;
  const double txx = 2 * muw * (dudx - divv / 3);
  const double txy = muw * (dudy + dvdx);
  const double txz = muw * (dudz + dwdx);
  const double tyy = 2 * muw * (dvdy - divv / 3);
  const double tyz = muw * (dvdz + dwdy);
  const double tzz = 2 * muw * (dwdz - divv / 3);
  const double mr = p.machfs / p.reyno;
  t[6] = mr * ds * (txx * kx + txy * ky + txz * kz);
  t[7] = mr * ds * (txy * kx + tyy * ky + tyz * kz);
  t[8] = mr * ds * (txz * kx + tyz * ky + tzz * kz);
  t[9] = mr * ds * (-(txy * kx + tyy * ky + tyz * kz) * zarm + (txz * kx + tyz * ky + tzz * kz) * yarm);
  t[10] = mr * ds * (-(txz * kx + tyz * ky + tzz * kz) * xarm + (txx * kx + txy * ky + txz * kz) * zarm);
  t[11] = mr * ds * (-(txx * kx + txy * ky + txz * kz) * yarm + (txy * kx + tyy * ky + tyz * kz) * xarm);
}

// fp(1:3) -= t, fp(4:6) += t, fv(1:6) += t over the faces of block b that
// belong to surface s, in run (patch) order.
__global__ void __launch_bounds__(64) k_force_sums(int nblocks, int nsnap, int nsurface,
                                                   const WallRun* __restrict__ runs, int nruns, int nf,
                                                   const double* __restrict__ T, double* __restrict__ S) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= nblocks * nsnap * nsurface * 12) return;
  const int cpt = gid % 12;
  int rest = gid / 12;
  const int s = rest % nsurface + 1;
  rest /= nsurface;
  const int n = rest % nsnap, b = rest / nsnap;
  double acc = 0.0;
  for (int r = 0; r < nruns; r++) {
    const WallRun w = runs[r];
    if (w.block != b || w.surface != s) continue;
    if (cpt >= 6 && !w.visc) continue;
    const double* t = T + ((ll)n * nf + w.first) * 12 + cpt;
    if (cpt < 3) {
      for (int k = 0; k < w.count; k++) acc = acc - t[(ll)k * 12];
    } else {
      for (int k = 0; k < w.count; k++) acc = acc + t[(ll)k * 12];
    }
  }
  S[(((ll)b * nsnap + n) * nsurface + (s - 1)) * 12 + cpt] = acc;
}

// Coefficients per block (src/output.f90:4857-4880), summed over the
// blocks in order (src/output.f90:4310-4325).  g = {aax, aay, aaz, ssx,
// ssy, ssz, bbx, bby, bbz, lngth(1:3), mach}.
__global__ void __launch_bounds__(64) k_force_coef(int nblocks, int nsnap, int nsurface, const double* __restrict__ S,
                                                   const double* __restrict__ g, double* __restrict__ coef) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= nsnap * nsurface) return;
  const int n = gid % nsnap, s = gid / nsnap;
  const double aax = g[0], aay = g[1], aaz = g[2], ssx = g[3], ssy = g[4], ssz = g[5], bbx = g[6], bby = g[7],
               bbz = g[8], l1 = g[9], l2 = g[10], l3 = g[11], mach = g[12];
  double tot[12];
#pragma unroll
  for (int m = 0; m < 12; m++) tot[m] = 0.0;
  for (int b = 0; b < nblocks; b++) {
    const double* fp = S + (((ll)b * nsnap + n) * nsurface + s) * 12;
    const double* fv = fp + 6;
    double cb[12];
    cb[0] = (fp[0] * aax + fp[1] * aay + fp[2] * aaz) * 2 / (l1 * l2 * (mach * mach));
    cb[1] = (fp[0] * ssx + fp[1] * ssy + fp[2] * ssz) * 2 / (l1 * l2 * (mach * mach));
    cb[2] = (fp[0] * bbx + fp[1] * bby + fp[2] * bbz) * 2 / (l1 * l2 * (mach * mach));
    cb[3] = fp[3] * 2 / (l1 * l2 * l3 * (mach * mach));
    cb[4] = fp[4] * 2 / (l1 * l2 * l3 * (mach * mach));
    cb[5] = fp[5] * 2 / (l1 * l2 * l3 * (mach * mach));
    cb[6] = (fv[0] * aax + fv[1] * aay + fv[2] * aaz) * 2 / (l1 * l2 * (mach * mach));
    cb[7] = (fv[0] * ssx + fv[1] * ssy + fv[2] * ssz) * 2 / (l1 * l2 * (mach * mach));
    cb[8] = (fv[0] * bbx + fv[1] * bby + fv[2] * bbz) * 2 / (l1 * l2 * (mach * mach));
    cb[9] = fv[3] * 2 / (l1 * l2 * l3 * (mach * mach));
    cb[10] = fv[4] * 2 / (l1 * l2 * l3 * (mach * mach));
    cb[11] = fv[5] * 2 / (l1 * l2 * l3 * (mach * mach));
#pragma unroll
    for (int m = 0; m < 12; m++) tot[m] = tot[m] + cb[m];
  }
  double* o = coef + ((ll)n + (ll)nsnap * s) * 12;
#pragma unroll
  for (int m = 0; m < 12; m++) o[m] = tot[m];
}

hipError_t launch_force(Solve& s, Level& L, int nsurface, const ForceBufs& fb) {
  const Par& p = s.par;
  const ll nt = (ll)fb.nfaces * p.nsnap;
  if (fb.nruns && nt)
    hipLaunchKernelGGL(k_force_terms, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, s.stream, p, L.dblk, fb.runs,
                       fb.nruns, fb.nfaces, L.f[COSAHIP_Q], L.f[COSAHIP_X], L.f[COSAHIP_Y], L.f[COSAHIP_Z],
                       L.f[COSAHIP_XDOT], L.f[COSAHIP_YDOT], L.f[COSAHIP_ZDOT], L.f[COSAHIP_SI], L.f[COSAHIP_SJ],
                       L.f[COSAHIP_SK], L.f[COSAHIP_XIDERI], L.f[COSAHIP_ETADERJ], L.f[COSAHIP_ZETADERK], fb.momr,
                       fb.terms);
  const int ns = L.nblocks * p.nsnap * nsurface * 12;
  hipLaunchKernelGGL(k_force_sums, dim3((ns + 63) / 64), dim3(64), 0, s.stream, L.nblocks, p.nsnap, nsurface,
                     fb.runs, fb.nruns, fb.nfaces, fb.terms, fb.sums);
  const int nc = p.nsnap * nsurface;
  hipLaunchKernelGGL(k_force_coef, dim3((nc + 63) / 64), dim3(64), 0, s.stream, L.nblocks, p.nsnap, nsurface,
                     fb.sums, fb.geom, fb.coef);
  return hipGetLastError();
}

}  // namespace cosahip
