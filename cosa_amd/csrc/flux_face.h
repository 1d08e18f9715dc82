// flux_face.h -- per-face arithmetic of COSA's inviscid residual, shared by
// the residual kernels: Roe flux-difference splitting (roflux_b,
// src/flux.f90:45-387), the limiters of limit_i/j/k (src/flux.f90:3472-3888)
// and the MUSCL face states of muscl_b* (src/flux.f90:1221-1812).  Each
// function follows the reference's operations in order, so a translation
// unit built with -ffp-contract=off reproduces the CPU path's rounding.
//
// Compile-time options (default -1: read the run-time Par flag):
//   ILIM  limiter id 0..4 of the mean-flow variables,
//   FO    1: first order (the level is at or below nl_crs), 0: MUSCL,
//   EF    0: no entropy fix and no eigenvalue floor.
#pragma once
#include "cosahip_internal.h"

namespace cosahip {

// ------------------------------------------------------------------------
// Roe flux-difference splitting for one face (mean-flow variables).
template <int NV, typename T = double, int EF = -1>
__device__ __forceinline__ void roe_face(const Par& p, const T* qr, const T* ql, T kx, T ky, T kz, T ds, T ubn,
                                         T* F, T* h0lr = nullptr) {
  const T gamma = (T)p.gamma;
  // gamma / (gamma - 1) of every face: in FP64 the host's quotient (the
  // same correctly rounded value, one division less per face)
  const T ggm1 = sizeof(T) == 8 ? (T)p.gam_gm1 : gamma / (gamma - 1);
  T rhor = qr[0], ur = qr[1], vr = qr[2], wr = qr[3], pr = qr[4];
  T unr = ur * kx + vr * ky + wr * kz - ubn;
  T q2r = ur * ur + vr * vr + wr * wr;
  T tker = NV > 5 ? qr[5] : T(0);
  T h0r = ggm1 * pr / rhor + q2r / 2 + tker;
  T rh0r = rhor * h0r;
  T fr0 = rhor * unr;
  T fr1 = (rhor * ur * unr + pr * kx);
  T fr2 = (rhor * vr * unr + pr * ky);
  T fr3 = (rhor * wr * unr + pr * kz);
  T fr4 = (rh0r * unr + pr * ubn);
  T rhol = ql[0], ul = ql[1], vl = ql[2], wl = ql[3], pl = ql[4];
  T unl = ul * kx + vl * ky + wl * kz - ubn;
  T q2l = ul * ul + vl * vl + wl * wl;
  T tkel = NV > 5 ? ql[5] : T(0);
  T h0l = ggm1 * pl / rhol + q2l / 2 + tkel;
  if (h0lr) {
    h0lr[0] = h0l;
    h0lr[1] = h0r;
  }
  T rh0l = rhol * h0l;
  T fl0 = rhol * unl;
  T fl1 = (rhol * ul * unl + pl * kx);
  T fl2 = (rhol * vl * unl + pl * ky);
  T fl3 = (rhol * wl * unl + pl * kz);
  T fl4 = (rh0l * unl + pl * ubn);
  T drho = rhor - rhol, dp = pr - pl, du = ur - ul, dv = vr - vl, dw = wr - wl, dun = unr - unl;
  T dtke = tker - tkel;
  T rhols = sqrt(rhol), rhors = sqrt(rhor);
  T rho = rhols * rhors;
  T denom = rhols + rhors;
  T u = (rhols * ul + rhors * ur) / denom;
  T v = (rhols * vl + rhors * vr) / denom;
  T w = (rhols * wl + rhors * wr) / denom;
  T un = u * kx + v * ky + w * kz;
  T h0 = (rhols * h0l + rhors * h0r) / denom;
  // mean flow alone: tke = (rhols 0 + rhors 0) / denom is +0 for the
  // positive finite weights, so it is taken as the constant
  T tke = NV > 5 ? (rhols * tkel + rhors * tker) / denom : T(0);
  T q2 = u * u + v * v + w * w;
  T a2 = (gamma - 1) * (h0 - q2 / 2 - tke);
  T a = sqrt(a2);
  T rm123 = fabs(un - ubn), rm4 = fabs(un + a - ubn), rm5 = fabs(un - a - ubn);
  if (EF != 0 && p.efx_setup_f) {
    T rdelta = (T)p.cntrpy_f * (p.entfx_f_w[0] * (T)p.entfxctf_f + p.entfx_f_w[1] * (sqrt(u * u + v * v + w * w) + a) +
                                  p.entfx_f_w[2] * T(0.2) * (fabs(un) + a + fabs(ubn)));
    rdelta = fmax(rdelta, T(1.e-12));
    if (rm123 <= rdelta && p.efx_all_f) rm123 = T(0.5) * (rm123 * rm123 + rdelta * rdelta) / rdelta;
    if (rm4 <= rdelta) rm4 = T(0.5) * (rm4 * rm4 + rdelta * rdelta) / rdelta;
    if (rm5 <= rdelta) rm5 = T(0.5) * (rm5 * rm5 + rdelta * rdelta) / rdelta;
  }
  if (EF != 0 && p.eig_limit_f) {
    if (rm123 <= (T)p.eig_cutoff_f) rm123 = (T)p.eig_cutoff_f;
    if (p.eig_lim_all_f && rm4 <= (T)p.eig_cutoff_f) rm4 = (T)p.eig_cutoff_f;
    if (p.eig_lim_all_f && rm5 <= (T)p.eig_cutoff_f) rm5 = (T)p.eig_cutoff_f;
  }
  T al1 = rm123 * (drho - dp / a2);
  T al2 = rm123 * rho;
  T al3 = rm4 * (dp / a2 + rho * dun / a) / 2;
  T al4 = rm5 * (dp / a2 - rho * dun / a) / 2;
  T df0 = al1 + al3 + al4;
  T df1 = al1 * u + al2 * (du - dun * kx) + al3 * (u + a * kx) + al4 * (u - a * kx);
  T df2 = al1 * v + al2 * (dv - dun * ky) + al3 * (v + a * ky) + al4 * (v - a * ky);
  T df3 = al1 * w + al2 * (dw - dun * kz) + al3 * (w + a * kz) + al4 * (w - a * kz);
  T df4 = al1 * (q2 / 2 + tke) + al2 * (u * du + v * dv + w * dw - un * dun + dtke) + al3 * (h0 + a * un) +
               al4 * (h0 - a * un);
  F[0] = (fr0 + fl0 - df0) / 2 * ds;
  F[1] = (fr1 + fl1 - df1) / 2 * ds;
  F[2] = (fr2 + fl2 - df2) / 2 * ds;
  F[3] = (fr3 + fl3 - df3) / 2 * ds;
  F[4] = (fr4 + fl4 - df4) / 2 * ds;
}

// Turbulence part of roflux_b (src/flux.f90:250-383): k and omega upwinded
// with the mean-flow waves.  h0l/h0r are those of the block's last face of
// the mean-flow loop (the reference reuses the variables), from h0tab.
template <typename T = double>
__device__ __forceinline__ void roe_turb(const Par& p, const T* qr, const T* ql, T kx, T ky, T kz, T ds, T ubn,
                                         T h0l, T h0r, T* F) {
  const T gamma = (T)p.gamma;
  T rhor = qr[0], ur = qr[1], vr = qr[2], wr = qr[3], pr = qr[4], omer = qr[6];
  T unr = ur * kx + vr * ky + wr * kz - ubn;
  T tker = qr[5];
  T fr1 = rhor * unr * tker, fr2 = rhor * unr * omer;
  T rhol = ql[0], ul = ql[1], vl = ql[2], wl = ql[3], pl = ql[4], omel = ql[6];
  T unl = ul * kx + vl * ky + wl * kz - ubn;
  T tkel = ql[5];
  T fl1 = rhol * unl * tkel, fl2 = rhol * unl * omel;
  T drho = rhor - rhol, dp = pr - pl, dun = unr - unl, dtke = tker - tkel, dome = omer - omel;
  T rhols = sqrt(rhol), rhors = sqrt(rhor);
  T rho = rhols * rhors;
  T denom = rhols + rhors;
  T u = (rhols * ul + rhors * ur) / denom;
  T v = (rhols * vl + rhors * vr) / denom;
  T w = (rhols * wl + rhors * wr) / denom;
  T un = u * kx + v * ky + w * kz;
  T h0 = (rhols * h0l + rhors * h0r) / denom;
  T tke = (rhols * tkel + rhors * tker) / denom;
  T ome = (rhols * omel + rhors * omer) / denom;
  T q2 = u * u + v * v + w * w;
  T a2 = (gamma - 1) * (h0 - q2 / 2 - tke);
  T a = sqrt(a2);
  T rm123 = fabs(un - ubn), rm4 = fabs(un + a - ubn), rm5 = fabs(un - a - ubn);
  if (p.efx_setup_t) {
    T rdelta = (T)p.cntrpy_t * (p.entfx_t_w[0] * (T)p.entfxctf_t + p.entfx_t_w[1] * (sqrt(u * u + v * v + w * w) + a) +
                                  p.entfx_t_w[2] * T(0.2) * (fabs(un) + a + fabs(ubn)));
    rdelta = fmax(rdelta, T(1.e-12));
    if (rm123 <= rdelta && p.efx_all_t) rm123 = T(0.5) * (rm123 * rm123 + rdelta * rdelta) / rdelta;
    if (rm4 <= rdelta) rm4 = T(0.5) * (rm4 * rm4 + rdelta * rdelta) / rdelta;
    if (rm5 <= rdelta) rm5 = T(0.5) * (rm5 * rm5 + rdelta * rdelta) / rdelta;
  }
  if (p.eig_limit_t) {
    if (rm123 <= (T)p.eig_cutoff_t) rm123 = (T)p.eig_cutoff_t;
    if (p.eig_lim_all_t && rm4 <= (T)p.eig_cutoff_t) rm4 = (T)p.eig_cutoff_t;
    if (p.eig_lim_all_t && rm5 <= (T)p.eig_cutoff_t) rm5 = (T)p.eig_cutoff_t;
  }
  T al1 = rm123 * (drho - dp / a2);
  T al2 = rm123 * rho;
  T al3 = rm4 * (dp / a2 + rho * dun / a) / 2;
  T al4 = rm5 * (dp / a2 - rho * dun / a) / 2;
  T df1 = (al1 + al3 + al4) * tke + al2 * dtke;
  T df2 = (al1 + al3 + al4) * ome + al2 * dome;
  F[5] = (fr1 + fl1 - df1) / 2 * ds;
  F[6] = (fr2 + fl2 - df2) / 2 * ds;
}

// ------------------------------------------------------------------------
// Van Albada slope of the cell between differences d1 (upper face) and d2
// (lower face), epsq = epslim^2 (limit_i ilim 4, src/flux.f90:3584-3604).
template <typename T = double>
__device__ __forceinline__ T va_slope(T epsq, T d1, T d2) {
  return (d1 * (d2 * d2 + epsq) + d2 * (d1 * d1 + epsq)) / (d1 * d1 + d2 * d2 + 2 * epsq);
}

// Limited slopes of one variable at face a: dqp(a), dqm(a) from
// dm1 = dq(a-1), d0 = dq(a), dp1 = dq(a+1)   (limit_i/j/k).
template <typename T = double>
__device__ __forceinline__ void limit_face(int ilim, T eps, T rkap, T dm1, T d0, T dp1, T& dqp, T& dqm) {
  if (ilim == 0) {
    T rkp = 1 + rkap, rkm = 1 - rkap;
    dqp = -(rkp * d0 + rkm * dp1) / 4;
    dqm = (rkm * dm1 + rkp * d0) / 4;
  } else if (ilim == 1) {
    T s = (2 * dp1 * d0 + eps) / (dp1 * dp1 + d0 * d0 + eps);
    T rkp = 1 + rkap * s, rkm = 1 - rkap * s;
    dqp = -s * (rkp * d0 + rkm * dp1) / 4;
    s = (2 * d0 * dm1 + eps) / (d0 * d0 + dm1 * dm1 + eps);
    rkp = 1 + rkap * s;
    rkm = 1 - rkap * s;
    dqm = s * (rkm * dm1 + rkp * d0) / 4;
  } else if (ilim == 2) {
    T rkp = 1 + rkap, rkm = 1 - rkap, b = (T(3) - rkap) / (T(1) - rkap);
    // dqp(a): loop index a with dq(a), dq(a+1)
    T dqmx = (copysign(T(1), d0) + copysign(T(1), dp1)) / 2 * fmin(fabs(d0), fabs(dp1) * b);
    T dqpx = (copysign(T(1), dp1) + copysign(T(1), d0)) / 2 * fmin(fabs(dp1), fabs(d0) * b);
    dqp = -(rkp * dqmx + rkm * dqpx) / 4;
    // dqm(a): loop index a-1 with dq(a-1), dq(a)
    dqmx = (copysign(T(1), dm1) + copysign(T(1), d0)) / 2 * fmin(fabs(dm1), fabs(d0) * b);
    dqpx = (copysign(T(1), d0) + copysign(T(1), dm1)) / 2 * fmin(fabs(d0), fabs(dm1) * b);
    dqm = (rkm * dqmx + rkp * dqpx) / 4;
  } else if (ilim == 3) {
    T epsq = eps * eps;
    T d1 = dp1, d2 = d0;
    T dqpx = (d2 * (d1 * d1 + 2 * epsq) + d1 * (2 * d2 * d2 + epsq)) / (2 * d2 * d2 - d2 * d1 + 2 * d1 * d1 + 3 * epsq);
    dqp = -dqpx / 2;
    d1 = d0;
    d2 = dm1;
    T dqmx = (d1 * (d2 * d2 + 2 * epsq) + d2 * (2 * d1 * d1 + epsq)) / (2 * d1 * d1 - d1 * d2 + 2 * d2 * d2 + 3 * epsq);
    dqm = dqmx / 2;
  } else {
    const T epsq = eps * eps;
    dqp = -va_slope<T>(epsq, dp1, d0) / 2;
    dqm = va_slope<T>(epsq, d0, dm1) / 2;
  }
}

// Face states qp (right), qm (left) of face a from q(a-2..a+1) of one line.
// fin/fout: the line's inner (a = 1) / outer (a = L) face is a strong BC.
// k, omega (v >= 5) use the turbulence limiter and are first order with
// foturb (muscl_b*, src/flux.f90:1221-1416).
template <int NV, typename T = double, int ILIM = -1, int FO = -1>
__device__ __forceinline__ void muscl_face(const Par& p, int nl, int a, int L, bool fin, bool fout, const T* qm2,
                                           const T* qm1, const T* q0, const T* qp1, T* QR, T* QL) {
  if (FO < 0 ? nl >= p.nl_crs : FO == 1) {
#pragma unroll
    for (int v = 0; v < NV; v++) {
      QR[v] = q0[v];
      QL[v] = qm1[v];
    }
  } else {
    bool sm1 = (a - 1 == 1 && fin) || (a - 1 == L && fout);
    bool s0 = (a == 1 && fin) || (a == L && fout);
    bool sp1 = (a + 1 == 1 && fin) || (a + 1 == L && fout);
#pragma unroll
    for (int v = 0; v < NV; v++) {
      if (v >= 5 && p.foturb) {
        QR[v] = q0[v];
        QL[v] = qm1[v];
        continue;
      }
      T dm1 = qm1[v] - qm2[v];
      T d0 = q0[v] - qm1[v];
      T dp1 = qp1[v] - q0[v];
      if (sm1) dm1 = (qm1[v] - qm2[v]) * 2;
      if (s0) d0 = (q0[v] - qm1[v]) * 2;
      if (sp1) dp1 = (qp1[v] - q0[v]) * 2;
      T dqp, dqm;
      if (v < 5) limit_face<T>(ILIM < 0 ? p.ilimf : ILIM, (T)p.epslimf, (T)p.rkap, dm1, d0, dp1, dqp, dqm);
      else limit_face<T>(p.ilimt, (T)p.epslimt, (T)p.rkap, dm1, d0, dp1, dqp, dqm);
      QR[v] = q0[v] + dqp;
      QL[v] = qm1[v] + dqm;
    }
  }
  if (a == 1 && fin) {
#pragma unroll
    for (int v = 0; v < NV; v++) QR[v] = QL[v] = qm1[v];
  }
  if (a == L && fout) {
#pragma unroll
    for (int v = 0; v < NV; v++) QR[v] = QL[v] = q0[v];
  }
}

template <typename T = double>
__device__ __forceinline__ void face_metric(const Par& p, const T* s, T& kx, T& ky, T& kz, T& ds, T& ubn) {
  kx = s[0];
  ky = s[1];
  kz = s[2];
  ds = s[3];
  ubn = T(0);
  if (p.moving) ubn = s[4] * kx + s[5] * ky + s[6] * kz;
}

}  // namespace cosahip
