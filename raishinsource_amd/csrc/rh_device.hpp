// Device-side pointwise physics of the RAISHIN GRMHD step (gfx950, FP64).
//
// Every function reproduces the reference arithmetic operation for operation
// (same association, same single-precision literals) so that, compiled with
// -ffp-contract=off, results are bit-identical to the Fortran reference.
// Fortran intrinsic semantics: MAX(a,b) = a>b?a:b, MIN(a,b) = a<b?a:b,
// SIGN(1d0,x) = copysign(1,x).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define RH_DEV __device__ __forceinline__

namespace rh {

RH_DEV double fmaxf_(double a, double b) { return a > b ? a : b; }
RH_DEV double fminf_(double a, double b) { return a < b ? a : b; }
RH_DEV double fsign1(double x) { return copysign(1.0, x); }

// Geometry of one model on the device.
struct Geo {
  int ni, nj, nk;        // imax, jmax, kmax (ghosts included)
  int nm0, nm1;          // ghost widths (main.f90:179-211)
  long long N;           // ni*nj*nk
  long long P2;          // ni*nk: plane size of the axisymmetric (i,k) tables
};

// One metric location (centre or a face) from the (i,k) tables:
// planes 0..15 gcov(m,n) at m+4n, 16..31 gcon, 32 sqrt(-g).
struct Met {
  const double* __restrict__ b;
  long long P2, q;
  RH_DEV double g(int m, int n) const { return b[(long long)(m + 4 * n) * P2 + q]; }
  RH_DEV double gi(int m, int n) const { return b[(long long)(16 + m + 4 * n) * P2 + q]; }
  RH_DEV double detg() const { return b[32LL * P2 + q]; }
};

// physfunc.f90 helpers -------------------------------------------------

// calgfl: physfunc.f90:2-41
RH_DEV double calgfl(const double u1, const double u2, const double u3, const Met& M) {
  const double ut[4] = {0.0, u1, u2, u3};
  double utsq = 0.0;
#pragma unroll
  for (int m = 1; m <= 3; m++)
#pragma unroll
    for (int n = 1; n <= 3; n++) utsq = utsq + M.g(m, n) * ut[m] * ut[n];
  if (utsq < 0.0) {
    if (fabs(utsq) <= 1.0e-10) utsq = fabs(utsq);
    else utsq = 1.0e-10;
  }
  if (fabs(utsq) > 5.0e2) utsq = 5.0e2;
  return sqrt(1.0 + utsq);
}

// calucon: physfunc.f90:44-74
RH_DEV void calucon(const double ut[4], double ucon[4], double gfl, const Met& M) {
  const double alpha1 = 1.0 / sqrt(-M.gi(0, 0));
  const double b1 = alpha1 * alpha1 * M.gi(0, 1);
  const double b2 = alpha1 * alpha1 * M.gi(0, 2);
  const double b3 = alpha1 * alpha1 * M.gi(0, 3);
  ucon[0] = gfl * (1.0 / alpha1);
  ucon[1] = ut[1] - gfl * b1 * (1.0 / alpha1);
  ucon[2] = ut[2] - gfl * b2 * (1.0 / alpha1);
  ucon[3] = ut[3] - gfl * b3 * (1.0 / alpha1);
}

// lower: physfunc.f90:158-192
RH_DEV void lower(const double v[4], double w[4], const Met& M) {
#pragma unroll
  for (int m = 0; m < 4; m++)
    w[m] = M.g(m, 0) * v[0] + M.g(m, 1) * v[1] + M.g(m, 2) * v[2] + M.g(m, 3) * v[3];
}
// upper: physfunc.f90:195-230
RH_DEV void upper(const double v[4], double w[4], const Met& M) {
#pragma unroll
  for (int m = 0; m < 4; m++)
    w[m] = M.gi(m, 0) * v[0] + M.gi(m, 1) * v[1] + M.gi(m, 2) * v[2] + M.gi(m, 3) * v[3];
}

// calbbcon: physfunc.f90:77-95
RH_DEV void calbbcon(const double bcon[4], double bb[4], const double ucov[4], const double ucon[4]) {
  bb[0] = bcon[1] * ucov[1] + bcon[2] * ucov[2] + bcon[3] * ucov[3];
  bb[1] = (bcon[1] + bb[0] * ucon[1]) * (1.0 / ucon[0]);
  bb[2] = (bcon[2] + bb[0] * ucon[2]) * (1.0 / ucon[0]);
  bb[3] = (bcon[3] + bb[0] * ucon[3]) * (1.0 / ucon[0]);
}

// cal3met: coord.f90:970-1013; g3/gi3 hold gcov3a/gcon3a(m,n) at (m-1)+3(n-1)
RH_DEV void cal3met(const Met& M, double g3[9], double gi3[9]) {
  const double alpha1 = 1.0 / sqrt(-M.gi(0, 0));
  double beta1[4];
  beta1[1] = alpha1 * alpha1 * M.gi(0, 1);
  beta1[2] = alpha1 * alpha1 * M.gi(0, 2);
  beta1[3] = alpha1 * alpha1 * M.gi(0, 3);
  double oncon[4], oncov[4];
  oncon[0] = 1.0 / alpha1;
  oncon[1] = beta1[1] / alpha1;
  oncon[2] = beta1[2] / alpha1;
  oncon[3] = beta1[3] / alpha1;
  oncov[0] = -alpha1; oncov[1] = 0.0; oncov[2] = 0.0; oncov[3] = 0.0;
#pragma unroll
  for (int m = 1; m <= 3; m++)
#pragma unroll
    for (int n = 1; n <= 3; n++) {
      g3[(m - 1) + 3 * (n - 1)] = M.g(m, n) + (oncov[m] * oncov[n]);
      gi3[(m - 1) + 3 * (n - 1)] = M.gi(m, n) + (oncon[m] * oncon[n]);
    }
}

// Shared prelude of caluu1/caluu3/calflx3*/caltenr/calcha4*:
// primitives -> enthalpy, Lorentz factor, u^mu, u_mu, b^mu, b_mu, b^2.
struct Prel {
  double de, pr, roh, gfl, bbsq;
  double ut[4], bcon[4], ucon[4], ucov[4], bb[4], bbcov[4];
};

RH_DEV void prelude(const double w[9], double gam, const Met& M, Prel& s) {
  s.de = w[0];
  s.ut[0] = 0.0; s.ut[1] = w[1]; s.ut[2] = w[2]; s.ut[3] = w[3];
  s.pr = w[4];
  s.bcon[0] = 0.0; s.bcon[1] = w[6]; s.bcon[2] = w[7]; s.bcon[3] = w[8];
  s.roh = s.de + (gam / (gam - 1.0)) * s.pr;   // caluu.f90:52 (ieos=0)
  s.gfl = calgfl(w[1], w[2], w[3], M);
  calucon(s.ut, s.ucon, s.gfl, M);
  lower(s.ucon, s.ucov, M);
  calbbcon(s.bcon, s.bb, s.ucov, s.ucon);
  lower(s.bb, s.bbcov, M);
  s.bbsq = s.bb[0] * s.bbcov[0] + s.bb[1] * s.bbcov[1] + s.bb[2] * s.bbcov[2] + s.bb[3] * s.bbcov[3];
}

// Row d of sqrt(-g) * (rho u^d, T^d_i, T^d_t + rho u^d, rho s u^d, induction):
// d=0 is caluu1/caluu3 (caluu.f90:85-103), d=1..3 calflx3i/j/k (calflx.f90:468-507,
// 608-646, 744-782). powdg = de**gam computed once by the caller.
RH_DEV void tmunu_row(const Prel& s, int d, double detg, double powdg, double out[9]) {
  double delta[4] = {0.0, 0.0, 0.0, 0.0};
  delta[d] = 1.0;
  const double rb = s.roh + s.bbsq, pb = s.pr + 0.5 * s.bbsq;
  const double ud = s.ucon[d];
  out[0] = s.de * ud;
  out[1] = rb * ud * s.ucov[1] + pb * delta[1] - s.bb[d] * s.bbcov[1];
  out[2] = rb * ud * s.ucov[2] + pb * delta[2] - s.bb[d] * s.bbcov[2];
  out[3] = rb * ud * s.ucov[3] + pb * delta[3] - s.bb[d] * s.bbcov[3];
  out[4] = rb * ud * s.ucov[0] + pb * delta[0] - s.bb[d] * s.bbcov[0] + s.de * ud;
  out[5] = s.de * (s.pr / powdg) * ud;
  if (d == 0) {
    out[6] = s.bcon[1]; out[7] = s.bcon[2]; out[8] = s.bcon[3];
  } else {
    const double r0 = 1.0 / s.ucon[0];
#pragma unroll
    for (int c = 1; c <= 3; c++)
      out[5 + c] = (c == d) ? 0.0 : (ud * r0) * s.bcon[c] - (s.ucon[c] * r0) * s.bcon[d];
  }
#pragma unroll
  for (int n = 0; n < 9; n++) out[n] = detg * out[n];
}

// Fast magnetosonic speeds (icha=2, Del Zanna et al. 2007) along direction d
// for a face state: calcha4i/j/k (calcha.f90:370-906).  The three directions
// write cs^2 and vA^2 differently (calcha.f90:480,648,826 and 486,654,832).
RH_DEV void face_speed(const Prel& s, int d, double gam, const Met& M, double& vfp, double& vfm) {
  const double alpha1 = 1.0 / sqrt(-M.gi(0, 0));
  const double beta_d = alpha1 * alpha1 * M.gi(0, d);
  double g3[9], gi3[9];
  cal3met(M, g3, gi3);
  double vcon[4];
  vcon[1] = s.ut[1] * (1.0 / s.gfl);
  vcon[2] = s.ut[2] * (1.0 / s.gfl);
  vcon[3] = s.ut[3] * (1.0 / s.gfl);
  double vsq = 0.0;
#pragma unroll
  for (int m = 1; m <= 3; m++)
#pragma unroll
    for (int n = 1; n <= 3; n++) vsq = vsq + g3[(m - 1) + 3 * (n - 1)] * vcon[m] * vcon[n];
  double cssq, vasq;
  if (d == 3) cssq = gam * s.pr * (1.0 / s.roh);
  else cssq = gam * s.pr / s.roh;
  if (d == 2) vasq = s.bbsq / (s.roh + s.bbsq);
  else vasq = s.bbsq * (1.0 / (s.roh + s.bbsq));
  const double omsq = vasq + cssq * (1.0 - vasq);
  const double f1 = vcon[d] * (1.0 - omsq);
  const double f2 = 1.0 - vsq * omsq;
  const double f3 = omsq * (1.0 - vsq) * ((1.0 - vsq * omsq) * gi3[(d - 1) * 4] - (1.0 - omsq) * vcon[d] * vcon[d]);
  double p = (f1 + sqrt(f3)) * (1.0 / f2);
  double m = (f1 - sqrt(f3)) * (1.0 / f2);
  p = alpha1 * p - beta_d;
  m = alpha1 * m - beta_d;
  if (fabs(p) > 1.0) p = 1.0;
  if (fabs(m) > 1.0) m = -1.0;
  vfp = p; vfm = m;
}

// Centre speeds for the CFL condition: calcha4 (calcha.f90:129-367).
RH_DEV void centre_speeds(const Prel& s, double gam, const Met& M, double vp[4], double vm[4]) {
  const double alpha1 = 1.0 / sqrt(-M.gi(0, 0));
  double beta1[4];
  beta1[1] = alpha1 * alpha1 * M.gi(0, 1);
  beta1[2] = alpha1 * alpha1 * M.gi(0, 2);
  beta1[3] = alpha1 * alpha1 * M.gi(0, 3);
  double g3[9], gi3[9];
  cal3met(M, g3, gi3);
  double vcon[4];
  vcon[1] = s.ut[1] * (1.0 / s.gfl);
  vcon[2] = s.ut[2] * (1.0 / s.gfl);
  vcon[3] = s.ut[3] * (1.0 / s.gfl);
  double vsq = 0.0;
#pragma unroll
  for (int m = 1; m <= 3; m++)
#pragma unroll
    for (int n = 1; n <= 3; n++) vsq = vsq + g3[(m - 1) + 3 * (n - 1)] * vcon[m] * vcon[n];
  const double cssq = gam * s.pr / s.roh;
  const double vasq = s.bbsq * (1.0 / (s.roh + s.bbsq));
  const double omsq = vasq + cssq * (1.0 - vasq);
  const double f2 = 1.0 - vsq * omsq;
#pragma unroll
  for (int d = 1; d <= 3; d++) {
    const double f1 = vcon[d] * (1.0 - omsq);
    const double f3 = omsq * (1.0 - vsq) * ((1.0 - vsq * omsq) * gi3[(d - 1) * 4] - (1.0 - omsq) * vcon[d] * vcon[d]);
    vp[d] = (f1 + sqrt(f3)) * (1.0 / f2);
    vm[d] = (f1 - sqrt(f3)) * (1.0 / f2);
  }
#pragma unroll
  for (int d = 1; d <= 3; d++) {
    vp[d] = alpha1 * vp[d] - beta1[d];
    vm[d] = alpha1 * vm[d] - beta1[d];
    if (fabs(vp[d]) > 1.0) vp[d] = 1.0;
    if (fabs(vm[d]) > 1.0) vm[d] = -1.0;
  }
}

// Reconstruction ---------------------------------------------------------

RH_DEV double dmm(double x, double y) {
  return 0.5 * (fsign1(x) + fsign1(y)) * fminf_(fabs(x), fabs(y));
}
RH_DEV double dm4(double w, double x, double y, double z) {
  return 0.125 * (fsign1(w) + fsign1(x)) * fabs(fsign1(w) + fsign1(y)) * (fsign1(w) + fsign1(z)) *
         fminf_(fminf_(fminf_(fabs(w), fabs(x)), fabs(y)), fabs(z));
}

// MP5 (Suresh & Huynh), rec.f90:2059-2343.  s[0..4] = u(c-2..c+2).
// Right edge of cell c (uri*l(c)): rec.f90:2113-2145.
RH_DEV double mp5_right(const double s[5]) {
  const double tmp1 = 1.0 / 60.0, tmp2 = 4.0 / 3.0, alpha = 4.0, eps1 = 1.0e-10;
  const double u1 = tmp1 * (2.0 * s[0] - 13.0 * s[1] + 47.0 * s[2] + 27.0 * s[3] - 3.0 * s[4]);
  const double ump = s[2] + dmm(s[3] - s[2], alpha * (s[2] - s[1]));
  if ((u1 - s[2]) * (u1 - ump) <= eps1) return u1;
  const double dim1 = s[0] - 2.0 * s[1] + s[2];
  const double di1 = s[1] - 2.0 * s[2] + s[3];
  const double dip1 = s[2] - 2.0 * s[3] + s[4];
  const double dph = dm4(4.0 * di1 - dip1, 4.0 * dip1 - di1, di1, dip1);
  const double dmh = dm4(4.0 * di1 - dim1, 4.0 * dim1 - di1, di1, dim1);
  const double ul = s[2] + alpha * (s[2] - s[1]);
  const double av = 0.5 * (s[2] + s[3]);
  const double md = av - 0.5 * dph;
  const double lc = s[2] + 0.5 * (s[2] - s[1]) + tmp2 * dmh;
  const double umax = fmaxf_(fminf_(fminf_(s[2], s[3]), md), fminf_(fminf_(s[2], ul), lc));
  const double umin = fminf_(fmaxf_(fmaxf_(s[2], s[3]), md), fmaxf_(fmaxf_(s[2], ul), lc));
  return u1 + dmm(umin - u1, umax - u1);
}
// Left edge of cell c (uri*r(c-1)): rec.f90:2149-2186.
RH_DEV double mp5_left(const double s[5]) {
  const double tmp1 = 1.0 / 60.0, tmp2 = 4.0 / 3.0, alpha = 4.0, eps1 = 1.0e-10;
  const double u1 = tmp1 * (2.0 * s[4] - 13.0 * s[3] + 47.0 * s[2] + 27.0 * s[1] - 3.0 * s[0]);
  const double ump = s[2] + dmm(s[1] - s[2], alpha * (s[2] - s[3]));
  if ((u1 - s[2]) * (u1 - ump) <= eps1) return u1;
  const double dim1 = s[0] - 2.0 * s[1] + s[2];
  const double di1 = s[1] - 2.0 * s[2] + s[3];
  const double dip1 = s[2] - 2.0 * s[3] + s[4];
  const double dph = dm4(4.0 * di1 - dip1, 4.0 * dip1 - di1, di1, dip1);
  const double dmh = dm4(4.0 * di1 - dim1, 4.0 * dim1 - di1, di1, dim1);
  const double ul = s[2] + alpha * (s[2] - s[3]);
  const double av = 0.5 * (s[2] + s[1]);
  const double md = av + 0.5 * dmh;
  const double lc = s[2] + 0.5 * (s[2] - s[3]) - tmp2 * dph;
  const double umax = fmaxf_(fminf_(fminf_(s[2], s[1]), md), fminf_(fminf_(s[2], ul), lc));
  const double umin = fminf_(fmaxf_(fmaxf_(s[2], s[1]), md), fmaxf_(fmaxf_(s[2], ul), lc));
  return u1 + dmm(umin - u1, umax - u1);
}

// WENO5 (Jiang & Shu), rec.f90:2373-2642; 13./12. is a single-precision literal.
RH_DEV double weno_edge(double a0, double a1, double a2, double a3, double a4) {
  // a0..a4 = stencil ordered from the far upwind side to the far downwind side
  const double c1312 = (double)(13.0f / 12.0f), c14 = 0.25, epsi = 1.0e-6;
  const double cd0 = 1.0 / 10.0, cd1 = 6.0 / 10.0, cd2 = 3.0 / 10.0;
  const double c00 = 2.0 / 6.0, c01 = -7.0 / 6.0, c02 = 11.0 / 6.0, c10 = -1.0 / 6.0, c11 = 5.0 / 6.0,
               c12 = 2.0 / 6.0, c20 = 2.0 / 6.0, c21 = 5.0 / 6.0, c22 = -1.0 / 6.0;
  double a, b;
  a = a0 - 2.0 * a1 + a2; b = a0 - 4.0 * a1 + 3.0 * a2;
  const double b0 = c1312 * (a * a) + c14 * (b * b) + epsi;
  a = a1 - 2.0 * a2 + a3; b = a1 - a3;
  const double b1 = c1312 * (a * a) + c14 * (b * b) + epsi;
  a = a2 - 2.0 * a3 + a4; b = 3.0 * a2 - 4.0 * a3 + a4;
  const double b2 = c1312 * (a * a) + c14 * (b * b) + epsi;
  const double w0 = cd0 / (b0 * b0), w1 = cd1 / (b1 * b1), w2 = cd2 / (b2 * b2);
  const double wt = w0 + w1 + w2;
  const double o0 = w0 / wt, o1 = w1 / wt, o2 = w2 / wt;
  const double r0 = c00 * a0 + c01 * a1 + c02 * a2;
  const double r1 = c10 * a1 + c11 * a2 + c12 * a3;
  const double r2 = c20 * a2 + c21 * a3 + c22 * a4;
  return o0 * r0 + o1 * r1 + o2 * r2;
}

}  // namespace rh
